"""Build the native parts of fk_mc_amd in-tree.

  fk_mc_amd/lib/libfkgpu.so    HIP kernels + C ABI (hipcc, gfx950)
  fk_mc_amd/lib/libfkmc_host.so  C++ host layer mirroring the reference classes
  fk_mc_amd/bin/fk_mc_gpu      the reference's Monte Carlo program on a batch of chains
  fk_mc_amd/lib/libfkmc_h5.so  its HDF5 output (built where an HDF5 C library is found:
                               $FK_HDF5_ROOT, /opt/conda, /usr, /usr/lib/x86_64-linux-gnu/hdf5/serial)

Every call recompiles (the build check must exercise the compiler).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
ARCH = os.environ.get("FK_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # packed FP32 VALU ops are kept out of every code object
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

GPU_SOURCES = ["csrc/fk_kernels.hip", "csrc/fk_vectors.hip", "csrc/fk_measures.hip", "csrc/fk_cheb.hip",
               "csrc/fk_mixed.hip", "csrc/fk_update.hip", "csrc/fk_green.hip", "csrc/fk_noalloc.hip", "csrc/fk_kpm_local.hip",
               "csrc/fkgpu.cpp"]
HOST_SOURCES = ["host/lattice.cpp", "host/configuration.cpp", "host/moves.cpp",
                "host/mc_metropolis.cpp", "host/measures.cpp", "host/capi.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=HERE)


def build_gpu():
    os.makedirs(LIB, exist_ok=True)
    srcs = [s for s in GPU_SOURCES if os.path.exists(os.path.join(HERE, s))]
    out = os.path.join(LIB, "libfkgpu.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + srcs)
    os.replace(tmp, out)
    return out


def build_host():
    srcs = [s for s in HOST_SOURCES if os.path.exists(os.path.join(HERE, s))]
    if not srcs:
        return None
    out = os.path.join(LIB, "libfkmc_host.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(HERE, "host"), "-o", tmp] + srcs +
         ["-L", LIB, "-lfkgpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_driver():
    src = "host/fk_mc_gpu.cpp"
    if not os.path.exists(os.path.join(HERE, src)):
        return None
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    out = os.path.join(HERE, "bin", "fk_mc_gpu")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", tmp, src,
          "-L", LIB, "-lfkgpu", "-ldl", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, out)
    return out


def find_hdf5():
    roots = [os.environ.get("FK_HDF5_ROOT", ""), "/opt/conda", "/usr", "/usr/lib/x86_64-linux-gnu/hdf5/serial"]
    for r in roots:
        if r and os.path.exists(os.path.join(r, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(r, "lib", "libhdf5.so")):
            return r
    return None


def build_h5():
    out = os.path.join(LIB, "libfkmc_h5.so")
    root = find_hdf5()
    if root is None:
        print("no HDF5 C library found: libfkmc_h5.so not built (fk_mc_gpu then writes .npy "
              "directories only)", flush=True)
        if os.path.exists(out):
            os.remove(out)
        return None
    os.makedirs(LIB, exist_ok=True)
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-fPIC", "-shared", "-I", os.path.join(root, "include"), "-o", tmp,
          "host/h5archive.c", "-L", os.path.join(root, "lib"), "-lhdf5", "-Wl,-rpath," + os.path.join(root, "lib")])
    os.replace(tmp, out)
    return out


def build_all():
    build_gpu()
    build_host()
    build_driver()
    build_h5()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
