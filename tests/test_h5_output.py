"""HDF5 output of fk_mc_gpu (prog/data_save.hxx:32-151: /parameters, /mc_data, /stats,
/binning in output.h5).  The file is read back with h5py from the image's second Python
(/opt/conda/bin/python3.9, h5py 3.3 over HDF5 1.10.6); the reads of /stats follow the
reference's scripts/parse/parse_thermod.py:49-51, parse_uscan.py:50-52 and
parse_ipr.py:70-72."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBH5 = os.path.join(ROOT, "fk_mc_amd", "lib", "libfkmc_h5.so")
EXE = os.path.join(ROOT, "fk_mc_amd", "bin", "fk_mc_gpu")
H5PY = "/opt/conda/bin/python3.9"

# every dataset of the file into an .npz (key = path with '/' as '|'), strings into a JSON side file
_DUMP = r"""
import sys, json, numpy as np, h5py
src, dst = sys.argv[1], sys.argv[2]
arrays, strings, kinds = {}, {}, {}
def visit(name, obj):
    if isinstance(obj, h5py.Dataset):
        v = obj[()]
        kinds[name] = obj.dtype.kind
        if isinstance(v, bytes):
            strings[name] = v.decode()
        elif obj.dtype.kind == 'O':
            strings[name] = v.decode() if isinstance(v, bytes) else str(v)
        else:
            arrays[name.replace('/', '|')] = np.asarray(v)
with h5py.File(src, 'r') as f:
    f.visititems(visit)
    # the parse scripts' reads
    parse = {}
    for obs in ('cv', 'energy', 'dos0', 'ipr0'):
        if obs in f['stats']:
            (nbins, value, disp, error) = f['stats'][obs]
            parse[obs] = [float(nbins), float(value), float(disp), float(error)]
    for obs in ('dos_err', 'ipr_err'):
        if obs in f['stats']:
            d = f['stats'][obs]
            parse[obs] = [d[:, 0].tolist(), d[:, 1].tolist(), d[:, 2].tolist()]
np.savez(dst + '.npz', **arrays)
json.dump({'strings': strings, 'kinds': kinds, 'parse': parse}, open(dst + '.json', 'w'))
"""


def _h5py_available():
    if not os.path.exists(H5PY):
        return False
    return subprocess.run([H5PY, "-c", "import h5py"], capture_output=True, timeout=120).returncode == 0


needs_h5py = pytest.mark.skipif(not _h5py_available(), reason="no h5py to read the files back (" + H5PY + ")")


def _dump(path, tmp_path):
    dst = str(tmp_path / "dump")
    r = subprocess.run([H5PY, "-c", _DUMP, str(path), dst], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with np.load(dst + ".npz", allow_pickle=False) as z:
        arrays = {k.replace("|", "/"): z[k] for k in z.files}
    meta = json.load(open(dst + ".json"))
    return arrays, meta


def _lib():
    if not os.path.exists(LIBH5):
        pytest.fail(f"{LIBH5} not built although an HDF5 library is installed")
    lib = ctypes.CDLL(LIBH5)
    lib.fkh5_create.restype = ctypes.c_void_p
    lib.fkh5_create.argtypes = [ctypes.c_char_p]
    for fn, t in (("fkh5_write_f64", ctypes.c_double), ("fkh5_write_i64", ctypes.c_int64)):
        getattr(lib, fn).argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(t), ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_uint64)]
    lib.fkh5_write_str.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
    lib.fkh5_close.argtypes = [ctypes.c_void_p]
    lib.fkh5_error.restype = ctypes.c_char_p
    return lib


def _put(lib, h, name, a):
    a = np.require(a, requirements="C")  # keeps rank 0, unlike ascontiguousarray
    dims = (ctypes.c_uint64 * max(a.ndim, 1))(*a.shape)
    if a.dtype == np.int64:
        return lib.fkh5_write_i64(h, name.encode(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.ndim, dims)
    return lib.fkh5_write_f64(h, name.encode(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.ndim, dims)


@needs_h5py
def test_archive_roundtrip(tmp_path):
    """Datasets of rank 0-4 in nested groups, int64 scalars and strings read back exactly;
    writing a path twice and creating a file in a missing directory are errors."""
    lib = _lib()
    path = tmp_path / "a.h5"
    h = lib.fkh5_create(str(path).encode())
    assert h
    rng = np.random.default_rng(3)
    data = {"x": np.array(2.5), "mc_data/energies": rng.standard_normal(7), "stats/cv": rng.standard_normal(4),
            "binning/energy": rng.standard_normal((5, 5)), "mc_data/spectrum_history": rng.standard_normal((2, 3, 4)),
            "mc_data/eig_history": rng.standard_normal((1, 2, 3, 3)), "parameters/L": np.array(16, np.int64)}
    for k, v in data.items():
        assert _put(lib, h, k, v) == 0, lib.fkh5_error()
    assert lib.fkh5_write_str(h, b"parameters/lattice", "cubic2d µ".encode()) == 0
    assert _put(lib, h, "stats/cv", data["stats/cv"]) != 0  # exists already
    assert b"stats/cv" in lib.fkh5_error()
    assert lib.fkh5_close(h) == 0
    assert not lib.fkh5_create(str(tmp_path / "no" / "such" / "dir.h5").encode())
    assert b"cannot create" in lib.fkh5_error()
    arrays, meta = _dump(path, tmp_path)
    assert set(arrays) == set(data)
    for k, v in data.items():
        assert arrays[k].dtype == v.dtype and arrays[k].shape == v.shape, k
        np.testing.assert_array_equal(arrays[k], v)
    assert meta["strings"] == {"parameters/lattice": "cubic2d µ"}


@pytest.mark.gpu
@needs_h5py
@pytest.mark.parametrize("chains", [1, 3])
def test_driver_h5_equals_npy(tmp_path, chains):
    """The same run written as output.h5 and as a .npy directory: every dataset equal (the
    chain axis of mc_data/ dropped in the file of one chain), /parameters typed, and the
    reads of the reference's parse scripts find the statistics."""
    args = ["--lattice", "cubic2d", "--L", 4, "--beta", 2.0, "--U", 2.0, "--mu_c", 1.0, "--mu_f", 1.0, "--nsweeps", 24,
            "--sweep_len", 4, "--ntherm_sweeps", 2, "--chains", chains, "--seed", 5, "--measure_ipr=true",
            "--measure_history=true", "--measure_stiffness=true", "--dos_npts", 16, "--show_output", "false"]
    outs = {}
    for name in ("run", "output.h5"):
        d = tmp_path / f"{name}_dir"
        d.mkdir()
        cmd = [EXE] + [str(a) for a in args] + ["--output", str(d / name)]
        print(f"case driver chains={chains} output={name}", flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[name] = d / name
    npy_root = outs["run"]
    npys = {os.path.relpath(os.path.join(dp, f), npy_root)[:-4]: np.load(os.path.join(dp, f))
            for dp, _, fs in os.walk(npy_root) for f in fs if f.endswith(".npy")}
    assert os.path.exists(outs["output.h5"]) and os.path.isfile(outs["output.h5"])
    assert os.path.exists(tmp_path / "output.h5_dir" / "stats.txt")
    arrays, meta = _dump(outs["output.h5"], tmp_path)
    data = {k: v for k, v in arrays.items() if not k.startswith("parameters/")}
    assert set(data) == set(npys)
    for k in ("mc_data/energies", "mc_data/spectrum", "mc_data/ipr_history", "mc_data/cond_history", "stats/cv",
              "stats/energy", "binning/energy", "stats/dos_err", "stats/ipr_err", "dos"):
        assert k in data, k
    for k, v in npys.items():
        if k.startswith("mc_data/") and chains == 1:
            v = v[0]
        assert data[k].shape == v.shape, k
        np.testing.assert_array_equal(data[k], v, err_msg=k)
    # parameters: the same values as the directory run's parameters.txt, typed
    ptxt = dict(line.rstrip("\n").split(" = ", 1) for line in open(npy_root / "parameters.txt"))
    pars = {k[len("parameters/"):]: v for k, v in arrays.items() if k.startswith("parameters/")}
    pars.update({k[len("parameters/"):]: v for k, v in meta["strings"].items()})
    ptxt["output"] = str(outs["output.h5"])
    assert set(pars) == set(ptxt)
    assert pars["lattice"] == "cubic2d" and int(pars["L"]) == 4 and pars["L"].dtype == np.int64
    assert float(pars["beta"]) == 2.0 and pars["beta"].dtype == np.float64
    assert int(pars["measure_ipr"]) == 1 and int(pars["chains"]) == chains
    # parse_thermod / parse_uscan: (nbins, value, disp, error) = h5data["stats"][obs]
    for obs in ("cv", "energy", "dos0", "ipr0"):
        np.testing.assert_array_equal(meta["parse"][obs], npys["stats/" + obs])
    # parse_ipr: columns w, value, error of stats/dos_err and stats/ipr_err
    for obs in ("dos_err", "ipr_err"):
        np.testing.assert_array_equal(np.array(meta["parse"][obs]).T, npys["stats/" + obs])


def test_h5_output_without_library_is_refused(tmp_path):
    """Without libfkmc_h5.so beside the program an .h5 output ends the run before any device
    work, with an error that names the way out (a directory output).  A copy of the program
    whose ../lib holds only libfkgpu.so; runs without a GPU."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built")
    r = subprocess.run([EXE, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "HDF5 file" in r.stdout.split("--output ")[1].split("\n")[0]
    (tmp_path / "bin").mkdir()
    (tmp_path / "lib").mkdir()
    exe = tmp_path / "bin" / "fk_mc_gpu"
    exe.write_bytes(open(EXE, "rb").read())
    exe.chmod(0o755)
    (tmp_path / "lib" / "libfkgpu.so").symlink_to(os.path.join(ROOT, "fk_mc_amd", "lib", "libfkgpu.so"))
    r = subprocess.run([str(exe), "--output", str(tmp_path / "out.h5"), "--nsweeps", "1"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "libfkmc_h5.so" in r.stderr and "directory name" in r.stderr, r.stderr
    assert not (tmp_path / "out.h5").exists()
