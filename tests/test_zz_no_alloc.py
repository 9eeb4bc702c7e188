"""fkgpu_eig_logz_device_noalloc against fkgpu_eig_logz_device, through the C ABI.

The allocation-free entry point runs the same kernels with their HBM workspace in
the library's static device scratch instead of a hipMalloc'd buffer; on the same
inputs it must give the same eigenvalues and logZ bit for bit.  Two calls in a row,
the second with a smaller batch (the scratch keeps the first call's contents), for
the register-resident solver (N=256, no workspace) and the HBM-workspace solvers
(triangular N=400 and cubic3d N=512); a batch beyond the scratch is refused.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # lattice kind, D, L, first batch, second batch
    (0, 2, 16, 96, 17),
    (1, 2, 20, 150, 23),
    (0, 3, 8, 100, 37),
]


class _Hip:
    """Device buffers through the HIP runtime the library itself is linked against."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def check(self, r, what):
        assert r == 0, f"{what} failed with HIP error {r}"

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.check(self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)), "hipMalloc")
        self.check(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def to_host(self, p, like):
        out = np.empty_like(like)
        self.check(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out


def _call(fk, ctx, fn, hip, f, beta, U, mu):
    B, N = f.shape
    bufs = [hip.to_device(x) for x in (f, beta, U, mu, np.full((B, N), np.nan), np.full(B, np.nan))]
    try:
        ctx._check(getattr(fk.lib(), fn)(ctx._h, B, *bufs))
        hip.check(hip.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")  # also waits for the context stream
        return hip.to_host(bufs[4], np.zeros((B, N))), hip.to_host(bufs[5], np.zeros(B))
    finally:
        for p in bufs:
            hip.rt.hipFree(p)


def _inputs(N, B, seed):
    rng = np.random.default_rng(seed)
    f = (rng.random((B, N)) < rng.uniform(0.3, 0.7, (B, 1))).astype(np.int32)
    beta = rng.uniform(0.5, 10.0, B)
    U = rng.choice([1.0, 2.0, 3.0, 5.0], B)
    return f, beta, U, U / 2


@pytest.mark.parametrize("kind,D,L,B1,B2", CASES)
def test_noalloc_matches_allocating_entry(fk, orc, kind, D, L, B1, B2):
    hip = _Hip()
    print(f"case noalloc kind={kind} D={D} L={L} B={B1},{B2}", flush=True)
    N = L ** D
    ctx = fk.Context(0).set_lattice(kind, D, L, 1.0, 1.0 if kind == 1 else 0.0)
    for call, B in enumerate((B1, B2)):
        f, beta, U, mu = _inputs(N, B, 100 * L + call)
        ev_a, lz_a = _call(fk, ctx, "fkgpu_eig_logz_device", hip, f, beta, U, mu)
        ev_n, lz_n = _call(fk, ctx, "fkgpu_eig_logz_device_noalloc", hip, f, beta, U, mu)
        assert np.array_equal(ev_n, ev_a), f"call {call}: eigenvalues differ"
        assert np.array_equal(lz_n, lz_a), f"call {call}: logZ differs"
        # and both are the reference's calc_ed of the configuration
        ref = orc.calc_ed(kind, D, L, f[0], beta[0], U[0], mu[0], tp=1.0 if kind == 1 else 0.0)
        assert np.abs(ev_n[0] - ref["evals"]).max() < 1e-12 * np.abs(ref["evals"]).max()
        assert abs(lz_n[0] - ref["logZ"]) < 1e-12 * abs(ref["logZ"])
    ctx.close()


def test_noalloc_refuses_batch_beyond_scratch(fk):
    hip = _Hip()
    print("case noalloc refusal", flush=True)
    L, N = 8, 512
    ctx = fk.Context(0).set_lattice(0, 3, L, 1.0, 0.0)
    B = 400  # 400 x 2 MiB of workspace > 256 MiB
    f, beta, U, mu = _inputs(N, B, 1)
    with pytest.raises(fk.FkgpuError, match="static scratch"):
        _call(fk, ctx, "fkgpu_eig_logz_device_noalloc", hip, f, beta, U, mu)
    # the context still works afterwards
    f, beta, U, mu = _inputs(N, 4, 2)
    ev, lz = _call(fk, ctx, "fkgpu_eig_logz_device_noalloc", hip, f, beta, U, mu)
    assert np.isfinite(ev).all() and np.isfinite(lz).all()
    ctx.close()
