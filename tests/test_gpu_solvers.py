"""Dense eigensolver variants (register-resident one-stage, unblocked HBM,
blocked with MFMA trailing updates in four tile shapes, one-wave) against the oracle,
and the autotuner."""
import numpy as np
import pytest

from test_gpu_chains import compare_chain

pytestmark = pytest.mark.gpu

VARIANT_NMAX = [64, 128, 256, 2048, 1024, 256, 64, 1024, 1024, 1024, 256]
CASES = [(0, 2, 8), (0, 2, 16), (1, 2, 18), (0, 3, 7), (0, 3, 8), (1, 2, 24), (0, 2, 32), (0, 1, 3), (0, 1, 5),
         (0, 1, 64), (1, 2, 7)]


@pytest.mark.parametrize("kind,D,L", CASES, ids=lambda c: str(c))
def test_variants_match_oracle(fk, orc, kind, D, L):
    n = L ** D
    rng = np.random.default_rng(n)
    B = 3
    f = rng.integers(0, 2, (B, n)).astype(np.int32)
    beta, U, mu = np.array([0.5, 3.0, 20.0]), np.array([1.0, 4.0, 8.0]), np.array([0.5, 2.0, 4.0])
    ref = [orc.calc_ed(kind, D, L, f[b], beta[b], U[b], mu[b]) for b in range(B)]
    ctx = fk.Context(0).set_lattice(kind, D, L, 1.0, 1.0)
    for v in range(fk.SOLVER_VARIANTS):
        if n > VARIANT_NMAX[v] or (v == 3 and n < 300 and n != 64):
            continue
        print(f"case solver variant {v} kind={kind} D={D} L={L}", flush=True)
        ctx.set_solver(v)
        assert ctx.get_solver() == v
        g = ctx.eig_logz(f, beta, U, mu)
        for b in range(B):
            scale = max(1.0, np.abs(ref[b]["evals"]).max())
            assert np.abs(g["evals"][b] - ref[b]["evals"]).max() <= 1e-12 * scale, (v, b)
            assert abs(g["logZ"][b] - ref[b]["logZ"]) <= 1e-12 * abs(ref[b]["logZ"]), (v, b)
    ctx.close()


def test_blocked_eigenvectors(fk, orc):
    """The blocked solver's reflectors feed the eigenvector back-transform (IPR)."""
    kind, D, L = 1, 2, 18
    n = L * L
    f = np.random.default_rng(5).integers(0, 2, (2, n)).astype(np.int32)
    ctx = fk.Context(0).set_lattice(kind, D, L, 1.0, 1.0).set_solver(4)
    print("case blocked eigenvectors", flush=True)
    ev, ipr = ctx.measure_ipr(f, 2.0, 3.0, 1.5)
    for b in range(2):
        r = orc.calc_ed(kind, D, L, f[b], 2.0, 3.0, 1.5, vectors=True)
        assert np.abs(ev[b] - r["evals"]).max() < 1e-11
        ref = orc.measure_ipr(r["evecs"])
        gaps = np.diff(r["evals"])
        ok = np.ones(n, bool)  # IPR is defined per eigenvector only for simple eigenvalues
        ok[:-1] &= gaps > 1e-8
        ok[1:] &= gaps > 1e-8
        assert np.abs(ipr[b][ok] - ref[ok]).max() < 1e-8
    ctx.close()


@pytest.mark.parametrize("variant", [4, 3, 8])
def test_chain_with_hbm_solvers(fk, orc, variant):
    """Chains at N = 343 through the HBM-workspace solvers reproduce the oracle chain."""
    D, L = 3, 7
    N = L ** D
    print(f"case chain variant {variant}", flush=True)
    rng = np.random.default_rng(11)
    B = 2
    f0 = rng.integers(0, 2, (B, N)).astype(np.int32)
    seeds = np.array([77, 78], np.uint32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0).set_solver(variant)
    ch = fk.Chains(ctx, f0, 2.0, 5.0, 2.5, 2.5, seeds)
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    tr = ch.sweep(60, trace=True)
    for c in range(B):
        trc = {k: v[:, c] for k, v in tr.items()}
        compare_chain(orc, orc.KIND_HYPERCUBIC, D, L, 2.0, 5.0, 2.5, 2.5, int(seeds[c]), f0[c],
                      [(orc.MOVE_ADDREMOVE, 0.5), (orc.MOVE_HOP, 0.5)], trc, tp=0.0)
    ch.close()
    ctx.close()


def test_autotune(fk):
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, 8, 1.0, 0.0)
    print("case autotune N=64", flush=True)
    best, ms = ctx.autotune(64)
    assert 0 <= best < fk.SOLVER_VARIANTS
    assert (ms >= 0).all()  # every variant fits N = 64
    assert ms[best] == ms.min()
    assert ctx.get_solver() == best
    ctx.set_lattice(fk.LATTICE_HYPERCUBIC, 3, 8, 1.0, 0.0)
    best, ms = ctx.autotune(16)
    assert list(ms[:3]) == [-1.0, -1.0, -1.0] and ms[5] == -1.0 and ms[10] == -1.0 and best in (3, 4, 7, 8, 9)
    ctx.close()
