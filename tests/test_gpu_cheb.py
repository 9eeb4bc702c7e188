"""Chebyshev paths on the GPU.

Exact trace (configuration_t::calc_chebyshev, src/configuration.cpp:94-206)
against the oracle's restatement on the same configuration (moments 1e-10,
logZ 1e-10 relative) and, as the reference's own test does
(fast_update_test.cpp:77), against ED within 5e-2.  The stochastic trace
(KPM) against the exact spectral moments within the estimator's error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("D,L,U,beta", [(2, 8, 1.0, 1.0), (2, 24, 8.0, 1 / 0.16), (1, 50, 2.0, 3.0)])
def test_exact_trace_vs_oracle(fk, orc, D, L, U, beta):
    print(f"case cheb exact D={D} L={L}", flush=True)
    n = L ** D
    f = orc.randomize_f(n, 32167, n // 2)
    # reference defaults: cheb_size = int(2.2 ln N) rounded to even, grid = max(2 M, 10) (fk_mc.hxx:61-64)
    M = int(np.log(n) * 2.2)
    M += M % 2
    G = max(2 * M, 10)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    g = ctx.cheb_logz(f, beta, U, U / 2, M, G)
    r = orc.calc_chebyshev(0, D, L, f, beta, U, U / 2, M, G, tp=0.0)
    assert abs(g["a"] - r["a"]) < 1e-12 * abs(r["a"]) and abs(g["b"] - r["b"]) < 1e-11
    assert np.abs(g["moments"] - r["moments"]).max() < 1e-10
    assert abs(g["logZ"] - r["logZ"]) < 1e-10 * abs(r["logZ"])
    ed = orc.calc_ed(0, D, L, f, beta, U, U / 2)
    if L == 24:
        assert abs((g["logZ"] - ed["logZ"]) / g["logZ"]) < 5e-2
    ctx.close()


def _dense_h(L, f, U, mu):
    """cubic2d (t=1, periodic) Hamiltonian as a dense matrix (numpy)."""
    n = L * L
    H = np.zeros((n, n))
    idx = np.arange(n).reshape(L, L)
    for shift in ((1, 0), (0, 1)):
        j = np.roll(idx, (-shift[0], -shift[1]), axis=(0, 1)).ravel()
        H[idx.ravel(), j] -= 1.0
        H[j, idx.ravel()] -= 1.0
    H[np.arange(n), np.arange(n)] = -mu + U * f
    return H


def test_exact_trace_at_scale(fk):
    """calc_chebyshev's algorithm above the dense solver's range: cubic2d L=64 (N=4096) with
    the reference's moment count int(2.2 ln N) rounded to even; the spectral bounds from the
    device Lanczos (ARPACK in the reference, src/configuration.cpp:99-100) equal LAPACK's
    extreme eigenvalues, and logZ is within 5e-2 of ED (fast_update_test.cpp:77).  L=128
    (N=16384) runs too (no dense solve to compare with there)."""
    print("case cheb exact L=64", flush=True)
    L, U, beta = 64, 2.0, 10.0
    n = L * L
    f = np.random.default_rng(64).integers(0, 2, n).astype(np.int32)
    M = int(np.log(n) * 2.2)
    M += M % 2
    G = max(2 * M, 10)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    g = ctx.cheb_logz(f, beta, U, U / 2, M, G)
    ev = np.linalg.eigvalsh(_dense_h(L, f, U, U / 2))
    print("lanczos bounds", g["e_min"], g["e_max"], "lapack", ev[0], ev[-1], flush=True)
    assert abs(g["e_min"] - ev[0]) < 1e-10 and abs(g["e_max"] - ev[-1]) < 1e-10
    lz = float(np.logaddexp(0.0, -beta * ev).sum())
    print("cheb logZ", g["logZ"], "ed", lz, flush=True)
    assert abs(g["logZ"] - lz) / lz < 5e-2
    x = (ev - g["b"]) / g["a"]
    exact = np.array([np.mean(np.cos(m * np.arccos(np.clip(x, -1, 1)))) for m in range(M)])
    assert np.abs(g["moments"] - exact).max() < 1e-9
    ctx.close()
    print("case cheb exact L=128", flush=True)
    L = 128
    n = L * L
    f = np.random.default_rng(128).integers(0, 2, n).astype(np.int32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    assert ctx.N == n
    g = ctx.cheb_logz(f, beta, U, U / 2, 22, 44)
    assert -5.0 - 1e-9 <= g["e_min"] < g["e_max"] <= 5.0 + 1e-9  # inside the Gershgorin discs
    assert abs(g["moments"][0] - 1.0) < 1e-14 and np.all(np.abs(g["moments"]) <= 1.0 + 1e-12)
    assert np.isfinite(g["logZ"]) and g["logZ"] > 0
    ctx.close()


def test_kpm_moments_and_logz(fk, orc):
    print("case kpm", flush=True)
    L, U, beta = 32, 2.0, 10.0
    n = L * L
    rng = np.random.default_rng(3)
    f = rng.integers(0, 2, (2, n)).astype(np.int32)
    R, M = 64, 256
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    out = ctx.kpm_logz(f, beta, U, U / 2, R=R, M=M, seed=7, jackson=True)
    again = ctx.kpm_logz(f, beta, U, U / 2, R=R, M=M, seed=7, jackson=True)
    assert np.array_equal(out["logZ"], again["logZ"])  # deterministic
    for b in range(2):
        ed = ctx.eig_logz(f[b:b + 1], beta, U, U / 2)
        x = (ed["evals"][0] - out["b"][b]) / out["a"][b]
        assert np.abs(x).max() < 1.0
        exact = np.array([np.mean(np.cos(m * np.arccos(x))) for m in range(M)])
        # Rademacher trace estimator: sd <= sqrt(2/(N R)) per moment
        tol = 6 * np.sqrt(2.0 / (n * R))
        assert np.abs(out["moments"][b] - exact).max() < tol
        assert abs(out["moments"][b][0] - 1.0) < 1e-15
        # logZ: Jackson-damped expansion + stochastic error; the estimator's sd for
        # Tr F(X) is sqrt(2 sum_{i!=j} F_ij^2 / R), about 1% of logZ here
        print("kpm logZ", out["logZ"][b], "ed", ed["logZ"][0], flush=True)
        assert abs(out["logZ"][b] - ed["logZ"][0]) / ed["logZ"][0] < 3e-2
    # a single-site flip with the same vectors: the noise largely cancels in the difference
    g = f[0:1].copy()
    g[0, 17] ^= 1
    pair = np.concatenate([f[0:1], g])
    kp = ctx.kpm_logz(pair, beta, U, U / 2, R=R, M=M, seed=7, jackson=True)
    d_kpm = kp["logZ"][1] - kp["logZ"][0]
    d_ed = ctx.eig_logz(g, beta, U, U / 2)["logZ"][0] - ctx.eig_logz(f[0:1], beta, U, U / 2)["logZ"][0]
    print("kpm dlogZ", d_kpm, "ed", d_ed, flush=True)
    assert abs(d_kpm - d_ed) < 0.1 * abs(d_ed) + 0.2
    ctx.close()


def _csr_hop(L, diag, seed):
    """Periodic L x L graph, nearest neighbours (and the two diagonals when diag) with
    random symmetric hopping values: a lattice whose hopping differs from site to site."""
    rng = np.random.default_rng(seed)
    n = L * L
    h = np.zeros((n, n))
    offs = [(1, 0), (0, 1)] + ([(1, 1), (1, -1)] if diag else [])
    for x in range(L):
        for y in range(L):
            for dx, dy in offs:
                i, j = x * L + y, ((x + dx) % L) * L + (y + dy) % L
                h[i, j] = h[j, i] = -rng.uniform(0.5, 1.5)
    return h


@pytest.mark.parametrize("kind", ["csr4", "csr8", "triangular", "cubic3d"])
def test_kpm_moments_lattices(fk, kind):
    """Stochastic-trace moments on lattices with 4, 6 and 8 neighbour slots, with the same
    hopping at every site and with hopping that varies from site to site: against the exact
    spectral moments within the estimator's error."""
    print("case kpm moments", kind, flush=True)
    ctx = fk.Context(0)
    if kind == "csr4":
        ctx.set_lattice_csr(_csr_hop(32, False, 1))
    elif kind == "csr8":
        ctx.set_lattice_csr(_csr_hop(32, True, 2))
    elif kind == "triangular":
        ctx.set_lattice(fk.LATTICE_TRIANGULAR, 2, 32, 1.0, 0.6)
    else:
        ctx.set_lattice(fk.LATTICE_HYPERCUBIC, 3, 10, 1.0, 0.0)
    n = ctx.N
    rng = np.random.default_rng(8)
    f = rng.integers(0, 2, (2, n)).astype(np.int32)
    R, M, beta, U = 64, 128, 4.0, 2.0
    out = ctx.kpm_logz(f, beta, U, U / 2, R=R, M=M, seed=11, jackson=True)
    tol = 6 * np.sqrt(2.0 / (n * R))
    for b in range(2):
        ed = ctx.eig_logz(f[b:b + 1], beta, U, U / 2)
        x = (ed["evals"][0] - out["b"][b]) / out["a"][b]
        assert np.abs(x).max() < 1.0
        exact = np.array([np.mean(np.cos(m * np.arccos(x))) for m in range(M)])
        assert abs(out["moments"][b][0] - 1.0) < 1e-15
        assert np.abs(out["moments"][b] - exact).max() < tol
        assert abs(out["logZ"][b] - ed["logZ"][0]) / ed["logZ"][0] < 3e-2
    ctx.close()


def test_kpm_chains(fk, orc):
    """Chebyshev-move chains (src/moves_chebyshev.cpp) weighed by the stochastic trace
    (FKGPU_KPM_DELTA_STOCHASTIC): draws follow the reference order, dlogZ of each solved
    proposal agrees with dense ED within the estimator's noise, graph replay equals eager
    launches."""
    print("case kpm chains", flush=True)
    L, U, beta = 32, 2.0, 10.0
    n = L * L
    rng = np.random.default_rng(11)
    B = 4
    f0 = rng.integers(0, 2, (B, n)).astype(np.int32)
    seeds = np.arange(B, dtype=np.uint32) + 500
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
    ch.set_moves([(fk.MOVE_ADDREMOVE, 1.0)])
    ch.set_kpm(R=64, M=256, seed=3, jackson=True)
    ch.set_kpm_delta(fk.KPM_DELTA_STOCHASTIC)
    lz0 = ch.state()["logZ"].copy()
    tr = ch.sweep(12, trace=True)
    # first move of every chain: same site and uniform as the reference chain
    for c in range(B):
        ref = orc.run_chain(0, 2, L, beta, U, U / 2, U / 2, int(seeds[c]), f0[c], [(orc.MOVE_ADDREMOVE, 1.0)], 1,
                            tp=0.0)
        assert tr["a"][0, c] == ref["a"][0] and tr["u"][0, c] == ref["u"][0]
        # the chain's estimates are the one-shot estimate with the chain's probe vectors
        # (seed + golden*(c+1)) and bounds (those of any configuration holding both occupations)
        g = f0[c:c + 1].copy()
        g[0, tr["a"][0, c]] ^= 1
        seed_c = (3 + 0x9E3779B97F4A7C15 * (c + 1)) % 2 ** 64
        one = ctx.kpm_logz(np.concatenate([f0[c:c + 1], g]), beta, U, U / 2, R=64, M=256, seed=seed_c, jackson=True)
        assert abs(one["logZ"][0] - lz0[c]) <= 1e-12 * lz0[c]
        assert abs(one["logZ"][1] - tr["logz"][0, c]) <= 1e-12 * lz0[c]
        # against dense ED: the estimator's noise on a difference is a few units at beta=10, R=64
        d_ed = ctx.eig_logz(g, beta, U, U / 2)["logZ"][0] - ctx.eig_logz(f0[c:c + 1], beta, U, U / 2)["logZ"][0]
        d_kpm = tr["logz"][0, c] - lz0[c]
        print("chain", c, "dlogZ kpm", d_kpm, "ed", d_ed, flush=True)
        assert abs(d_kpm - d_ed) < 10.0
    # graph replay vs eager from identical starting points
    a = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
    b = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
    for x in (a, b):
        x.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
        x.set_kpm(R=16, M=64, seed=9, jackson=True)
        x.set_kpm_delta(fk.KPM_DELTA_STOCHASTIC)
    for rep in range(2):
        a.sweep(8)
        b.sweep_graph(8)
        sa, sb = a.state(), b.state()
        assert np.array_equal(sa["f"], sb["f"]) and np.array_equal(sa["logZ"], sb["logZ"]), rep
    for x in (ch, a, b):
        x.close()
    ctx.close()


def _exact_logz_2d(L, f, beta, U, mu):
    """logZ of cubic2d (t=1, periodic) from a dense FP64 eigensolve (LAPACK via numpy)."""
    n = L * L
    H = np.zeros((n, n))
    idx = np.arange(n).reshape(L, L)
    for shift in ((1, 0), (0, 1)):
        j = np.roll(idx, (-shift[0], -shift[1]), axis=(0, 1)).ravel()
        H[idx.ravel(), j] -= 1.0
        H[j, idx.ravel()] -= 1.0
    H[np.arange(n), np.arange(n)] = -mu + U * f
    return float(np.logaddexp(0.0, -beta * np.linalg.eigvalsh(H)).sum())


def test_kpm_chain_large_lattice(fk):
    """cubic2d L=64 (N=4096), beta=10, U=2 with proposals weighed by the stochastic trace
    (above the dense range).  dOmega of the first proposals against an exact FP64 dense
    eigensolve: the chain's estimate within the estimator's spread (sd of dlogZ about 1.3
    at R=64, measured over independent probe sets), and the mean over 16 independent probe
    sets within 4 standard errors (the estimator is unbiased up to the Jackson smoothing)."""
    print("case kpm L=64", flush=True)
    L = 64
    n = L * L
    beta, U = 10.0, 2.0
    rng = np.random.default_rng(2)
    f0 = rng.integers(0, 2, (2, n)).astype(np.int32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, [1, 2])
    with pytest.raises(fk.FkgpuError):
        ch.sweep(1)  # no dense solver at this size
    ch.set_moves([(fk.MOVE_ADDREMOVE, 1.0)])
    ch.set_kpm(R=64, M=256, seed=5, jackson=True)
    ch.set_kpm_delta(fk.KPM_DELTA_STOCHASTIC)
    lz0 = ch.state()["logZ"].copy()
    tr = ch.sweep(4, trace=True)
    st = ch.state()
    assert np.all(np.isfinite(st["logZ"])) and np.all(st["logZ"] > 0)
    assert np.all(tr["logz"][tr["kind"] == 0] != 0.0)
    ch.close()
    ex0 = [_exact_logz_2d(L, f0[c], beta, U, U / 2) for c in range(2)]
    for c in range(2):
        g = f0[c].copy()
        g[tr["a"][0, c]] ^= 1
        d_ex = _exact_logz_2d(L, g, beta, U, U / 2) - ex0[c]
        d_kpm = tr["logz"][0, c] - lz0[c]
        print("chain", c, "dOmega kpm", -d_kpm / beta, "exact", -d_ex / beta, flush=True)
        assert abs(d_kpm - d_ex) / beta < 1.0
        if c == 0:
            pair = np.stack([f0[0], g]).astype(np.int32)
            ds = np.array([np.diff(ctx.kpm_logz(pair, beta, U, U / 2, R=64, M=256, seed=sd, jackson=True)["logZ"])[0]
                           for sd in range(100, 116)])
            se = ds.std(ddof=1) / np.sqrt(len(ds))
            print("16 probe sets: mean dlogZ", ds.mean(), "sd", ds.std(ddof=1), "exact", d_ex, flush=True)
            assert 0.3 < ds.std(ddof=1) < 4.0
            assert abs(ds.mean() - d_ex) < 4 * se + 0.05
    ctx.close()


# ---- local estimator of a proposal's dlogZ (FKGPU_KPM_DELTA_LOCAL) -------------------

NEAR_TIE = 1e-9


def _compare_local_chain(orc, kind, D, L, beta, U, mu, seed, f0, moves, tr, tp, W=None):
    """Device chain with locally estimated dlogZ against the oracle's dense-ED chain:
    draws bit for bit, log-weights within 1e-9, decisions equal except at near-ties
    (where the oracle follows the device).  Returns (#near-ties, max |dlogw|)."""
    nsteps = len(tr["u"])
    forced = -np.ones(nsteps, np.int32)
    orc.set_chain_ff(W)
    orc.set_flip_ff(1)  # the Chebyshev move_flip weighs dE_ff (src/moves_chebyshev.cpp:22)
    try:
        for _attempt in range(8):
            ref = orc.run_chain(kind, D, L, beta, U, mu, mu, seed, f0, moves, nsteps, tp=tp, forced=forced)
            assert np.array_equal(tr["kind"], ref["kind"]), "move choice differs"
            assert np.array_equal(tr["a"], ref["a"]), "site draw differs"
            assert np.array_equal(tr["b"], ref["b"]), "second site differs"
            assert np.array_equal(tr["u"], ref["u"]), "uniform differs"
            fin = np.isfinite(ref["logw"])
            assert np.array_equal(fin, np.isfinite(tr["logw"]))
            dw = np.abs(tr["logw"][fin] - ref["logw"][fin]).max(initial=0.0)
            assert dw < 1e-9, f"log-weight differs by {dw}"
            mism = np.nonzero(tr["acc"] != ref["acc"])[0]
            if len(mism) == 0:
                return int(np.sum(forced >= 0)), dw
            s = mism[0]
            gap = abs(np.log(ref["u"][s]) - ref["logw"][s])
            assert gap < NEAR_TIE, f"accept differs at move {s}: |log u - log w| = {gap}"
            forced[s] = tr["acc"][s]
    finally:
        orc.set_chain_ff(None)
        orc.set_flip_ff(0)
    raise AssertionError("too many near-ties")


LOCAL_CASES = [
    # name, kind, D, L, beta, U, moves, W
    ("square8", 0, 2, 8, 10.0, 2.0, "ah", None),
    ("square12_flip", 0, 2, 12, 5.0, 4.0, "afh", None),
    ("triangular6", 1, 2, 6, 8.0, 3.0, "ah", None),
    ("cubic3d4", 0, 3, 4, 2.0, 5.0, "ah", None),
    ("chain24_ff", 0, 1, 24, 3.0, 2.0, "ah", [0.0, 0.5, 0.2]),
    ("chain24_flip_ff", 0, 1, 24, 3.0, 2.0, "af", [0.0, 0.5, 0.2]),
]


@pytest.mark.parametrize("name,kind,D,L,beta,U,mv,W", LOCAL_CASES)
def test_kpm_local_chains_match_oracle(fk, orc, name, kind, D, L, beta, U, mv, W):
    """Chains on the Chebyshev path with the local dlogZ estimator reproduce the
    reference's ED chains (src/moves.cpp, src/mc_metropolis.cpp:34-52) of the same seeds."""
    print(f"case kpm local {name}", flush=True)
    n = L ** D
    tp = 1.0 if kind == 1 else 0.0
    table = {"a": (fk.MOVE_ADDREMOVE, orc.MOVE_ADDREMOVE), "f": (fk.MOVE_FLIP, orc.MOVE_FLIP),
             "h": (fk.MOVE_HOP, orc.MOVE_HOP)}
    w = 1.0 / len(mv)
    moves_fk = [(table[c][0], w) for c in mv]
    moves_orc = [(table[c][1], w) for c in mv]
    rng = np.random.default_rng(n + 3)
    B = 3
    f0 = rng.integers(0, 2, (B, n)).astype(np.int32)
    seeds = np.arange(B, dtype=np.uint32) + 900
    ctx = fk.Context(0).set_lattice(kind, D, L, 1.0, tp)
    ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
    ch.set_moves(moves_fk)
    if W is not None:
        ch.set_ff(W)
    ch.set_kpm(R=8, M=64, seed=1, jackson=True)
    info = ch.kpm_delta()
    assert info["mode"] == fk.KPM_DELTA_LOCAL and info["moments"] >= 64 and info["frequencies"] == 512
    tr = ch.sweep(150, trace=True)
    worst = 0.0
    for c in range(B):
        trc = {k: v[:, c] for k, v in tr.items()}
        _, dw = _compare_local_chain(orc, kind, D, L, beta, U, U / 2, int(seeds[c]), f0[c], moves_orc, trc, tp, W)
        worst = max(worst, dw)
    print(name, "moments", info["moments"], "max |dlogw| vs oracle", worst, flush=True)
    ch.close()
    ctx.close()


def test_kpm_local_graph_replay_and_refusals(fk):
    """Graph replay equals eager launches; randomize is refused with the local estimator
    and runs with the stochastic one."""
    print("case kpm local graph", flush=True)
    L, n = 16, 256
    rng = np.random.default_rng(4)
    f0 = rng.integers(0, 2, (4, n)).astype(np.int32)
    seeds = np.arange(4, dtype=np.uint32) + 70
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    a = fk.Chains(ctx, f0, 6.0, 2.0, 1.0, 1.0, seeds)
    b = fk.Chains(ctx, f0, 6.0, 2.0, 1.0, 1.0, seeds)
    for x in (a, b):
        x.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
        x.set_kpm(R=8, M=64, seed=2, jackson=True)
    for rep in range(3):
        a.sweep(8)
        b.sweep_graph(8)
        sa, sb = a.state(), b.state()
        assert np.array_equal(sa["f"], sb["f"]) and np.array_equal(sa["logZ"], sb["logZ"]), rep
    a.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_RANDOMIZE, 0.5)])
    with pytest.raises(fk.FkgpuError, match="randomize"):
        a.sweep(2)
    a.set_kpm_delta(fk.KPM_DELTA_STOCHASTIC)
    a.sweep(2)
    with pytest.raises(fk.FkgpuError):
        a.set_kpm_delta(fk.KPM_DELTA_LOCAL, moments=1)
    a.close()
    b.close()
    ctx.close()


def test_kpm_local_large_lattice(fk):
    """cubic2d L=64 (N=4096), beta=10, U=2, addremove / exchange: dOmega of the traced
    proposals against exact FP64 dense eigensolves of both configurations (LAPACK via
    numpy) within 1e-9; the reference's own Chebyshev weights are held to 6e-2
    (benchmark/fast_update.cpp:67-72)."""
    print("case kpm local L=64", flush=True)
    L = 64
    n = L * L
    beta, U = 10.0, 2.0
    rng = np.random.default_rng(2)
    B = 2
    f0 = rng.integers(0, 2, (B, n)).astype(np.int32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, [1, 2])
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    ch.set_kpm(R=64, M=256, seed=5, jackson=True)
    lz0 = ch.state()["logZ"].copy()
    tr = ch.sweep(3, trace=True)
    checked = 0
    for c in range(B):
        f = f0[c].copy()
        lz = lz0[c]
        ex_cur = _exact_logz_2d(L, f, beta, U, U / 2)
        for s in range(3):
            a, b = tr["a"][s, c], tr["b"][s, c]
            g = f.copy()
            if tr["kind"][s, c] == fk.MOVE_HOP:
                g[a], g[b] = f[b], f[a]
            else:
                g[a] ^= 1
            solved = tr["logz"][s, c] != 0.0
            ex_new = None
            if solved:
                ex_new = _exact_logz_2d(L, g, beta, U, U / 2)
                d_loc = -(tr["logz"][s, c] - lz) / beta
                d_ex = -(ex_new - ex_cur) / beta
                print("chain", c, "step", s, "dOmega local", d_loc, "exact", d_ex, flush=True)
                assert abs(d_loc - d_ex) < 1e-9
                checked += 1
            if tr["acc"][s, c]:
                f = g
                if solved:
                    lz = tr["logz"][s, c]
                    ex_cur = ex_new
    assert checked >= 2
    ch.close()
    ctx.close()


def test_occupation_only_measurements(fk):
    """FKGPU_MEASURE_NO_SPECTRUM: occupations recorded without an eigensolve, also above
    the dense range (cubic2d L=64 on the Chebyshev path); at a size with a dense solver
    the occupation records equal those of a fully measured twin chain set."""
    print("case occupation-only measures", flush=True)
    for L, kpm in ((64, True), (12, False)):
        n = L * L
        rng = np.random.default_rng(L)
        f0 = rng.integers(0, 2, (3, n)).astype(np.int32)
        ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
        sets = []
        for flags in (fk.MEASURE_FOCC_HISTORY | fk.MEASURE_NO_SPECTRUM,
                      fk.MEASURE_FOCC_HISTORY | fk.MEASURE_SPECTRUM_HISTORY):
            if kpm and not (flags & fk.MEASURE_NO_SPECTRUM):
                continue
            ch = fk.Chains(ctx, f0, 4.0, 2.0, 1.0, 1.0, [3, 4, 5])
            ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
            if kpm:
                ch.set_kpm(R=8, M=32, seed=1, jackson=False)
            ch.measure_enable(6, flags)
            ch.sweep_measure(6, 2)
            c, e, _, _ = ch.energies()
            h = ch.histories()
            st = ch.state()
            assert c == 6
            np.testing.assert_array_equal(h["focc"][:, -1], st["f"])
            if flags & fk.MEASURE_NO_SPECTRUM:
                assert np.all(e == 0.0)
            sets.append((h["focc"], ch.means()["focc"]))
            ch.close()
        if len(sets) == 2:
            np.testing.assert_array_equal(sets[0][0], sets[1][0])
            np.testing.assert_array_equal(sets[0][1], sets[1][1])
        with pytest.raises(fk.FkgpuError):
            bad = fk.Chains(ctx, f0, 4.0, 2.0, 1.0, 1.0, [3, 4, 5])
            try:
                bad.measure_enable(4, fk.MEASURE_NO_SPECTRUM | fk.MEASURE_SPECTRUM_HISTORY)
            finally:
                bad.close()
        ctx.close()
