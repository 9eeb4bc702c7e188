"""Mixed-precision path: FP32 Householder reduction (MFMA trailing updates) with FP64
refinement of the eigenvalues, against the FP64 oracle; mixed-precision chains against
FP64 chains on the same seeds (|dOmega| differences and accept-sequence mismatches)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(2)
    out = []
    for kind, D, L, U, beta in [(0, 2, 8, 2.0, 2.0), (0, 2, 16, 1.0, 10.0), (0, 2, 16, 8.0, 0.5),
                                (0, 3, 8, 2.0, 2.0), (0, 3, 8, 5.0, 10.0), (1, 2, 24, 3.0, 5.0)]:
        n = L ** D
        filling = rng.uniform(0.3, 0.7)
        f = (rng.random((3, n)) < filling).astype(np.int32)
        out.append((kind, D, L, U, beta, f))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: f"k{c[0]}D{c[1]}L{c[2]}U{c[3]}b{c[4]}")
def test_mixed_eigenvalues_vs_oracle(fk, orc, case):
    kind, D, L, U, beta, f = case
    print(f"case mixed eig kind={kind} D={D} L={L} U={U} beta={beta}", flush=True)
    ctx = fk.Context(0).set_lattice(kind, D, L, 1.0, 1.0)
    g = ctx.eig_logz_mixed(f, beta, U, U / 2)
    om = []
    for b in range(f.shape[0]):
        r = orc.calc_ed(kind, D, L, f[b], beta, U, U / 2)
        scale = max(1.0, np.abs(r["evals"]).max())
        err = np.abs(g["evals"][b] - r["evals"]).max()
        print(f"  max eigenvalue error {err:.2e}, fallbacks {g['nfallback']}", flush=True)
        assert err <= 1e-11 * scale
        assert abs(g["logZ"][b] - r["logZ"]) <= 1e-12 * abs(r["logZ"])
        om.append((-g["logZ"][b] / beta, -r["logZ"] / beta))
    om = np.array(om)
    d_mixed = np.diff(om[:, 0])
    d_ref = np.diff(om[:, 1])
    assert np.abs(d_mixed - d_ref).max() <= 1e-10
    ctx.close()


@pytest.mark.parametrize("mode", ["uniform", "checkerboard", "empty"])
def test_mixed_degenerate_spectra(fk, orc, mode):
    """Exactly degenerate levels (lattice symmetry): clusters are diagonalised exactly or
    handed to the FP64 solver; either way the result is FP64-accurate."""
    D, L, U = 2, 16, 2.0
    n = L * L
    if mode == "uniform":
        f = np.ones(n, np.int32)
    elif mode == "empty":
        f = np.zeros(n, np.int32)
    else:
        x, y = np.divmod(np.arange(n), L)
        f = ((x + y) % 2).astype(np.int32)
    print(f"case mixed degenerate {mode}", flush=True)
    ctx = fk.Context(0).set_lattice(0, D, L, 1.0, 0.0)
    g = ctx.eig_logz_mixed(f[None, :], 3.0, U, U / 2)
    r = orc.calc_ed(0, D, L, f, 3.0, U, U / 2)
    assert np.abs(g["evals"][0] - r["evals"]).max() <= 1e-11 * max(1.0, np.abs(r["evals"]).max())
    assert abs(g["logZ"][0] - r["logZ"]) <= 1e-12 * abs(r["logZ"])
    ctx.close()


@pytest.mark.parametrize("D,L,B,steps", [(2, 16, 8, 40), (3, 8, 4, 12)])
def test_mixed_chains_vs_fp64(fk, D, L, B, steps):
    n = L ** D
    rng = np.random.default_rng(n)
    f0 = (rng.random((B, n)) < rng.uniform(0.3, 0.7, (B, 1))).astype(np.int32)
    beta = np.geomspace(0.5, 20.0, B)
    U = np.array([[1.0, 2.0, 4.0, 8.0][c % 4] for c in range(B)])
    seeds = np.arange(B, dtype=np.uint32) + 500
    ctx = fk.Context(0).set_lattice(0, D, L, 1.0, 0.0)
    print(f"case mixed chains D={D} L={L}", flush=True)
    traces = []
    for prec in (fk.PRECISION_FP64, fk.PRECISION_MIXED):
        ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
        ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
        ch.set_precision(prec)
        traces.append((ch.sweep(steps, trace=True), ch.state(), ch.fallbacks()))
        ch.close()
    (t64, s64, _), (tmx, smx, nfb) = traces
    print(f"  fallbacks {nfb}", flush=True)
    ties = 0
    for c in range(B):
        for s in range(steps):
            if t64["acc"][s, c] != tmx["acc"][s, c]:
                gap = abs(np.log(t64["u"][s, c]) - t64["logw"][s, c])
                assert gap < 1e-9, f"accept differs at chain {c} move {s}, |log u - log w| = {gap}"
                ties += 1
                break  # the chains follow different paths afterwards
            fin = np.isfinite(t64["logw"][s, c])
            if fin:
                assert abs(tmx["logw"][s, c] - t64["logw"][s, c]) <= 1e-10 * beta[c] + 1e-12
        else:
            assert np.array_equal(s64["f"][c], smx["f"][c])
            assert abs(s64["logZ"][c] - smx["logZ"][c]) <= 1e-12 * abs(s64["logZ"][c])
            assert np.abs(s64["spectrum"][c] - smx["spectrum"][c]).max() <= 1e-11 * np.abs(s64["spectrum"][c]).max()
    ctx.close()


def test_mixed_chain_graph_replay(fk):
    """The mixed step captured in a hipGraph gives the eager result, bit for bit."""
    D, L, B = 2, 8, 6
    n = L * L
    f0 = (np.random.default_rng(4).random((B, n)) < 0.5).astype(np.int32)
    seeds = np.arange(B, dtype=np.uint32) + 9
    ctx = fk.Context(0).set_lattice(0, D, L, 1.0, 0.0)
    print("case mixed graph replay", flush=True)
    states = []
    for graph in (False, True):
        ch = fk.Chains(ctx, f0, 2.0, 4.0, 2.0, 2.0, seeds)
        ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
        ch.set_precision(fk.PRECISION_MIXED)
        for rep in range(3):
            if graph:
                ch.sweep_graph(5)
            else:
                ch.sweep(5)
        states.append(ch.state())
        ch.close()
    assert np.array_equal(states[0]["f"], states[1]["f"])
    assert np.array_equal(states[0]["logZ"], states[1]["logZ"])
    ctx.close()
