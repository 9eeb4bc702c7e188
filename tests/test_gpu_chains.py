"""GPU Metropolis chains against the oracle chain (src/mc_metropolis.cpp:34-52,
src/moves.cpp).

For a fixed seed the device chain must draw the same sites and uniforms as the
reference chain, and take the same accept/reject decisions except at near-ties
|log u - log w| < 1e-9 (counted).  After a near-tie decided differently the
oracle is re-run following the device's decision there, so the comparison
continues over the whole chain.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NEAR_TIE = 1e-9


def compare_chain(orc, kind, D, L, beta, U, mu_c, mu_f, seed, f0, moves, tr, t=1.0, tp=1.0):
    """Compare one device chain trace with the oracle; returns #near-ties."""
    nsteps = len(tr["u"])
    forced = -np.ones(nsteps, np.int32)
    for _attempt in range(8):
        ref = orc.run_chain(kind, D, L, beta, U, mu_c, mu_f, seed, f0, moves, nsteps, t=t, tp=tp, forced=forced)
        assert np.array_equal(tr["kind"], ref["kind"]), "move choice differs"
        assert np.array_equal(tr["a"], ref["a"]), "site draw differs"
        assert np.array_equal(tr["b"], ref["b"]), "second site differs"
        assert np.array_equal(tr["u"], ref["u"]), "uniform differs"
        solved = ref["logz"] != 0.0
        dl = np.abs(tr["logz"][solved] - ref["logz"][solved])
        assert dl.max(initial=0.0) <= 1e-10 * np.abs(ref["logz"][solved]).max(initial=1.0), dl.max()
        mism = np.nonzero(tr["acc"] != ref["acc"])[0]
        if len(mism) == 0:
            return int(np.sum(forced >= 0))
        s = mism[0]
        gap = abs(np.log(ref["u"][s]) - ref["logw"][s])
        assert gap < NEAR_TIE, f"accept differs at move {s}: |log u - log w| = {gap}"
        forced[s] = tr["acc"][s]
    raise AssertionError("too many near-ties")


@pytest.mark.parametrize("D,L,nsteps", [(2, 8, 2000), (1, 30, 500), (2, 16, 150)])
def test_addremove_chain(fk, orc, D, L, nsteps):
    print(f"case addremove D={D} L={L}", flush=True)
    N = L ** D
    rng = np.random.default_rng(L)
    B = 4
    f0 = rng.integers(0, 2, size=(B, N)).astype(np.int32)
    beta, U = 1.0, 1.0
    seeds = np.arange(B, dtype=np.uint32) + 1000
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
    ch.set_moves([(fk.MOVE_ADDREMOVE, 1.0)])
    tr = ch.sweep(nsteps, trace=True)
    for c in range(B):
        trc = {k: v[:, c] for k, v in tr.items()}
        compare_chain(orc, orc.KIND_HYPERCUBIC, D, L, beta, U, U / 2, U / 2, int(seeds[c]), f0[c],
                      [(orc.MOVE_ADDREMOVE, 1.0)], trc, tp=0.0)
    st = ch.state()
    ref = orc.run_chain(orc.KIND_HYPERCUBIC, D, L, beta, U, U / 2, U / 2, int(seeds[0]), f0[0],
                        [(orc.MOVE_ADDREMOVE, 1.0)], nsteps, tp=0.0, forced=None)
    if np.array_equal(tr["acc"][:, 0], ref["acc"]):
        assert np.array_equal(st["f"][0], ref["f_final"])
        assert np.abs(st["spectrum"][0] - ref["spectrum_final"]).max() < 1e-10 * np.abs(ref["spectrum_final"]).max()
    ch.close()
    ctx.close()


def test_mixed_moves_chain(fk, orc):
    """addremove + global exchange + nearest-neighbour exchange + randomize."""
    print("case mixed moves", flush=True)
    D, L = 2, 6
    N = L ** D
    rng = np.random.default_rng(3)
    B = 3
    f0 = rng.integers(0, 2, size=(B, N)).astype(np.int32)
    f0[2] = 0  # empty configuration: flip returns weight 0 without drawing
    beta = np.array([0.7, 2.0, 5.0])
    U = np.array([1.0, 2.0, 4.0])
    seeds = np.array([11, 12, 13], np.uint32)
    moves_fk = [(fk.MOVE_ADDREMOVE, 0.4), (fk.MOVE_FLIP, 0.3), (fk.MOVE_HOP, 0.25), (fk.MOVE_RANDOMIZE, 0.05)]
    moves_orc = [(orc.MOVE_ADDREMOVE, 0.4), (orc.MOVE_FLIP, 0.3), (orc.MOVE_HOP, 0.25), (orc.MOVE_RANDOMIZE, 0.05)]
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
    ch.set_moves(moves_fk)
    tr = ch.sweep(600, trace=True)
    for c in range(B):
        trc = {k: v[:, c] for k, v in tr.items()}
        compare_chain(orc, orc.KIND_HYPERCUBIC, D, L, beta[c], U[c], U[c] / 2, U[c] / 2, int(seeds[c]), f0[c],
                      moves_orc, trc, tp=0.0)
    ch.close()
    ctx.close()


def test_graph_replay_matches_eager(fk, orc):
    """hipGraph-captured move steps give the same chain state as eager launches."""
    print("case graph replay", flush=True)
    D, L = 2, 8
    N = L ** D
    rng = np.random.default_rng(9)
    f0 = rng.integers(0, 2, size=(5, N)).astype(np.int32)
    seeds = np.arange(5, dtype=np.uint32) + 77
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    a = fk.Chains(ctx, f0, 2.0, 2.0, 1.0, 1.0, seeds)
    b = fk.Chains(ctx, f0, 2.0, 2.0, 1.0, 1.0, seeds)
    for ch in (a, b):
        ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    for rep in range(3):
        a.sweep(16)
        b.sweep_graph(16)
        sa, sb = a.state(), b.state()
        assert np.array_equal(sa["f"], sb["f"]), rep
        assert np.array_equal(sa["logZ"], sb["logZ"]), rep
        assert np.array_equal(sa["naccept"], sb["naccept"]), rep
    a.close()
    b.close()
    ctx.close()


@pytest.mark.parametrize("D,L,B", [(2, 8, 3), (2, 8, 64), (2, 16, 64), (2, 16, 300), (1, 300, 2), (1, 1030, 2), (2, 6, 1)])
def test_steps_per_launch_agree(fk, D, L, B):
    """Several move steps per launch (one workgroup per (step, chain) item, each chain's
    steps in order) draw, decide and end exactly as one launch per step: traces and final
    states bit-identical, eager and graph-replayed, for every dense solver kind."""
    print(f"case steps per launch D={D} L={L} B={B}", flush=True)
    N = L ** D
    rng = np.random.default_rng(L + B)
    f0 = rng.integers(0, 2, size=(B, N)).astype(np.int32)
    beta = np.linspace(1.0, 8.0, B)
    U = np.linspace(1.0, 4.0, B)
    seeds = np.arange(B, dtype=np.uint32) + 77
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    nsteps = 12 if N > 1000 else 40
    res = []
    for spl in (1, 3, 16):
        ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
        ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_FLIP, 0.2), (fk.MOVE_HOP, 0.3)])
        ch.set_steps_per_launch(spl)
        tr = ch.sweep(nsteps, trace=True)
        ch.sweep_graph(7)
        ch.sweep_graph(7)
        res.append((tr, ch.state()))
        ch.close()
    for tr, st in res[1:]:
        for k in tr:
            assert np.array_equal(tr[k], res[0][0][k]), k
        for k in ("f", "nf", "logZ", "spectrum", "naccept", "nsolve"):
            assert np.array_equal(st[k], res[0][1][k]), k
    ctx.close()


def test_randomize_f_from_chain_generator(fk, orc):
    """fkgpu_chains_randomize_f draws configuration_t::randomize_f from each chain's own
    generator (src/configuration.cpp:47-56): the oracle's randomize_f of the same seed,
    including the drawn count for nf_start = 0; then the chain continues exactly like an
    oracle chain that started from that configuration with the advanced generator."""
    print("case randomize_f", flush=True)
    D, L, B = 2, 6, 4
    N = L * L
    seeds = np.array([5, 6, 7, 8], np.uint32)
    nf = np.array([0, 1, N // 2, N], np.int32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    ch = fk.Chains(ctx, np.zeros((B, N), np.int32), 2.0, 3.0, 1.5, 1.5, seeds)
    ch.randomize_f(nf)
    st = ch.state()
    for c in range(B):
        ref = orc.randomize_f(N, int(seeds[c]), int(nf[c]))
        np.testing.assert_array_equal(st["f"][c], ref)
        assert st["nf"][c] == ref.sum()
        ed = orc.calc_ed(orc.KIND_HYPERCUBIC, D, L, ref, 2.0, 3.0, 1.5, tp=0.0)
        np.testing.assert_allclose(st["spectrum"][c], ed["evals"], rtol=0, atol=1e-11)
        np.testing.assert_allclose(st["logZ"][c], ed["logZ"], rtol=1e-12)
    with pytest.raises(fk.FkgpuError):
        ch.randomize_f(N + 1)
    ch.close()
    ctx.close()


def test_queue_wait_bound_reports_error(fk):
    """The move-step queue's bounded wait, forced from the host: with a zero bound a
    workgroup that finds its slot not yet filled gives up at once, every later one leaves,
    the launch ends normally and the call reports the error (and keeps reporting it)."""
    print("case queue wait bound", flush=True)
    L = 8
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    ch = fk.Chains(ctx, np.ones((1, L * L), np.int32), 2.0, 2.0, 1.0, 1.0, np.array([3], np.uint32))
    ch.set_moves([(fk.MOVE_ADDREMOVE, 1.0)])
    ch.set_queue_timeout(0.0)
    # one chain, 16 steps in one launch: steps 1..15 wait for their predecessors
    with pytest.raises(fk.FkgpuError):
        ch.sweep(16)
    with pytest.raises(fk.FkgpuError):
        ch.sweep_graph(16)
    ch.close()
    # a fresh chain set with the default bound runs normally
    ch = fk.Chains(ctx, np.ones((1, L * L), np.int32), 2.0, 2.0, 1.0, 1.0, np.array([3], np.uint32))
    ch.set_moves([(fk.MOVE_ADDREMOVE, 1.0)])
    ch.sweep(16)
    ch.close()
    ctx.close()
