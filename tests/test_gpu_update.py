"""Rank-one update chains (secular equation + Gu-Eisenstat eigenvectors, V <- V Q on
acceptance) against the oracle chain: same sites, uniforms and decisions except at
counted near-ties, proposal logZ within 1e-10 of the oracle's dense FP64 eigensolve."""
import numpy as np
import pytest

from test_gpu_chains import compare_chain

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("D,L,moves,steps,refresh", [
    (2, 8, "addremove", 300, 0),
    (2, 8, "mixed", 300, 50),
    (2, 16, "mixed", 120, 0),
    (1, 40, "mixed", 200, 0),
    (3, 4, "addremove", 200, 64),
])
def test_rank1_chain_vs_oracle(fk, orc, D, L, moves, steps, refresh):
    N = L ** D
    print(f"case rank1 chain D={D} L={L} {moves} refresh={refresh}", flush=True)
    rng = np.random.default_rng(N + steps)
    B = 4
    f0 = rng.integers(0, 2, (B, N)).astype(np.int32)
    beta = np.array([0.5, 2.0, 8.0, 20.0])
    U = np.array([1.0, 2.0, 4.0, 8.0])
    seeds = np.arange(B, dtype=np.uint32) + 40
    if moves == "addremove":
        mv_fk, mv_orc = [(fk.MOVE_ADDREMOVE, 1.0)], [(orc.MOVE_ADDREMOVE, 1.0)]
    else:
        mv_fk = [(fk.MOVE_ADDREMOVE, 0.4), (fk.MOVE_HOP, 0.4), (fk.MOVE_FLIP, 0.2)]
        mv_orc = [(orc.MOVE_ADDREMOVE, 0.4), (orc.MOVE_HOP, 0.4), (orc.MOVE_FLIP, 0.2)]
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
    ch.set_moves(mv_fk)
    ch.set_update(fk.UPDATE_RANK1, refresh)
    tr = ch.sweep(steps, trace=True)
    ties = 0
    for c in range(B):
        trc = {k: v[:, c] for k, v in tr.items()}
        ties += compare_chain(orc, orc.KIND_HYPERCUBIC, D, L, beta[c], U[c], U[c] / 2, U[c] / 2, int(seeds[c]),
                              f0[c], mv_orc, trc, tp=0.0)
    st = ch.state()
    ref = ctx.eig_logz(st["f"], beta, U, U / 2)
    scale = np.abs(ref["evals"]).max()
    assert np.abs(st["spectrum"] - ref["evals"]).max() < 1e-11 * scale
    assert np.abs(st["logZ"] - ref["logZ"]).max() < 1e-11 * np.abs(ref["logZ"]).max()
    print(f"  near-ties {ties}", flush=True)
    ch.close()
    ctx.close()


def test_rank1_degenerate_start(fk, orc):
    """Empty and full lattices: highly degenerate spectra exercise both deflations."""
    D, L = 2, 8
    N = L * L
    f0 = np.stack([np.zeros(N), np.ones(N)]).astype(np.int32)
    seeds = np.array([3, 4], np.uint32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    print("case rank1 degenerate", flush=True)
    ch = fk.Chains(ctx, f0, 3.0, 2.0, 1.0, 1.0, seeds)
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    ch.set_update(fk.UPDATE_RANK1, 0)
    tr = ch.sweep(150, trace=True)
    for c in range(2):
        trc = {k: v[:, c] for k, v in tr.items()}
        compare_chain(orc, orc.KIND_HYPERCUBIC, D, L, 3.0, 2.0, 1.0, 1.0, int(seeds[c]), f0[c],
                      [(orc.MOVE_ADDREMOVE, 0.5), (orc.MOVE_HOP, 0.5)], trc, tp=0.0)
    ch.close()
    ctx.close()


def test_rank1_graph_replay(fk):
    D, L, B = 2, 8, 8
    N = L * L
    f0 = np.random.default_rng(9).integers(0, 2, (B, N)).astype(np.int32)
    seeds = np.arange(B, dtype=np.uint32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    print("case rank1 graph replay", flush=True)
    out = []
    for graph in (False, True):
        ch = fk.Chains(ctx, f0, 2.0, 4.0, 2.0, 2.0, seeds)
        ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
        ch.set_update(fk.UPDATE_RANK1, 10)
        for _ in range(3):
            if graph:
                ch.sweep_graph(10)
            else:
                ch.sweep(10)
        out.append(ch.state())
        ch.close()
    assert np.array_equal(out[0]["f"], out[1]["f"])
    assert np.array_equal(out[0]["logZ"], out[1]["logZ"])
    ctx.close()


def test_rank1_rejects_randomize(fk):
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, 4, 1.0, 0.0)
    ch = fk.Chains(ctx, np.zeros((1, 16), np.int32), 1.0, 1.0, 0.5, 0.5, np.array([1], np.uint32))
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_RANDOMIZE, 0.5)])
    with pytest.raises(fk.FkgpuError):
        ch.set_update(fk.UPDATE_RANK1, 0)
    ch.close()
    ctx.close()


def test_rank1_refresh_period_across_calls(fk):
    """The re-decomposition period is counted on the device: 16 steps in one call, in four
    calls of 4 steps and in graph replays give bit-identical chains (re-decomposed after
    steps 8 and 16 in every case)."""
    D, L, B = 2, 8, 4
    N = L * L
    rng = np.random.default_rng(77)
    f0 = rng.integers(0, 2, (B, N)).astype(np.int32)
    beta = np.array([0.5, 2.0, 8.0, 20.0])
    U = np.array([1.0, 2.0, 4.0, 8.0])
    seeds = np.arange(B, dtype=np.uint32) + 90
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, D, L, 1.0, 0.0)
    print("case rank1 refresh period", flush=True)
    states = []
    for mode in ("one", "four", "graph"):
        ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
        ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
        ch.set_update(fk.UPDATE_RANK1, 8)
        if mode == "one":
            ch.sweep(16)
        elif mode == "four":
            for _ in range(4):
                ch.sweep(4)
        else:
            for _ in range(4):
                ch.sweep_graph(4)
        states.append(ch.state())
        ch.close()
    for k in ("f", "logZ", "spectrum", "naccept"):
        assert np.array_equal(states[0][k], states[1][k]), k
        assert np.array_equal(states[0][k], states[2][k]), k
    # and the last re-decomposition (after step 16) left the spectrum of a fresh solve
    ref = ctx.eig_logz(states[0]["f"], beta, U, U / 2)
    assert np.array_equal(states[0]["spectrum"], ref["evals"])
    ctx.close()
