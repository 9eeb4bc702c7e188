"""Chains on several lattices in one chain set (fkgpu_chains_create_multi): every chain's
move steps (mc_metropolis::update, src/mc_metropolis.cpp:34-52, with the moves of
src/moves.cpp) run in the same persistent launch, each with its own lattice size and
neighbour table.

* Each chain reproduces the oracle's chain of its lattice, seed and start (draws bit for
  bit, decisions except at counted near-ties).
* Each chain is bit-identical to the same chain in a chain set of its own lattice run with
  the same solver, so the per-chain lattice plumbing changes nothing.
* 16 steps per launch equal one launch per step.
"""
import numpy as np
import pytest

from test_gpu_chains import compare_chain

pytestmark = pytest.mark.gpu

# (kind, D, L, t, tp)
LATS_SMALL = [(0, 2, 8, 1.0, 0.0), (0, 2, 12, 1.0, 0.0), (0, 2, 16, 1.0, 0.0), (1, 2, 10, 1.0, 1.0),
              (0, 1, 30, 1.0, 0.0)]
LATS_LARGE = [(0, 2, 16, 1.0, 0.0), (1, 2, 20, 1.0, 1.0), (0, 3, 8, 1.0, 0.0)]


def _setup(fk, lats, B, seed):
    ctxs = [fk.Context(0).set_lattice(k, D, L, t, tp) for (k, D, L, t, tp) in lats]
    Ns = [c.N for c in ctxs]
    nmax = max(Ns)
    rng = np.random.default_rng(seed)
    lat = np.arange(B) % len(lats)
    f = np.zeros((B, nmax), np.int32)
    for c in range(B):
        f[c, :Ns[lat[c]]] = rng.integers(0, 2, Ns[lat[c]])
    beta = np.geomspace(0.5, 12.0, B)
    U = np.array([1.0, 2.0, 4.0, 3.0])[np.arange(B) % 4]
    seeds = (seed + np.arange(B)).astype(np.uint32)
    return ctxs, Ns, lat, f, beta, U, seeds


MOVES = [(0, 0.5), (2, 0.3), (1, 0.2)]  # addremove, hop, flip


@pytest.mark.parametrize("lats,B,nsteps", [(LATS_SMALL, 20, 40), (LATS_LARGE, 9, 12)])
def test_multi_lattice_chains(fk, orc, lats, B, nsteps):
    print(f"case multi-lattice {len(lats)} lattices B={B}", flush=True)
    ctxs, Ns, lat, f, beta, U, seeds = _setup(fk, lats, B, 500 + B)
    runs = {}
    for spl in (16, 1):
        ch = fk.Chains(ctxs, f, beta, U, U / 2, U / 2, seeds, lattice=lat)
        ch.set_moves(MOVES)
        ch.set_steps_per_launch(spl)
        runs[spl] = (ch.sweep(nsteps, trace=True), ch.state())
        ch.close()
    for k in runs[1][0]:
        assert np.array_equal(runs[16][0][k], runs[1][0][k]), k
    for k in ("f", "nf", "logZ", "spectrum", "naccept", "nsolve"):
        assert np.array_equal(runs[16][1][k], runs[1][1][k]), k
    tr, st = runs[16]
    # rows are zero past each chain's lattice size
    for c in range(B):
        assert not st["f"][c, Ns[lat[c]]:].any() and not st["spectrum"][c, Ns[lat[c]]:].any()
    # the same chains, one lattice per chain set, the same solver
    solver = 10 if max(Ns) <= 256 else 8
    for li, (k, D, L, t, tp) in enumerate(lats):
        idx = np.nonzero(lat == li)[0]
        c1 = fk.Context(0).set_lattice(k, D, L, t, tp)
        c1.set_solver(solver)
        one = fk.Chains(c1, f[idx, :Ns[li]], beta[idx], U[idx], U[idx] / 2, U[idx] / 2, seeds[idx])
        one.set_moves(MOVES)
        t1 = one.sweep(nsteps, trace=True)
        s1 = one.state()
        for key in t1:
            assert np.array_equal(t1[key], tr[key][:, idx]), (li, key)
        assert np.array_equal(s1["f"], st["f"][idx, :Ns[li]]) and np.array_equal(s1["logZ"], st["logZ"][idx])
        one.close()
        c1.close()
        # and the oracle's chain of the first of them
        c = idx[0]
        trc = {key: v[:, c] for key, v in tr.items()}
        compare_chain(orc, k, D, L, beta[c], U[c], U[c] / 2, U[c] / 2, int(seeds[c]), f[c, :Ns[li]],
                      [(orc.MOVE_ADDREMOVE, 0.5), (orc.MOVE_HOP, 0.3), (orc.MOVE_FLIP, 0.2)], trc, t=t, tp=tp)
    for c in ctxs:
        c.close()


def test_multi_lattice_refusals(fk):
    print("case multi-lattice refusals", flush=True)
    ctxs, Ns, lat, f, beta, U, seeds = _setup(fk, LATS_SMALL[:2], 4, 9)
    ch = fk.Chains(ctxs, f, beta, U, U / 2, U / 2, seeds, lattice=lat)
    for call in (lambda: ch.set_kpm(R=4, M=8), lambda: ch.measure_enable(4, 0),
                 lambda: ch.set_update(fk.UPDATE_RANK1, 8), lambda: ch.randomize_f(3)):
        with pytest.raises(fk.FkgpuError):
            call()
    ch.sweep(2)
    ch.close()
    bad = lat.copy()
    bad[0] = 7
    with pytest.raises(fk.FkgpuError):
        fk.Chains(ctxs, f, beta, U, U / 2, U / 2, seeds, lattice=bad)
    for c in ctxs:
        c.close()
