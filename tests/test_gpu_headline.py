"""The batched FP64 workload at its full shape: cubic2d L=16 (N=256), 1024 chains,
beta_c = 0.5 * 40**(c/1023), U_c = {1,2,4,8}[c mod 4], mu_c = mu_f = U/2,
f ~ Bernoulli(0.5), addremove / nearest-neighbour exchange at 0.5 / 0.5
(mc_metropolis::update, src/mc_metropolis.cpp:34-52; src/moves.cpp:52-67).

* 16 move steps per launch (one workgroup per (step, chain) item handed over
  between XCDs through the ready-chain queue) give traces and final states
  bit-identical to one launch per step over 48 steps (three full launches).
* Eight chains spread over beta and U reproduce the oracle chain of the same
  seed and start (draws bit for bit, decisions except at counted near-ties).
"""
import numpy as np
import pytest

from test_gpu_chains import compare_chain

pytestmark = pytest.mark.gpu

L, B, NSTEPS = 16, 1024, 48


def _workload(seed=20240601):
    N = L * L
    rng = np.random.default_rng(seed)
    f = (rng.random((B, N)) < 0.5).astype(np.int32)
    c = np.arange(B)
    beta = 0.5 * 40.0 ** (c / (B - 1))
    U = np.array([1.0, 2.0, 4.0, 8.0])[c % 4]
    seeds = (seed + c).astype(np.uint32)
    return N, f, beta, U, seeds


@pytest.fixture(scope="module")
def headline_runs(fk):
    N, f, beta, U, seeds = _workload()
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    out = {}
    for spl in (16, 1):
        print(f"case headline B={B} N={N} steps_per_launch={spl}", flush=True)
        ch = fk.Chains(ctx, f, beta, U, U / 2, U / 2, seeds)
        ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
        ch.set_steps_per_launch(spl)
        tr = ch.sweep(NSTEPS, trace=True)
        out[spl] = (tr, ch.state())
        ch.close()
    ctx.close()
    return (N, f, beta, U, seeds), out


def test_headline_steps_per_launch_bit_identical(headline_runs):
    _, out = headline_runs
    tr16, st16 = out[16]
    tr1, st1 = out[1]
    for k in tr1:
        a, b = tr16[k], tr1[k]
        same = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
        if not same.all():
            s, c = np.argwhere(~same)[0]
            raise AssertionError(f"trace field {k} differs first at step {s}, chain {c}")
    for k in ("f", "nf", "logZ", "spectrum", "naccept", "nsolve"):
        assert np.array_equal(st16[k], st1[k]), k
    # the workload really moved: about 3/4 of the proposals need a solve
    assert 0.6 * B * NSTEPS < st1["nsolve"].sum() < 0.9 * B * NSTEPS


@pytest.mark.parametrize("c", [0, 129, 258, 387, 516, 645, 894, 1023])
def test_headline_chains_match_oracle(fk, orc, headline_runs, c):
    (N, f, beta, U, seeds), out = headline_runs
    tr16, _ = out[16]
    print(f"case headline oracle chain {c} beta={beta[c]:.3f} U={U[c]}", flush=True)
    trc = {k: v[:, c] for k, v in tr16.items()}
    compare_chain(orc, orc.KIND_HYPERCUBIC, 2, L, beta[c], U[c], U[c] / 2, U[c] / 2, int(seeds[c]), f[c],
                  [(orc.MOVE_ADDREMOVE, 0.5), (orc.MOVE_HOP, 0.5)], trc, tp=0.0)


def test_hop_rejected_on_isolated_site(fk):
    """A CSR lattice with an isolated site has no neighbour to exchange with: the
    exchange move is refused when the moves are set, not drawn into an undefined slot."""
    n = 9
    hop = np.zeros((n, n))
    for i in range(n - 2):  # a path 0-1-...-7, site 8 isolated
        hop[i, i + 1] = hop[i + 1, i] = -1.0
    ctx = fk.Context(0).set_lattice_csr(hop)
    ch = fk.Chains(ctx, np.zeros((1, n), np.int32), 1.0, 1.0, 0.5, 0.5, np.array([1], np.uint32))
    with pytest.raises(fk.FkgpuError):
        ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    ch.set_moves([(fk.MOVE_ADDREMOVE, 1.0)])
    ch.sweep(4)
    ch.close()
    ctx.close()
