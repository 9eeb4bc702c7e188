"""The HIP path against the committed golden fixtures (LAPACK eigenvalues, closed
forms, stiffness KATs and the 10^4-move config-1 chain)."""
import json
import os

import numpy as np
import pytest

from test_golden import GOLD, load_chain, load_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["eigs_lapack.npz", "closed_forms.npz"])
def test_gpu_eigs_vs_fixture(fk, name):
    for kind, D, L, U, mu, f, ev in load_cases(name):
        print(f"case {name} kind={kind} D={D} L={L} U={U}", flush=True)
        ctx = fk.Context(0).set_lattice(kind, D, L, 1.0, 1.0 if kind == 1 else 0.0)
        g = ctx.eig_logz(f[None, :], 1.0, U, mu)
        scale = max(1.0, np.abs(ev).max())
        assert np.abs(g["evals"][0] - ev).max() <= 1e-12 * scale
        direct = np.sum(np.log1p(np.exp(-ev)))
        assert abs(g["logZ"][0] - direct) <= 1e-12 * abs(direct)
        ctx.close()


def test_gpu_stiffness_fixture(fk):
    for kat in json.load(open(os.path.join(GOLD, "stiffness.json"))):
        L = kat["L"]
        ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
        st, _ = ctx.measure_stiffness(np.array(kat["f"], np.int32)[None, :], kat["beta"], kat["U"], kat["mu"],
                                      wgrid=[0.0], offset=0.05)
        s = st[0]
        assert abs(s - kat["reference_value"]) < kat["reference_tol"]
        assert abs(s - kat["lapack_value"]) < 1e-8
        ctx.close()


def test_gpu_chain_fixture(fk):
    """Config 1 (cubic2d L=8, 10^4 addremove moves): the device chain reproduces the
    fixture's sites and decisions, except at near-ties |log u - log w| < 1e-9 (none here)."""
    c = load_chain()
    N = c["L"] ** c["D"]
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, c["D"], c["L"], 1.0, 0.0)
    ch = fk.Chains(ctx, c["f0"][None, :], c["beta"], c["U"], c["U"] / 2, c["U"] / 2,
                   np.array([c["seed"]], np.uint32))
    ch.set_moves([(fk.MOVE_ADDREMOVE, 1.0)])
    tr = ch.sweep(c["nmoves"], trace=True)
    ties = np.abs(np.log(tr["u"][:, 0]) - c["logw"]) < 1e-9
    assert not ties.any()
    assert np.array_equal(tr["a"][:, 0], c["site"])
    assert np.array_equal(tr["acc"][:, 0], c["acc"])
    np.testing.assert_allclose(tr["logw"][:, 0], c["logw"], rtol=0, atol=1e-10)
    st = ch.state()
    assert np.array_equal(st["f"][0], c["f_final"])
    assert np.abs(st["spectrum"][0] - c["spectrum_final"]).max() < 1e-12 * N
    ch.close()
    ctx.close()


def test_host_chain_fixture(fk):
    """The same chain through the C++ host layer (configuration_t / move_addremove / mc_metropolis)."""
    from fk_mc_amd import host_api
    c = load_chain()
    n = 2000
    m = host_api.model(fk.LATTICE_HYPERCUBIC, c["D"], c["L"], c["beta"], c["U"], c["U"] / 2, c["U"] / 2, tp=0.0)
    tr = host_api.run_chain(m, c["seed"], c["f0"], [(fk.MOVE_ADDREMOVE, 1.0)], n)
    assert np.array_equal(tr["a"], c["site"][:n])
    assert np.array_equal(tr["acc"].astype(np.int8), c["acc"][:n])
    np.testing.assert_allclose(np.log(tr["w"]), c["logw"][:n], rtol=0, atol=1e-10)
    np.testing.assert_allclose(tr["logz_cur"][::100], c["logz_cur"][:n // 100], rtol=1e-12)
