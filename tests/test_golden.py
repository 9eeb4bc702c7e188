"""The oracle against the committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py from LAPACK, closed forms and the reference's own KATs)."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_cases(name):
    z = np.load(os.path.join(GOLD, name))
    out = []
    for c in range(int(z["ncases"])):
        kind, D, L, U, mu = z[f"case{c}_meta"]
        out.append((int(kind), int(D), int(L), float(U), float(mu), z[f"case{c}_f"].astype(np.int32),
                    z[f"case{c}_evals"]))
    return out


@pytest.mark.parametrize("case", load_cases("eigs_lapack.npz"), ids=lambda c: f"k{c[0]}D{c[1]}L{c[2]}U{c[3]}")
def test_oracle_eigs_vs_lapack_fixture(orc, case):
    kind, D, L, U, mu, f, ev = case
    r = orc.calc_ed(kind, D, L, f, 1.0, U, mu)
    assert np.abs(r["evals"] - ev).max() <= 1e-12 * max(1.0, np.abs(ev).max())


@pytest.mark.parametrize("case", load_cases("closed_forms.npz"), ids=lambda c: f"D{c[1]}L{c[2]}U{c[3]}")
def test_oracle_closed_form_fixture(orc, case):
    kind, D, L, U, mu, f, ev = case
    r = orc.calc_ed(kind, D, L, f, 1.0, U, mu)
    assert np.abs(r["evals"] - ev).max() <= 1e-12 * max(1.0, np.abs(ev).max())


def test_fixture_generator_checks_itself():
    """The LAPACK fixture and the closed forms agree where both apply (U=0 spectrum)."""
    cf = load_cases("closed_forms.npz")
    empty = [c for c in cf if c[3] == 0.0]
    assert len(empty) == 4
    for kind, D, L, U, mu, f, ev in empty:
        assert f.sum() == 0 and abs(np.sum(ev) + mu * L ** D) < 1e-9  # Tr H = -N mu_c


@pytest.mark.parametrize("kat", json.load(open(os.path.join(GOLD, "stiffness.json"))), ids=lambda k: f"U{k['U']}")
def test_oracle_stiffness_fixture(orc, kat):
    L = kat["L"]
    r = orc.calc_ed(0, 2, L, np.array(kat["f"], np.int32), kat["beta"], kat["U"], kat["mu"], vectors=True)
    s = orc.measure_stiffness(2, L, r["evecs"], r["evals"], r["cfermi"], wgrid=[0.0], offset=0.05)["stiffness"]
    assert abs(s - kat["reference_value"]) < kat["reference_tol"]
    assert abs(s - kat["lapack_value"]) < 1e-10


def test_oracle_cheb_quadrature_fixture(orc):
    g = json.load(open(os.path.join(GOLD, "cheb_quadrature.json")))
    fn_id = {"one": 0, "x^2": 1, "sin": 2}
    for c in g["cases"]:
        v = orc.cheb_moment_f(g["max_moment"], g["grid"], fn_id[c["fn"]], c["order"])
        assert abs(v - c["value"]) < c["tol"]
        if "reference_value" in c:
            assert abs(v - c["reference_value"]) < c["tol"]


def load_chain():
    z = np.load(os.path.join(GOLD, "chain_c1.npz"))
    D, L, U, beta, seed, nmoves = z["meta"]
    n = int(nmoves)
    return dict(D=int(D), L=int(L), U=float(U), beta=float(beta), seed=int(seed), nmoves=n,
                f0=z["f0"].astype(np.int32), site=z["site"].astype(np.int32),
                acc=np.unpackbits(z["acc"])[:n].astype(np.int8), logw=z["logw"], logz_cur=z["logz_cur"],
                f_final=z["f_final"].astype(np.int32), spectrum_final=z["spectrum_final"])


def test_oracle_chain_fixture(orc):
    c = load_chain()
    tr = orc.run_chain(orc.KIND_HYPERCUBIC, c["D"], c["L"], c["beta"], c["U"], c["U"] / 2, c["U"] / 2, c["seed"],
                       c["f0"], [(orc.MOVE_ADDREMOVE, 1.0)], c["nmoves"], tp=0.0)
    assert np.array_equal(tr["a"], c["site"])
    assert np.array_equal(tr["acc"].astype(np.int8), c["acc"])
    assert np.array_equal(tr["logw"], c["logw"])
    assert np.array_equal(tr["f_final"], c["f_final"])
