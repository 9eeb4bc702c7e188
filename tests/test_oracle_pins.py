"""Pins of two oracle functions that no reference vector covers.

measure_energy (src/measures/energy.cpp:6-26) is checked against the
thermodynamic identities it must satisfy at a fixed occupation f, with logZ
from the oracle's own calc_ed (src/configuration.cpp:208-246, itself pinned to
LAPACK and closed forms in test_oracle.py):
    E_c = sum_k e_k f(e_k)       = -d logZ / d beta
    d2E = sum_k e_k^2 f(1 - f)   =  d^2 logZ / d beta^2
    E   = E_c - mu_f N_f + E_ff  = -d/d beta [logZ + beta mu_f N_f - beta E_ff]
by Richardson-extrapolated central differences in beta (the spectrum does not
depend on beta).  Mutated energy formulas (a sign slip in the Fermi factor, a
lost factor 1/2, a wrong sign of mu_f N_f) are shown to fail the same check.

The nearest-neighbour exchange ("f-hop", no reference counterpart) is checked by
exact enumeration: the oracle chain's histogram over all 2^9 configurations of
3x3 square, 3x3 triangular and 9-site 1-D lattices (with an f-f interaction)
against exp(logZ(f) + beta mu_f N_f - beta E_ff(f)), for the hop alone (inside
a fixed-N_f sector, which it conserves) and mixed with addremove.  Mistakes
planted in the oracle's hop (oracle.set_mutation) are shown to fail it.
"""
import itertools

import numpy as np
import pytest

# ---------------------------------------------------------------------------
# measure_energy against derivatives of logZ in beta
# ---------------------------------------------------------------------------


def _logz(orc, kind, D, L, f, beta, U, mu):
    return orc.calc_ed(kind, D, L, f, beta, U, mu, tp=1.0 if kind == 1 else 0.0)["logZ"]


def _derivs(orc, kind, D, L, f, beta, U, mu):
    """First and second beta-derivatives of logZ by central differences with one
    Richardson step (error O(h^4))."""
    def d1(h):
        return (_logz(orc, kind, D, L, f, beta + h, U, mu) - _logz(orc, kind, D, L, f, beta - h, U, mu)) / (2 * h)

    def d2(h):
        return (_logz(orc, kind, D, L, f, beta + h, U, mu) - 2 * _logz(orc, kind, D, L, f, beta, U, mu)
                + _logz(orc, kind, D, L, f, beta - h, U, mu)) / h ** 2
    h1 = 1e-3 * beta
    h2 = 2e-3 * beta
    first = (4 * d1(h1) - d1(h2)) / 3
    second = (4 * d2(h1) - d2(h2)) / 3
    return first, second


def _oracle_energy(orc, spectrum, cexp, mu_f, nf, e_ff):
    e = orc.measure_energy(spectrum, cexp, mu_f, nf, e_ff)
    return e["c_energy"], e["d2energy"], e["energy"]


def _energy_identities_hold(energy_fn, orc, kind, D, L, f, beta, U, mu_c, mu_f, W=None):
    tp = 1.0 if kind == 1 else 0.0
    r = orc.calc_ed(kind, D, L, f, beta, U, mu_c, tp=tp)
    nf = int(np.sum(f))
    e_ff = orc.ff_energy(D, f, W) if W is not None else 0.0
    ec, d2e, etot = energy_fn(orc, r["evals"], r["cexp"], mu_f, nf, e_ff)
    d1, d2 = _derivs(orc, kind, D, L, f, beta, U, mu_c)
    scale = np.sum(np.abs(r["evals"]))
    ok_c = abs(ec - (-d1)) <= 1e-7 * scale
    ok_2 = abs(d2e - d2) <= 1e-6 * max(scale, 1.0)
    # the whole weight logZ + beta mu_f N_f - beta E_ff
    ok_t = abs(etot - (-d1 - mu_f * nf + e_ff)) <= 1e-7 * scale
    return ok_c and ok_2 and ok_t


ENERGY_CASES = [
    # kind, D, L, beta, U, mu_c, mu_f, W
    (0, 2, 8, 1.0, 1.0, 0.5, 0.5, None),
    (0, 2, 6, 7.5, 4.0, 2.0, 1.7, None),
    (1, 2, 6, 5.0, 3.0, 1.5, 1.5, None),
    (0, 3, 4, 2.0, 5.0, 2.5, 0.3, None),
    (0, 1, 12, 3.0, 2.0, 1.0, 0.8, [0.0, 0.6, 0.25]),
]


@pytest.mark.parametrize("kind,D,L,beta,U,mu_c,mu_f,W", ENERGY_CASES)
def test_measure_energy_is_beta_derivative_of_logz(orc, kind, D, L, beta, U, mu_c, mu_f, W):
    rng = np.random.default_rng(L * 31 + D + kind)
    n = L ** D
    for _ in range(2):
        f = rng.integers(0, 2, n).astype(np.int32)
        assert _energy_identities_hold(_oracle_energy, orc, kind, D, L, f, beta, U, mu_c, mu_f, W)


def test_measure_energy_u0_closed_form(orc):
    """U = 0 (f plays no role): E_c = sum over the plane-wave levels
    e_k = -2t(cos kx + cos ky) - mu_c of e_k / (1 + e^{beta e_k})."""
    L, beta, mu = 8, 2.5, 0.3
    k = 2 * np.pi * np.arange(L) / L
    ek = (-2 * (np.cos(k)[:, None] + np.cos(k)[None, :])).ravel() - mu
    f = np.random.default_rng(5).integers(0, 2, L * L).astype(np.int32)
    r = orc.calc_ed(0, 2, L, f, beta, 0.0, mu)
    e = orc.measure_energy(r["evals"], r["cexp"], 0.0, int(f.sum()))
    assert abs(e["c_energy"] - np.sum(ek / (1 + np.exp(beta * ek)))) < 1e-12 * np.abs(ek).sum()
    fk = 1 / (1 + np.exp(beta * ek))
    assert abs(e["d2energy"] - np.sum(ek ** 2 * fk * (1 - fk))) < 1e-12 * np.sum(ek ** 2)


def _mutant_fermi_sign(orc, spectrum, cexp, mu_f, nf, e_ff):
    ec = np.sum(spectrum / (1 + 1 / cexp))            # f(-e) instead of f(e)
    d2 = np.sum(spectrum ** 2 / (1 + 0.5 * (cexp + 1 / cexp))) / 2
    return ec, d2, ec - mu_f * nf + e_ff


def _mutant_no_half(orc, spectrum, cexp, mu_f, nf, e_ff):
    ec, d2, et = _oracle_energy(orc, spectrum, cexp, mu_f, nf, e_ff)
    return ec, 2 * d2, et


def _mutant_mu_f_sign(orc, spectrum, cexp, mu_f, nf, e_ff):
    ec, d2, et = _oracle_energy(orc, spectrum, cexp, mu_f, nf, e_ff)
    return ec, d2, ec + mu_f * nf + e_ff


@pytest.mark.parametrize("mutant", [_mutant_fermi_sign, _mutant_no_half, _mutant_mu_f_sign])
def test_energy_identity_check_catches_mutants(orc, mutant):
    rng = np.random.default_rng(17)
    f = rng.integers(0, 2, 64).astype(np.int32)
    assert _energy_identities_hold(_oracle_energy, orc, 0, 2, 8, f, 2.0, 3.0, 1.5, 1.1)
    assert not _energy_identities_hold(mutant, orc, 0, 2, 8, f, 2.0, 3.0, 1.5, 1.1)


# ---------------------------------------------------------------------------
# Nearest-neighbour exchange: detailed balance by exact enumeration
# ---------------------------------------------------------------------------

BETA, U, MU = 2.0, 3.0, 1.5
LATTICES = {
    "square3x3": (0, 2, 3, None),
    "triangular3x3": (1, 2, 3, None),
    "chain9_ff": (0, 1, 9, [0.0, 0.7, 0.4]),
}
_EXACT = {}


def _exact_logw(orc, name):
    if name not in _EXACT:
        kind, D, L, W = LATTICES[name]
        n = L ** D
        confs = np.array(list(itertools.product([0, 1], repeat=n)), np.int32)
        lw = np.array([_logz(orc, kind, D, L, c, BETA, U, MU) + BETA * MU * c.sum()
                       - BETA * (orc.ff_energy(D, c, W) if W is not None else 0.0) for c in confs])
        _EXACT[name] = (confs, lw)
    return _EXACT[name]


def _visited(orc, name, moves, nmoves, seed, mutation=0):
    """Run the oracle chain and rebuild the index of the configuration after every move."""
    kind, D, L, W = LATTICES[name]
    n = L ** D
    f0 = np.zeros(n, np.int32)
    f0[:4] = 1
    orc.set_chain_ff(W)
    orc.set_mutation(mutation)
    try:
        tr = orc.run_chain(kind, D, L, BETA, U, MU, MU, seed, f0, moves, nmoves, tp=1.0 if kind == 1 else 0.0)
    finally:
        orc.set_mutation(0)
        orc.set_chain_ff(None)
    cur = f0.copy()
    idx = np.zeros(nmoves, np.int64)
    weights = 1 << np.arange(n - 1, -1, -1)
    for s in range(nmoves):
        if tr["acc"][s]:
            if tr["kind"][s] == orc.MOVE_ADDREMOVE:
                cur[tr["a"][s]] ^= 1
            elif tr["a"][s] >= 0 and tr["b"][s] >= 0:
                a, b = tr["a"][s], tr["b"][s]
                cur[a], cur[b] = cur[b], cur[a]
        idx[s] = int(cur @ weights)
    assert np.array_equal(cur, tr["f_final"])
    return idx[5000:]


def _tv(orc, name, moves, nmoves, seed, sector=None, mutation=0):
    confs, lw = _exact_logw(orc, name)
    sel = np.ones(len(confs), bool) if sector is None else confs.sum(axis=1) == sector
    p = np.where(sel, np.exp(lw - lw[sel].max()), 0.0)
    p /= p.sum()
    idx = _visited(orc, name, moves, nmoves, seed, mutation)
    hist = np.bincount(idx, minlength=len(confs)) / len(idx)
    if sector is not None:
        assert hist[~sel].sum() == 0.0  # the exchange conserves N_f
    return 0.5 * np.abs(p - hist).sum()


HOP_ONLY = [(2, 1.0)]
MIXED = [(0, 0.2), (2, 0.8)]


@pytest.mark.parametrize("name", list(LATTICES))
def test_hop_detailed_balance_in_sector(orc, name):
    """Hop alone samples exp(logZ - beta E_ff) inside the N_f = 4 sector; an
    inverted dlogZ or an f-dependent neighbour choice is caught."""
    assert _tv(orc, name, HOP_ONLY, 200000, 11, sector=4) < 0.05
    assert _tv(orc, name, HOP_ONLY, 200000, 11, sector=4, mutation=1) > 0.08
    assert _tv(orc, name, HOP_ONLY, 200000, 11, sector=4, mutation=3) > 0.08
    if LATTICES[name][3] is not None:
        assert _tv(orc, name, HOP_ONLY, 200000, 11, sector=4, mutation=2) > 0.08


@pytest.mark.parametrize("name", list(LATTICES))
def test_addremove_hop_detailed_balance(orc, name):
    """addremove + hop (0.2 / 0.8) samples exp(logZ + beta mu_f N_f - beta E_ff)
    over all 512 configurations; an inverted dlogZ (or f-f term) in the hop is caught."""
    assert _tv(orc, name, MIXED, 400000, 23) < 0.05
    mutation = 2 if LATTICES[name][3] is not None else 1
    assert _tv(orc, name, MIXED, 400000, 23, mutation=mutation) > 0.08


def test_chebyshev_flip_weighs_ff_energy(orc):
    """The global exchange of src/moves_chebyshev.cpp:22 weighs -beta dE_ff, so on a 1-D
    chain with f-f interaction flip + addremove samples exp(logZ + beta mu_f N_f - beta E_ff);
    the ED move_flip (src/moves.cpp:19-20) omits the term and samples something else."""
    name = "chain9_ff"
    moves = [(orc.MOVE_ADDREMOVE, 0.3), (orc.MOVE_FLIP, 0.7)]
    orc.set_flip_ff(1)
    try:
        tv_cheb = _tv(orc, name, moves, 300000, 31)
    finally:
        orc.set_flip_ff(0)
    tv_ed = _tv(orc, name, moves, 300000, 31)
    assert tv_cheb < 0.05
    assert tv_ed > 0.08
