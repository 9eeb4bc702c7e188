"""Pin the CPU oracle (the checker) before trusting it.

Each check is chosen so that a plausible mistake in the restatement fails it:
LAPACK eigenvalues (numpy eigvalsh) on seeded configurations, closed-form
spectra, the reference's own known-answer tests (stiffness_test.cpp:60-63,
fast_update_test.cpp:62-65,77-78, config_test.cpp:46-72), and detailed balance
of the Metropolis chain by exact enumeration on a tiny lattice.
"""
import itertools

import numpy as np
import pytest


def test_lattice_structure(orc):
    """Square lattice: 4 neighbours, symmetric, index order last-coordinate fastest
    (src/lattice/hypercubic.cpp:31-51,115-131)."""
    hop, nbr, z = orc.lattice_hop(orc.KIND_HYPERCUBIC, 2, 5, 1.0, 0.0)
    assert np.array_equal(hop, hop.T)
    assert np.all(z == 4)
    assert np.allclose(hop.sum(axis=1), -4.0)
    # site (x=1,y=2) -> index 1*5+2 = 7; neighbours in neighbor_index order
    assert list(nbr[7]) == [2, 12, 6, 8]
    tri, nbr3, z3 = orc.lattice_hop(orc.KIND_TRIANGULAR, 2, 4, 1.0, 0.5)
    assert np.all(z3 == 6)
    assert np.allclose(tri.sum(axis=1), -4.0 - 1.0)
    # triangular diagonal: (x-1,y-1) and (x+1,y+1)
    assert tri[5, 0] == -0.5 and tri[5, 10] == -0.5
    with pytest.raises(ValueError):
        orc.lattice_hop(orc.KIND_HYPERCUBIC, 2, 2)


@pytest.mark.parametrize("kind,D,L,U", [(0, 1, 9, 1.0), (0, 2, 8, 1.0), (0, 2, 16, 4.0), (1, 2, 6, 3.0),
                                        (0, 3, 4, 5.0)])
def test_eigenvalues_vs_lapack(orc, kind, D, L, U):
    rng = np.random.default_rng(L * 7 + D)
    hop, _, _ = orc.lattice_hop(kind, D, L, 1.0, 1.0)
    n = hop.shape[0]
    for _ in range(3):
        f = rng.integers(0, 2, n)
        H = hop + np.diag(-U / 2 + U * f)
        ev = orc.eigh(H)
        ref = np.linalg.eigvalsh(H)
        assert np.abs(ev - ref).max() < 1e-12 * max(1, np.abs(ref).max())


def test_eigenvectors_residual_and_orthogonality(orc):
    rng = np.random.default_rng(3)
    hop, _, _ = orc.lattice_hop(0, 2, 8, 1.0, 0.0)
    f = rng.integers(0, 2, 64)
    H = hop + np.diag(-1.0 + 2.0 * f)
    ev, V = orc.eigh(H, vectors=True)
    assert np.abs(H @ V - V * ev).max() < 1e-12
    assert np.abs(V.T @ V - np.eye(64)).max() < 1e-12


def test_closed_form_spectra(orc):
    L = 8
    k = 2 * np.pi * np.arange(L) / L
    ek = (-2 * (np.cos(k)[:, None] + np.cos(k)[None, :])).ravel()
    r = orc.calc_ed(0, 2, L, np.zeros(L * L, np.int32), 1.0, 0.0, 0.3)
    assert np.abs(r["evals"] - np.sort(ek - 0.3)).max() < 1e-12
    # checkerboard f: U/2 - mu +- sqrt(ek^2 + U^2/4), each level once per (k, k+Q) pair
    U, mu = 3.0, 0.7
    x, y = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    f = ((x + y) % 2).ravel().astype(np.int32)
    r = orc.calc_ed(0, 2, L, f, 1.0, U, mu)
    lev = np.sort(np.concatenate([U / 2 - mu + np.sqrt(ek ** 2 + U ** 2 / 4), U / 2 - mu - np.sqrt(ek ** 2 + U ** 2 / 4)]))
    assert np.abs(r["evals"] - lev[::2]).max() < 1e-12
    # triangular, U=0: -2t(cos kx + cos ky) - 2tp cos(kx+ky)
    t, tp = 1.0, 0.6
    ekt = (-2 * t * (np.cos(k)[:, None] + np.cos(k)[None, :]) - 2 * tp * np.cos(k[:, None] + k[None, :])).ravel()
    r = orc.calc_ed(1, 2, L, np.zeros(L * L, np.int32), 1.0, 0.0, 0.0, t=t, tp=tp)
    assert np.abs(r["evals"] - np.sort(ekt)).max() < 1e-12


def test_logz_shifted_equals_direct(orc):
    """fast_update_test.cpp:78: the shifted log-sum equals the direct sum to 1e-15."""
    rng = np.random.default_rng(1)
    hop, _, _ = orc.lattice_hop(0, 2, 24, 1.0, 0.0)
    f = np.zeros(576, np.int32)
    f[rng.permutation(576)[:288]] = 1
    r = orc.calc_ed(0, 2, 24, f, 1 / 0.16, 8.0, 4.0)
    direct = np.sum(np.log(1 + np.exp(-(1 / 0.16) * r["evals"])))
    assert abs((r["logZ"] - direct) / r["logZ"]) <= 1e-15 * 4
    assert abs(orc.logz_stable(r["evals"], 1 / 0.16) - direct) <= 1e-13 * abs(direct)
    lz, ce, cf = orc.logz_reference(r["evals"], 1 / 0.16)
    assert np.allclose(cf, 1 / (1 + np.exp(r["evals"] / 0.16)))


# stiffness_test.cpp:60-63 (L=5 and L=7 square lattices, beta=1000, tol 1e-3)
STIFF_F5 = [0, 1, 0, 1, 1, 0, 1, 0, 0, 1, 0, 0, 1, 0, 1, 0, 1, 1, 1, 0, 1, 1, 0, 1, 1]
STIFF_F7 = [0, 1, 0, 0, 0, 0, 1, 1, 0, 1, 1, 0, 0, 1, 1, 0, 1, 1, 0, 1, 1, 0, 1, 0, 0, 0, 1, 1, 0, 1, 1, 1, 1, 1, 0,
            1, 1, 0, 1, 1, 1, 1, 0, 0, 0, 1, 0, 1, 1]
STIFF_KATS = [(0.0, STIFF_F5, 1.31597), (2.0, STIFF_F5, 0.287547), (6.0, STIFF_F5, 0.00260831),
              (0.37, STIFF_F7, 1.26046)]


@pytest.mark.parametrize("U,fconf,expected", STIFF_KATS)
def test_stiffness_kats(orc, U, fconf, expected):
    L = int(round(np.sqrt(len(fconf))))
    r = orc.calc_ed(0, 2, L, np.array(fconf, np.int32), 1000.0, U, U / 2, vectors=True)
    s = orc.measure_stiffness(2, L, r["evecs"], r["evals"], r["cfermi"], wgrid=[0.0], offset=0.05)
    assert abs(s["stiffness"] - expected) < 1e-3


def test_chebyshev_quadrature_kats(orc):
    """fast_update_test.cpp:62-65 with the evaluator of that test (12 moments, 12 points)."""
    M, G = 12, 12
    assert abs(orc.cheb_moment_f(M, G, 0, 0) - 1.0) < 1e-8
    assert abs(orc.cheb_moment_f(M, G, 1, 0) - 0.5) < 1e-8
    assert abs(orc.cheb_moment_f(M, G, 2, 1) - 0.440051) < 1e-6   # J_1(1)
    assert abs(orc.cheb_moment_f(M, G, 2, 5) - 0.000249758) < 1e-9  # J_5(1)


def test_chebyshev_logz_vs_ed(orc):
    """fast_update_test.cpp:77: exact-trace Chebyshev logZ within 5e-2 of ED
    (L=24, U=8, T=0.16, half filling, cheb_size = 2*int(ln N), grid = cheb_size)."""
    L = 24
    n = L * L
    f = orc.randomize_f(n, 32167, n // 2)
    assert f.sum() == n // 2
    beta = 1 / 0.16
    M = int(np.log(n)) * 2
    r = orc.calc_chebyshev(0, 2, L, f, beta, 8.0, 4.0, M, M)
    ed = orc.calc_ed(0, 2, L, f, beta, 8.0, 4.0)
    assert abs((r["logZ"] - ed["logZ"]) / r["logZ"]) < 5e-2
    assert abs(r["moments"][0] - 1.0) < 1e-15
    # moments of the rescaled spectrum: mu_m = mean_k T_m(x_k)
    x = (ed["evals"] - r["b"]) / r["a"]
    exact = [np.mean(np.cos(m * np.arccos(np.clip(x, -1, 1)))) for m in range(M)]
    assert np.abs(r["moments"] - exact).max() < 1e-10


def test_ff_energy_kat(orc):
    """config_test.cpp:46-72: 1-D L=12, W = {0,1,2}."""
    L = 12
    f = np.array([x % 2 for x in range(L)], np.int32)
    f[0] = 1
    W = [0.0, 1.0, 2.0]
    assert abs(orc.ff_energy(1, f, W) - (2 * L / 2 * W[2] + 4 * W[1])) < 1e-15
    f2 = np.ones(16, np.int32)
    assert orc.ff_energy(2, f2, W) == 0.0  # only 1-D lattices carry f-f energy


def test_randomize_f_counts(orc):
    for nf in (1, 5, 32):
        f = orc.randomize_f(64, 11 + nf, nf)
        assert f.sum() == nf


def test_chain_detailed_balance_enumeration(orc):
    """The oracle chain samples exp(logZ(f) + beta mu_f N_f) on a 3x3 lattice
    (all 512 configurations enumerated)."""
    L, D = 3, 2
    n = L ** D
    beta, U, mu = 1.5, 2.0, 1.0
    confs = np.array(list(itertools.product([0, 1], repeat=n)), np.int32)
    logw = np.array([orc.calc_ed(0, D, L, c, beta, U, mu)["logZ"] + beta * mu * c.sum() for c in confs])
    p = np.exp(logw - logw.max())
    p /= p.sum()
    nmoves = 60000
    tr = orc.run_chain(0, D, L, beta, U, mu, mu, 7, np.zeros(n, np.int32),
                       [(orc.MOVE_ADDREMOVE, 0.6), (orc.MOVE_FLIP, 0.4)], nmoves, tp=0.0)
    # reconstruct the visited configurations from the trace
    cur = np.zeros(n, np.int32)
    idx = np.zeros(nmoves, np.int64)
    weights = 1 << np.arange(n - 1, -1, -1)
    for s in range(nmoves):
        if tr["acc"][s]:
            if tr["kind"][s] == orc.MOVE_ADDREMOVE:
                cur[tr["a"][s]] ^= 1
            elif tr["kind"][s] == orc.MOVE_FLIP and tr["a"][s] >= 0:
                cur[tr["a"][s]] = 0
                cur[tr["b"][s]] = 1
        idx[s] = int(cur @ weights)
    assert np.array_equal(cur, tr["f_final"])
    hist = np.bincount(idx[2000:], minlength=2 ** n) / (nmoves - 2000)
    # compare the distribution of N_f (coarse, low-noise) and the total variation distance
    nf = confs.sum(axis=1)
    pn = np.array([p[nf == k].sum() for k in range(n + 1)])
    hn = np.array([hist[nf == k].sum() for k in range(n + 1)])
    assert np.abs(pn - hn).max() < 0.03
    assert 0.5 * np.abs(p - hist).sum() < 0.12
