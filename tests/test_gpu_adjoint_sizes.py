"""Adjoint / differentiation workload on triangular lattices of uneven sizes
(L in {12, 16, 20, 24}, t = tp = 1, U = 3, beta = 5): Omega and the adjoint gradients
dOmega/d eps_i, dOmega/dU, dOmega/dt, dOmega/dmu against central finite differences
(h = 1e-6, tolerance 1e-7) on 32 seeded configurations (8 per size), and mu_c
Newton-tuned to conduction filling 0.5 with dN_c/dmu_c from the adjoint.  The four
sizes are four contexts whose launches are issued back to back on one stream each."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = (12, 16, 20, 24)
U, BETA, MU = 3.0, 5.0, 1.5


def spectra(fk, L, f, U_=U, mu_=MU, t_=1.0, tp_=1.0):
    c = fk.Context(0).set_lattice(fk.LATTICE_TRIANGULAR, 2, L, t_, tp_)
    ev = c.eig_logz(f, BETA, U_, mu_)["evals"]
    c.close()
    return ev


def omega_diff(a, b):
    """Omega(a) - Omega(b) summed eigenvalue by eigenvalue (Omega ~ -700 here: differencing
    the totals would lose ~1e-13 absolute, 1e-7 after division by 2h), with
    -log1p(e^{-beta x})/beta differences formed through expm1."""
    a, b = np.asarray(a), np.asarray(b)
    d = np.empty_like(a)
    pos = b >= 0
    x = np.exp(-BETA * b[pos])
    d[pos] = -np.log1p(x * np.expm1(-BETA * (a[pos] - b[pos])) / (1 + x)) / BETA
    y = np.exp(BETA * b[~pos])
    d[~pos] = (a[~pos] - b[~pos]) - np.log1p(y * np.expm1(BETA * (a[~pos] - b[~pos])) / (1 + y)) / BETA
    return d.sum(axis=-1)


@pytest.mark.parametrize("L", SIZES)
def test_adjoint_vs_fd_uneven_sizes(fk, orc, L):
    n = L * L
    rng = np.random.default_rng(100 + L)
    f = (rng.random((8, n)) < 0.5).astype(np.int32)
    print(f"case adjoint triangular L={L}", flush=True)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_TRIANGULAR, 2, L, 1.0, 1.0)
    a = ctx.adjoint(f, BETA, U, MU)
    h = 1e-6
    fd = {}
    fd["U"] = omega_diff(spectra(fk, L, f, U_=U + h), spectra(fk, L, f, U_=U - h)) / (2 * h)
    fd["mu"] = omega_diff(spectra(fk, L, f, mu_=MU + h), spectra(fk, L, f, mu_=MU - h)) / (2 * h)
    fd["t"] = omega_diff(spectra(fk, L, f, t_=1.0 + h), spectra(fk, L, f, t_=1.0 - h)) / (2 * h)
    fd["tp"] = omega_diff(spectra(fk, L, f, tp_=1.0 + h), spectra(fk, L, f, tp_=1.0 - h)) / (2 * h)
    assert np.abs(a["dOmega_dU"] - fd["U"]).max() < 1e-7
    assert np.abs(a["dOmega_dmu"] - fd["mu"]).max() < 1e-7
    assert np.abs(a["dOmega_dt"] - fd["t"]).max() < 1e-7
    assert np.abs(a["dOmega_dtp"] - fd["tp"]).max() < 1e-7
    # local densities dOmega/d eps_i (a site energy added to H_ii): the eigenvector densities
    # sum_k f_k V_ik^2 from LAPACK eigenpairs (1e-12), and central differences of LAPACK spectra.
    # Each spectrum carries rounding of eps ||H|| per eigenvalue, which after division by 2h
    # leaves the difference quotient itself uncertain by ~ eps |Omega| / h (1e-7 at N = 576):
    # the tolerance is 1e-7 up to N = 400 and 3e-7 at N = 576.
    hop, _, _ = orc.lattice_hop(1, 2, L, 1.0, 1.0)
    tol_fd = 1e-7 if n <= 400 else 3e-7
    for b in range(2):
        H = hop + np.diag(-MU + U * f[b])
        ev, V = np.linalg.eigh(H)
        rho_ii = (V ** 2) @ (1.0 / (1.0 + np.exp(BETA * ev)))
        assert np.abs(a["dens"][b] - rho_ii).max() < 1e-12
        for i in (0, n // 3, n - 1):
            e = np.zeros(n)
            e[i] = h
            fd_i = omega_diff(np.linalg.eigvalsh(H + np.diag(e)), np.linalg.eigvalsh(H - np.diag(e))) / (2 * h)
            assert abs(a["dens"][b][i] - fd_i) < tol_fd
    ctx.close()


def test_mu_newton_tuning_uneven_sizes(fk, orc):
    print("case mu tuning", flush=True)
    for L in SIZES:
        n = L * L
        f = (np.random.default_rng(L).random((8, n)) < 0.5).astype(np.int32)
        ctx = fk.Context(0).set_lattice(fk.LATTICE_TRIANGULAR, 2, L, 1.0, 1.0)
        mu, its = ctx.tune_mu(f, BETA, U, mu0=0.0, filling=0.5, tol=1e-12, maxit=60)
        assert (its > 0).all(), its
        for b in range(2):
            r = orc.calc_ed(1, 2, L, f[b], BETA, U, mu[b], vectors=False)
            nc = np.sum(1.0 / (1.0 + np.exp(BETA * r["evals"])))
            assert abs(nc / n - 0.5) < 1e-10
        print(f"  L={L}: Newton steps {its.min()}..{its.max()}", flush=True)
        ctx.close()
