"""GPU eigenvectors, measures and adjoint against the oracle.

Eigenvector-basis independent checks (residual, orthogonality, projectors of
non-degenerate states) because eigenvectors inside a degenerate subspace are
not unique.  Stiffness against the reference KATs (stiffness_test.cpp:60-63,
tol 1e-3 as there) and against the oracle at 1e-8; the adjoint against
central finite differences of Omega (h = 1e-6, tolerance 1e-7).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from test_oracle import STIFF_KATS  # noqa: E402


def _rand_f(rng, n):
    return rng.integers(0, 2, n).astype(np.int32)


@pytest.mark.parametrize("kind,D,L", [(0, 2, 8), (0, 1, 30), (0, 2, 16), (1, 2, 12), (0, 3, 8), (1, 2, 24)])
def test_eigenvectors(fk, orc, kind, D, L):
    print(f"case eigenvectors kind={kind} D={D} L={L}", flush=True)
    rng = np.random.default_rng(L + 3 * D)
    n = L ** D
    f = np.stack([_rand_f(rng, n) for _ in range(3)])
    U = np.array([1.0, 3.0, 6.0])
    ctx = fk.Context(0).set_lattice(kind, D, L, 1.0, 1.0)
    out = ctx.eig_logz(f, 2.0, U, U / 2, vectors=True, with_exp=True)
    hop, _, _ = orc.lattice_hop(kind, D, L, 1.0, 1.0)
    for b in range(3):
        H = hop + np.diag(-U[b] / 2 + U[b] * f[b])
        V, ev = out["evecs"][b], out["evals"][b]
        assert np.abs(H @ V - V * ev).max() < 1e-9 * max(1, np.abs(ev).max())
        assert np.abs(V.T @ V - np.eye(n)).max() < 1e-9
        ref = orc.calc_ed(kind, D, L, f[b], 2.0, U[b], U[b] / 2, vectors=True)
        gaps = np.minimum(np.diff(ref["evals"], prepend=-np.inf), np.diff(ref["evals"], append=np.inf))
        nd = gaps > 1e-6
        # non-degenerate eigenvectors agree up to sign
        dots = np.abs(np.sum(V[:, nd] * ref["evecs"][:, nd], axis=0))
        assert np.abs(dots - 1).max() < 1e-8
    ctx.close()


def test_ipr(fk, orc):
    print("case ipr", flush=True)
    rng = np.random.default_rng(2)
    L = 8
    f = np.stack([_rand_f(rng, L * L) for _ in range(2)])
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    ev, ipr = ctx.measure_ipr(f, 1.0, 4.0, 2.0)
    for b in range(2):
        ref = orc.calc_ed(0, 2, L, f[b], 1.0, 4.0, 2.0, vectors=True)
        ripr = orc.measure_ipr(ref["evecs"])
        gaps = np.minimum(np.diff(ref["evals"], prepend=-np.inf), np.diff(ref["evals"], append=np.inf))
        nd = gaps > 1e-6
        assert np.abs(ipr[b][nd] - ripr[nd]).max() < 1e-9
    # closed form: the uniform k=0 state of U=0 (non-degenerate ground state) has
    # ||v||_4 / ||v||_2^2 = (N * N^-2)^(1/4) = N^(-1/4)
    f0 = np.zeros((1, L * L), np.int32)
    ev0, ipr0 = ctx.measure_ipr(f0, 1.0, 0.0, 0.0)
    assert abs(ipr0[0][0] - (L * L) ** -0.25) < 1e-10
    ctx.close()


@pytest.mark.parametrize("U,fconf,expected", STIFF_KATS)
def test_stiffness_kats(fk, orc, U, fconf, expected):
    print(f"case stiffness KAT U={U}", flush=True)
    L = int(round(np.sqrt(len(fconf))))
    f = np.array([fconf], np.int32)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    st, cond = ctx.measure_stiffness(f, 1000.0, U, U / 2, wgrid=[0.0], offset=0.05)
    assert abs(st[0] - expected) < 1e-3
    r = orc.calc_ed(0, 2, L, f[0], 1000.0, U, U / 2, vectors=True)
    s = orc.measure_stiffness(2, L, r["evecs"], r["evals"], r["cfermi"], wgrid=[0.0], offset=0.05)
    assert abs(st[0] - s["stiffness"]) < 1e-8
    ctx.close()


def test_stiffness_conductivity_grid(fk, orc):
    print("case stiffness grid", flush=True)
    rng = np.random.default_rng(4)
    L = 6
    f = np.stack([_rand_f(rng, L * L) for _ in range(2)])
    w = np.linspace(-2, 2, 9)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    st, cond = ctx.measure_stiffness(f, 5.0, 2.0, 1.0, wgrid=w, offset=0.05)
    for b in range(2):
        r = orc.calc_ed(0, 2, L, f[b], 5.0, 2.0, 1.0, vectors=True)
        s = orc.measure_stiffness(2, L, r["evecs"], r["evals"], r["cfermi"], wgrid=w, offset=0.05)
        assert abs(st[b] - s["stiffness"]) < 1e-8
        assert np.abs(cond[b] - s["cond"]).max() < 1e-8
    ctx.close()


def test_dos(fk):
    print("case dos", flush=True)
    rng = np.random.default_rng(5)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, 4, 1.0, 0.0)
    spec = np.sort(rng.normal(size=(3, 16)), axis=1)
    w = np.linspace(-3, 3, 50)
    dos = ctx.dos(spec, w, 0.05)
    ref = np.array([[np.mean(0.05 / np.pi / ((x - s) ** 2 + 0.05 ** 2)) for x in w] for s in spec])
    assert np.abs(dos - ref).max() < 1e-12
    ctx.close()


def test_ipr_w(fk):
    """ipr_f of save_ipr (prog/data_save.hxx:490-501): a Lorentzian-weighted mean of ipr^4,
    so constant for constant IPR, linear in ipr^4, tending to ipr_k^4 at an isolated
    eigenvalue as eta -> 0, and its weights are those of the DOS."""
    print("case ipr_w", flush=True)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, 4, 1.0, 0.0)
    rng = np.random.default_rng(12)
    S, N = 70000, 16  # more spectra than one grid extent of 65535
    spec = np.sort(rng.uniform(-4, 4, (S, N)), axis=1)
    w = np.array([-3.0, -0.5, 0.0, 1.7])
    out = ctx.ipr_w(spec, np.full((S, N), 0.5), w, 0.05)
    np.testing.assert_allclose(out, 0.0625, rtol=1e-13)
    a4 = rng.uniform(0.05, 1.0, (S, N))
    b4 = rng.uniform(0.05, 1.0, (S, N))
    oa = ctx.ipr_w(spec[:50], a4[:50] ** 0.25, w, 0.1)
    ob = ctx.ipr_w(spec[:50], b4[:50] ** 0.25, w, 0.1)
    oab = ctx.ipr_w(spec[:50], (2 * a4[:50] + 3 * b4[:50]) ** 0.25, w, 0.1)
    np.testing.assert_allclose(oab, 2 * oa + 3 * ob, rtol=1e-12)
    # at an eigenvalue with a tiny width the weight concentrates on it
    s0 = np.linspace(-3, 3, N)
    r0 = rng.uniform(0.2, 0.9, N)
    o = ctx.ipr_w(s0, r0, s0, 1e-7)[0]
    np.testing.assert_allclose(o, r0 ** 4, rtol=1e-6)
    # the denominator is pi N DOS(w): with ipr^4 = 1 at one level only, out = L_k / (pi N dos)
    r1 = np.zeros(N)
    r1[5] = 1.0
    o = ctx.ipr_w(s0, r1, w, 0.2)[0]
    dos = ctx.dos(s0, w, 0.2)[0]
    np.testing.assert_allclose(o, 0.2 / ((w - s0[5]) ** 2 + 0.04) / (np.pi * N * dos), rtol=1e-12)
    with pytest.raises(fk.FkgpuError):
        ctx.ipr_w(s0, r0, w, 0.0)
    ctx.close()


@pytest.mark.parametrize("kind,D,L", [(0, 2, 8), (1, 2, 6), (0, 3, 4)])
def test_adjoint_vs_finite_differences(fk, orc, kind, D, L):
    print(f"case adjoint kind={kind} L={L}", flush=True)
    rng = np.random.default_rng(7 + L)
    n = L ** D
    f = _rand_f(rng, n)[None, :]
    beta, U, mu, t, tp = 5.0, 3.0, 1.5, 1.0, 1.0
    h = 1e-6

    def omega(U_=U, mu_=mu, t_=t, tp_=tp, eps=None):
        c = fk.Context(0).set_lattice(kind, D, L, t_, tp_)
        if eps is None:
            lz = c.eig_logz(f, beta, U_, mu_)["logZ"][0]
        else:
            hop, _, _ = orc.lattice_hop(kind, D, L, t_, tp_)
            lz = orc.logz_stable(np.linalg.eigvalsh(hop + np.diag(-mu_ + U_ * f[0] + eps)), beta)
        c.close()
        return -lz / beta

    ctx = fk.Context(0).set_lattice(kind, D, L, t, tp)
    a = ctx.adjoint(f, beta, U, mu, full=True)
    fd_U = (omega(U_=U + h) - omega(U_=U - h)) / (2 * h)
    fd_mu = (omega(mu_=mu + h) - omega(mu_=mu - h)) / (2 * h)
    fd_t = (omega(t_=t + h) - omega(t_=t - h)) / (2 * h)
    assert abs(a["dOmega_dU"][0] - fd_U) < 1e-7
    assert abs(a["dOmega_dmu"][0] - fd_mu) < 1e-7
    assert abs(a["dOmega_dt"][0] - fd_t) < 1e-7
    if kind == 1:
        fd_tp = (omega(tp_=tp + h) - omega(tp_=tp - h)) / (2 * h)
        assert abs(a["dOmega_dtp"][0] - fd_tp) < 1e-7
    # local densities: d Omega / d eps_i at a few sites
    for i in (0, n // 2, n - 1):
        e = np.zeros(n)
        e[i] = h
        fd = (omega(eps=e) - omega(eps=-e)) / (2 * h)
        assert abs(a["dens"][0][i] - fd) < 1e-7
    # rho = V f(Lambda) V^T from the oracle's eigenpairs
    r = orc.calc_ed(kind, D, L, f[0], beta, U, mu, t=t, tp=tp, vectors=True)
    rho = (r["evecs"] * r["cfermi"]) @ r["evecs"].T
    assert np.abs(a["rho"][0] - rho).max() < 1e-9
    assert abs(a["Nc"][0] - r["cfermi"].sum()) < 1e-9
    assert abs(a["dNc_dmu"][0] - beta * np.sum(r["cfermi"] * (1 - r["cfermi"]))) < 1e-9
    ctx.close()
