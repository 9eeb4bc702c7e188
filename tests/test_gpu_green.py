"""Lattice Green's function and local densities of states over the recorded
eigen-decompositions of a chain set (fkgpu_chains_green, fkgpu_chains_local_dos), the
quantities of data_saver::save_gwr (prog/data_save.hxx:535-708).

The reference sums eigs[m] diag(1 / (z - e)) eigs[m]^T over the measurements; that equals
the resolvent (z - H_m)^{-1} of each recorded configuration, which the test builds from the
oracle's hopping matrix and the recorded occupations and inverts with numpy -- no
eigenvectors involved.  The site mean of the local DOS is the DOS of the recorded spectra
(fkgpu_dos, dos0_f of prog/data_save.hxx:264-345).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HIST, EIG = 1 | 2, 16


def _run(fk, D, L, B, n, seed, kind=None, tp=0.0):
    N = L ** D
    kind = fk.LATTICE_HYPERCUBIC if kind is None else kind
    rng = np.random.default_rng(seed)
    ctx = fk.Context(0).set_lattice(kind, D, L, 1.0, tp)
    beta = np.linspace(2.0, 10.0, B)
    U = np.linspace(1.0, 4.0, B)
    f0 = rng.integers(0, 2, size=(B, N)).astype(np.int32)
    ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, (np.arange(B) + 17 * seed).astype(np.uint32))
    ch.set_moves([(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)])
    ch.measure_enable(n, HIST | EIG)
    ch.sweep_measure(n, 3)
    return ctx, ch, U


def _resolvent_mean(orc, kind, D, L, focc, U, z, tp=0.0):
    hop = orc.lattice_hop(kind, D, L, 1.0, tp)[0]
    N = hop.shape[0]
    G = np.zeros((N, N), complex)
    cnt = 0
    for c in range(focc.shape[0]):
        for m in range(focc.shape[1]):
            H = hop + np.diag(-U[c] / 2 + U[c] * focc[c, m].astype(float))
            G += np.linalg.inv(z * np.eye(N) - H)
            cnt += 1
    return G / cnt


@pytest.mark.parametrize("D,L,B,n", [(2, 6, 3, 4), (2, 10, 2, 3), (1, 70, 2, 2)])
def test_green_is_mean_resolvent(fk, orc, D, L, B, n):
    print(f"case green D={D} L={L} B={B} n={n}", flush=True)
    ctx, ch, U = _run(fk, D, L, B, n, 3)
    focc = ch.histories()["focc"]
    for w, eta in ((0.0, 0.05), (0.7, 0.2)):
        re, im = ch.green(w, eta)
        ref = _resolvent_mean(orc, orc.KIND_HYPERCUBIC, D, L, focc, U, w + 1j * eta)
        scale = np.abs(ref).max()
        assert np.abs(re - ref.real).max() <= 1e-10 * scale, (w, eta)
        assert np.abs(im - ref.imag).max() <= 1e-10 * scale, (w, eta)
        np.testing.assert_allclose(re, re.T, atol=1e-12 * scale)
    ch.close()
    ctx.close()


def test_local_dos(fk, orc):
    print("case local dos", flush=True)
    D, L, B, n = 2, 6, 3, 4
    ctx, ch, U = _run(fk, D, L, B, n, 5)
    h = ch.histories()
    wgrid = np.linspace(-5.0, 5.0, 97)  # more than one 64-frequency tile
    eta = 0.05
    ld = ch.local_dos(wgrid, eta)
    N = L ** D
    assert ld.shape == (97, N)
    # the diagonal of -Im G / pi, frequency by frequency
    for q in (0, 40, 96):
        ref = _resolvent_mean(orc, orc.KIND_HYPERCUBIC, D, L, h["focc"], U, wgrid[q] + 1j * eta)
        want = -np.diag(ref).imag / np.pi
        assert np.abs(ld[q] - want).max() <= 1e-10 * max(1.0, want.max()), q
    _, im = ch.green(wgrid[40], eta)
    np.testing.assert_allclose(ld[40], -np.diag(im) / np.pi, rtol=1e-11, atol=1e-13)
    # site mean = DOS of the recorded spectra (dos0_f)
    dos = ctx.dos(h["spectrum"].reshape(-1, N), wgrid, eta).mean(axis=0)
    np.testing.assert_allclose(ld.mean(axis=1), dos, rtol=1e-10, atol=1e-13)
    ch.close()
    ctx.close()


def test_green_triangular(fk, orc):
    print("case green triangular", flush=True)
    ctx, ch, U = _run(fk, 2, 6, 2, 3, 8, kind=fk.LATTICE_TRIANGULAR, tp=1.0)
    focc = ch.histories()["focc"]
    re, im = ch.green(-0.4, 0.1)
    ref = _resolvent_mean(orc, orc.KIND_TRIANGULAR, 2, 6, focc, U, -0.4 + 0.1j, tp=1.0)
    scale = np.abs(ref).max()
    assert np.abs(re - ref.real).max() <= 1e-10 * scale
    assert np.abs(im - ref.imag).max() <= 1e-10 * scale
    ch.close()
    ctx.close()


def test_green_errors(fk):
    print("case green errors", flush=True)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, 4, 1.0, 0.0)
    ch = fk.Chains(ctx, np.zeros((2, 16), np.int32), 1.0, 1.0, 0.5, 0.5, np.array([1, 2], np.uint32))
    ch.measure_enable(2, HIST)  # no eigenfunctions
    ch.sweep_measure(1, 1)
    with pytest.raises(fk.FkgpuError):
        ch.green(0.0, 0.05)
    ch.close()
    ch = fk.Chains(ctx, np.zeros((2, 16), np.int32), 1.0, 1.0, 0.5, 0.5, np.array([1, 2], np.uint32))
    ch.measure_enable(2, HIST | EIG)
    with pytest.raises(fk.FkgpuError):
        ch.green(0.0, 0.05)  # nothing recorded yet
    ch.sweep_measure(1, 1)
    with pytest.raises(fk.FkgpuError):
        ch.green(0.0, 0.0)
    with pytest.raises(fk.FkgpuError):
        ch.local_dos([0.0, np.nan], 0.05)
    re, im = ch.green(0.0, 0.05)
    assert np.isfinite(re).all() and np.isfinite(im).all()
    ch.close()
    ctx.close()


def test_green_headline_size(fk):
    """At the headline lattice (cubic2d L=16, N=256: 4 x 4 output tiles, records split over the
    grid) through size-independent identities: G is symmetric, -Im G_ii / pi is the local DOS,
    and the site mean of the local DOS is the DOS of the recorded spectra."""
    print("case green headline size L=16", flush=True)
    ctx, ch, U = _run(fk, 2, 16, 32, 2, 21)
    N = 256
    h = ch.histories()
    w = np.array([-2.0, 0.0, 0.35, 3.1])
    ld = ch.local_dos(w, 0.05)
    dos = ctx.dos(h["spectrum"].reshape(-1, N), w, 0.05).mean(axis=0)
    np.testing.assert_allclose(ld.mean(axis=1), dos, rtol=1e-10, atol=1e-13)
    assert (ld > 0).all()
    re, im = ch.green(0.35, 0.05)
    scale = np.abs(re).max() + np.abs(im).max()
    assert np.abs(re - re.T).max() <= 1e-12 * scale
    assert np.abs(im - im.T).max() <= 1e-12 * scale
    np.testing.assert_allclose(-np.diag(im) / np.pi, ld[2], rtol=1e-10, atol=1e-13)
    ch.close()
    ctx.close()
