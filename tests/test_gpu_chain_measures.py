"""On-device measurements of chain sets (fkgpu_chains_measure_enable, fkgpu_chains_measure,
fkgpu_sweep_measure and the readers) against the oracle:

* measure_energy (src/measures/energy.cpp:6-26) through the oracle's measure_energy on
  the recorded spectrum, with the f-f energy of 1-D lattices (src/configuration.cpp:59-77);
* measure_spectrum's running mean (src/measures/spectrum.cpp:13-21), bit for bit with the
  reference's loop (no fused operations on either side);
* the spectrum and occupation histories (spectrum_history.cpp:13-19,
  focc_history.cpp:8-13): each recorded spectrum is the oracle's calc_ed of the
  recorded occupations;
* measure_ipr (include/fk_mc/measures/ipr.hpp:39-56) on the oracle's eigenvectors of the
  recorded occupations, for eigenvalues separated from their neighbours (the IPR of a
  degenerate eigenspace depends on the basis chosen inside it);
* measure_stiffness (include/fk_mc/measures/stiffness.hpp:129-187) against the
  single-configuration entry point and the oracle;
* measure_eigenfunctions (src/measures/eigenfunctions.cpp:12-18): each recorded basis is
  orthonormal, diagonalises the Hamiltonian of the recorded occupations with the recorded
  spectrum, and spans the oracle's eigenvector for every eigenvalue separated from its
  neighbours (signs and bases inside degenerate eigenspaces are free: the reference's
  Eigen bits are parity unpinned).

Measuring must not change the chains: a measured chain set ends bit-identical to an
unmeasured one with the same seeds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALL = 1 | 2 | 4


def _running_mean(hist):
    """The reference's running mean, (avg * Z + e) / (Z + 1), one measurement at a time."""
    avg = np.zeros(hist.shape[-1])
    for z, s in enumerate(hist):
        avg = (avg * float(z) + s) / float(z + 1)
    return avg


def _chains(fk, D, L, B, seed, beta, U, moves, kind=None):
    N = L ** D
    rng = np.random.default_rng(seed)
    f0 = rng.integers(0, 2, size=(B, N)).astype(np.int32)
    kind = fk.LATTICE_HYPERCUBIC if kind is None else kind
    ctx = fk.Context(0).set_lattice(kind, D, L, 1.0, 1.0 if kind == fk.LATTICE_TRIANGULAR else 0.0)
    beta = np.broadcast_to(np.asarray(beta, float), (B,)).copy()
    U = np.broadcast_to(np.asarray(U, float), (B,)).copy()
    seeds = (np.arange(B) + 100 * seed).astype(np.uint32)
    ch = fk.Chains(ctx, f0, beta, U, U / 2, U / 2, seeds)
    ch.set_moves(moves)
    return ctx, ch, beta, U


def _check_records(orc, D, L, ch, beta, U, W=None, ipr_tol=1e-9, spec_tol=1e-10, kind=None, tp=0.0):
    """Every recorded measurement of every chain against the oracle."""
    kind = orc.KIND_HYPERCUBIC if kind is None else kind
    count, e, d2, ce = ch.energies()
    h = ch.histories()
    mn = ch.means()
    assert h["count"] == count == mn["count"]
    n = e.shape[1]
    checked_ipr = 0
    for c in range(ch.B):
        for m in range(n):
            fo = h["focc"][c, m]
            spec = h["spectrum"][c, m]
            want_vec = h["ipr"] is not None
            ed = orc.calc_ed(kind, D, L, fo, beta[c], U[c], U[c] / 2, tp=tp, vectors=want_vec)
            scale = max(1.0, np.abs(ed["evals"]).max())
            assert np.abs(spec - ed["evals"]).max() <= spec_tol * scale, (c, m)
            eff = orc.ff_energy(D, fo, W) if W is not None else 0.0
            ref = orc.measure_energy(spec, np.exp(beta[c] * spec), U[c] / 2, int(fo.sum()), eff)
            for got, key in ((e[c, m], "energy"), (d2[c, m], "d2energy"), (ce[c, m], "c_energy")):
                assert abs(got - ref[key]) <= 1e-12 * max(1.0, abs(ref[key])) * len(spec), (c, m, key, got, ref[key])
            if want_vec:
                ipr = orc.measure_ipr(ed["evecs"])
                gap = np.diff(ed["evals"])
                iso = np.ones(len(gap) + 1, bool)
                iso[:-1] &= gap > 1e-6
                iso[1:] &= gap > 1e-6
                got = h["ipr"][c, m]
                assert np.abs(got[iso] - ipr[iso]).max(initial=0.0) <= ipr_tol, (c, m)
                checked_ipr += int(iso.sum())
        if n == count:
            assert np.array_equal(mn["spectrum"][c], _running_mean(h["spectrum"][c]))
            assert np.array_equal(mn["focc"][c], h["focc"][c].astype(float).sum(0) / count)
            assert abs(mn["energy"][c] - e[c].mean()) <= 1e-13 * max(1.0, abs(e[c]).max())
            assert abs(mn["d2energy"][c] - d2[c].mean()) <= 1e-13 * max(1.0, abs(d2[c]).max())
    if h["ipr"] is not None:
        assert checked_ipr > 0
    return count, n


def test_measures_dense_chains(fk, orc):
    print("case chain measures dense L=6", flush=True)
    D, L, B = 2, 6, 8
    beta = np.geomspace(0.5, 20.0, B)
    U = np.array([1.0, 2.0, 4.0, 8.0] * 2)
    moves = [(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)]
    ctx, ch, beta, U = _chains(fk, D, L, B, 1, beta, U, moves)
    ch.measure_enable(6, ALL)
    ch.sweep_measure(4, 5)
    ch.measure()
    count, n = _check_records(orc, D, L, ch, beta, U)
    assert count == 5 and n == 5
    st = ch.state()
    h = ch.histories()
    assert np.array_equal(h["focc"][:, -1], st["f"])
    assert np.array_equal(h["spectrum"][:, -1], st["spectrum"])
    # the same chains without measurements end in the same state
    ctx2, ch2, _, _ = _chains(fk, D, L, B, 1, beta, U, moves)
    ch2.sweep_graph(20)
    st2 = ch2.state()
    for k in ("f", "logZ", "spectrum", "naccept"):
        assert np.array_equal(st[k], st2[k]), k
    for o in (ch, ch2, ctx, ctx2):
        o.close()


def test_measure_schedules_agree(fk, orc):
    """One graph of 3 measured sweeps, three graphs of one, and eager sweeps + measure
    record the same bits; records beyond the capacity still enter the means."""
    print("case chain measures schedules", flush=True)
    D, L, B = 2, 6, 4
    moves = [(fk.MOVE_ADDREMOVE, 0.6), (fk.MOVE_FLIP, 0.2), (fk.MOVE_HOP, 0.2)]
    runs = []
    for mode in range(3):
        ctx, ch, beta, U = _chains(fk, D, L, B, 2, 2.0, 3.0, moves)
        ch.measure_enable(2 if mode == 0 else 3, ALL)
        if mode == 0:
            ch.sweep_measure(3, 4)
        elif mode == 1:
            for _ in range(3):
                ch.sweep_measure(1, 4)
        else:
            for _ in range(3):
                ch.sweep(4)
                ch.measure()
        runs.append((ch.energies(), ch.means(), ch.histories(), ch.state()))
        ch.close()
        ctx.close()
    (c0, e0, d0, ce0), m0, h0, s0 = runs[0]
    assert c0 == 3 and e0.shape == (B, 2)
    for (c, e, d, ce), m, h, s in runs[1:]:
        assert c == 3 and e.shape == (B, 3)
        assert np.array_equal(e[:, :2], e0) and np.array_equal(d[:, :2], d0) and np.array_equal(ce[:, :2], ce0)
        for k in ("energy", "d2energy", "spectrum", "focc"):
            assert np.array_equal(m[k], m0[k]), k
        for k in ("spectrum", "focc", "ipr"):
            assert np.array_equal(h[k][:, :2], h0[k]), k
        assert np.array_equal(s["f"], s0["f"]) and np.array_equal(s["logZ"], s0["logZ"])
    assert np.array_equal(runs[1][2]["ipr"], runs[2][2]["ipr"])


def test_measures_rank1_chains(fk, orc):
    """The rank-one mode measures from its resident eigenvectors."""
    print("case chain measures rank-one L=8", flush=True)
    D, L, B = 2, 8, 4
    moves = [(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)]
    ctx, ch, beta, U = _chains(fk, D, L, B, 3, [1.0, 5.0, 10.0, 20.0], [1.0, 2.0, 4.0, 8.0], moves)
    ch.set_update(fk.UPDATE_RANK1, 16)
    ch.measure_enable(4, ALL)
    ch.sweep_measure(4, 8)
    count, n = _check_records(orc, D, L, ch, beta, U, ipr_tol=1e-8, spec_tol=1e-9)
    assert count == 4 and n == 4
    ch.close()
    ctx.close()


def test_measures_mixed_and_kpm_chains(fk, orc):
    """Mixed precision keeps a refined FP64 spectrum; stochastic-trace chains are
    diagonalised at each measurement."""
    print("case chain measures mixed / kpm L=8", flush=True)
    D, L, B = 2, 8, 4
    moves = [(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)]
    ctx, ch, beta, U = _chains(fk, D, L, B, 4, 2.0, [1.0, 2.0, 4.0, 8.0], moves)
    ch.set_precision(fk.PRECISION_MIXED)
    ch.measure_enable(3, ALL)
    ch.sweep_measure(3, 4)
    assert _check_records(orc, D, L, ch, beta, U, spec_tol=1e-10) == (3, 3)
    ch.close()
    ctx.close()
    ctx, ch, beta, U = _chains(fk, D, L, B, 5, 2.0, 2.0, moves)
    ch.set_kpm(R=8, M=64, seed=3)
    ch.measure_enable(3, 1 | 2)
    ch.sweep_measure(3, 4)
    assert _check_records(orc, D, L, ch, beta, U) == (3, 3)
    ch.close()
    ctx.close()


def test_measures_1d_ff_energy(fk, orc):
    """E includes the f-f energy of 1-D lattices (configuration_t::calc_ff_energy)."""
    print("case chain measures 1-D ff", flush=True)
    D, L, B = 1, 30, 4
    W = np.array([0.0, 0.5, 0.25])
    moves = [(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_FLIP, 0.5)]
    ctx, ch, beta, U = _chains(fk, D, L, B, 6, 3.0, 2.0, moves)
    ch.set_ff(W)
    ch.measure_enable(5, 1 | 2)
    ch.sweep_measure(5, 3)
    assert _check_records(orc, D, L, ch, beta, U, W=W) == (5, 5)
    ch.close()
    ctx.close()


def test_measures_triangular_ipr(fk, orc):
    """Triangular lattice (no bipartite symmetry): IPR from an FP64 decomposition of the
    dense chains' configurations."""
    print("case chain measures triangular L=6", flush=True)
    D, L, B = 2, 6, 3
    moves = [(fk.MOVE_ADDREMOVE, 1.0)]
    ctx, ch, beta, U = _chains(fk, D, L, B, 7, 4.0, 3.0, moves, kind=fk.LATTICE_TRIANGULAR)
    ch.measure_enable(3, ALL)
    ch.sweep_measure(3, 6)
    assert _check_records(orc, D, L, ch, beta, U, kind=orc.KIND_TRIANGULAR, tp=1.0) == (3, 3)
    ch.close()
    ctx.close()


def test_measure_errors(fk):
    print("case chain measures errors", flush=True)
    ctx, ch, _, _ = _chains(fk, 2, 6, 2, 8, 1.0, 1.0, [(fk.MOVE_ADDREMOVE, 1.0)])
    with pytest.raises(fk.FkgpuError):
        ch.measure()  # not enabled
    with pytest.raises(fk.FkgpuError):
        ch.measure_enable(-1, 0)
    with pytest.raises(fk.FkgpuError):
        ch.measure_enable(4, 64)  # no such flag
    with pytest.raises(fk.FkgpuError):
        ch.measure_conductivity([0.0], -1.0)
    ch.measure_enable(4, 1)
    with pytest.raises(fk.FkgpuError):
        ch.stiffness()  # not enabled
    with pytest.raises(fk.FkgpuError):
        ch.eigenfunctions()  # not enabled
    ch.measure()
    ch._mflags = 2  # ask for a history that was not enabled
    with pytest.raises(fk.FkgpuError):
        ch.histories()
    with pytest.raises(fk.FkgpuError):
        ch.sweep_measure(1, 0)
    ch.close()
    ctx.close()
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 1, 1100, 1.0, 0.0)
    ch = fk.Chains(ctx, np.zeros((1, 1100), np.int32), 1.0, 1.0, 0.5, 0.5, np.array([1], np.uint32))
    with pytest.raises(fk.FkgpuError):
        ch.measure_enable(2, 4)  # eigenvectors implemented for N <= 1024
    ch.close()
    ctx.close()
    ctx, ch, _, _ = _chains(fk, 1, 12, 2, 8, 1.0, 1.0, [(fk.MOVE_ADDREMOVE, 1.0)])
    with pytest.raises(fk.FkgpuError):
        ch.measure_enable(2, 8)  # stiffness: hypercubic lattices of dimension >= 2
    ch.close()
    ctx.close()


@pytest.mark.parametrize("mode", ["dense", "rank1"])
def test_measures_stiffness(fk, orc, mode):
    """Stiffness and conductivity records (measure_stiffness::accumulate,
    include/fk_mc/measures/stiffness.hpp:129-187) equal the single-configuration entry point
    fkgpu_measure_stiffness on the recorded occupations, and the oracle on one of them; from
    an FP64 decomposition per measurement (dense chains) or the resident eigenvectors
    (rank-one chains).  Means are the means of the records."""
    print(f"case chain stiffness {mode}", flush=True)
    D, L, B = 2, 6, 4
    wgrid = np.array([0.0, 0.3, 1.0, 2.5])
    moves = [(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)]
    ctx, ch, beta, U = _chains(fk, D, L, B, 9, [1.0, 3.0, 10.0, 30.0], 2.0, moves)
    if mode == "rank1":
        ch.set_update(fk.UPDATE_RANK1, 16)
    ch.measure_conductivity(wgrid, 0.05)
    ch.measure_enable(4, 2 | 8)
    ch.sweep_measure(4, 6)
    count, st, cd, sm, cm = ch.stiffness()
    h = ch.histories()
    assert count == 4 and st.shape == (B, 4) and cd.shape == (B, 4, len(wgrid))
    for c in range(B):
        ref_st, ref_cd = ctx.measure_stiffness(h["focc"][c], beta[c], U[c], U[c] / 2, wgrid, 0.05)
        np.testing.assert_allclose(st[c], ref_st, rtol=1e-9, atol=1e-11)
        np.testing.assert_allclose(cd[c], ref_cd, rtol=1e-9, atol=1e-11)
        np.testing.assert_allclose(sm[c], st[c].mean(), rtol=1e-13)
        np.testing.assert_allclose(cm[c], cd[c].mean(axis=0), rtol=1e-13)
    f = h["focc"][0, -1].astype(np.int32)
    ed = orc.calc_ed(orc.KIND_HYPERCUBIC, D, L, f, beta[0], U[0], U[0] / 2, tp=0.0, vectors=True)
    o = orc.measure_stiffness(D, L, ed["evecs"], ed["evals"], ed["cfermi"], wgrid, 0.05)
    np.testing.assert_allclose(st[0, -1], o["stiffness"], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(cd[0, -1], o["cond"], rtol=1e-9, atol=1e-11)
    ch.close()
    ctx.close()


@pytest.mark.parametrize("mode", ["dense", "rank1"])
def test_measures_eigenfunctions(fk, orc, mode):
    """Eigenvector records (measure_eigenfunctions::accumulate,
    src/measures/eigenfunctions.cpp:12-18) from an FP64 decomposition per measurement
    (dense chains) or the resident eigenvectors (rank-one chains)."""
    print(f"case chain eigenfunctions {mode}", flush=True)
    D, L, B = 2, 6, 3
    N = L ** D
    moves = [(fk.MOVE_ADDREMOVE, 0.5), (fk.MOVE_HOP, 0.5)]
    ctx, ch, beta, U = _chains(fk, D, L, B, 11, [2.0, 5.0, 20.0], [1.0, 3.0, 6.0], moves)
    if mode == "rank1":
        ch.set_update(fk.UPDATE_RANK1, 16)
    ch.measure_enable(3, 1 | 2 | 16)
    ch.sweep_measure(3, 5)
    count, V = ch.eigenfunctions()
    h = ch.histories()
    assert count == 3 and V.shape == (B, 3, N, N)
    hop = orc.lattice_hop(orc.KIND_HYPERCUBIC, D, L, 1.0, 0.0)[0]
    checked = 0
    for c in range(B):
        for m in range(3):
            v = V[c, m]
            ev = h["spectrum"][c, m]
            H = hop + np.diag(-U[c] / 2 + U[c] * h["focc"][c, m].astype(float))
            np.testing.assert_allclose(v.T @ v, np.eye(N), atol=1e-10)
            np.testing.assert_allclose(H @ v, v * ev[None, :], atol=1e-9)
            ed = orc.calc_ed(orc.KIND_HYPERCUBIC, D, L, h["focc"][c, m], beta[c], U[c], U[c] / 2, tp=0.0,
                             vectors=True)
            gap = np.diff(ed["evals"])
            iso = np.ones(N, bool)
            iso[:-1] &= gap > 1e-6
            iso[1:] &= gap > 1e-6
            ov = np.abs(np.sum(v * ed["evecs"], axis=0))
            assert np.abs(ov[iso] - 1.0).max(initial=0.0) <= 1e-9, (c, m)
            checked += int(iso.sum())
    assert checked > 0
    ch.close()
    ctx.close()
