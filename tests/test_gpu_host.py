"""The C++ host layer (configuration_t / move_* / mc_metropolis / measure_*, driven
through include/fkmc_host.h) against the oracle: the plain per-move path of the
reference program, with every diagonalisation on the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HC, TRI = 0, 1
ADD, FLIP, HOP, RAND = 0, 1, 2, 3


@pytest.fixture(scope="module")
def host():
    from fk_mc_amd import host_api
    return host_api


def _compare(tr, ref, moves, rtol=1e-9):
    # decisions: identical unless a Metropolis comparison is a near-tie (must be none here)
    ties = np.abs(np.log(np.maximum(ref["u"], 1e-300)) - ref["logw"]) < 1e-9
    assert not ties.any(), "near-tie in the oracle trace: pick another seed"
    kinds = np.array([k for k, _ in moves])[tr["kind"]]  # host trace: add_move index
    np.testing.assert_array_equal(kinds, ref["kind"])
    np.testing.assert_array_equal(tr["a"], ref["a"])
    np.testing.assert_array_equal(tr["b"], ref["b"])
    np.testing.assert_array_equal(tr["acc"], ref["acc"])
    np.testing.assert_array_equal(tr["u"], ref["u"])
    w = tr["w"]
    lw = np.where(w > 0, np.log(np.maximum(w, 1e-300)), -np.inf)
    # move_randomize returns its clamped weight (src/moves.cpp:46-48); the oracle traces the raw log-weight
    fin = np.isfinite(ref["logw"]) & (ref["logw"] > -600) & (kinds != RAND)
    np.testing.assert_allclose(lw[fin], ref["logw"][fin], rtol=0, atol=1e-8 * (1 + np.abs(ref["logw"][fin]).max()))
    np.testing.assert_allclose(tr["logz_cur"], ref["logz_cur"], rtol=rtol)


@pytest.mark.parametrize("kind,D,L,moves,seed", [
    (HC, 2, 8, [(ADD, 0.5), (HOP, 0.5)], 11),
    (HC, 1, 24, [(FLIP, 1.0), (ADD, 2.0)], 5),
    (TRI, 2, 6, [(ADD, 0.5), (HOP, 0.5)], 3),
    (HC, 3, 4, [(ADD, 1.0)], 9),
    (HC, 2, 6, [(ADD, 0.4), (FLIP, 0.3), (RAND, 0.3)], 21),
])
def test_plain_chain_matches_oracle(host, orc, kind, D, L, moves, seed):
    beta, U, mu = 3.0, 4.0, 2.0
    n = L ** D
    f0 = orc.randomize_f(n, seed, n // 2)
    nmoves = 400
    m = host.model(kind, D, L, beta, U, mu, mu)
    tr = host.run_chain(m, seed, f0, moves, nmoves)
    ref = orc.run_chain(kind, D, L, beta, U, mu, mu, seed, f0, moves, nmoves)
    print(f"host chain kind={kind} D={D} L={L} seed={seed} accepted {tr['naccept']}/{nmoves}", flush=True)
    _compare(tr, ref, moves)
    assert tr["naccept"] == ref["naccept"]
    np.testing.assert_array_equal(tr["f_final"], ref["f_final"])
    np.testing.assert_allclose(tr["spectrum_final"], ref["spectrum_final"], rtol=0, atol=1e-11)


def test_chebyshev_chain_consistent(host, orc):
    """Chebyshev moves: every weight equals the oracle's exact-trace logZ ratio."""
    kind, D, L, seed = HC, 2, 6, 4
    beta, U, mu = 2.0, 2.0, 1.0
    n = L ** D
    M, G = 32, 64
    f0 = orc.randomize_f(n, seed, n // 2)
    m = host.model(kind, D, L, beta, U, mu, mu, cheb_size=M, cheb_grid=G)
    tr = host.run_chain(m, seed, f0, [(ADD, 1.0)], 60)
    f = f0.copy()
    lz = orc.calc_chebyshev(kind, D, L, f, beta, U, mu, M, G)["logZ"]
    for s in range(60):
        a = tr["a"][s]
        g = f.copy()
        g[a] = 1 - g[a]
        lz_new = orc.calc_chebyshev(kind, D, L, g, beta, U, mu, M, G)["logZ"]
        sgn = 1 if g[a] else -1
        np.testing.assert_allclose(np.log(tr["w"][s]), lz_new - lz + sgn * beta * mu, rtol=0, atol=1e-8)
        if tr["acc"][s]:
            f, lz = g, lz_new
        np.testing.assert_allclose(tr["logz_cur"][s], lz, rtol=1e-10)
    np.testing.assert_array_equal(tr["f_final"], f)


def test_run_mc_measures(host, orc):
    kind, D, L, seed = HC, 2, 6, 17
    beta, U, mu = 2.0, 4.0, 2.0
    n = L ** D
    m = host.model(kind, D, L, beta, U, mu, mu)
    out = host.run_mc(m, seed, [(ADD, 1.0)], nsweeps=12, sweep_len=8, ntherm=2)
    spectra = []
    for i in range(12):
        f = out["focc"][i].astype(np.int32)
        ed = orc.calc_ed(kind, D, L, f, beta, U, mu)
        e = orc.measure_energy(ed["evals"], ed["cexp"], mu, int(f.sum()))
        np.testing.assert_allclose(out["energies"][i], e["energy"], rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(out["d2energies"][i], e["d2energy"], rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(out["c_energies"][i], e["c_energy"], rtol=1e-10, atol=1e-10)
        spectra.append(ed["evals"])
    np.testing.assert_allclose(out["avg_spectrum"], np.mean(spectra, axis=0), rtol=0, atol=1e-11)
    assert 0.0 < out["acceptance"] <= 1.0


def test_measure_config(host, orc):
    kind, D, L = HC, 2, 8
    beta, U, mu = 5.0, 2.0, 1.0
    n = L ** D
    f = orc.randomize_f(n, 8, n // 3)
    wgrid = np.linspace(0.0, 3.0, 7)
    m = host.model(kind, D, L, beta, U, mu, mu)
    out = host.measure_config(m, f, wgrid=wgrid, offset=0.05, eta=0.05)
    ed = orc.calc_ed(kind, D, L, f, beta, U, mu, vectors=True)
    np.testing.assert_allclose(out["spectrum"], ed["evals"], rtol=0, atol=1e-11)
    e = orc.measure_energy(ed["evals"], ed["cexp"], mu, int(f.sum()))
    np.testing.assert_allclose(out["energy"], [e["energy"], e["d2energy"], e["c_energy"]], rtol=1e-10)
    V = out["evecs"]
    H = orc.lattice_hop(kind, D, L)[0] + np.diag(-mu + U * f)
    assert np.abs(H @ V - V * out["spectrum"]).max() < 1e-10
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-10
    np.testing.assert_allclose(out["ipr"], orc.measure_ipr(V), rtol=1e-9)
    st = orc.measure_stiffness(D, L, ed["evecs"], ed["evals"], ed["cfermi"], wgrid, 0.05)
    np.testing.assert_allclose(out["stiffness"][0], st["stiffness"], rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(out["cond"], st["cond"], rtol=1e-7, atol=1e-9)
    dos = [(0.05 / np.pi / ((w - ed["evals"]) ** 2 + 0.05 ** 2)).mean() for w in wgrid]
    np.testing.assert_allclose(out["dos"], dos, rtol=1e-12)


def test_errors_are_reported(host):
    from fk_mc_amd import FkgpuError
    m = host.model(HC, 4, 3, 1.0, 1.0, 0.5, 0.5)
    with pytest.raises(FkgpuError, match="D in 1..3"):
        host.run_chain(m, 1, np.zeros(81, dtype=np.int32), [(ADD, 1.0)], 1)
    m = host.model(TRI, 2, 6, 1.0, 1.0, 0.5, 0.5)
    with pytest.raises(FkgpuError, match="hypercubic"):
        host.measure_config(m, np.zeros(36, dtype=np.int32))
