"""The batched Monte Carlo program fk_mc_amd/bin/fk_mc_gpu -- the reference's
fk_mc_<lattice> program (prog/fk_mc_exec.cpp, include/fk_mc/fk_mc.hxx) with its parameter
names and defaults -- against the host layer's single-chain run of the same seeds
(fkmc_run_mc: configuration_t::randomize_f, move_addremove, mc_metropolis and the
reference's measures, one chain per seed as one MPI rank each) and against the oracle's
calc_ed / measure_energy / measure_ipr of the recorded occupations."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fk_mc_amd", "bin", "fk_mc_gpu")


def _run(tmp_path, *args, output="run"):
    out = tmp_path / output
    cmd = [EXE] + [str(a) for a in args] + ["--output", str(out), "--show_output", "false"]
    print("case driver", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def _params(out):
    p = {}
    for line in open(out / "parameters.txt"):
        k, _, v = line.partition(" = ")
        p[k.strip()] = v.strip()
    return p


@pytest.mark.gpu
def test_driver_matches_host_layer(tmp_path, orc):
    from fk_mc_amd import host_api as host
    L, R, seed, beta, U, mu = 4, 3, 7, 2.0, 2.0, 1.0
    ns, sl, nt = 12, 4, 2
    N = L * L
    out = _run(tmp_path, "--lattice", "cubic2d", "--L", L, "--beta", beta, "--U", U, "--mu_c", mu, "--mu_f", mu,
               "--Nf_start", N // 2, "--nsweeps", ns, "--sweep_len", sl, "--ntherm_sweeps", nt, "--chains", R,
               "--seed", seed, "--measure_ipr=true")
    md = out / "mc_data"
    e, d2, ce = (np.load(md / f"{k}.npy") for k in ("energies", "d2energies", "c_energies"))
    fh, sh, ih = (np.load(md / f"{k}.npy") for k in ("focc_history", "spectrum_history", "ipr_history"))
    spec = np.load(md / "spectrum.npy")
    assert e.shape == (R, ns) and fh.shape == (R, ns, N) and spec.shape == (R, N)
    m = host.model(0, 2, L, beta, U, mu, mu)
    checked = 0
    for r in range(R):
        # rank r of the reference: seed + r, randomize_f(msize / 2) from the Metropolis generator
        ref = host.run_mc(m, seed + r, [(0, 1.0)], nsweeps=ns, sweep_len=sl, ntherm=nt)
        np.testing.assert_array_equal(fh[r], ref["focc"])
        np.testing.assert_allclose(e[r], ref["energies"], rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(d2[r], ref["d2energies"], rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(ce[r], ref["c_energies"], rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(spec[r], ref["avg_spectrum"], rtol=0, atol=1e-11)
        for i in range(ns):
            f = fh[r, i].astype(np.int32)
            ed = orc.calc_ed(orc.KIND_HYPERCUBIC, 2, L, f, beta, U, mu, tp=0.0, vectors=True)
            np.testing.assert_allclose(sh[r, i], ed["evals"], rtol=0, atol=1e-10)
            en = orc.measure_energy(ed["evals"], ed["cexp"], mu, int(f.sum()))
            np.testing.assert_allclose(e[r, i], en["energy"], rtol=1e-10, atol=1e-10)
            ipr = orc.measure_ipr(ed["evecs"])
            gap = np.diff(ed["evals"])
            iso = np.ones(N, bool)
            iso[:-1] &= gap > 1e-6
            iso[1:] &= gap > 1e-6
            np.testing.assert_allclose(ih[r, i][iso], ipr[iso], rtol=0, atol=1e-9)
            checked += int(iso.sum())
    assert checked > 0
    # DOS of the recorded spectra on the reference's grid (prog/data_save.hxx:67-69)
    dg = np.load(out / "dos.npy")
    grid = -6.0 + 12.0 * np.arange(240) / 240.0
    np.testing.assert_allclose(dg[:, 0], grid, rtol=0, atol=1e-14)
    eta = 0.05
    lor = (eta / np.pi) / ((grid[None, :] - sh.reshape(-1)[:, None]) ** 2 + eta ** 2)
    np.testing.assert_allclose(dg[:, 1], lor.sum(0) / (R * ns * N), rtol=1e-10, atol=1e-13)
    p = _params(out)
    assert p["seed"] == str(seed) and p["chains"] == str(R) and p["mc_add_remove"] == "1"
    stats = open(out / "stats.txt").read()
    assert "energy/N" in stats and "cv" in stats


@pytest.mark.gpu
def test_driver_moves_ini_and_1d(tmp_path, orc):
    """An ini file as the first argument, flip + add_remove moves registered in the
    reference's order, the 1-D f-f term W, no history (energies only)."""
    from fk_mc_amd import host_api as host
    ini = tmp_path / "p.ini"
    ini.write_text("# parameters\nL = 12\nbeta = 3\nU = 2\nmu_c = 1\nmu_f = 1\nmc_flip = 0.3\n"
                   "mc_add_remove = 0.7\nW = 0.0, 0.5\nNf_start = 6\nmeasure_history = false\n")
    out = _run(tmp_path, ini, "--lattice", "cubic1d", "--nsweeps", 8, "--sweep_len", 5, "--ntherm_sweeps", 1,
               "--chains", 2, "--seed", 11)
    md = out / "mc_data"
    e = np.load(md / "energies.npy")
    assert not (md / "spectrum_history.npy").exists()
    m = host.model(0, 1, 12, 3.0, 2.0, 1.0, 1.0, W=[0.0, 0.5])
    for r in range(2):
        # moves in the reference's registration order: flip, then add_remove
        ref = host.run_mc(m, 11 + r, [(1, 0.3), (0, 0.7)], nsweeps=8, sweep_len=5, ntherm=1)
        np.testing.assert_allclose(e[r], ref["energies"], rtol=1e-10, atol=1e-10)


def test_driver_parameters_without_device(tmp_path):
    """Parameter handling ends before any device call: runs without a GPU."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built")
    r = subprocess.run([EXE, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    for name, default in (("beta", "10"), ("U", "1"), ("mu_c", "0.5"), ("mc_add_remove", "1"), ("Nf_start", "5"),
                          ("nsweeps", "1024"), ("sweep_len", "16"), ("ntherm_sweeps", "1"), ("L", "4"),
                          ("dos_npts", "240")):
        assert f"--{name} " in r.stdout and f"(default: {default})" in r.stdout.split(f"--{name} ")[1].split("\n")[0]
    r = subprocess.run([EXE, "--no_such_parameter", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "unknown parameter" in r.stderr
    r = subprocess.run([EXE, "--cheb_moves", "true", "--output", str(tmp_path / "x")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0
    # prog/fk_mc_exec.cpp:97-117: `exit` prints the parameter summary and ends before any
    # device work or output; the reference's other program parameters are accepted
    r = subprocess.run([EXE, "--exit", "true", "--L", "6", "--U", "2.5", "--nsweeps", "7", "--maxtime", "10",
                        "--h5lattice", "x.h5", "--max_time", "100", "--output", str(tmp_path / "dry.h5")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "System size 6" in r.stdout and "U 2.5" in r.stdout and "sweeps 7" in r.stdout
    assert not (tmp_path / "dry.h5").exists()


@pytest.mark.gpu
def test_driver_update_modes_agree(tmp_path):
    """--update rank1 and --precision mixed run the same chains as the dense FP64 default:
    the same occupations and energies within 1e-10 (decisions can only part at near-ties)."""
    common = ["--lattice", "cubic2d", "--L", 6, "--beta", 4, "--U", 2, "--mu_c", 1, "--mu_f", 1, "--Nf_start", 18,
              "--mc_add_remove", 0.5, "--mc_hop", 0.5, "--nsweeps", 10, "--sweep_len", 8, "--ntherm_sweeps", 1,
              "--chains", 8, "--seed", 21]
    runs = {}
    for name, extra in (("dense", []), ("rank1", ["--update", "rank1", "--refresh", 16]),
                        ("mixed", ["--precision", "mixed"])):
        d = tmp_path / name
        d.mkdir()
        out = _run(d, *(common + extra))
        runs[name] = {k: np.load(out / "mc_data" / f"{k}.npy") for k in ("energies", "focc_history")}
    for name in ("rank1", "mixed"):
        np.testing.assert_array_equal(runs[name]["focc_history"], runs["dense"]["focc_history"])
        np.testing.assert_allclose(runs[name]["energies"], runs["dense"]["energies"], rtol=1e-10, atol=1e-10)


@pytest.mark.gpu
def test_driver_stiffness(tmp_path):
    """--measure_stiffness: the stiffness and conductivity records on the reference's
    conductivity grid (prog/fk_mc_exec.cpp:127-139) equal fkgpu_measure_stiffness of the
    recorded occupations."""
    import fk_mc_amd as fk
    L, U, beta, mu = 6, 2.0, 5.0, 1.0
    out = _run(tmp_path, "--lattice", "cubic2d", "--L", L, "--beta", beta, "--U", U, "--mu_c", mu, "--mu_f", mu,
               "--Nf_start", 18, "--nsweeps", 3, "--sweep_len", 6, "--ntherm_sweeps", 1, "--chains", 2,
               "--seed", 5, "--measure_stiffness", "true", "--cond_npoints", 4, "--cond_offset", 0.05)
    md = out / "mc_data"
    st, cd = np.load(md / "stiffness.npy"), np.load(md / "cond_history.npy")
    fh = np.load(md / "focc_history.npy").astype(np.int32)
    w = np.load(out / "wgrid_conductivity.npy")
    w2 = 8.0 * np.exp(np.linspace(-15.0, 0.0, 4))
    np.testing.assert_allclose(w, np.concatenate([-w2[::-1], [0.0], w2]), rtol=1e-15)
    assert st.shape == (2, 3) and cd.shape == (2, 3, 9)
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 2, L, 1.0, 0.0)
    for r in range(2):
        ref_st, ref_cd = ctx.measure_stiffness(fh[r], beta, U, mu, w, 0.05)
        np.testing.assert_allclose(st[r], ref_st, rtol=1e-9, atol=1e-11)
        np.testing.assert_allclose(cd[r], ref_cd, rtol=1e-9, atol=1e-11)
    ctx.close()
    assert "stiffness" in open(out / "stats.txt").read()


@pytest.mark.gpu
def test_driver_eigenfunctions_and_gwr(tmp_path, orc):
    """--measure_eigenfunctions (switching the histories on, prog/fk_mc_exec.cpp:62) with
    --save_eigenfunctions: eig_history, and save_gwr's outputs (prog/data_save.hxx:535-708)
    -- G(0 + i dos_offset) as the mean resolvent of the recorded configurations, its
    translation average and numpy's 2-D FFT of it, and the DOS / typical DOS rows from the
    recorded spectra."""
    L, U, beta, mu, R, ns, off = 4, 2.0, 3.0, 1.0, 2, 3, 0.1
    N = L * L
    out = _run(tmp_path, "--lattice", "cubic2d", "--L", L, "--beta", beta, "--U", U, "--mu_c", mu, "--mu_f", mu,
               "--Nf_start", 8, "--nsweeps", ns, "--sweep_len", 4, "--ntherm_sweeps", 1, "--chains", R,
               "--seed", 3, "--measure_history", "false", "--measure_eigenfunctions", "true",
               "--save_eigenfunctions", "true", "--dos_offset", off, "--dos_npts", 20, "--dos_width", 4)
    md, sd = out / "mc_data", out / "stats"
    V = np.load(md / "eig_history.npy")
    fh = np.load(md / "focc_history.npy")  # histories on despite --measure_history false
    sh = np.load(md / "spectrum_history.npy")
    assert V.shape == (R, ns, N, N)
    hop = orc.lattice_hop(orc.KIND_HYPERCUBIC, 2, L, 1.0, 0.0)[0]
    G = np.zeros((N, N), complex)
    for r in range(R):
        for m in range(ns):
            H = hop + np.diag(-mu + U * fh[r, m])
            np.testing.assert_allclose(V[r, m].T @ V[r, m], np.eye(N), atol=1e-10)
            np.testing.assert_allclose(H @ V[r, m], V[r, m] * sh[r, m][None, :], atol=1e-9)
            G += np.linalg.inv(1j * off * np.eye(N) - H)
    G /= R * ns
    ws = "0._0.1"
    gre, gim = np.load(sd / f"gr_full_w{ws}_re.npy"), np.load(sd / f"gr_full_w{ws}_im.npy")
    np.testing.assert_allclose(gre, G.real, atol=1e-10 * np.abs(G).max())
    np.testing.assert_allclose(gim, G.imag, atol=1e-10 * np.abs(G).max())
    r2 = np.load(sd / f"gr_w{ws}_re.npy") + 1j * np.load(sd / f"gr_w{ws}_im.npy")
    gk = np.load(sd / f"gk_w{ws}_re.npy") + 1j * np.load(sd / f"gk_w{ws}_im.npy")
    assert abs(r2[0, 0] - np.trace(G) / N) <= 1e-12 * np.abs(G).max()
    assert abs(r2.sum() - G.sum() / N) <= 1e-11 * np.abs(G).sum()
    np.testing.assert_allclose(gk, np.fft.fft2(r2), atol=1e-11 * np.abs(gk).max())
    ld0 = -np.diag(G).imag / np.pi
    t0 = np.load(sd / f"tdos{ws}.npy")
    np.testing.assert_allclose(t0[[0, 1, 3, 5, 7]], [0.0, off, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(t0[4], ld0.mean(), rtol=1e-10)
    np.testing.assert_allclose(t0[2], np.exp(np.log(ld0).mean()), rtol=1e-10)
    tv = np.load(sd / "tdos_gwr.npy")
    grid = -4.0 + 8.0 * np.arange(20) / 20
    assert tv.shape == (20, 8)
    np.testing.assert_allclose(tv[:, 0], grid, rtol=1e-15)
    lor = (off / np.pi) / ((grid[:, None] - sh.reshape(-1)[None, :]) ** 2 + off * off)
    np.testing.assert_allclose(tv[:, 4], lor.sum(axis=1) / (R * ns * N), rtol=1e-10)
    assert (tv[:, 2] <= tv[:, 4] * (1 + 1e-12)).all()  # geometric mean <= arithmetic mean
    np.testing.assert_allclose(tv[:, 6], tv[:, 2] / tv[:, 4], rtol=1e-14)


@pytest.mark.gpu
def test_driver_error_analysis(tmp_path):
    """binning/ and stats/ as data_saver writes them (prog/data_save.hxx:158-345,486-532):
    the series of all chains in chain order, binned from the last measurement backwards,
    each equal to the host library's binning / jackknife of the recorded series; dos_err and
    ipr_err on the DOS grid at the bin that dos0 / ipr0 pick."""
    from fk_mc_amd import host_api as host
    L, R, ns, beta, U, off = 4, 3, 40, 4.0, 2.0, 0.05
    N = L * L
    out = _run(tmp_path, "--lattice", "cubic2d", "--L", L, "--beta", beta, "--U", U, "--mu_c", 1, "--mu_f", 1,
               "--Nf_start", 8, "--nsweeps", ns, "--sweep_len", 2, "--ntherm_sweeps", 1, "--chains", R,
               "--seed", 9, "--measure_ipr", "true", "--dos_npts", 16, "--dos_width", 4, "--dos_offset", off)
    md, sd, bd = out / "mc_data", out / "stats", out / "binning"
    n = R * ns
    depth = host.max_bin(n)
    rev = lambda a: np.ascontiguousarray(a.reshape(-1)[::-1])  # noqa: E731
    for name in ("energies", "d2energies", "c_energies"):
        t, b = host.binning(rev(np.load(md / f"{name}.npy")), depth)
        key = name[:-3] + "y" if name != "energies" else "energy"
        np.testing.assert_array_equal(np.load(bd / f"{key}.npy"), t)
        np.testing.assert_array_equal(np.load(sd / f"{key}.npy"), t[b, :4])
    e, d2 = rev(np.load(md / "energies.npy")), rev(np.load(md / "d2energies.npy"))
    t, b = host.jackknife_cv(e, d2, beta, N, depth)
    np.testing.assert_allclose(np.load(sd / "cv.npy"), t[b], rtol=1e-12)
    sh = np.load(md / "spectrum_history.npy").reshape(n, N)[::-1]
    ih = np.load(md / "ipr_history.npy").reshape(n, N)[::-1]
    grid = -4.0 + 8.0 * np.arange(16) / 16
    lor = off / ((grid[:, None, None] - sh[None]) ** 2 + off * off)  # [w][record][k]
    dos_rec = lor.sum(axis=2) / (np.pi * N)
    ipr_rec = (lor * ih[None] ** 4).sum(axis=2) / lor.sum(axis=2)
    for name, rec in (("dos", dos_rec), ("ipr", ipr_rec)):
        s0 = np.load(sd / f"{name}0.npy")
        tab = np.load(sd / f"{name}_err.npy")
        assert tab.shape == (16, 3)
        np.testing.assert_allclose(tab[:, 0], grid, rtol=1e-15)
        bt = np.load(bd / f"{name}0.npy")
        b = int(np.flatnonzero((bt[:, :4] == s0).all(axis=1))[0])  # the level estimate_bin chose
        used = int(bt[b, 0]) << b  # the records that level covers
        np.testing.assert_allclose(tab[:, 1], rec[:, :used].mean(axis=1), rtol=1e-10)
        for q in (0, 7, 15):
            tq, _ = host.binning(rec[q], b)
            np.testing.assert_allclose(tab[q, 2], tq[b, 3], rtol=1e-9)
    assert "cv = " in open(out / "stats.txt").read()


@pytest.mark.gpu
def test_driver_gwr_cubic3d(tmp_path, orc):
    """save_gwr on a 3-D lattice: G(r_j - r_i) is averaged over the first two lattice axes
    (the site index runs fastest along the last axis), so gr_w holds sums of G over site
    pairs with the same (dx, dy) offset; its 2-D transform is numpy's fft2 of it."""
    L, U, R, ns, off = 4, 3.0, 2, 2, 0.1
    N = L ** 3
    out = _run(tmp_path, "--lattice", "cubic3d", "--L", L, "--beta", 2, "--U", U, "--mu_c", U / 2, "--mu_f",
               U / 2, "--Nf_start", 20, "--nsweeps", ns, "--sweep_len", 2, "--ntherm_sweeps", 1, "--chains", R,
               "--seed", 4, "--measure_eigenfunctions", "true", "--dos_offset", off, "--dos_npts", 8)
    md, sd = out / "mc_data", out / "stats"
    fh = np.load(md / "focc_history.npy")
    hop = orc.lattice_hop(orc.KIND_HYPERCUBIC, 3, L, 1.0, 0.0)[0]
    G = np.zeros((N, N), complex)
    for r in range(R):
        for m in range(ns):
            G += np.linalg.inv(1j * off * np.eye(N) - (hop + np.diag(-U / 2 + U * fh[r, m])))
    G /= R * ns
    ws = "0._0.1"
    g = np.load(sd / f"gr_full_w{ws}_re.npy") + 1j * np.load(sd / f"gr_full_w{ws}_im.npy")
    np.testing.assert_allclose(g, G, atol=1e-10 * np.abs(G).max())
    r2 = np.load(sd / f"gr_w{ws}_re.npy") + 1j * np.load(sd / f"gr_w{ws}_im.npy")
    assert r2.shape == (L, L)
    idx = np.arange(N)
    x, y = idx // (L * L), (idx // L) % L
    want = np.zeros((L, L), complex)
    np.add.at(want, ((x[None, :] - x[:, None]) % L, (y[None, :] - y[:, None]) % L), G / N)
    np.testing.assert_allclose(r2, want, atol=1e-12 * np.abs(want).max())
    gk = np.load(sd / f"gk_w{ws}_re.npy") + 1j * np.load(sd / f"gk_w{ws}_im.npy")
    np.testing.assert_allclose(gk, np.fft.fft2(r2), atol=1e-11 * np.abs(gk).max())
    assert not (md / "eig_history.npy").exists()  # only with save_eigenfunctions


@pytest.mark.gpu
def test_driver_cheb_moves(tmp_path, fk):
    """cheb_moves (include/fk_mc/fk_mc.hxx:59-78): the moves run on the Chebyshev path, the
    global exchange with its f-f term (src/moves_chebyshev.cpp:22), and without measure_ipr
    only occupations are recorded (fk_mc.hxx:115-124).  Each chain equals a library chain set
    built the same way: randomize_f(Nf_start) from the chain's generator, the moves in the
    reference's order, Chebyshev weights, occupation-only measurements."""
    L, R, seed, beta, U, mu = 24, 3, 11, 3.0, 2.0, 1.0
    ns, sl, nt = 10, 4, 2
    W = [0.0, 0.5, 0.2]
    out = _run(tmp_path, "--lattice", "cubic1d", "--L", L, "--beta", beta, "--U", U, "--mu_c", mu, "--mu_f", mu,
               "--Nf_start", L // 2, "--nsweeps", ns, "--sweep_len", sl, "--ntherm_sweeps", nt, "--chains", R,
               "--seed", seed, "--cheb_moves", "true", "--mc_flip", 0.5, "--mc_add_remove", 0.5,
               "--measure_history", "true", "--W", ",".join(str(x) for x in W))
    md = out / "mc_data"
    fh = np.load(md / "focc_history.npy")
    assert fh.shape == (R, ns, L)
    assert not (md / "energies.npy").exists() and not (md / "spectrum_history.npy").exists()
    ctx = fk.Context(0).set_lattice(fk.LATTICE_HYPERCUBIC, 1, L, 1.0, 0.0)
    ch = fk.Chains(ctx, np.zeros((R, L), np.int32), beta, U, mu, mu, np.arange(R, dtype=np.uint32) + seed)
    ch.randomize_f(np.full(R, L // 2, np.int32))
    ch.set_moves([(fk.MOVE_FLIP, 0.5), (fk.MOVE_ADDREMOVE, 0.5)])
    ch.set_ff(W)
    ch.set_kpm(R=8, M=8, seed=1, jackson=False)
    ch.measure_enable(ns, fk.MEASURE_FOCC_HISTORY | fk.MEASURE_NO_SPECTRUM)
    ch.sweep_graph(nt * sl)
    ch.sweep_measure(ns, sl)
    np.testing.assert_array_equal(ch.histories()["focc"], fh.astype(np.int32))
    ch.close()
    ctx.close()
    # the mixed / rank-one modes and the reshuffle move are refused with Chebyshev moves
    for extra in (["--mc_reshuffle", 0.1], ["--precision", "mixed"]):
        cmd = [EXE, "--lattice", "cubic1d", "--L", L, "--cheb_moves", "true", "--nsweeps", 2, "--chains", 1,
               "--output", str(tmp_path / "bad.h5"), "--show_output", "false"] + extra
        r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and "cheb_moves" in r.stderr


@pytest.mark.gpu
def test_driver_max_time(tmp_path):
    """max_time ends the measured sweeps after the block in which it ran out (the reference's
    alps::stop_callback): with max_time 0 one block of nsweeps / 10 sweeps is measured and every
    series has that length; the chains are the same as in a full run over those sweeps."""
    L, R, ns = 4, 2, 40
    common = ["--lattice", "cubic2d", "--L", L, "--beta", 2.0, "--U", 2.0, "--mu_c", 1.0, "--mu_f", 1.0, "--nsweeps", ns,
              "--sweep_len", 4, "--ntherm_sweeps", 1, "--chains", R, "--seed", 3, "--measure_history", "true"]
    out = _run(tmp_path, *(common + ["--max_time", 0]), output="short")
    e = np.load(out / "mc_data" / "energies.npy")
    fh = np.load(out / "mc_data" / "focc_history.npy")
    assert e.shape == (R, ns // 10) and fh.shape == (R, ns // 10, L * L)
    full = _run(tmp_path, *common, output="full")
    np.testing.assert_array_equal(np.load(full / "mc_data" / "energies.npy")[:, :ns // 10], e)
    np.testing.assert_array_equal(np.load(full / "mc_data" / "focc_history.npy")[:, :ns // 10], fh)
