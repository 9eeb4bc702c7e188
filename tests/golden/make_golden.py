"""Writes the golden fixtures of this directory.

Run from the repository root:  python tests/golden/make_golden.py

Every fixture is data (inputs and expected outputs).  Sources of truth:
  eigs_lapack.npz     LAPACK (numpy.linalg.eigvalsh, dsyevd) on H built here in
                      numpy from the lattice definition (src/lattice/hypercubic.cpp:115-155)
                      and H = hop + diag(-mu_c + U f) (src/configuration.cpp:79-91)
  closed_forms.npz    analytic spectra: U = 0 plane waves, uniform f, and the
                      checkerboard f on even-L bipartite lattices
  stiffness.json      test/stiffness_test.cpp:60-63 configurations with the
                      reference's quoted values (tol 1e-3) and the value from a
                      numpy restatement of stiffness.hpp:129-187 on LAPACK eigenpairs
  cheb_quadrature.json  test/fast_update_test.cpp:62-65 with scipy Bessel values
  chain_c1.npz        10^4-move addremove chain, cubic2d L=8, U=1, beta=1, mu=1/2,
                      f ~ Bernoulli(0.5) (numpy default_rng(2024)), mt19937 seed 12345,
                      produced by the oracle chain (the CPU restatement of
                      src/mc_metropolis.cpp:34-52 and src/moves.cpp:52-67)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def hop_matrix(kind, D, L, t=1.0, tp=1.0):
    """Hopping matrix; index = sum_d p_d L^(D-1-d) (last coordinate fastest)."""
    n = L ** D
    H = np.zeros((n, n))
    coords = np.array(np.unravel_index(np.arange(n), (L,) * D)).T
    for i, p in enumerate(coords):
        for d in range(D):
            for s in (-1, 1):
                q = p.copy()
                q[d] = (q[d] + s) % L
                H[i, np.ravel_multi_index(q, (L,) * D)] += -t
        if kind == 1:
            for s in (-1, 1):
                q = (p + s) % L
                H[i, np.ravel_multi_index(q, (L,) * D)] += -tp
    return H


def hamiltonian(kind, D, L, f, U, mu_c, t=1.0, tp=1.0):
    return hop_matrix(kind, D, L, t, tp) + np.diag(-mu_c + U * np.asarray(f, float))


def eigs_fixture():
    cases = []
    rng = np.random.default_rng(7)
    for i in range(8):  # N = 64
        U = [0.5, 1.0, 2.0, 4.0, 8.0, 1.0, 3.0, 6.0][i]
        cases.append((0, 2, 8, U, U / 2, rng.integers(0, 2, 64)))
    for U in (2.0, 8.0):  # N = 256
        cases.append((0, 2, 16, U, U / 2, rng.integers(0, 2, 256)))
    cases.append((1, 2, 6, 3.0, 1.5, rng.integers(0, 2, 36)))   # triangular
    cases.append((0, 3, 4, 5.0, 2.5, rng.integers(0, 2, 64)))   # cubic3d
    cases.append((0, 1, 40, 2.0, 1.0, rng.integers(0, 2, 40)))  # chain
    out = {}
    for c, (kind, D, L, U, mu, f) in enumerate(cases):
        ev = np.linalg.eigvalsh(hamiltonian(kind, D, L, f, U, mu))
        out[f"case{c}_meta"] = np.array([kind, D, L, U, mu])
        out[f"case{c}_f"] = f.astype(np.int8)
        out[f"case{c}_evals"] = ev
    out["ncases"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "eigs_lapack.npz"), **out)


def closed_forms_fixture():
    out = {}
    c = 0
    for D, L in ((1, 8), (2, 4), (2, 8), (3, 4)):
        ks = np.array(np.meshgrid(*[2 * np.pi * np.arange(L) / L] * D, indexing="ij")).reshape(D, -1)
        eps = -2.0 * np.cos(ks).sum(axis=0)
        for U, mu, mode in ((0.0, 0.3, "empty"), (2.0, 1.0, "full"), (4.0, 2.0, "checkerboard")):
            n = L ** D
            coords = np.array(np.unravel_index(np.arange(n), (L,) * D))
            if mode == "empty":
                f = np.zeros(n, np.int8)
                ev = np.sort(eps - mu)
            elif mode == "full":
                f = np.ones(n, np.int8)
                ev = np.sort(eps - mu + U)
            else:
                f = (coords.sum(axis=0) % 2).astype(np.int8)
                # U f = U/2 - (U/2)(-1)^{sum p} couples k with k+Q (eps_{k+Q} = -eps_k):
                # levels U/2 - mu -+ sqrt(eps_k^2 + U^2/4), each pair (k, k+Q) listed twice
                r = np.sqrt(eps ** 2 + U * U / 4)
                ev = np.sort(np.concatenate([U / 2 - mu - r, U / 2 - mu + r]))[::2]
            out[f"case{c}_meta"] = np.array([0, D, L, U, mu])
            out[f"case{c}_f"] = f
            out[f"case{c}_evals"] = ev
            c += 1
    out["ncases"] = np.array(c)
    np.savez_compressed(os.path.join(HERE, "closed_forms.npz"), **out)


STIFF_F5 = [0, 1, 0, 1, 1, 0, 1, 0, 0, 1, 0, 0, 1, 0, 1, 0, 1, 1, 1, 0, 1, 1, 0, 1, 1]
STIFF_F7 = [0, 1, 0, 0, 0, 0, 1, 1, 0, 1, 1, 0, 0, 1, 1, 0, 1, 1, 0, 1, 1, 0, 1, 0, 0, 0, 1, 1, 0, 1, 1, 1, 1, 1, 0,
            1, 1, 0, 1, 1, 1, 1, 0, 0, 0, 1, 0, 1, 1]


def stiffness_numpy(L, f, U, beta):
    """stiffness.hpp:129-187 restated in numpy: (T + V)/N along direction 0."""
    n = L * L
    H = hamiltonian(0, 2, L, f, U, U / 2)
    ev, V = np.linalg.eigh(H)
    with np.errstate(over="ignore"):
        fermi = 1.0 / (1.0 + np.exp(beta * ev))
    T = np.zeros((n, n))
    J = np.zeros((n, n))
    for i in range(n):
        x, y = divmod(i, L)
        left, right = ((x - 1) % L) * L + y, ((x + 1) % L) * L + y
        T[i, left] += -1.0
        T[i, right] += -1.0
        J[i, left] += -1.0
        J[i, right] += 1.0
    tk = np.einsum("ik,ij,jk->k", V, T, V)
    Tval = -np.pi * np.sum(tk * fermi)
    mJ = V.T @ J @ V
    Vval = 0.0
    for k in range(n):
        for l in range(k):
            de = ev[k] - ev[l]
            s = np.pi * (fermi[l] - fermi[k]) * mJ[l, k] * mJ[k, l]
            if abs(de) > 1e-12 and abs(s) > 1e-13:
                Vval += 2.0 * s / de
    return (Tval + Vval) / n


def stiffness_fixture():
    kats = [(0.0, STIFF_F5, 1.31597), (2.0, STIFF_F5, 0.287547), (6.0, STIFF_F5, 0.00260831),
            (0.37, STIFF_F7, 1.26046)]
    out = []
    for U, f, ref in kats:
        L = int(round(np.sqrt(len(f))))
        out.append(dict(L=L, U=U, mu=U / 2, beta=1000.0, f=f, reference_value=ref, reference_tol=1e-3,
                        lapack_value=float(stiffness_numpy(L, np.array(f), U, 1000.0))))
    json.dump(out, open(os.path.join(HERE, "stiffness.json"), "w"), indent=1)


def cheb_fixture():
    from scipy.special import jv
    out = dict(max_moment=12, grid=12, cases=[
        dict(fn="one", order=0, value=1.0, tol=1e-8),
        dict(fn="x^2", order=0, value=0.5, tol=1e-8),
        dict(fn="sin", order=1, value=float(jv(1, 1.0)), reference_value=0.440051, tol=1e-6),
        dict(fn="sin", order=5, value=float(jv(5, 1.0)), reference_value=0.000249758, tol=1e-9),
    ])
    json.dump(out, open(os.path.join(HERE, "cheb_quadrature.json"), "w"), indent=1)


def chain_fixture():
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    D, L, U, beta, seed, nmoves = 2, 8, 1.0, 1.0, 12345, 10000
    f0 = np.random.default_rng(2024).integers(0, 2, L ** D).astype(np.int32)
    tr = orc.run_chain(orc.KIND_HYPERCUBIC, D, L, beta, U, U / 2, U / 2, seed, f0, [(orc.MOVE_ADDREMOVE, 1.0)],
                       nmoves, tp=0.0)
    np.savez_compressed(os.path.join(HERE, "chain_c1.npz"), meta=np.array([D, L, U, beta, seed, nmoves]),
                        f0=f0.astype(np.int8), site=tr["a"].astype(np.uint8), acc=np.packbits(tr["acc"] > 0),
                        logw=tr["logw"], logz_cur=tr["logz_cur"][::100],
                        f_final=tr["f_final"].astype(np.int8), spectrum_final=tr["spectrum_final"])


if __name__ == "__main__":
    eigs_fixture()
    closed_forms_fixture()
    stiffness_fixture()
    cheb_fixture()
    chain_fixture()
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))
