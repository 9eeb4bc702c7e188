"""ctypes wrapper of the CPU oracle (oracle/libfkoracle.so).

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import it; the product
package fk_mc_amd never does.

Each function restates one piece of the reference path; the C++ source
(oracle/fkmc_oracle.cpp) cites the reference file:line it follows.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KIND_HYPERCUBIC = 0
KIND_TRIANGULAR = 1
MOVE_ADDREMOVE, MOVE_FLIP, MOVE_HOP, MOVE_RANDOMIZE = 0, 1, 2, 3

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)


def build():
    """Compile the oracle with its Makefile (g++ only)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libfkoracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_lattice_hop.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _d, _i, _i]
        L.orc_eigh.argtypes = [C.c_int, _d, _d, _d]
        L.orc_tridiagonalize.argtypes = [C.c_int, _d, _d, _d]
        L.orc_logz_reference.argtypes = [C.c_int, _d, C.c_double, _d, _d]
        L.orc_logz_reference.restype = C.c_double
        L.orc_logz_stable.argtypes = [C.c_int, _d, C.c_double]
        L.orc_logz_stable.restype = C.c_double
        L.orc_calc_ed.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _i, C.c_double, C.c_double,
                                  C.c_double, C.c_int, _d, _d, _d, _d, _d]
        L.orc_randomize_f.argtypes = [C.c_int, C.c_uint32, C.c_int, _i]
        L.orc_ff_energy.argtypes = [C.c_int, C.c_int, _i, C.c_int, _d]
        L.orc_ff_energy.restype = C.c_double
        L.orc_run_chain.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint32,
                                    _i, C.c_int, _i, _d, C.c_int, _i, C.c_int,
                                    _i, _i, _i, _d, _d, _d, _i, _i, _d, _i, _d]
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_set_mutation.restype = None
        L.orc_set_flip_ff.argtypes = [C.c_int]
        L.orc_set_flip_ff.restype = None
        L.orc_set_chain_ff.argtypes = [C.c_int, _d]
        L.orc_set_chain_ff.restype = None
        L.orc_rng_draws.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_int, _d, C.c_int, _d]
        L.orc_measure_energy.argtypes = [C.c_int, _d, _d, C.c_double, C.c_int, C.c_double, _d, _d, _d]
        L.orc_measure_ipr.argtypes = [C.c_int, _d, _d]
        L.orc_measure_stiffness.argtypes = [C.c_int, C.c_int, C.c_int, _d, _d, _d, C.c_int, _d, C.c_double,
                                            _d, _d, _d, _d]
        L.orc_cheb_moment_f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_cheb_moment_f.restype = C.c_double
        L.orc_cheb_logz_coeffs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _d]
        L.orc_calc_chebyshev.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _i, C.c_double,
                                         C.c_double, C.c_double, C.c_int, C.c_int, _d, _d, _d]
        L.orc_bench_chains.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_uint32, C.c_int, C.c_int]
        L.orc_bench_chains.restype = C.c_long
        L.orc_bench_chain_set.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _d, _d, _d,
                                          _i, C.POINTER(C.c_uint32), C.c_int, _i, _d, C.c_int,
                                          C.POINTER(C.c_long)]
        L.orc_bench_chain_set.restype = C.c_long
        _LIB = L
    return _LIB


def _pd(a):
    return a.ctypes.data_as(_d) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(_i) if a is not None else None


def lattice_hop(kind, D, L, t=1.0, tp=1.0):
    """Dense hopping matrix and neighbour lists (src/lattice/hypercubic.cpp:115-155)."""
    n = L ** D
    hop = np.zeros((n, n), dtype=np.float64)
    zmax = 2 * D + (2 if kind == KIND_TRIANGULAR else 0)
    nbr = np.zeros((n, zmax), dtype=np.int32)
    z = np.zeros(n, dtype=np.int32)
    r = lib().orc_lattice_hop(kind, D, L, t, tp, _pd(hop), _pi(nbr), _pi(z))
    if r < 0:
        raise ValueError("invalid lattice")
    return hop, nbr, z


def eigh(H, vectors=False):
    H = np.ascontiguousarray(H, dtype=np.float64)
    n = H.shape[0]
    ev = np.zeros(n)
    V = np.zeros((n, n)) if vectors else None
    if lib().orc_eigh(n, _pd(H), _pd(ev), _pd(V)) != 0:
        raise RuntimeError("oracle eigensolve failed")
    return (ev, V) if vectors else ev


def tridiagonalize(H):
    H = np.ascontiguousarray(H, dtype=np.float64)
    n = H.shape[0]
    d = np.zeros(n)
    e = np.zeros(max(n - 1, 1))
    lib().orc_tridiagonalize(n, _pd(H), _pd(d), _pd(e))
    return d, e[: n - 1]


def logz_reference(ev, beta):
    ev = np.ascontiguousarray(ev, dtype=np.float64)
    ce = np.zeros_like(ev)
    cf = np.zeros_like(ev)
    lz = lib().orc_logz_reference(len(ev), _pd(ev), beta, _pd(ce), _pd(cf))
    return lz, ce, cf


def logz_stable(ev, beta):
    ev = np.ascontiguousarray(ev, dtype=np.float64)
    return lib().orc_logz_stable(len(ev), _pd(ev), beta)


def calc_ed(kind, D, L, f, beta, U, mu_c, t=1.0, tp=1.0, vectors=False):
    """calc_hamiltonian + calc_ed of one configuration (src/configuration.cpp:79-91,208-246)."""
    n = L ** D
    f = np.ascontiguousarray(f, dtype=np.int32)
    ev = np.zeros(n)
    V = np.zeros((n, n)) if vectors else None
    lz = np.zeros(1)
    ce = np.zeros(n)
    cf = np.zeros(n)
    r = lib().orc_calc_ed(kind, D, L, t, tp, _pi(f), beta, U, mu_c, int(vectors), _pd(ev), _pd(V), _pd(lz),
                          _pd(ce), _pd(cf))
    if r != 0:
        raise RuntimeError("oracle calc_ed failed")
    return dict(evals=ev, evecs=V, logZ=float(lz[0]), cexp=ce, cfermi=cf)


def randomize_f(n, seed, nf=0):
    f = np.zeros(n, dtype=np.int32)
    lib().orc_randomize_f(n, seed, nf, _pi(f))
    return f


def ff_energy(D, f, W):
    f = np.ascontiguousarray(f, dtype=np.int32)
    W = np.ascontiguousarray(W, dtype=np.float64)
    return lib().orc_ff_energy(D, len(f), _pi(f), len(W), _pd(W))


def run_chain(kind, D, L, beta, U, mu_c, mu_f, seed, f_init, moves, nmoves, t=1.0, tp=1.0, forced=None,
              stable_logz=False):
    """Metropolis chain with the reference's draw order (src/mc_metropolis.cpp:34-52).

    moves: list of (kind, probability) in add_move order.
    Returns a dict of per-move trace arrays plus the final state.
    """
    n = L ** D
    kinds = np.array([m[0] for m in moves], dtype=np.int32)
    probs = np.array([m[1] for m in moves], dtype=np.float64)
    f0 = np.ascontiguousarray(f_init, dtype=np.int32)
    tr = {k: np.zeros(nmoves, dtype=np.int32) for k in ("kind", "a", "b", "acc", "nf")}
    tr.update({k: np.zeros(nmoves) for k in ("u", "logw", "logz", "logz_cur")})
    ffin = np.zeros(n, dtype=np.int32)
    spec = np.zeros(n)
    forced_arr = None if forced is None else np.ascontiguousarray(forced, dtype=np.int32)
    nacc = lib().orc_run_chain(kind, D, L, t, tp, beta, U, mu_c, mu_f, seed, _pi(f0), len(kinds), _pi(kinds),
                               _pd(probs), nmoves, _pi(forced_arr), int(stable_logz),
                               _pi(tr["kind"]), _pi(tr["a"]), _pi(tr["b"]), _pd(tr["u"]), _pd(tr["logw"]),
                               _pd(tr["logz"]), _pi(tr["acc"]), _pi(tr["nf"]), _pd(tr["logz_cur"]),
                               _pi(ffin), _pd(spec))
    if nacc < 0:
        raise RuntimeError("oracle chain failed")
    tr["naccept"] = nacc
    tr["f_final"] = ffin
    tr["spectrum_final"] = spec
    return tr


def set_mutation(m):
    """Plant (m > 0) or remove (m = 0) a deliberate mistake in the nearest-neighbour
    exchange, for tests that show their checks catch it (see g_mutation in
    fkmc_oracle.cpp)."""
    lib().orc_set_mutation(int(m))


def set_flip_ff(on):
    """Global exchange with (1) or without (0, default) the f-f term in its weight: the
    Chebyshev move_flip (src/moves_chebyshev.cpp:22) or the ED one (src/moves.cpp:19-20)."""
    lib().orc_set_flip_ff(int(on))


def set_chain_ff(W):
    """f-f interaction W (1-D lattices) of the chains run_chain builds from now on."""
    W = np.ascontiguousarray(W if W is not None else [], dtype=np.float64)
    lib().orc_set_chain_ff(len(W), _pd(W) if len(W) else None)


def rng_draws(seed, mode, count, hi=0, probs=None):
    out = np.zeros(count)
    p = None if probs is None else np.ascontiguousarray(probs, dtype=np.float64)
    lib().orc_rng_draws(seed, mode, hi, 0 if p is None else len(p), _pd(p), count, _pd(out))
    return out


def measure_energy(spectrum, cexp, mu_f, nf, e_ff=0.0):
    s = np.ascontiguousarray(spectrum, dtype=np.float64)
    ce = np.ascontiguousarray(cexp, dtype=np.float64)
    o = np.zeros(3)
    lib().orc_measure_energy(len(s), _pd(s), _pd(ce), mu_f, nf, e_ff, _pd(o[0:]), _pd(o[1:]), _pd(o[2:]))
    return dict(energy=o[0], d2energy=o[1], c_energy=o[2])


def measure_ipr(V):
    V = np.ascontiguousarray(V, dtype=np.float64)
    out = np.zeros(V.shape[0])
    lib().orc_measure_ipr(V.shape[0], _pd(V), _pd(out))
    return out


def measure_stiffness(D, L, V, evals, fermi, wgrid=(0.0,), offset=0.05):
    V = np.ascontiguousarray(V, dtype=np.float64)
    ev = np.ascontiguousarray(evals, dtype=np.float64)
    fe = np.ascontiguousarray(fermi, dtype=np.float64)
    wg = np.ascontiguousarray(wgrid, dtype=np.float64)
    cond = np.zeros(len(wg))
    out = np.zeros(3)
    lib().orc_measure_stiffness(D, L, V.shape[0], _pd(V), _pd(ev), _pd(fe), len(wg), _pd(wg), offset,
                                _pd(out[0:]), _pd(cond), _pd(out[1:]), _pd(out[2:]))
    return dict(stiffness=out[0], T=out[1], V=out[2], cond=cond)


def cheb_moment_f(max_moment, grid, fn, order):
    return lib().orc_cheb_moment_f(max_moment, grid, fn, order)


def cheb_logz_coeffs(max_moment, grid, n, beta, a, b):
    g = np.zeros(max_moment + 2)
    M = lib().orc_cheb_logz_coeffs(max_moment, grid, n, beta, a, b, _pd(g))
    return g[:M]


def calc_chebyshev(kind, D, L, f, beta, U, mu_c, max_moment, grid, t=1.0, tp=1.0):
    f = np.ascontiguousarray(f, dtype=np.int32)
    mom = np.zeros(max_moment + 2)
    lz = np.zeros(1)
    ab = np.zeros(4)
    M = lib().orc_calc_chebyshev(kind, D, L, t, tp, _pi(f), beta, U, mu_c, max_moment, grid, _pd(mom), _pd(lz),
                                 _pd(ab))
    if M < 0:
        raise RuntimeError("oracle calc_chebyshev failed")
    return dict(moments=mom[:M], logZ=float(lz[0]), a=ab[0], b=ab[1], e_min=ab[2], e_max=ab[3])


def bench_chains(kind, D, L, beta, U, mu, seed, nthreads, nmoves, t=1.0, tp=1.0):
    return lib().orc_bench_chains(kind, D, L, t, tp, beta, U, mu, seed, nthreads, nmoves)


def bench_chain_set(kind, D, L, beta, U, mu, f0, seeds, moves, nmoves, t=1.0, tp=1.0):
    """One chain per thread (chain k: beta[k], U[k], mu[k], f0[k], seeds[k]), nmoves moves
    each; returns (total moves, eigensolves)."""
    n = len(beta)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    u = np.ascontiguousarray(U, dtype=np.float64)
    m = np.ascontiguousarray(mu, dtype=np.float64)
    f = np.ascontiguousarray(f0, dtype=np.int32)
    sd = np.ascontiguousarray(seeds, dtype=np.uint32)
    kinds = np.array([x[0] for x in moves], dtype=np.int32)
    probs = np.array([x[1] for x in moves], dtype=np.float64)
    ns = C.c_long(0)
    tot = lib().orc_bench_chain_set(kind, D, L, t, tp, n, _pd(b), _pd(u), _pd(m), _pi(f),
                                    sd.ctypes.data_as(C.POINTER(C.c_uint32)), len(kinds), _pi(kinds), _pd(probs),
                                    nmoves, C.byref(ns))
    if tot < 0:
        raise RuntimeError("oracle bench failed")
    return int(tot), int(ns.value)
