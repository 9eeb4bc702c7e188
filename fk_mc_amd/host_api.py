"""ctypes binding of the C++ host layer (include/fkmc_host.h, lib/libfkmc_host.so).

The host layer keeps the reference's configuration_t / move_* / mc_metropolis /
measure_* classes and sends every diagonalisation to libfkgpu.so; there is no
CPU eigensolver behind it, so a call without a HIP device fails with FkgpuError.
"""
import ctypes as C
import os

import numpy as np

from .fkgpu import FkgpuError, lib as gpu_lib

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfkmc_host.so")

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)


class Model(C.Structure):
    """fkmc_model: lattice, couplings and (cheb_size > 0) Chebyshev weights of one chain."""
    _fields_ = [("lattice_kind", C.c_int), ("D", C.c_int), ("L", C.c_int), ("t", C.c_double),
                ("tp", C.c_double), ("beta", C.c_double), ("U", C.c_double), ("mu_c", C.c_double),
                ("mu_f", C.c_double), ("nW", C.c_int), ("W", _d), ("cheb_size", C.c_int),
                ("cheb_grid", C.c_int)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        gpu_lib()  # libfkgpu.so first: the host library links against it
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"host library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(Model)
        L.fkmc_host_last_error.restype = C.c_char_p
        L.fkmc_run_chain.argtypes = [P, C.c_uint32, _i32, C.c_int, _i32, _d, C.c_int, _i32, _i32, _i32, _i32,
                                     _d, _d, _d, _i32, _d]
        L.fkmc_run_mc.argtypes = [P, C.c_uint32, _i32, C.c_int, _i32, _d, C.c_int, C.c_int, C.c_int, _d, _d, _d,
                                  _d, _d, _d]
        L.fkmc_measure_config.argtypes = [P, _i32, _d, _d, _d, _d, C.c_int, _d, C.c_double, _d, _d, C.c_double, _d]
        L.fkmc_binning.argtypes = [C.c_int, _d, C.c_int, _d]
        L.fkmc_max_bin.argtypes = [C.c_int]
        L.fkmc_jackknife_cv.argtypes = [C.c_int, _d, _d, C.c_double, C.c_int, C.c_int, _d]
        _LIB = L
    return _LIB


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def model(kind, D, L, beta, U, mu_c, mu_f, t=1.0, tp=1.0, W=None, cheb_size=0, cheb_grid=0):
    m = Model(kind, D, L, t, tp, beta, U, mu_c, mu_f, 0, None, cheb_size, cheb_grid)
    if W is not None and len(W):
        w = np.ascontiguousarray(W, dtype=np.float64)
        m.nW, m.W = len(w), _p(w, _d)
        m._keep = w
    return m


def _check(r):
    if r < 0:
        raise FkgpuError(r, lib().fkmc_host_last_error().decode())
    return r


def run_chain(m, seed, f_init, moves, nmoves):
    """Plain per-move chain through configuration_t / move_* / mc_metropolis."""
    n = m.L ** m.D
    f0 = np.ascontiguousarray(f_init, dtype=np.int32)
    assert f0.shape == (n,)
    kinds = np.array([k for k, _ in moves], dtype=np.int32)
    probs = np.array([p for _, p in moves], dtype=np.float64)
    tr = {k: np.zeros(nmoves, dtype=np.int32) for k in ("kind", "a", "b", "acc")}
    tr.update({k: np.zeros(nmoves) for k in ("w", "u", "logz_cur")})
    ffin = np.zeros(n, dtype=np.int32)
    spec = np.zeros(n)
    nacc = _check(lib().fkmc_run_chain(C.byref(m), seed, _p(f0, _i32), len(kinds), _p(kinds, _i32),
                                       _p(probs, _d), nmoves, _p(tr["kind"], _i32), _p(tr["a"], _i32),
                                       _p(tr["b"], _i32), _p(tr["acc"], _i32), _p(tr["w"], _d), _p(tr["u"], _d),
                                       _p(tr["logz_cur"], _d), _p(ffin, _i32), _p(spec, _d)))
    tr.update(naccept=nacc, f_final=ffin, spectrum_final=spec)
    return tr


def run_mc(m, seed, moves, nsweeps, sweep_len, ntherm, f_init=None):
    n = m.L ** m.D
    kinds = np.array([k for k, _ in moves], dtype=np.int32)
    probs = np.array([p for _, p in moves], dtype=np.float64)
    f0 = None if f_init is None else np.ascontiguousarray(f_init, dtype=np.int32)
    out = {k: np.zeros(nsweeps) for k in ("energies", "d2energies", "c_energies")}
    out["avg_spectrum"] = np.zeros(n)
    out["focc"] = np.zeros((nsweeps, n))
    acc = np.zeros(1)
    _check(lib().fkmc_run_mc(C.byref(m), seed, _p(f0, _i32), len(kinds), _p(kinds, _i32), _p(probs, _d), nsweeps,
                             sweep_len, ntherm, _p(out["energies"], _d), _p(out["d2energies"], _d),
                             _p(out["c_energies"], _d), _p(out["avg_spectrum"], _d), _p(out["focc"], _d),
                             _p(acc, _d)))
    out["acceptance"] = float(acc[0])
    return out


def measure_config(m, f, wgrid=(0.0,), offset=0.05, eta=0.05, evecs=True, stiffness=True):
    n = m.L ** m.D
    fa = np.ascontiguousarray(f, dtype=np.int32)
    wg = np.ascontiguousarray(wgrid, dtype=np.float64)
    out = dict(energy=np.zeros(3), spectrum=np.zeros(n), ipr=np.zeros(n),
               evecs=np.zeros((n, n)) if evecs else None, stiffness=np.zeros(1) if stiffness else None,
               cond=np.zeros(len(wg)), dos=np.zeros(len(wg)))
    _check(lib().fkmc_measure_config(C.byref(m), _p(fa, _i32), _p(out["energy"], _d), _p(out["spectrum"], _d),
                                     _p(out["ipr"], _d), _p(out["evecs"], _d), len(wg), _p(wg, _d), offset,
                                     _p(out["stiffness"], _d), _p(out["cond"], _d), eta, _p(out["dos"], _d)))
    return out


def binning(x, depth):
    """Binning table [depth + 1, 5] (size, mean, variance, error, correlation length) of the
    series x in the given order, and the level estimate_bin picks."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    t = np.zeros((depth + 1, 5))
    b = _check(lib().fkmc_binning(len(x), _p(x, _d), int(depth), _p(t, _d)))
    return t, b


def max_bin(nmeasures):
    return _check(lib().fkmc_max_bin(int(nmeasures)))


def jackknife_cv(e, d2e, beta, volume, depth):
    """Jackknife table [depth + 1, 4] (size, value, variance, error) of the specific heat and
    the level estimate_bin picks."""
    e = np.ascontiguousarray(e, dtype=np.float64).ravel()
    d = np.ascontiguousarray(d2e, dtype=np.float64).ravel()
    t = np.zeros((depth + 1, 4))
    b = _check(lib().fkmc_jackknife_cv(len(e), _p(e, _d), _p(d, _d), float(beta), int(volume), int(depth), _p(t, _d)))
    return t, b
