"""ctypes binding of the fkgpu C ABI (include/fkgpu.h, fk_mc_amd/lib/libfkgpu.so).

There is no CPU fallback: importing this module without the built HIP library
raises, and every call reports a non-zero status as FkgpuError.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfkgpu.so")

OK, EINVAL, ENOLATTICE, EUNSUPPORTED, EHIP, ENOMEM = 0, -1, -2, -3, -4, -5
LATTICE_HYPERCUBIC, LATTICE_TRIANGULAR = 0, 1
MOVE_ADDREMOVE, MOVE_FLIP, MOVE_HOP, MOVE_RANDOMIZE = 0, 1, 2, 3
SOLVER_AUTO, SOLVER_VARIANTS = -1, 11
PRECISION_FP64, PRECISION_MIXED = 0, 1
UPDATE_DENSE, UPDATE_RANK1 = 0, 1
KPM_DELTA_LOCAL, KPM_DELTA_STOCHASTIC = 0, 1
MEASURE_SPECTRUM_HISTORY, MEASURE_FOCC_HISTORY, MEASURE_IPR, MEASURE_STIFFNESS = 1, 2, 4, 8
MEASURE_EIGENFUNCTIONS = 16
MEASURE_NO_SPECTRUM = 32

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_u32 = C.POINTER(C.c_uint32)
_i64 = C.POINTER(C.c_int64)
_i8 = C.POINTER(C.c_int8)


class FkgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fkgpu error {code}: {msg}")
        self.code = code


class _Trace(C.Structure):
    _fields_ = [("kind", _i32), ("site_a", _i32), ("site_b", _i32), ("u", _d), ("logw", _d),
                ("logz_new", _d), ("accepted", _i8)]


_LIB = None
_LIVE = []  # weak references to open handles, closed children-first at exit


def _close_all():
    for kind in ("chains", "ctx"):
        for ref in list(_LIVE):
            obj = ref()
            if obj is not None and obj._kind == kind:
                try:
                    obj.close()
                except Exception:
                    pass


atexit.register(_close_all)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.fkgpu_create.argtypes = [C.c_int, C.POINTER(P)]
        L.fkgpu_destroy.argtypes = [P]
        L.fkgpu_destroy.restype = None
        L.fkgpu_last_error.argtypes = [P]
        L.fkgpu_last_error.restype = C.c_char_p
        L.fkgpu_stream.argtypes = [P]
        L.fkgpu_stream.restype = P
        L.fkgpu_set_lattice.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
        L.fkgpu_set_lattice_csr.argtypes = [P, C.c_int, _i32, _i32, _d]
        L.fkgpu_msize.argtypes = [P]
        L.fkgpu_eig_logz.argtypes = [P, C.c_int, _i32, _d, _d, _d, C.c_int, _d, _d, _d, _d, _d]
        L.fkgpu_eig_logz_device.argtypes = [P, C.c_int, P, P, P, P, P, P]
        L.fkgpu_eig_logz_device_noalloc.argtypes = [P, C.c_int, P, P, P, P, P, P]
        L.fkgpu_chains_create.argtypes = [P, C.c_int, _i32, _d, _d, _d, _d, _u32, C.POINTER(P)]
        L.fkgpu_chains_create_multi.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.c_int, _i32, _i32, _d, _d, _d, _d,
                                                _u32, C.POINTER(P)]
        L.fkgpu_chains_destroy.argtypes = [P]
        L.fkgpu_chains_destroy.restype = None
        L.fkgpu_chains_set_moves.argtypes = [P, C.c_int, _i32, _d]
        L.fkgpu_chains_set_ff.argtypes = [P, C.c_int, _d]
        L.fkgpu_sweep.argtypes = [P, C.c_int, C.POINTER(_Trace)]
        L.fkgpu_chains_set_kpm.argtypes = [P, C.c_int, C.c_int, C.c_uint64, C.c_int]
        L.fkgpu_chains_set_kpm_delta.argtypes = [P, C.c_int, C.c_int, C.c_int]
        L.fkgpu_chains_kpm_delta.argtypes = [P, _i32, _i32, _i32]
        L.fkgpu_sweep_graph.argtypes = [P, C.c_int]
        L.fkgpu_chains_last_ms.argtypes = [P, _d]
        L.fkgpu_chains_state.argtypes = [P, _i32, _i32, _d, _d, _i64, _i64]
        L.fkgpu_rng_draws.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_int, _d, C.c_int, _d]
        L.fkgpu_measure_ipr.argtypes = [P, C.c_int, _i32, _d, _d, _d, _d, _d]
        L.fkgpu_measure_stiffness.argtypes = [P, C.c_int, _i32, _d, _d, _d, C.c_int, _d, C.c_double, _d, _d]
        L.fkgpu_dos.argtypes = [P, C.c_int, _d, C.c_int, _d, C.c_double, _d]
        L.fkgpu_ipr_w.argtypes = [P, C.c_int, _d, _d, C.c_int, _d, C.c_double, _d]
        L.fkgpu_adjoint.argtypes = [P, C.c_int, _i32, _d, _d, _d, _d, _d, _d, _d]
        L.fkgpu_cheb_logz.argtypes = [P, _i32, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _d, _d, _d]
        L.fkgpu_eig_logz_mixed.argtypes = [P, C.c_int, _i32, _d, _d, _d, _d, _d, _i64]
        L.fkgpu_chains_set_precision.argtypes = [P, C.c_int]
        L.fkgpu_chains_set_update.argtypes = [P, C.c_int, C.c_int]
        L.fkgpu_chains_fallbacks.argtypes = [P, _i64]
        L.fkgpu_tune_mu.argtypes = [P, C.c_int, _i32, _d, _d, _d, C.c_double, C.c_double, C.c_int, _i32]
        L.fkgpu_set_solver.argtypes = [P, C.c_int]
        L.fkgpu_get_solver.argtypes = [P]
        L.fkgpu_autotune.argtypes = [P, C.c_int, _d]
        L.fkgpu_kpm_logz.argtypes = [P, C.c_int, _i32, _d, _d, _d, C.c_int, C.c_int, C.c_uint64, C.c_int, _d, _d, _d]
        L.fkgpu_chains_set_steps_per_launch.argtypes = [P, C.c_int]
        L.fkgpu_chains_randomize_f.argtypes = [P, _i32]
        L.fkgpu_chains_set_queue_timeout.argtypes = [P, C.c_double]
        L.fkgpu_chains_measure_enable.argtypes = [P, C.c_int, C.c_int]
        L.fkgpu_chains_measure.argtypes = [P]
        L.fkgpu_sweep_measure.argtypes = [P, C.c_int, C.c_int]
        L.fkgpu_chains_energies.argtypes = [P, _i64, _d, _d, _d]
        L.fkgpu_chains_means.argtypes = [P, _i64, _d, _d, _d, _d]
        L.fkgpu_chains_histories.argtypes = [P, _i64, _d, _i32, _d]
        L.fkgpu_chains_measure_conductivity.argtypes = [P, C.c_int, _d, C.c_double]
        L.fkgpu_chains_stiffness.argtypes = [P, _i64, _d, _d, _d, _d]
        L.fkgpu_chains_eigenfunctions.argtypes = [P, _i64, _d]
        L.fkgpu_chains_green.argtypes = [P, C.c_double, C.c_double, _d, _d]
        L.fkgpu_chains_local_dos.argtypes = [P, C.c_int, _d, C.c_double, _d]
        _LIB = L
    return _LIB


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Context:
    """A device context holding one lattice (fkgpu_ctx)."""

    _kind = "ctx"

    def __init__(self, device=0):
        self._h = C.c_void_p()
        _LIVE.append(weakref.ref(self))
        r = lib().fkgpu_create(device, C.byref(self._h))
        if r != OK:
            raise FkgpuError(r, "fkgpu_create failed (no HIP device?)")
        self.N = 0

    def close(self):
        if self._h:
            lib().fkgpu_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, r):
        if r != OK:
            raise FkgpuError(r, lib().fkgpu_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        return lib().fkgpu_stream(self._h)

    def set_lattice(self, kind, ndim, L, t=1.0, tp=1.0):
        self._check(lib().fkgpu_set_lattice(self._h, kind, ndim, L, t, tp))
        self.N = lib().fkgpu_msize(self._h)
        return self

    def set_lattice_csr(self, hop_dense):
        h = np.asarray(hop_dense, dtype=np.float64)
        n = h.shape[0]
        rp = [0]
        cols, vals = [], []
        for i in range(n):
            nz = np.nonzero(h[i])[0]
            cols.extend(nz.tolist())
            vals.extend(h[i, nz].tolist())
            rp.append(len(cols))
        rp = np.array(rp, dtype=np.int32)
        cols = np.array(cols if cols else [0], dtype=np.int32)
        vals = np.array(vals if vals else [0.0], dtype=np.float64)
        self._check(lib().fkgpu_set_lattice_csr(self._h, n, _p(rp, _i32), _p(cols, _i32), _p(vals, _d)))
        self.N = lib().fkgpu_msize(self._h)
        return self

    def eig_logz(self, f, beta, U, mu_c, vectors=False, with_exp=False):
        """calc_hamiltonian + calc_ed for a batch of configurations f[B, N]."""
        f = np.ascontiguousarray(np.atleast_2d(f), dtype=np.int32)
        B, N = f.shape
        bet = np.ascontiguousarray(np.broadcast_to(np.asarray(beta, dtype=np.float64), (B,)))
        uu = np.ascontiguousarray(np.broadcast_to(np.asarray(U, dtype=np.float64), (B,)))
        mc = np.ascontiguousarray(np.broadcast_to(np.asarray(mu_c, dtype=np.float64), (B,)))
        ev = np.zeros((B, N))
        lz = np.zeros(B)
        V = np.zeros((B, N, N)) if vectors else None
        ce = np.zeros((B, N)) if with_exp else None
        cf = np.zeros((B, N)) if with_exp else None
        self._check(lib().fkgpu_eig_logz(self._h, B, _p(f, _i32), _p(bet, _d), _p(uu, _d), _p(mc, _d),
                                         int(vectors), _p(ev, _d), _p(V, _d), _p(lz, _d), _p(ce, _d), _p(cf, _d)))
        out = dict(evals=ev, logZ=lz)
        if vectors:
            out["evecs"] = V
        if with_exp:
            out["cexp"] = ce
            out["cfermi"] = cf
        return out


def _batch(f, *params):
    f = np.ascontiguousarray(np.atleast_2d(f), dtype=np.int32)
    B = f.shape[0]
    return (f, B) + tuple(np.ascontiguousarray(np.broadcast_to(np.asarray(p, dtype=np.float64), (B,)))
                          for p in params)


def _measure_ipr(self, f, beta, U, mu_c):
    """measure_ipr for a batch: returns (evals[B,N], ipr[B,N])."""
    f, B, b, u, m = _batch(f, beta, U, mu_c)
    ev = np.zeros((B, self.N))
    ipr = np.zeros((B, self.N))
    self._check(lib().fkgpu_measure_ipr(self._h, B, _p(f, _i32), _p(b, _d), _p(u, _d), _p(m, _d), _p(ev, _d),
                                        _p(ipr, _d)))
    return ev, ipr


def _measure_stiffness(self, f, beta, U, mu_c, wgrid=(0.0,), offset=0.05):
    f, B, b, u, m = _batch(f, beta, U, mu_c)
    wg = np.ascontiguousarray(wgrid, dtype=np.float64)
    st = np.zeros(B)
    cond = np.zeros((B, max(len(wg), 1)))
    self._check(lib().fkgpu_measure_stiffness(self._h, B, _p(f, _i32), _p(b, _d), _p(u, _d), _p(m, _d), len(wg),
                                              _p(wg, _d), offset, _p(st, _d), _p(cond, _d)))
    return st, cond[:, :len(wg)]


def _dos(self, spectra, wgrid, eta=0.05):
    sp = np.ascontiguousarray(np.atleast_2d(spectra), dtype=np.float64)
    wg = np.ascontiguousarray(wgrid, dtype=np.float64)
    out = np.zeros((sp.shape[0], len(wg)))
    self._check(lib().fkgpu_dos(self._h, sp.shape[0], _p(sp, _d), len(wg), _p(wg, _d), eta, _p(out, _d)))
    return out


def _ipr_w(self, spectra, ipr, wgrid, eta=0.05):
    sp = np.ascontiguousarray(np.atleast_2d(spectra), dtype=np.float64)
    ip = np.ascontiguousarray(np.atleast_2d(ipr), dtype=np.float64)
    if ip.shape != sp.shape:
        raise ValueError("spectra and ipr shapes differ")
    wg = np.ascontiguousarray(wgrid, dtype=np.float64)
    out = np.zeros((sp.shape[0], len(wg)))
    self._check(lib().fkgpu_ipr_w(self._h, sp.shape[0], _p(sp, _d), _p(ip, _d), len(wg), _p(wg, _d), eta,
                                  _p(out, _d)))
    return out


def _adjoint(self, f, beta, U, mu_c, full=False):
    f, B, b, u, m = _batch(f, beta, U, mu_c)
    lz = np.zeros(B)
    dens = np.zeros((B, self.N))
    gr = np.zeros((B, 8))
    rho = np.zeros((B, self.N, self.N)) if full else None
    self._check(lib().fkgpu_adjoint(self._h, B, _p(f, _i32), _p(b, _d), _p(u, _d), _p(m, _d), _p(lz, _d),
                                    _p(dens, _d), _p(gr, _d), _p(rho, _d)))
    out = dict(logZ=lz, dens=dens, dOmega_dU=gr[:, 0], dOmega_dt=gr[:, 1], dOmega_dtp=gr[:, 2],
               dOmega_dmu=gr[:, 3], Nc=gr[:, 4], dNc_dmu=gr[:, 5])
    if full:
        out["rho"] = rho
    return out


def _cheb_logz(self, f, beta, U, mu_c, cheb_size, grid_size):
    """calc_chebyshev (exact trace) of one configuration."""
    f = np.ascontiguousarray(np.ravel(f), dtype=np.int32)
    M = cheb_size + cheb_size % 2
    mom = np.zeros(M)
    lz = np.zeros(1)
    ab = np.zeros(4)
    self._check(lib().fkgpu_cheb_logz(self._h, _p(f, _i32), beta, U, mu_c, cheb_size, grid_size, _p(mom, _d),
                                      _p(lz, _d), _p(ab, _d)))
    return dict(moments=mom, logZ=float(lz[0]), a=ab[0], b=ab[1], e_min=ab[2], e_max=ab[3])


def _kpm_logz(self, f, beta, U, mu_c, R=64, M=256, seed=1, jackson=True):
    f, B, b, u, m = _batch(f, beta, U, mu_c)
    lz = np.zeros(B)
    mom = np.zeros((B, M))
    ab = np.zeros((B, 2))
    self._check(lib().fkgpu_kpm_logz(self._h, B, _p(f, _i32), _p(b, _d), _p(u, _d), _p(m, _d), R, M, seed,
                                     int(jackson), _p(lz, _d), _p(mom, _d), _p(ab, _d)))
    return dict(logZ=lz, moments=mom, a=ab[:, 0], b=ab[:, 1])


def _set_solver(self, variant):
    """Force a dense eigensolver variant (-1: by size)."""
    self._check(lib().fkgpu_set_solver(self._h, variant))
    return self


def _get_solver(self):
    return lib().fkgpu_get_solver(self._h)


def _autotune(self, batch=256):
    """Time every solver variant that fits N; keeps and returns the fastest and the times."""
    ms = np.zeros(SOLVER_VARIANTS)
    r = lib().fkgpu_autotune(self._h, batch, _p(ms, _d))
    if r < 0:
        self._check(r)
    return r, ms


def _eig_logz_mixed(self, f, beta, U, mu_c):
    """Mixed-precision calc_ed (eigenvalues): returns dict(evals, logZ, nfallback)."""
    f, B, b, u, m = _batch(f, beta, U, mu_c)
    ev = np.zeros((B, self.N))
    lz = np.zeros(B)
    nfb = np.zeros(1, np.int64)
    self._check(lib().fkgpu_eig_logz_mixed(self._h, B, _p(f, _i32), _p(b, _d), _p(u, _d), _p(m, _d), _p(ev, _d),
                                           _p(lz, _d), _p(nfb, _i64)))
    return dict(evals=ev, logZ=lz, nfallback=int(nfb[0]))


def _tune_mu(self, f, beta, U, mu0=0.0, filling=0.5, tol=1e-12, maxit=50):
    """Newton-tune mu_c to a conduction filling with the adjoint's dN_c/dmu_c."""
    f, B, b, u, m = _batch(f, beta, U, mu0)
    m = m.copy()
    its = np.zeros(B, np.int32)
    self._check(lib().fkgpu_tune_mu(self._h, B, _p(f, _i32), _p(b, _d), _p(u, _d), _p(m, _d), filling, tol, maxit,
                                    _p(its, _i32)))
    return m, its


Context.tune_mu = _tune_mu
Context.eig_logz_mixed = _eig_logz_mixed
Context.set_solver = _set_solver
Context.get_solver = _get_solver
Context.autotune = _autotune
Context.cheb_logz = _cheb_logz
Context.kpm_logz = _kpm_logz
Context.measure_ipr = _measure_ipr
Context.measure_stiffness = _measure_stiffness
Context.dos = _dos
Context.ipr_w = _ipr_w
Context.adjoint = _adjoint


class Chains:
    """Batched persistent Metropolis chains (fkgpu_chains)."""

    _kind = "chains"

    def __init__(self, ctx, f_init, beta, U, mu_c, mu_f, seeds, lattice=None):
        self.ctx = ctx
        self._h = C.c_void_p()
        _LIVE.append(weakref.ref(self))
        f = np.ascontiguousarray(np.atleast_2d(f_init), dtype=np.int32)
        self.B, self.N = f.shape
        self.steps_per_launch = 16  # the library's default
        self._mcap, self._mflags = 0, 0  # measurement records (measure_enable)
        B = self.B

        def vec(x, dt=np.float64):
            return np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=dt), (B,)))

        self._keep = [vec(beta), vec(U), vec(mu_c), vec(mu_f), vec(seeds, np.uint32)]
        self._h = C.c_void_p()
        if lattice is None:
            ctx._check(lib().fkgpu_chains_create(ctx.handle, B, _p(f, _i32), _p(self._keep[0], _d),
                                                 _p(self._keep[1], _d), _p(self._keep[2], _d), _p(self._keep[3], _d),
                                                 _p(self._keep[4], _u32), C.byref(self._h)))
        else:
            # chains on several lattices: ctx is the list of contexts, lattice[c] indexes it
            self.contexts = list(ctx)
            self.ctx = self.contexts[0]
            arr = (C.c_void_p * len(self.contexts))(*[C.c_void_p(c.handle.value) for c in self.contexts])
            lat = np.ascontiguousarray(lattice, dtype=np.int32)
            self._keep += [arr, lat]
            self.ctx._check(lib().fkgpu_chains_create_multi(len(self.contexts), arr, B, _p(lat, _i32), _p(f, _i32),
                                                            _p(self._keep[0], _d), _p(self._keep[1], _d),
                                                            _p(self._keep[2], _d), _p(self._keep[3], _d),
                                                            _p(self._keep[4], _u32), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().fkgpu_chains_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_moves(self, moves):
        kinds = np.array([m[0] for m in moves], dtype=np.int32)
        w = np.array([m[1] for m in moves], dtype=np.float64)
        self.ctx._check(lib().fkgpu_chains_set_moves(self._h, len(kinds), _p(kinds, _i32), _p(w, _d)))

    def set_kpm(self, R=64, M=256, seed=1, jackson=True):
        self.ctx._check(lib().fkgpu_chains_set_kpm(self._h, R, M, seed, int(jackson)))

    def set_kpm_delta(self, mode, moments=0, frequencies=0):
        self.ctx._check(lib().fkgpu_chains_set_kpm_delta(self._h, mode, moments, frequencies))

    def kpm_delta(self):
        v = np.zeros(3, np.int32)
        self.ctx._check(lib().fkgpu_chains_kpm_delta(self._h, _p(v[0:], _i32), _p(v[1:], _i32), _p(v[2:], _i32)))
        return dict(mode=int(v[0]), moments=int(v[1]), frequencies=int(v[2]))

    def set_ff(self, W):
        W = np.ascontiguousarray(W, dtype=np.float64)
        self.ctx._check(lib().fkgpu_chains_set_ff(self._h, len(W), _p(W, _d) if len(W) else None))

    def sweep(self, nsteps, trace=False):
        if not trace:
            self.ctx._check(lib().fkgpu_sweep(self._h, nsteps, None))
            return None
        n = nsteps * self.B
        tr = dict(kind=np.zeros(n, np.int32), a=np.zeros(n, np.int32), b=np.zeros(n, np.int32),
                  u=np.zeros(n), logw=np.zeros(n), logz=np.zeros(n), acc=np.zeros(n, np.int8))
        t = _Trace(_p(tr["kind"], _i32), _p(tr["a"], _i32), _p(tr["b"], _i32), _p(tr["u"], _d),
                   _p(tr["logw"], _d), _p(tr["logz"], _d), _p(tr["acc"], _i8))
        self.ctx._check(lib().fkgpu_sweep(self._h, nsteps, C.byref(t)))
        return {k: v.reshape(nsteps, self.B) for k, v in tr.items()}

    def randomize_f(self, nf_start):
        nf = np.ascontiguousarray(np.broadcast_to(np.asarray(nf_start, dtype=np.int32), (self.B,)))
        self.ctx._check(lib().fkgpu_chains_randomize_f(self._h, _p(nf, _i32)))

    def set_queue_timeout(self, seconds):
        self.ctx._check(lib().fkgpu_chains_set_queue_timeout(self._h, seconds))

    def set_steps_per_launch(self, steps):
        self.ctx._check(lib().fkgpu_chains_set_steps_per_launch(self._h, steps))
        self.steps_per_launch = steps

    def set_precision(self, precision):
        self.ctx._check(lib().fkgpu_chains_set_precision(self._h, precision))

    def set_update(self, mode, refresh=256):
        self.ctx._check(lib().fkgpu_chains_set_update(self._h, mode, refresh))

    def fallbacks(self):
        n = np.zeros(1, np.int64)
        self.ctx._check(lib().fkgpu_chains_fallbacks(self._h, _p(n, _i64)))
        return int(n[0])

    def sweep_graph(self, nsteps):
        self.ctx._check(lib().fkgpu_sweep_graph(self._h, nsteps))

    def last_ms(self):
        ms = C.c_double()
        self.ctx._check(lib().fkgpu_chains_last_ms(self._h, C.byref(ms)))
        return ms.value

    def state(self):
        B, N = self.B, self.N
        f = np.zeros((B, N), np.int32)
        nf = np.zeros(B, np.int32)
        lz = np.zeros(B)
        sp = np.zeros((B, N))
        na = np.zeros(B, np.int64)
        ns = np.zeros(B, np.int64)
        self.ctx._check(lib().fkgpu_chains_state(self._h, _p(f, _i32), _p(nf, _i32), _p(lz, _d), _p(sp, _d),
                                                 _p(na, _i64), _p(ns, _i64)))
        return dict(f=f, nf=nf, logZ=lz, spectrum=sp, naccept=na, nsolve=ns)

    # ---- on-device measurements (fkgpu_chains_measure_enable and friends) ----
    def measure_enable(self, capacity, flags=0):
        self.ctx._check(lib().fkgpu_chains_measure_enable(self._h, capacity, flags))
        self._mcap, self._mflags = capacity, flags

    def measure_conductivity(self, wgrid, offset=0.05):
        w = np.ascontiguousarray(wgrid, dtype=np.float64)
        self.ctx._check(lib().fkgpu_chains_measure_conductivity(self._h, len(w), _p(w, _d), offset))
        self._nw = len(w)

    def measure(self):
        self.ctx._check(lib().fkgpu_chains_measure(self._h))

    def sweep_measure(self, nsweeps, sweep_len):
        self.ctx._check(lib().fkgpu_sweep_measure(self._h, nsweeps, sweep_len))

    def _mcount(self):
        c = np.zeros(1, np.int64)
        self.ctx._check(lib().fkgpu_chains_energies(self._h, _p(c, _i64), None, None, None))
        return int(c[0])

    def energies(self):
        """(count, energies, d2energies, c_energies), each series [B, min(count, capacity)]."""
        n = min(self._mcount(), self._mcap)
        e, d2, ce = (np.zeros((self.B, n)) for _ in range(3))
        c = np.zeros(1, np.int64)
        self.ctx._check(lib().fkgpu_chains_energies(self._h, _p(c, _i64), _p(e, _d), _p(d2, _d), _p(ce, _d)))
        return int(c[0]), e, d2, ce

    def means(self):
        B, N = self.B, self.N
        c = np.zeros(1, np.int64)
        e, d2 = np.zeros(B), np.zeros(B)
        sp, fo = np.zeros((B, N)), np.zeros((B, N))
        self.ctx._check(lib().fkgpu_chains_means(self._h, _p(c, _i64), _p(e, _d), _p(d2, _d), _p(sp, _d), _p(fo, _d)))
        return dict(count=int(c[0]), energy=e, d2energy=d2, spectrum=sp, focc=fo)

    def stiffness(self):
        """(count, stiffness [B, n], cond [B, n, nw], stiffness_mean [B], cond_mean [B, nw])."""
        B, nw = self.B, getattr(self, "_nw", 1)
        n = min(self._mcount(), self._mcap)
        st, cd = np.zeros((B, n)), np.zeros((B, n, nw))
        sm, cm = np.zeros(B), np.zeros((B, nw))
        c = np.zeros(1, np.int64)
        self.ctx._check(lib().fkgpu_chains_stiffness(self._h, _p(c, _i64), _p(st, _d), _p(cd, _d), _p(sm, _d),
                                                     _p(cm, _d)))
        return int(c[0]), st, cd, sm, cm

    def eigenfunctions(self):
        """(count, evecs [B, n, N, N])."""
        n = min(self._mcount(), self._mcap)
        V = np.zeros((self.B, n, self.N, self.N))
        c = np.zeros(1, np.int64)
        self.ctx._check(lib().fkgpu_chains_eigenfunctions(self._h, _p(c, _i64), _p(V, _d)))
        return int(c[0]), V

    def green(self, w, offset):
        """G(w + i offset) averaged over every record of every chain: (re [N, N], im [N, N])."""
        re = np.zeros((self.N, self.N))
        im = np.zeros((self.N, self.N))
        self.ctx._check(lib().fkgpu_chains_green(self._h, float(w), float(offset), _p(re, _d), _p(im, _d)))
        return re, im

    def local_dos(self, wgrid, offset):
        """Local densities of states [len(wgrid), N] averaged over every record of every chain."""
        w = np.ascontiguousarray(wgrid, dtype=np.float64).ravel()
        out = np.zeros((len(w), self.N))
        self.ctx._check(lib().fkgpu_chains_local_dos(self._h, len(w), _p(w, _d), float(offset), _p(out, _d)))
        return out

    def histories(self):
        """Enabled histories, each [B, min(count, capacity), N] (None where not enabled)."""
        B, N = self.B, self.N
        n = min(self._mcount(), self._mcap)
        sp = np.zeros((B, n, N)) if self._mflags & MEASURE_SPECTRUM_HISTORY else None
        fo = np.zeros((B, n, N), np.int32) if self._mflags & MEASURE_FOCC_HISTORY else None
        ip = np.zeros((B, n, N)) if self._mflags & MEASURE_IPR else None
        c = np.zeros(1, np.int64)
        self.ctx._check(lib().fkgpu_chains_histories(self._h, _p(c, _i64), _p(sp, _d), _p(fo, _i32), _p(ip, _d)))
        return dict(count=int(c[0]), spectrum=sp, focc=fo, ipr=ip)


def rng_draws(seed, mode, count, hi=0, probs=None):
    """Draws from the library's port of the reference's generator (host side)."""
    out = np.zeros(count)
    p = None if probs is None else np.ascontiguousarray(probs, dtype=np.float64)
    r = lib().fkgpu_rng_draws(seed, mode, hi, 0 if p is None else len(p), _p(p, _d), count, _p(out, _d))
    if r != OK:
        raise FkgpuError(r, "fkgpu_rng_draws failed")
    return out
