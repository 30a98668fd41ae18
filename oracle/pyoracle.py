"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker / the timed CPU
baseline.  The product package ``dnsmc_amd`` never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
MAXLEV = 6

_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int)
_u32p = C.POINTER(C.c_uint32)


def build(force=False):
    """Compile oracle/_build/liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class OrcProblem(C.Structure):
    _fields_ = [
        ("nlev", C.c_int),
        ("n", (C.c_int * 3) * MAXLEV),
        ("face", (_f32p * 3) * MAXLEV),
        ("centre", _f32p * 3),
        ("T", _f32p * MAXLEV),
        ("maxs", C.c_int * MAXLEV),
        ("bc", (C.c_int * 2) * 3),
        ("polar_axis", C.c_int),
        ("tau_cut", C.c_float),
        ("spectral", C.c_int),
        ("nB", C.c_int), ("nQ", C.c_int), ("nT", C.c_int),
        ("inv_dT", C.c_float), ("mT0", C.c_float),
        ("kq", _f32p), ("Ib", _f32p), ("Ibw", _f32p),
        ("kappa_grey", C.c_float),
        ("cdf_band", _f32p), ("cdf_g", _f32p), ("inv_emax", _f32p),
        ("Ibmax", C.c_float),
        ("Ibmax_d", C.c_double),
        ("Tmax", C.c_float),
        ("g_search", C.c_int),
        ("u_off", C.c_float),
    ]


class OrcRayOut(C.Structure):
    _fields_ = [("steps", C.c_int), ("end_type", C.c_int), ("end_cell", C.c_int * 3),
                ("end_level", C.c_int), ("hash", C.c_uint32), ("D", C.c_float), ("G", C.c_float)]


class OrcXorwow(C.Structure):
    _fields_ = [("x", C.c_uint32 * 5), ("d", C.c_uint32)]


RAY_DESC = np.dtype([("w0", "<u4"), ("w1", "<u4"), ("w", "<f4"), ("winv", "<f4")])


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    def __init__(self, path=None):
        self.lib = C.CDLL(path or build())
        L = self.lib
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_xorwow_next.restype = C.c_uint32
        L.orc_xorwow_next.argtypes = [C.POINTER(OrcXorwow)]
        L.orc_xorwow_seed.argtypes = [C.POINTER(OrcXorwow), C.c_uint64]
        L.orc_xorwow_skip_brute.argtypes = [C.POINTER(OrcXorwow), C.c_uint64]
        L.orc_xorwow_init.argtypes = [C.POINTER(OrcXorwow), C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_u01_incl.restype = C.c_float
        L.orc_u01_incl.argtypes = [C.c_uint32]
        L.orc_u01_open.restype = C.c_float
        L.orc_u01_open.argtypes = [C.c_uint32]
        L.orc_sobol_dirs.argtypes = [C.c_int, _u32p]
        L.orc_sobol_point.restype = C.c_uint32
        L.orc_sobol_point.argtypes = [_u32p, C.c_uint64]
        L.orc_sincos2pi.argtypes = [C.c_float, _f32p, _f32p]
        L.orc_cdf_search.restype = C.c_int
        L.orc_cdf_search.argtypes = [_f32p, C.c_int, C.c_float]
        L.orc_grid_ref_quad.argtypes = [C.c_int, C.c_int, C.c_float, _f32p, _f32p]
        L.orc_grid_uniform.argtypes = [C.c_int, C.c_float, _f32p, _f32p]
        L.orc_grid_tanh.argtypes = [C.c_int, C.c_float, C.c_double, _f32p, _f32p]
        L.orc_plant_mistake.restype = None
        L.orc_plant_mistake.argtypes = [C.c_int]
        L.orc_planck_band.restype = C.c_float
        L.orc_planck_band.argtypes = [C.c_float, C.c_float]
        L.orc_interp3d.argtypes = [_i32p, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p, _f32p, _f32p]
        L.orc_expint_En.restype = C.c_double
        L.orc_expint_En.argtypes = [C.c_int, C.c_double]
        L.orc_slab_grey_Q.argtypes = [C.c_int, _f64p, _f64p, C.c_double, C.c_double, C.c_double, _f64p]
        L.orc_slab_exchange.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_double, C.c_double, _f64p]
        L.orc_slab_exchange_cut.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                            C.c_double, C.c_int, _f64p]
        L.orc_trace_ray_f32.argtypes = [C.POINTER(OrcProblem), C.c_int, C.c_int, C.c_int, C.c_float,
                                        C.c_float, C.c_int, C.c_int, C.POINTER(OrcRayOut)]
        L.orc_gen_philox.argtypes = [C.POINTER(OrcProblem), C.c_uint64, C.c_uint32, C.c_int, C.c_int,
                                     C.c_int, C.c_void_p]
        L.orc_gen_compat.argtypes = [C.POINTER(OrcProblem), C.c_uint64, C.c_uint64, C.c_int, C.c_uint64,
                                     C.c_int, C.c_int, C.c_void_p]
        L.orc_gen_compat_sorted.argtypes = [C.POINTER(OrcProblem), C.c_uint64, C.c_uint64, C.c_int, C.c_uint64,
                                            C.c_int, C.c_int, _i32p, C.c_void_p]
        L.orc_sort_idx.argtypes = [_f32p, C.c_int, _i32p]
        L.orc_trace_descs.argtypes = [C.POINTER(OrcProblem), C.c_int, C.c_void_p, C.c_int, _f64p, _f64p,
                                      _f64p, C.POINTER(C.c_long)]
        for fn in (L.orc_mc_philox, L.orc_mc_philox_f64):
            fn.argtypes = [C.POINTER(OrcProblem), C.c_uint64, C.c_uint32, C.c_int, _i32p, C.c_int,
                           _f64p, _f64p, _f64p, C.POINTER(C.c_long), C.c_int]
        L.orc_lerp_f32.restype = C.c_float
        L.orc_lerp_f32.argtypes = [_f32p, C.c_int, C.c_float, C.c_float, C.c_float]
        L.orc_build_proposal.argtypes = ([C.c_int] * 4 + [_f32p] * 8 + [_f64p, C.c_float, C.c_int]
                                         + [_f32p] * 5 + [_f64p, _f32p])
        L.orc_mlp_forward.argtypes = [C.c_int] + [_f64p] * 7 + [C.c_int] * 3 + [_f64p] * 3
        L.orc_mlp_forward_bf16.argtypes = L.orc_mlp_forward.argtypes
        L.orc_bf16_round.argtypes = [C.c_double]
        L.orc_bf16_round.restype = C.c_double
        L.orc_mlp_loss_grad.argtypes = ([C.c_int] + [_f64p] * 7 + [C.c_int] * 4 + [_i32p, _i32p, _f64p]
                                        + [_f64p] * 7)
        L.orc_mlp_loss_grad_bf16.argtypes = L.orc_mlp_loss_grad.argtypes
        L.orc_learned_draw.argtypes = ([C.c_int] * 6 + [_f32p] * 3 + [_i32p, C.c_uint32, C.c_uint64, C.c_float]
                                       + [_u32p, _u32p, _f32p])

    # ---------------------------------------------------------------- RNG --
    def philox(self, ctr, key):
        c = np.asarray(ctr, dtype=np.uint32)
        k = np.asarray(key, dtype=np.uint32)
        o = np.zeros(4, np.uint32)
        self.lib.orc_philox4x32_10(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(o, _u32p))
        return o

    def xorwow_state(self, seed, subseq=0, offset=0):
        s = OrcXorwow()
        self.lib.orc_xorwow_init(C.byref(s), seed, subseq, offset)
        return s

    def xorwow_draws(self, s, n):
        return np.array([self.lib.orc_xorwow_next(C.byref(s)) for _ in range(n)], np.uint32)

    def sobol_dirs(self, dim):
        v = np.zeros(32, np.uint32)
        self.lib.orc_sobol_dirs(dim, _ptr(v, _u32p))
        return v

    def sobol_points(self, dim, n0, n):
        v = self.sobol_dirs(dim)
        return np.array([self.lib.orc_sobol_point(_ptr(v, _u32p), n0 + i) for i in range(n)], np.uint32)

    def sincos2pi(self, u):
        s = np.zeros(1, np.float32)
        c = np.zeros(1, np.float32)
        self.lib.orc_sincos2pi(C.c_float(u), _ptr(s, _f32p), _ptr(c, _f32p))
        return float(s[0]), float(c[0])

    def cdf_search(self, cdf, R):
        cdf = np.ascontiguousarray(cdf, np.float32)
        return self.lib.orc_cdf_search(_ptr(cdf, _f32p), len(cdf), C.c_float(R))

    # -------------------------------------------------------------- grids --
    def grid(self, kind, n, L, n0=None, beta=1.5):
        f = np.zeros(n + 1, np.float32)
        c = np.zeros(n + 2, np.float32)
        if kind == "ref_quad":
            self.lib.orc_grid_ref_quad(n, n0 or n, C.c_float(L), _ptr(f, _f32p), _ptr(c, _f32p))
        elif kind == "uniform":
            self.lib.orc_grid_uniform(n, C.c_float(L), _ptr(f, _f32p), _ptr(c, _f32p))
        elif kind == "tanh":
            self.lib.orc_grid_tanh(n, C.c_float(L), beta, _ptr(f, _f32p), _ptr(c, _f32p))
        else:
            raise ValueError(kind)
        return f, c

    def planck_band(self, T, nu):
        return self.lib.orc_planck_band(C.c_float(T), C.c_float(nu))

    def interp3d(self, nf, xf, yf, zf, Tf, nc, xc, yc, zc):
        nf_ = np.asarray(nf, np.int32)
        nc_ = np.asarray(nc, np.int32)
        arrs = [np.ascontiguousarray(a, np.float32) for a in (xf, yf, zf, Tf, xc, yc, zc)]
        out = np.zeros(int(np.prod(nc)), np.float32)
        self.lib.orc_interp3d(_ptr(nf_, _i32p), *[_ptr(a, _f32p) for a in arrs[:4]], _ptr(nc_, _i32p),
                              *[_ptr(a, _f32p) for a in arrs[4:]], _ptr(out, _f32p))
        return out

    def En(self, n, x):
        return self.lib.orc_expint_En(n, x)

    def slab_grey_Q(self, faces, T, kappa, Tw_lo, Tw_hi):
        f = np.ascontiguousarray(faces, np.float64)
        t = np.ascontiguousarray(T, np.float64)
        q = np.zeros(len(t))
        self.lib.orc_slab_grey_Q(len(t), _ptr(f, _f64p), _ptr(t, _f64p), kappa, Tw_lo, Tw_hi, _ptr(q, _f64p))
        return q

    def slab_exchange(self, faces, kap, Ib, Ib_lo, Ib_hi, tau_cut=None, npanel=2000):
        f = np.ascontiguousarray(faces, np.float64)
        k = np.ascontiguousarray(kap, np.float64)
        b = np.ascontiguousarray(Ib, np.float64)
        F = np.zeros(len(k))
        if tau_cut is None:
            self.lib.orc_slab_exchange(len(k), _ptr(f, _f64p), _ptr(k, _f64p), _ptr(b, _f64p),
                                       Ib_lo, Ib_hi, _ptr(F, _f64p))
        else:
            self.lib.orc_slab_exchange_cut(len(k), _ptr(f, _f64p), _ptr(k, _f64p), _ptr(b, _f64p),
                                           Ib_lo, Ib_hi, tau_cut, npanel, _ptr(F, _f64p))
        return F

    def plant_mistake(self, k):
        """Build the plain proposal with mistake k (1: width of the next band,
        2: g weights reversed, 3: g CDF shifted by one; 0: correct)."""
        self.lib.orc_plant_mistake(int(k))

    def build_proposal(self, mode, t, Tmax, R, wq):
        """t: dict Tnb, kP, wvl, wvr, wvc, kq[nB][nT][nQ], prob[nT][nB], probg[nT][nB][nQ]."""
        a = {k: np.ascontiguousarray(t[k], np.float32) for k in ("Tnb", "kP", "wvl", "wvr", "wvc", "kq", "prob", "probg")}
        nB, nT, nQ = a["kq"].shape
        wq = np.ascontiguousarray(wq, np.float64)
        Ib = np.zeros(nB * nT, np.float32)
        kqrow = np.zeros(nB * nQ * nT, np.float32)
        cb = np.zeros(nB, np.float32)
        cg = np.zeros(nB * nQ, np.float32)
        ie = np.zeros(nB * nQ, np.float32)
        ib = np.zeros(1)
        km = np.zeros(1, np.float32)
        self.lib.orc_build_proposal(mode, nB, nQ, nT, *[_ptr(a[k], _f32p) for k in ("Tnb", "kP", "wvl", "wvr", "wvc", "kq", "prob", "probg")],
                                    _ptr(wq, _f64p), C.c_float(Tmax), R, _ptr(Ib, _f32p), _ptr(kqrow, _f32p),
                                    _ptr(cb, _f32p), _ptr(cg, _f32p), _ptr(ie, _f32p), _ptr(ib, _f64p), _ptr(km, _f32p))
        return dict(Ib=Ib, kq=kqrow, cdf_band=cb, cdf_g=cg, inv_emax=ie, Ibmax_d=float(ib[0]), kappamax=float(km[0]))

    # --------------------------------------------------------------- MC ----
    def problem(self, d):
        """Build an OrcProblem from a dict of numpy arrays / scalars (see
        dnsmc-style problem dicts produced by tests/problems.py). The returned
        object keeps its arrays alive."""
        P = OrcProblem()
        keep = []

        def f32(a):
            a = np.ascontiguousarray(a, np.float32)
            keep.append(a)
            return _ptr(a, _f32p)

        P.nlev = len(d["levels"])
        for l, lev in enumerate(d["levels"]):
            for a in range(3):
                P.n[l][a] = int(lev["n"][a])
                P.face[l][a] = f32(lev["faces"][a])
            P.T[l] = f32(lev["T"])
            P.maxs[l] = int(lev["maxs"])
        for a in range(3):
            P.centre[a] = f32(d["levels"][0]["centres"][a])
            for s in range(2):
                P.bc[a][s] = int(d["bc"][a][s])
        P.polar_axis = int(d["polar_axis"])
        P.tau_cut = float(d["tau_cut"])
        P.spectral = int(d["spectral"])
        P.nB, P.nQ, P.nT = int(d["nB"]), int(d["nQ"]), int(d["nT"])
        P.inv_dT = float(d.get("inv_dT", 0.0))
        P.mT0 = float(d.get("mT0", 0.0))
        P.kq = f32(d.get("kq", np.zeros(2)))
        P.Ib = f32(d.get("Ib", np.zeros(2)))
        P.Ibw = f32(d["Ibw"])
        P.kappa_grey = float(d.get("kappa_grey", 0.0))
        P.cdf_band = f32(d["cdf_band"])
        P.cdf_g = f32(d["cdf_g"])
        P.inv_emax = f32(d["inv_emax"])
        P.Ibmax = float(d["Ibmax"])
        P.Ibmax_d = float(d.get("Ibmax_d", d["Ibmax"]))
        P.Tmax = float(d["Tmax"])
        P.g_search = int(d["g_search"])
        P.u_off = float(d["u_off"])
        P._keep = keep
        return P

    def trace_ray(self, P, i, j, k, ut, up, band, g):
        o = OrcRayOut()
        self.lib.orc_trace_ray_f32(C.byref(P), i, j, k, C.c_float(ut), C.c_float(up), band, g, C.byref(o))
        return o

    def gen_philox(self, P, seed, snap, cell, r0, nr):
        out = np.zeros(nr, RAY_DESC)
        self.lib.orc_gen_philox(C.byref(P), seed, snap, cell, r0, nr, out.ctypes.data)
        return out

    def gen_compat(self, P, seed, subseq, use_sobol, sobol_offset, cell, nr):
        out = np.zeros(nr, RAY_DESC)
        self.lib.orc_gen_compat(C.byref(P), seed, subseq, int(use_sobol), sobol_offset, cell, nr,
                                out.ctypes.data)
        return out

    def gen_compat_sorted(self, P, seed, subseq, use_sobol, sobol_offset, cell, nr, order):
        out = np.zeros(nr, RAY_DESC)
        order = np.ascontiguousarray(order, np.int32)
        self.lib.orc_gen_compat_sorted(C.byref(P), seed, subseq, int(use_sobol), sobol_offset, cell, nr,
                                       _ptr(order, _i32p), out.ctypes.data)
        return out

    def sort_idx(self, kavg):
        kavg = np.ascontiguousarray(kavg, np.float32)
        order = np.zeros(len(kavg), np.int32)
        self.lib.orc_sort_idx(_ptr(kavg, _f32p), len(kavg), _ptr(order, _i32p))
        return order

    def trace_descs(self, P, cell, descs):
        d = np.ascontiguousarray(descs, RAY_DESC)
        E, A, Q2 = (np.zeros(1) for _ in range(3))
        ns = C.c_long(0)
        self.lib.orc_trace_descs(C.byref(P), cell, d.ctypes.data, len(d), _ptr(E, _f64p), _ptr(A, _f64p),
                                 _ptr(Q2, _f64p), C.byref(ns))
        return float(E[0]), float(A[0]), float(Q2[0]), ns.value

    def mc_philox(self, P, seed, snap, R, cells=None, nthreads=0, f64=False):
        ntot = P.n[0][0] * P.n[0][1] * P.n[0][2]
        if cells is None:
            cp, nc = None, ntot
        else:
            cells = np.ascontiguousarray(cells, np.int32)
            cp, nc = _ptr(cells, _i32p), len(cells)
        E, A, Q2 = (np.zeros(nc) for _ in range(3))
        ns = C.c_long(0)
        fn = self.lib.orc_mc_philox_f64 if f64 else self.lib.orc_mc_philox
        fn(C.byref(P), seed, snap, R, cp, nc, _ptr(E, _f64p), _ptr(A, _f64p), _ptr(Q2, _f64p),
           C.byref(ns), nthreads)
        return E, A, Q2, ns.value

    # -------------------------------------------------------------- MLP ----
    def bf16_round(self, x):
        return self.lib.orc_bf16_round(float(x))

    def mlp_loss_grad(self, X, W1, b1, W2, b2, W3, b3, ndir, td, tb, wt, bf16=False):
        """Loss and gradients; bf16=True rounds the product operands as the
        kernels do (orc_mlp_loss_grad_bf16)."""
        X, W1, b1, W2, b2, W3, b3, wt = (np.ascontiguousarray(a, np.float64) for a in (X, W1, b1, W2, b2, W3, b3, wt))
        td = np.ascontiguousarray(td, np.int32)
        tb = np.ascontiguousarray(tb, np.int32)
        nb, nin = X.shape
        nh = W1.shape[1]
        nout = W3.shape[1]
        g = [np.zeros_like(a) for a in (W1, b1, W2, b2, W3, b3)]
        loss = np.zeros(1)
        fn = self.lib.orc_mlp_loss_grad_bf16 if bf16 else self.lib.orc_mlp_loss_grad
        fn(nb, *[_ptr(a, _f64p) for a in (X, W1, b1, W2, b2, W3, b3)], nin, nh, nout,
           ndir, _ptr(td, _i32p), _ptr(tb, _i32p), _ptr(wt, _f64p), _ptr(loss, _f64p),
           *[_ptr(a, _f64p) for a in g])
        return float(loss[0]), g

    def learned_draw(self, U, pcdf, p_bg, cells, R, ndir, nQ, snap, seed, eps):
        """Descriptors (w0, w1) and p/q of the learned draw for the slots whose
        cumulative-weight rows U[nslots, noutp] are given (orc_learned_draw)."""
        U = np.ascontiguousarray(U, np.float32)
        pcdf = np.ascontiguousarray(pcdf, np.float32)
        p_bg = np.ascontiguousarray(p_bg, np.float32)
        cells = np.ascontiguousarray(cells, np.int32)
        nslots, noutp = U.shape
        n = nslots * R
        w0 = np.zeros(n, np.uint32)
        w1 = np.zeros(n, np.uint32)
        ratio = np.zeros(n, np.float32)
        self.lib.orc_learned_draw(nslots, R, ndir, len(pcdf), nQ, noutp, _ptr(U, _f32p), _ptr(pcdf, _f32p),
                                  _ptr(p_bg, _f32p), _ptr(cells, _i32p), snap, seed, eps, _ptr(w0, _u32p),
                                  _ptr(w1, _u32p), _ptr(ratio, _f32p))
        return w0, w1, ratio

    def mlp_forward(self, X, W1, b1, W2, b2, W3, b3, bf16=False):
        X, W1, b1, W2, b2, W3, b3 = (np.ascontiguousarray(a, np.float64) for a in (X, W1, b1, W2, b2, W3, b3))
        nb, nin = X.shape
        nh = W1.shape[1]
        nout = W3.shape[1]
        H1 = np.zeros((nb, nh)); H2 = np.zeros((nb, nh)); Z = np.zeros((nb, nout))
        fn = self.lib.orc_mlp_forward_bf16 if bf16 else self.lib.orc_mlp_forward
        fn(nb, *[_ptr(a, _f64p) for a in (X, W1, b1, W2, b2, W3, b3)], nin, nh, nout,
           _ptr(H1, _f64p), _ptr(H2, _f64p), _ptr(Z, _f64p))
        return H1, H2, Z


_ORACLE = None


def get():
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = Oracle()
    return _ORACLE
