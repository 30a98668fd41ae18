"""Host-side mirror of the reference's radiation interface.

``mc_gpu(T, ystart)`` / ``get_results(Q, V)`` keep the argument meaning of
the reference's Fortran-callable pair (code/mc_gpu.cu:440, :640): T is the
ghost-inclusive (imax+2, jmax+2, kmax+2) field in Fortran order, the call
returns before the GPU finishes, and get_results fills the interior of Q and
of the standard-error array.  ``Solver`` exposes the extended C API
(runtime configuration, several snapshots per launch, FP64 results).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import Config, Stats, check, lib


def default_config():
    c = Config()
    lib().dnsmc_default_config(C.byref(c))
    return c


def parse_config(text, base=None):
    c = base if base is not None else default_config()
    check(lib().dnsmc_parse_config(text.encode(), C.byref(c)))
    return c


def config_from_dict(d, base=None):
    """Build a Config from a dict of keys of include/dnsmc_rad.h:dnsmc_config."""
    c = base if base is not None else default_config()
    for k, v in d.items():
        if k == "table_dir":
            c.table_dir = str(v).encode()
        elif k in ("n", "L", "stretch", "stretch_beta", "maxs"):
            arr = getattr(c, k)
            for i, x in enumerate(v):
                arr[i] = x
        elif k in ("bc", "Tw"):
            arr = getattr(c, k)
            for a in range(3):
                for s in range(2):
                    arr[a][s] = v[a][s]
        else:
            setattr(c, k, v)
    return c


def probe_tables(cfg):
    """Read the spectral tables cfg names on the host only (no device):
    returns (nB, nQ, nT) or raises DnsmcError (DNSMC_EIO for a missing, short
    or inconsistent file)."""
    nB, nQ, nT = C.c_int(0), C.c_int(0), C.c_int(0)
    check(lib().dnsmc_probe_tables(C.byref(cfg), C.byref(nB), C.byref(nQ), C.byref(nT)))
    return nB.value, nQ.value, nT.value


def _f32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """One solver instance on the current HIP device."""

    def __init__(self, cfg):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().dnsmc_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.n = tuple(cfg.n)
        self.nemit = self.n[0] * self.n[1] * self.n[2]

    @property
    def rays_per_slot(self):
        return self.cfg.photmax if self.cfg.mode == _lib.MODE_COMPAT else self.cfg.rays_per_cell

    @property
    def nslots(self):
        return self.nemit * (self.cfg.nvar if self.cfg.mode == _lib.MODE_COMPAT else 1)

    def close(self):
        if self.h:
            lib().dnsmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # tables ---------------------------------------------------------------
    def load_tables(self):
        check(lib().dnsmc_load_tables(self.h))

    def set_tables(self, t):
        """t: dict with Tnb, kP, wvl, wvr, wvc, kq[nB][nT][nQ], prob[nT][nB], probg[nT][nB][nQ] and
        optionally kavg[nB] (the sorted mode's band key)."""
        arrs = {k: np.ascontiguousarray(t[k], np.float32) for k in
                ("Tnb", "kP", "wvl", "wvr", "wvc", "kq", "prob", "probg")}
        kavg = np.ascontiguousarray(t["kavg"], np.float32) if t.get("kavg") is not None else None
        nB, nT, nQ = arrs["kq"].shape
        self._tables = dict(arrs, kavg=kavg)
        check(lib().dnsmc_set_tables(self.h, nB, nQ, nT, *[_f32(arrs[k]) for k in
                                                            ("Tnb", "kP", "wvl", "wvr", "wvc", "kq", "prob", "probg")],
                                     _f32(kavg) if kavg is not None else None))

    # field / emitters -----------------------------------------------------
    def set_field(self, snap, T_ghost):
        T = np.asfortranarray(T_ghost, dtype=np.float32)
        if T.shape != (self.n[0] + 2, self.n[1] + 2, self.n[2] + 2):
            raise ValueError("T must have shape (n0+2, n1+2, n2+2), got %s" % (T.shape,))
        check(lib().dnsmc_set_field(self.h, snap, T.ctypes.data, 0))

    def fix_temp_slab(self, snap, cnew, startc, endc, Tc, Tplus):
        """One rank's slab of the nondimensional temperature into the device
        field (fix_temp, code/fix_var.f:49-73, without the all-reduce): cnew is
        the rank's array cnew(0:n0+1, nj, nk) (Fortran order, float64), startc
        / endc the 1-based inclusive global ranges.  Call commit_field after
        the last slab."""
        st = np.ascontiguousarray(startc, np.int32)
        en = np.ascontiguousarray(endc, np.int32)
        c = np.asfortranarray(cnew, dtype=np.float64)
        shp = (self.n[0] + 2, int(en[1] - st[1] + 1), int(en[2] - st[2] + 1))
        if c.shape != shp:
            raise ValueError("cnew must have shape %s, got %s" % (shp, c.shape))
        i32p = C.POINTER(C.c_int)
        check(lib().dnsmc_fix_temp_slab(self.h, snap, _f64(c), st.ctypes.data_as(i32p), en.ctypes.data_as(i32p),
                                        float(Tc), float(Tplus)))

    def commit_field(self, snap):
        check(lib().dnsmc_commit_field(self.h, snap))

    def fix_rad_slab(self, snap, startc, endc, Tc):
        """divQ and vcore of one rank (fix_rad, code/fix_var.f:75-121): arrays
        (0:n0+1, n1/p_row, nk) in Fortran order, FP64."""
        st = np.ascontiguousarray(startc, np.int32)
        en = np.ascontiguousarray(endc, np.int32)
        shp = (self.n[0] + 2, self.n[1] // self.cfg.p_row, int(en[2] - st[2] + 1))
        dq = np.zeros(shp, np.float64, order="F")
        vc = np.zeros(shp, np.float64, order="F")
        i32p = C.POINTER(C.c_int)
        check(lib().dnsmc_fix_rad_slab(self.h, snap, st.ctypes.data_as(i32p), en.ctypes.data_as(i32p), float(Tc),
                                       _f64(dq), _f64(vc)))
        return dq, vc

    def set_emitters_slab(self, ystart):
        check(lib().dnsmc_set_emitters_slab(self.h, ystart))
        self.nemit = self.n[0] * (self.n[1] // self.cfg.p_row) * self.n[2]

    def set_emitters(self, cells):
        cells = np.ascontiguousarray(cells, np.int32)
        check(lib().dnsmc_set_emitters_list(self.h, cells.ctypes.data_as(C.POINTER(C.c_int)), len(cells)))
        self.nemit = len(cells)

    def set_mlp(self, W1, b1, W2, b2, W3, b3, mix_eps=0.1):
        """Learned proposal weights: W* as BF16 bit patterns (uint16, row-major
        [in][out]), b* float32; see include/dnsmc_rad.h:dnsmc_set_mlp."""
        a = [np.ascontiguousarray(x, np.uint16 if i % 2 == 0 else np.float32)
             for i, x in enumerate((W1, b1, W2, b2, W3, b3))]
        check(lib().dnsmc_set_mlp(self.h, *[x.ctypes.data for x in a], C.c_float(mix_eps)))

    def set_draw_log(self, on=True):
        check(lib().dnsmc_set_draw_log(self.h, 1 if on else 0))

    def set_record(self, on=True):
        check(lib().dnsmc_set_record(self.h, 1 if on else 0))

    # run ------------------------------------------------------------------
    def run(self):
        check(lib().dnsmc_run(self.h))

    def wait(self):
        check(lib().dnsmc_wait(self.h))

    def results(self, snap=0):
        out = {k: np.zeros(self.nemit) for k in ("Q", "E", "A", "sdev")}
        check(lib().dnsmc_get_results(self.h, snap, _f64(out["Q"]), _f64(out["E"]), _f64(out["A"]),
                                      _f64(out["sdev"])))
        return out

    def results_fortran(self, snap=0):
        shp = (self.n[0] + 2, self.n[1] // self.cfg.p_row + 2, self.n[2] + 2)
        Q = np.zeros(shp, np.float32, order="F")
        V = np.zeros(shp, np.float32, order="F")
        check(lib().dnsmc_get_results_fortran(self.h, snap, _f32(Q), _f32(V)))
        return Q, V

    def stats(self):
        s = Stats()
        check(lib().dnsmc_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    # inspection -----------------------------------------------------------
    def array(self, name, level=0):
        n = check(lib().dnsmc_get_array(self.h, name.encode(), level, None, 0))
        out = np.zeros(n, np.float32)
        check(lib().dnsmc_get_array(self.h, name.encode(), level, _f32(out), n))
        return out

    def Ibmax_d(self):
        out = np.zeros(1, np.float64)
        check(lib().dnsmc_get_array(self.h, b"Ibmax_d", 0, out.ctypes.data_as(C.POINTER(C.c_float)), 2))
        return float(out[0])

    def rays(self, n):
        w0 = np.zeros(n, np.uint32)
        w1 = np.zeros(n, np.uint32)
        w = np.zeros(n, np.float32)
        wi = np.zeros(n, np.float32)
        m = check(lib().dnsmc_get_rays(self.h, w0.ctypes.data, w1.ctypes.data, w.ctypes.data, wi.ctypes.data, n))
        out = np.zeros(m, dtype=[("w0", "<u4"), ("w1", "<u4"), ("w", "<f4"), ("winv", "<f4")])
        out["w0"], out["w1"], out["w"], out["winv"] = w0[:m], w1[:m], w[:m], wi[:m]
        return out

    def records(self, n):
        st = np.zeros(n, np.int32)
        en = np.zeros(n, np.int32)
        hs = np.zeros(n, np.uint32)
        D = np.zeros(n, np.float32)
        G = np.zeros(n, np.float32)
        m = check(lib().dnsmc_get_ray_records(self.h, st.ctypes.data, en.ctypes.data, hs.ctypes.data,
                                              D.ctypes.data, G.ctypes.data, n))
        return {"steps": st[:m], "end": en[:m], "hash": hs[:m], "D": D[:m], "G": G[:m]}


# ----------------------------------------------------- reference mirror ----
def mc_gpu(T, ystart=1):
    """code/mc_gpu.cu:440 mc_gpu_(T, ystart): asynchronous launch (copies T)."""
    T = np.asfortranarray(T, dtype=np.float32)
    ys = C.c_int(ystart)
    lib().mc_gpu_(_f32(T), C.byref(ys))


def get_results(Q, V):
    """code/mc_gpu.cu:640 get_results_(Q, V): wait and fill the interior."""
    if not (Q.flags.f_contiguous and V.flags.f_contiguous and Q.dtype == np.float32 and V.dtype == np.float32):
        raise ValueError("Q and V must be Fortran-ordered float32 arrays")
    lib().get_results_(_f32(Q), _f32(V))


def fix_temp_gpu(cnew, startc, endc, Tc, Tplus):
    """fix_temp_gpu_ (replaces code/fix_var.f:49-73): one rank's slab into the
    legacy solver's field; cnew(0:n0+1, nj, nk) float64, Fortran order."""
    c = np.asfortranarray(cnew, dtype=np.float64)
    st = (C.c_int * 3)(*[int(x) for x in startc])
    en = (C.c_int * 3)(*[int(x) for x in endc])
    lib().fix_temp_gpu_(_f64(c), st, en, C.byref(C.c_double(Tc)), C.byref(C.c_double(Tplus)))


def mc_gpu_commit(ystart=1):
    """mc_gpu_commit_: queue the solve of the field the slabs built."""
    lib().mc_gpu_commit_(C.byref(C.c_int(ystart)))


def fix_rad_gpu(startc, endc, Tc, shape):
    """fix_rad_gpu_ (replaces code/fix_var.f:75-121): one rank's divQ, vcore
    of the last solve, arrays of `shape` = (n0+2, n1/p_row, nk), float64."""
    dq = np.zeros(shape, np.float64, order="F")
    vc = np.zeros(shape, np.float64, order="F")
    st = (C.c_int * 3)(*[int(x) for x in startc])
    en = (C.c_int * 3)(*[int(x) for x in endc])
    lib().fix_rad_gpu_(_f64(dq), _f64(vc), st, en, C.byref(C.c_double(Tc)))
    return dq, vc


def sync():
    check(lib().dnsmc_sync())


__all__ = ["Solver", "default_config", "parse_config", "config_from_dict", "mc_gpu", "get_results", "sync",
           "fix_temp_gpu", "mc_gpu_commit", "fix_rad_gpu", "_lib"]
