"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

F64, F32, F16, BF16 = 0, 1, 2, 3

_lib = None


def build(force=False):
    """Compiles oracle.cpp with g++ (OpenMP) into oracle/liborc.so."""
    import subprocess
    src = os.path.join(HERE, "oracle.cpp")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-fopenmp", "-shared", "-fPIC",
                               "-o", LIB_PATH, src])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i, ll, f, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float, ctypes.c_double
        P = ctypes.POINTER
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp, i]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_model.argtypes = [vp, vp, vp, vp, vp]
        L.orc_set_source_point.argtypes = [vp, i, i, i, f]
        L.orc_set_source_table.argtypes = [vp, vp, vp, ll, i, f]
        L.orc_set_stations.argtypes = [vp, vp, i]
        L.orc_step.argtypes = [vp, i]
        for n in ("orc_get_stations", "orc_get_pgv", "orc_get_state", "orc_set_state",
                  "orc_get_cjm64", "orc_get_rdz", "orc_get_pml"):
            getattr(L, n).argtypes = [vp, vp]
        L.orc_current_step.argtypes = [vp]
        L.orc_debug_rhs.argtypes = [vp, i, i, i, vp]
        L.orc_debug_source.argtypes = [vp]
        L.orc_debug_aux_rhs.restype = ll
        L.orc_debug_aux_rhs.argtypes = [vp, i, vp, ll]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_apply_L_line.argtypes = [vp, i, i, d, vp]
        L.orc_wave_rk.argtypes = [i, vp, vp, vp, vp, ll]
        L.orc_source_function.restype = d
        L.orc_source_function.argtypes = [d, i, d]
        L.orc_gauss_weights.argtypes = [vp]
        L.orc_moment_tensor.argtypes = [d, d, d, d, d, vp]
        L.orc_round_f16.argtypes = [vp, vp, ll]
        L.orc_round_bf16.argtypes = [vp, vp, ll]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Oracle:
    """One CPU simulation; mirrors the product's domain handle call for call."""

    def __init__(self, desc, precision=F64):
        self.desc = desc
        self.prec = precision
        self.h = lib().orc_create(ctypes.byref(desc), precision)
        if not self.h:
            raise ValueError(lib().orc_last_error().decode())
        self.nx_, self.ny_, self.nz_ = desc.NX + 6, desc.NY + 6, desc.NZ + 6
        self.num = self.nx_ * self.ny_ * self.nz_

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}: {lib().orc_last_error().decode()}")

    def set_model(self, terrain, vs, vp, rho):
        t = None if terrain is None else np.ascontiguousarray(terrain, dtype=np.float64)
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (vs, vp, rho)]
        self._keep = (t, a)
        self._chk(lib().orc_set_model(self.h, _p(t), _p(a[0]), _p(a[1]), _p(a[2])))

    def set_source_point(self, X, Y, Z, fc):
        self._chk(lib().orc_set_source_point(self.h, X, Y, Z, fc))

    def set_source_table(self, xyz, rate, dt):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        rate = np.ascontiguousarray(rate, dtype=np.float32)
        nt, npts = rate.shape[0], rate.shape[1]
        self._chk(lib().orc_set_source_table(self.h, _p(xyz), _p(rate), npts, nt, dt))

    def set_stations(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        self.nst = xyz.shape[0]
        self._chk(lib().orc_set_stations(self.h, _p(xyz), self.nst))

    def step(self, n=1):
        self._chk(lib().orc_step(self.h, n))

    def stations(self):
        out = np.zeros((self.nst, self.desc.NT, 9), np.float32)
        self._chk(lib().orc_get_stations(self.h, _p(out)))
        return out

    def pgv(self):
        out = np.zeros((self.desc.NY, self.desc.NX, 2), np.float32)
        self._chk(lib().orc_get_pgv(self.h, _p(out)))
        return out

    def state(self):
        d = self.desc
        out = np.zeros((9, d.NZ, d.NY, d.NX), np.float64)
        self._chk(lib().orc_get_state(self.h, _p(out)))
        return out

    def set_state(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        self._chk(lib().orc_set_state(self.h, _p(w)))

    def cjm64(self):
        out = np.zeros((13, self.nz_, self.ny_, self.nx_), np.float64)
        self._chk(lib().orc_get_cjm64(self.h, _p(out)))
        return out

    def rdz(self):
        out = np.zeros((2, self.ny_, self.nx_, 9), np.float64)
        self._chk(lib().orc_get_rdz(self.h, _p(out)))
        return out

    def pml(self):
        out = np.zeros(3 * (self.nx_ + self.ny_ + self.nz_), np.float64)
        self._chk(lib().orc_get_pml(self.h, _p(out)))
        res, o = {}, 0
        for ax, n in zip("xyz", (self.nx_, self.ny_, self.nz_)):
            for name in ("alpha", "beta", "d"):
                res[name + "_" + ax] = out[o:o + n].copy()
                o += n
        return res

    @property
    def it(self):
        return lib().orc_current_step(self.h)

    def rhs(self, FB1, FB2, FB3):
        """One stage's increment h = DT*RHS on the interior (W and Aux unchanged)."""
        d = self.desc
        out = np.zeros((9, d.NZ, d.NY, d.NX), np.float64)
        self._chk(lib().orc_debug_rhs(self.h, FB1, FB2, FB3, _p(out)))
        return out

    def aux_rhs(self, slab):
        """Auxiliary increments h_Aux = DT*RHS of PML slab `slab` (0..5: x1 y1 z1 x2 y2 z2) as the
        last rhs() call left them, shaped (n2, n1, n0, 9) in slab-local (k0, j0, i0)."""
        d = self.desc
        n = d.nPML
        dims = [(d.NZ, d.NY, n), (d.NZ, n, d.NX), (n, d.NY, d.NX)][slab % 3]
        cap = 9 * dims[0] * dims[1] * dims[2]
        out = np.zeros(cap, np.float64)
        r = lib().orc_debug_aux_rhs(self.h, slab, _p(out), cap)
        if r < 0:
            raise RuntimeError(f"oracle error {r}: {lib().orc_last_error().decode()}")
        if r == 0:
            return None
        return out.reshape(dims + (9,))

    def inject_source(self):
        """Applies only this step's source injection into W."""
        self._chk(lib().orc_debug_source(self.h))


def apply_L_line(f, FB, rDH):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.full_like(f, np.nan)
    lib().orc_apply_L_line(_p(f), f.size, FB, rDH, _p(out))
    return out


def wave_rk(irk, h, W, t, m):
    for a in (h, W, t, m):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    lib().orc_wave_rk(irk, _p(h), _p(W), _p(t), _p(m), W.size)


def source_function(fc, it, DT):
    return lib().orc_source_function(fc, it, DT)


def gauss_weights():
    out = np.zeros(343, np.float64)
    lib().orc_gauss_weights(_p(out))
    return out.reshape(7, 7, 7)


def moment_tensor(strike, dip, rake, area=1.0, rate=1.0):
    out = np.zeros(6, np.float64)
    lib().orc_moment_tensor(strike, dip, rake, area, rate, _p(out))
    return out


def round_f16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    lib().orc_round_f16(_p(x), _p(out), x.size)
    return out


def round_bf16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    lib().orc_round_bf16(_p(x), _p(out), x.size)
    return out
