"""Python binding of the MI355X CGFDM C ABI (include/cgfdm.h) over ctypes.

The library (libcgfdm.so, built from csrc/ by build.py) holds every kernel; this module only
marshals host arrays. There is no fallback: if the HIP library is missing or cannot be
loaded, every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CGFDM_LIB") or os.path.join(HERE, "libcgfdm.so")

FP32, FP16, BF16 = 1, 2, 3
PRECISIONS = {"fp32": FP32, "fp16": FP16, "bf16": BF16}

STATUS = {0: "OK", -1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ECFL", -5: "EPML",
          -6: "ESTATE", -7: "EIO", -8: "ENONFINITE"}
STATUS_CODES = {v: k for k, v in STATUS.items()}


class Desc(ctypes.Structure):
    """Mirror of struct cgfdm_desc (include/cgfdm.h)."""
    _fields_ = [
        ("NX", ctypes.c_int), ("NY", ctypes.c_int), ("NZ", ctypes.c_int),
        ("nPML", ctypes.c_int),
        ("DH", ctypes.c_double), ("DT", ctypes.c_double),
        ("NT", ctypes.c_int), ("precision", ctypes.c_int),
        ("free_surface", ctypes.c_int), ("pml", ctypes.c_int),
        ("use_terrain", ctypes.c_int),
        ("depth", ctypes.c_double),
        ("centerX", ctypes.c_int), ("centerY", ctypes.c_int),
        ("Cv", ctypes.c_double), ("Cs", ctypes.c_double), ("Crho", ctypes.c_double),
        ("check_cfl", ctypes.c_int),
        ("chi_mode", ctypes.c_int),
    ]


def make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=100, nPML=12, precision=FP32,
              free_surface=1, pml=1, use_terrain=0, depth=None, centerX=None, centerY=None,
              Cv=0.0, Cs=0.0, Crho=0.0, check_cfl=0, chi_mode=0):
    if isinstance(precision, str):
        precision = PRECISIONS[precision]
    d = Desc()
    d.NX, d.NY, d.NZ, d.nPML = NX, NY, NZ, nPML
    d.DH, d.DT, d.NT, d.precision = DH, DT, NT, precision
    d.free_surface, d.pml, d.use_terrain = free_surface, pml, use_terrain
    d.depth = DH * NZ if depth is None else depth
    d.centerX = NX // 2 if centerX is None else centerX
    d.centerY = NY // 2 if centerY is None else centerY
    d.Cv, d.Cs, d.Crho, d.check_cfl, d.chi_mode = Cv, Cs, Crho, check_cfl, chi_mode
    return d


def default_scales(precision):
    """Reference scale factors (macro.h:65-73): 1 for FP32, (5e3, 1e-6, 1e6) otherwise."""
    if precision == FP32:
        return 1.0, 1.0, 1.0
    return 5.0e3, 1.0e-6, 1.0e6


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, ll, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float
        P = ctypes.POINTER
        L.cgfdm_last_error.restype = ctypes.c_char_p
        L.cgfdm_domain_create.argtypes = [vp, i, P(vp)]
        L.cgfdm_domain_destroy.argtypes = [vp]
        L.cgfdm_domain_bytes.argtypes = [vp]
        L.cgfdm_domain_bytes.restype = ll
        L.cgfdm_estimate_bytes.argtypes = [vp]
        L.cgfdm_estimate_bytes.restype = ll
        L.cgfdm_set_model.argtypes = [vp, vp, vp, vp, vp]
        L.cgfdm_set_source_point.argtypes = [vp, i, i, i, f]
        L.cgfdm_set_source_table.argtypes = [vp, vp, vp, ll, i, f]
        L.cgfdm_set_stations.argtypes = [vp, vp, i]
        L.cgfdm_step.argtypes = [vp, i]
        L.cgfdm_sync.argtypes = [vp]
        L.cgfdm_current_step.argtypes = [vp]
        L.cgfdm_set_check_interval.argtypes = [vp, i]
        L.cgfdm_nonfinite_step.argtypes = [vp, P(i)]
        for n in ("cgfdm_get_stations", "cgfdm_get_pgv", "cgfdm_get_state", "cgfdm_set_state",
                  "cgfdm_get_cjm"):
            getattr(L, n).argtypes = [vp, vp]
        L.cgfdm_get_slice.argtypes = [vp, i, i, i, vp]
        L.cgfdm_get_slice_noalloc.argtypes = [vp, i, i, i, vp]
        L.cgfdm_stage_timing.argtypes = [vp, i, P(ctypes.c_double), P(ll)]
        L.cgfdm_kernel_timing.argtypes = [vp, i, vp, P(ll)]
        L.cgfdm_enable_timing.argtypes = [vp, i]
        L.cgfdm_partition.argtypes = [vp, vp]
        L.cgfdm_uniform_metric_pair.argtypes = [vp, P(i)]
        L.cgfdm_decompose.argtypes = [vp, vp, i]
        L.cgfdm_set_model_layered.argtypes = [vp, vp, i, vp, vp, vp, vp]
        L.cgfdm_finite_fault_map.argtypes = [ctypes.c_char_p, vp, vp, ctypes.c_double, ctypes.c_double, vp, vp, ll,
                                             P(ll), P(i), P(f)]
        L.cgfdm_aeqd_forward.argtypes = [ctypes.c_double, ctypes.c_double, ll, vp, vp, vp, vp]
        L.cgfdm_point_source_scales.argtypes = [ctypes.c_double] * 4 + [P(ctypes.c_double)] * 2
        L.cgfdm_source_scales.argtypes = [ctypes.c_double, vp, vp, ll, i, f, ctypes.c_double, P(ctypes.c_double),
                                          P(ctypes.c_double)]
        L.cgfdm_server_create.argtypes = [vp, P(vp)]
        L.cgfdm_server_submit.argtypes = [vp, vp, P(ll)]
        L.cgfdm_server_wait.argtypes = [vp, vp, i, i]
        L.cgfdm_server_result.argtypes = [vp, ll, vp, vp, vp]
        L.cgfdm_server_stats_get.argtypes = [vp, vp]
        L.cgfdm_server_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class CgfdmError(RuntimeError):
    pass


def _chk(rc):
    if rc != 0:
        raise CgfdmError(f"cgfdm error {STATUS.get(rc, rc)}: {lib().cgfdm_last_error().decode()}")


class Domain:
    """One simulation on one GPU (a cgfdm_domain handle)."""

    def __init__(self, desc, device=0):
        self.desc = desc
        h = ctypes.c_void_p()
        _chk(lib().cgfdm_domain_create(ctypes.byref(desc), device, ctypes.byref(h)))
        self.h = h
        self.nx_, self.ny_, self.nz_ = desc.NX + 6, desc.NY + 6, desc.NZ + 6
        self.num = self.nx_ * self.ny_ * self.nz_
        self.nst = 0

    def close(self):
        if self.h:
            lib().cgfdm_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def bytes(self):
        return lib().cgfdm_domain_bytes(self.h)

    def set_model(self, terrain, vs, vp, rho):
        t = None if terrain is None else np.ascontiguousarray(terrain, dtype=np.float64)
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (vs, vp, rho)]
        _chk(lib().cgfdm_set_model(self.h, _p(t), _p(a[0]), _p(a[1]), _p(a[2])))

    def set_source_point(self, X, Y, Z, fc):
        _chk(lib().cgfdm_set_source_point(self.h, X, Y, Z, fc))

    def set_source_table(self, xyz, rate, dt):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        rate = np.ascontiguousarray(rate, dtype=np.float32)
        nt, npts = rate.shape[0], rate.shape[1]
        _chk(lib().cgfdm_set_source_table(self.h, _p(xyz), _p(rate), npts, nt, dt))

    def set_stations(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        self.nst = xyz.shape[0]
        _chk(lib().cgfdm_set_stations(self.h, _p(xyz), self.nst))

    def step(self, n=1):
        _chk(lib().cgfdm_step(self.h, n))

    def sync(self):
        _chk(lib().cgfdm_sync(self.h))

    @property
    def it(self):
        return lib().cgfdm_current_step(self.h)

    def set_check_interval(self, every):
        """Scan the wavefield for non-finite values every `every` steps (0: never)."""
        _chk(lib().cgfdm_set_check_interval(self.h, every))

    def nonfinite_step(self):
        """Step count after which the guard first found a non-finite value, or -1."""
        out = ctypes.c_int(0)
        _chk(lib().cgfdm_nonfinite_step(self.h, ctypes.byref(out)))
        return out.value

    def stations(self):
        out = np.zeros((self.nst, self.desc.NT, 9), np.float32)
        _chk(lib().cgfdm_get_stations(self.h, _p(out)))
        return out

    def pgv(self):
        out = np.zeros((self.desc.NY, self.desc.NX, 2), np.float32)
        _chk(lib().cgfdm_get_pgv(self.h, _p(out)))
        return out

    def slice(self, axis, index, var):
        d = self.desc
        shape = {0: (d.NZ, d.NY), 1: (d.NZ, d.NX), 2: (d.NY, d.NX)}[axis]
        out = np.zeros(shape, np.float32)
        _chk(lib().cgfdm_get_slice(self.h, axis, index, var, _p(out)))
        return out

    def slice_noalloc(self, axis, index, var):
        """As slice(), staged in the library's fixed device plane (cgfdm_get_slice_noalloc)."""
        d = self.desc
        shape = {0: (d.NZ, d.NY), 1: (d.NZ, d.NX), 2: (d.NY, d.NX)}[axis]
        out = np.zeros(shape, np.float32)
        _chk(lib().cgfdm_get_slice_noalloc(self.h, axis, index, var, _p(out)))
        return out

    def state(self):
        d = self.desc
        out = np.zeros((9, d.NZ, d.NY, d.NX), np.float32)
        _chk(lib().cgfdm_get_state(self.h, _p(out)))
        return out

    def set_state(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _chk(lib().cgfdm_set_state(self.h, _p(w)))

    def cjm(self):
        out = np.zeros((13, self.nz_, self.ny_, self.nx_), np.float32)
        _chk(lib().cgfdm_get_cjm(self.h, _p(out)))
        return out

    def enable_timing(self, on=True):
        _chk(lib().cgfdm_enable_timing(self.h, 1 if on else 0))

    def set_model_layered(self, terrain, depth, vs, vp, rho):
        t = None if terrain is None else np.ascontiguousarray(terrain, dtype=np.float64)
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (depth, vs, vp, rho)]
        _chk(lib().cgfdm_set_model_layered(self.h, _p(t), len(a[0]), *[_p(x) for x in a]))

    def kernel_timing(self, reset=False):
        """(stage, bulk kernel, boundary kernel) average ms per stage, and the stages timed."""
        out = np.zeros(3, np.float64)
        cnt = ctypes.c_longlong()
        _chk(lib().cgfdm_kernel_timing(self.h, 1 if reset else 0, _p(out), ctypes.byref(cnt)))
        return tuple(out), cnt.value

    def partition(self):
        """(bulk points, boundary points, tile shape, z-chunk) of a stage."""
        out = np.zeros(4, np.int64)
        _chk(lib().cgfdm_partition(self.h, _p(out)))
        return tuple(int(x) for x in out)

    def uniform_metric_pair(self):
        """True when the stage kernels take (xi_x, eta_y) from one stored copy."""
        out = ctypes.c_int(0)
        _chk(lib().cgfdm_uniform_metric_pair(self.h, ctypes.byref(out)))
        return bool(out.value)

    def stage_timing(self, reset=False):
        avg = ctypes.c_double()
        cnt = ctypes.c_longlong()
        _chk(lib().cgfdm_stage_timing(self.h, 1 if reset else 0, ctypes.byref(avg), ctypes.byref(cnt)))
        return avg.value, cnt.value


def decompose(desc):
    """[(i0, i1, j0, j1, k0, k1, kind, axis)] of the stage decomposition (no device needed)."""
    out = np.zeros((32, 8), np.int32)
    n = lib().cgfdm_decompose(ctypes.byref(desc), _p(out), 32)
    if n < 0:
        _chk(n)
    return [tuple(int(x) for x in r) for r in out[:n]]


def finite_fault_map(path, desc, terrain, lon0, lat0):
    """The reference's finite-fault file mapped onto the grid of `desc` (host only):
    (xyz int32 [n, 3] global interior indices, rate float32 [nt, n, 6], dt) as
    set_source_table takes them."""
    L = lib()
    n, nt, dt = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_float()
    tp = None if terrain is None else np.ascontiguousarray(terrain, np.float64)
    _chk(L.cgfdm_finite_fault_map(str(path).encode(), ctypes.byref(desc), _p(tp), lon0, lat0, None, None, 0,
                                  ctypes.byref(n), ctypes.byref(nt), ctypes.byref(dt)))
    xyz = np.zeros((max(n.value, 1), 3), np.int32)
    rate = np.zeros((nt.value, max(n.value, 1), 6), np.float32)
    _chk(L.cgfdm_finite_fault_map(str(path).encode(), ctypes.byref(desc), _p(tp), lon0, lat0, _p(xyz), _p(rate),
                                  n.value, ctypes.byref(n), ctypes.byref(nt), ctypes.byref(dt)))
    return xyz[:n.value], rate[:, :n.value], dt.value


def aeqd_forward(lon0, lat0, lon, lat):
    """Azimuthal equidistant projection on WGS84 centred on (lon0, lat0): degrees -> metres."""
    lon = np.ascontiguousarray(lon, np.float64)
    lat = np.ascontiguousarray(lat, np.float64)
    x, y = np.zeros_like(lon), np.zeros_like(lat)
    _chk(lib().cgfdm_aeqd_forward(lon0, lat0, lon.size, _p(lon), _p(lat), _p(x), _p(y)))
    return x, y


def estimate_bytes(desc):
    return lib().cgfdm_estimate_bytes(ctypes.byref(desc))


def source_scales(DH, xyz=None, table=None, dt=None, mu_max=None):
    """(Cv, Cs) of 16-bit storage chosen from the source (cgfdm_source_scales): the point source
    without a table, else the moment-rate table (nt x npts x 6) at the global interior indices
    xyz (npts x 3), sampled every dt s, in a medium whose largest shear modulus is mu_max."""
    cv, cs = ctypes.c_double(), ctypes.c_double()
    if table is None:
        _chk(lib().cgfdm_source_scales(float(DH), None, None, 0, 0, 0.0, 0.0, ctypes.byref(cv), ctypes.byref(cs)))
    else:
        t = np.ascontiguousarray(table, np.float32)
        x = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        if len(x) != t.shape[1]:
            raise ValueError("xyz and table disagree on the number of nodes")
        _chk(lib().cgfdm_source_scales(float(DH), _p(x), _p(t), t.shape[1], t.shape[0], float(dt), float(mu_max),
                                       ctypes.byref(cv), ctypes.byref(cs)))
    return cv.value, cs.value


def point_source_scales(DH, fc, mu_src, mu_max):
    """(Cv, Cs) of 16-bit storage for the explosive point source of peak frequency fc in a
    medium whose shear modulus is mu_src at the source and mu_max at most
    (cgfdm_point_source_scales)."""
    cv, cs = ctypes.c_double(), ctypes.c_double()
    _chk(lib().cgfdm_point_source_scales(float(DH), float(fc), float(mu_src), float(mu_max), ctypes.byref(cv),
                                         ctypes.byref(cs)))
    return cv.value, cs.value


class ServerConfig(ctypes.Structure):
    """Mirror of struct cgfdm_server_config."""
    _fields_ = [("device", ctypes.c_int), ("precision", ctypes.c_int),
                ("hbm_budget", ctypes.c_longlong), ("host_budget", ctypes.c_longlong),
                ("max_domains", ctypes.c_int), ("steps_per_launch", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("admit_order", ctypes.c_int),
                ("rescale_retries", ctypes.c_int)]


class Request(ctypes.Structure):
    """Mirror of struct cgfdm_request."""
    _fields_ = [("desc", Desc), ("terrain", ctypes.c_void_p), ("nlayer", ctypes.c_int),
                ("layer_depth", ctypes.c_void_p), ("layer_vs", ctypes.c_void_p),
                ("layer_vp", ctypes.c_void_p), ("layer_rho", ctypes.c_void_p),
                ("vs", ctypes.c_void_p), ("vp", ctypes.c_void_p), ("rho", ctypes.c_void_p),
                ("source", ctypes.c_int), ("src_xyz", ctypes.c_int * 3), ("fc", ctypes.c_float),
                ("tab_xyz", ctypes.c_void_p), ("tab_rate", ctypes.c_void_p),
                ("tab_npts", ctypes.c_longlong), ("tab_nt", ctypes.c_int), ("tab_dt", ctypes.c_float),
                ("st_xyz", ctypes.c_void_p), ("nst", ctypes.c_int)]


class RequestInfo(ctypes.Structure):
    """Mirror of struct cgfdm_request_info."""
    _fields_ = [("status", ctypes.c_int), ("NT", ctypes.c_int), ("nst", ctypes.c_int),
                ("NX", ctypes.c_int), ("NY", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("device_bytes", ctypes.c_longlong), ("error", ctypes.c_char * 160),
                ("Cv", ctypes.c_double), ("Cs", ctypes.c_double), ("rescaled", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    """Mirror of struct cgfdm_server_stats."""
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong),
                ("failed", ctypes.c_longlong), ("launches", ctypes.c_longlong),
                ("resident", ctypes.c_int), ("pending", ctypes.c_int), ("peak_resident", ctypes.c_int),
                ("graph", ctypes.c_int), ("tile_cfg", ctypes.c_int), ("bulk_blocks", ctypes.c_int),
                ("general_blocks", ctypes.c_int),
                ("arena_bytes", ctypes.c_longlong), ("arena_used", ctypes.c_longlong),
                ("arena_peak", ctypes.c_longlong), ("rescaled", ctypes.c_longlong)]


class Server:
    """Continuous-batching server (cgfdm_server): submit requests, wait for completions, fetch
    results. Requests are dicts with keys desc, and either layers=(depth, vs, vp, rho) or
    medium=(vs, vp, rho); terrain (with desc.use_terrain); point=(X, Y, Z, fc) or
    table=(xyz, rate[nt][npts][6], dt); stations (n x 3). With 16-bit storage a request whose
    results come back non-finite is run again with 8x smaller scales, up to rescale_retries
    times (0: reported at once as CGFDM_ENONFINITE)."""

    def __init__(self, precision=FP16, device=0, hbm_budget=0, host_budget=0, max_domains=16,
                 steps_per_launch=8, use_graph=True, least_work_first=False, rescale_retries=2):
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        c = ServerConfig(device, precision, hbm_budget, host_budget, max_domains, steps_per_launch,
                         1 if use_graph else 0, 1 if least_work_first else 0, rescale_retries)
        h = ctypes.c_void_p()
        _chk(lib().cgfdm_server_create(ctypes.byref(c), ctypes.byref(h)))
        self.h = h
        self.meta = {}

    def submit(self, req):
        r = Request()
        r.desc = req["desc"]
        keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(x, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        if req.get("terrain") is not None:
            r.terrain = arr(req["terrain"], np.float64)
        if "layers" in req:
            dep, vs, vp, rho = req["layers"]
            r.nlayer = len(dep)
            r.layer_depth, r.layer_vs = arr(dep, np.float32), arr(vs, np.float32)
            r.layer_vp, r.layer_rho = arr(vp, np.float32), arr(rho, np.float32)
        else:
            vs, vp, rho = req["medium"]
            r.vs, r.vp, r.rho = arr(vs, np.float32), arr(vp, np.float32), arr(rho, np.float32)
        if "point" in req:
            X, Y, Z, fc = req["point"]
            r.source = 1
            r.src_xyz[0], r.src_xyz[1], r.src_xyz[2] = X, Y, Z
            r.fc = fc
        else:
            xyz, rate, dt = req["table"]
            rate = np.asarray(rate, np.float32)
            r.source = 2
            r.tab_xyz, r.tab_rate = arr(xyz, np.int32), arr(rate, np.float32)
            r.tab_npts, r.tab_nt, r.tab_dt = rate.shape[1], rate.shape[0], dt
        st = np.ascontiguousarray(req.get("stations", np.zeros((0, 3))), dtype=np.int32)
        r.st_xyz = arr(st, np.int32)
        r.nst = st.shape[0]
        rid = ctypes.c_longlong()
        _chk(lib().cgfdm_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid)))
        d = req["desc"]
        self.meta[rid.value] = (r.nst, d.NT, d.NX, d.NY)
        return rid.value

    def wait(self, max_ids=64, timeout_ms=-1):
        ids = np.zeros(max_ids, np.int64)
        n = lib().cgfdm_server_wait(self.h, _p(ids), max_ids, timeout_ms)
        if n < 0:
            _chk(n)
        return [int(x) for x in ids[:n]]

    def result(self, rid):
        """(status, stations, pgv, info dict) of a completed request; the record is released."""
        nst, NT, NX, NY = self.meta.pop(rid)
        st = np.zeros((nst, NT, 9), np.float32)
        pgv = np.zeros((NY, NX, 2), np.float32)
        info = RequestInfo()
        rc = lib().cgfdm_server_result(self.h, rid, _p(st), _p(pgv), ctypes.byref(info))
        i = {"status": info.status, "t_submit": info.t_submit, "t_admit": info.t_admit,
             "t_done": info.t_done, "device_bytes": info.device_bytes,
             "error": info.error.decode(errors="replace"), "Cv": info.Cv, "Cs": info.Cs,
             "rescaled": info.rescaled}
        return rc, st, pgv, i

    def stats(self):
        s = ServerStats()
        _chk(lib().cgfdm_server_stats_get(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().cgfdm_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
