"""Seeded synthetic workloads: terrain height maps, layered and perturbed media, double-couple
source tables and station grids, generated slab by slab so that 1001x1001x401 grids fit in a few
GB of host memory. Used by bench.py, the tests and the CLI driver.
"""
import math

import numpy as np

PI_F = float(np.float32(3.141592657))  # the reference's PI (macro.h:60)


# ------------------------------------------------------------------ terrain
def gauss_hill(NX, NY, DH, h0=1500.0, width=6000.0, cx=None, cy=None):
    """z = h0*exp(-r^2/width^2) on the padded (NY+6, NX+6) grid, centred on (cx, cy)."""
    cx = NX // 2 if cx is None else cx
    cy = NY // 2 if cy is None else cy
    x = (np.arange(NX + 6) - 3 - cx) * DH
    y = (np.arange(NY + 6) - 3 - cy) * DH
    X, Y = np.meshgrid(x, y)
    return h0 * np.exp(-(X ** 2 + Y ** 2) / width ** 2)


def smooth_random_terrain(NX, NY, DH, rms=1000.0, corr=5000.0, seed=0):
    """Gaussian-filtered white noise on the padded grid, zero mean, scaled to `rms` metres."""
    rng = np.random.default_rng(seed)
    n1, n0 = NY + 6, NX + 6
    noise = rng.standard_normal((n1, n0))
    kx = np.fft.fftfreq(n0, d=DH)
    ky = np.fft.fftfreq(n1, d=DH)
    KX, KY = np.meshgrid(kx, ky)
    filt = np.exp(-(math.pi * corr) ** 2 * (KX ** 2 + KY ** 2))
    h = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))
    h -= h.mean()
    h *= rms / max(h.std(), 1e-30)
    return h


def terrain_z_slab(terrain, NZ, depth, k):
    """z at interior layer k of the terrain-following grid (terrain.cpp:930-937), (NY, NX)."""
    h = terrain[3:-3, 3:-3]
    D = abs(depth)
    return -D + (h + D) / (NZ - 1) * k


# ------------------------------------------------------------------ media
LAYERS_VS = (2000.0, 3000.0, 3500.0)
LAYERS_TOP_KM = (0.0, 3.0, 10.0)


def layered_vs(depth_below_surface_m, tops_km=LAYERS_TOP_KM, vs=LAYERS_VS):
    out = np.full(depth_below_surface_m.shape, vs[0], np.float64)
    for top, v in zip(tops_km[1:], vs[1:]):
        out = np.where(depth_below_surface_m >= top * 1e3, v, out)
    return out


def vp_rho_from_vs(vs, ratio=1.73):
    vp = ratio * vs
    rho = 1740.0 * (vp / 1000.0) ** 0.25
    return vp, rho


class CoarseNoise:
    """Trilinear interpolation of seeded white noise on a coarse grid of spacing `corr` cells:
    a smooth random field with about that correlation length, unit standard deviation."""

    def __init__(self, NX, NY, NZ, corr_cells, seed):
        rng = np.random.default_rng(seed)
        self.c = corr_cells
        self.g = rng.standard_normal((NZ // corr_cells + 2, NY // corr_cells + 2, NX // corr_cells + 2))
        self.Ax = self._interp_matrix(NX, self.g.shape[2])
        self.Ay = self._interp_matrix(NY, self.g.shape[1])
        # trilinear interpolation of unit white noise has variance (2/3)^3 at random offsets
        self.norm = 1.0 / math.sqrt((2.0 / 3.0) ** 3)

    def _interp_matrix(self, n, nc):
        A = np.zeros((n, nc))
        x = np.arange(n) / self.c
        i0 = np.floor(x).astype(int)
        f = x - i0
        A[np.arange(n), i0] = 1 - f
        A[np.arange(n), np.minimum(i0 + 1, nc - 1)] += f
        return A

    def slab(self, k):
        z = k / self.c
        k0 = int(math.floor(z))
        f = z - k0
        plane = (1 - f) * self.g[k0] + f * self.g[min(k0 + 1, self.g.shape[0] - 1)]
        return (self.Ay @ plane @ self.Ax.T) * self.norm


def medium_slab(terrain, NZ, depth, k, DH, flat=False, noise=None, pert=0.0):
    """Vs, Vp, rho of interior layer k: 3-layer model by depth below the local surface, with an
    optional multiplicative random perturbation of standard deviation `pert`."""
    if flat:
        NY, NX = terrain
        z = np.full((NY, NX), (k + 1) * DH - NZ * DH)
        surf = 0.0
    else:
        z = terrain_z_slab(terrain, NZ, depth, k)
        surf = terrain[3:-3, 3:-3]
    vs = layered_vs(surf - z)
    if noise is not None and pert > 0:
        vs = vs * (1.0 + pert * noise.slab(k))
    vp, rho = vp_rho_from_vs(vs)
    return vs.astype(np.float32), vp.astype(np.float32), rho.astype(np.float32)


def build_medium(NX, NY, NZ, DH, terrain=None, depth=None, pert=0.0, corr_m=1000.0, seed=0):
    """Full (NZ, NY, NX) float32 arrays, generated layer by layer."""
    vs = np.empty((NZ, NY, NX), np.float32)
    vp = np.empty_like(vs)
    rho = np.empty_like(vs)
    noise = CoarseNoise(NX, NY, NZ, max(1, int(round(corr_m / DH))), seed) if pert > 0 else None
    for k in range(NZ):
        if terrain is None:
            a, b, c = medium_slab((NY, NX), NZ, depth, k, DH, flat=True, noise=noise, pert=pert)
        else:
            a, b, c = medium_slab(terrain, NZ, depth, k, DH, noise=noise, pert=pert)
        vs[k], vp[k], rho[k] = a, b, c
    return vs, vp, rho


# ------------------------------------------------------------------ sources
def moment_tensor(strike, dip, rake):
    """(xx, yy, zz, xy, xz, yz) in x=East, y=North, z=Up of a unit double couple
    (solveMomentRate, multiSource.cpp:531-543, with the reference's float PI)."""
    d2r = PI_F / 180.0
    s, d, r = strike * d2r, dip * d2r, rake * d2r
    M11 = -(math.sin(d) * math.cos(r) * math.sin(2 * s) + math.sin(2 * d) * math.sin(r) * math.sin(s) ** 2)
    M22 = math.sin(d) * math.cos(r) * math.sin(2 * s) - math.sin(2 * d) * math.sin(r) * math.cos(s) ** 2
    M33 = -(M11 + M22)
    M12 = math.sin(d) * math.cos(r) * math.cos(2 * s) + 0.5 * math.sin(2 * d) * math.sin(r) * math.sin(2 * s)
    M13 = -(math.cos(d) * math.cos(r) * math.cos(s) + math.cos(2 * d) * math.sin(r) * math.sin(s))
    M23 = -(math.cos(d) * math.cos(r) * math.sin(s) - math.cos(2 * d) * math.sin(r) * math.cos(s))
    return np.array([M22, M11, M33, M12, -M23, -M13])


def ricker(t, fc, t0=None):
    t0 = 1.2 / fc if t0 is None else t0
    r = math.pi * fc * (t - t0)
    return (1 - 2 * r * r) * np.exp(-r * r)


def double_couple_table(M0, strike, dip, rake, fc, nt, dt, mu):
    """One-point moment-rate table [nt][1][6] such that after the library's mu*Cv/(J*DH^3)
    scaling the injected moment-rate tensor is M0 * ricker(t) * M(strike, dip, rake)."""
    t = np.arange(nt) * dt
    m = moment_tensor(strike, dip, rake)
    rate = (M0 / mu) * ricker(t, fc)[:, None] * m[None, :]
    return rate[:, None, :].astype(np.float32)


def surface_stations(NX, NY, NZ, n=21, margin=None):
    """An n x n grid of receivers on the free surface (Z = NZ-1)."""
    margin = max(2, NX // 8) if margin is None else margin
    xs = np.linspace(margin, NX - 1 - margin, n).astype(int)
    ys = np.linspace(margin, NY - 1 - margin, n).astype(int)
    return np.array([[x, y, NZ - 1] for y in ys for x in xs], np.int32)


# ------------------------------------------------------------------ serving request stream
def layered_profile(n_layers, rng, vs_range=(1500.0, 4000.0), max_depth_m=15000.0):
    """A random 1-D layered model: n_layers tops (first at 0 m, then increasing), Vs drawn in
    vs_range and sorted to increase with depth, Vp = 1.73 Vs, rho = 1740 (Vp/1000)^0.25."""
    tops = np.concatenate([[0.0], np.sort(rng.uniform(500.0, max_depth_m, n_layers - 1))])
    vs = np.sort(rng.uniform(*vs_range, n_layers))
    vp, rho = vp_rho_from_vs(vs)
    return (tops.astype(np.float32), vs.astype(np.float32), vp.astype(np.float32),
            rho.astype(np.float32))


def layer_mu_at(depth_m, layers):
    """mu = rho Vs^2 of the layer containing depth_m (below the surface)."""
    tops, vs, vp, rho = layers
    q = int(np.searchsorted(tops, depth_m, side="right")) - 1
    q = max(0, min(q, len(tops) - 1))
    return float(rho[q]) * float(vs[q]) ** 2


def serving_stream(n, seed, rate_hz=20.0, edge_xy=(101, 301), edge_z=(51, 151), layers=(2, 5),
                   vs_range=(1500.0, 4000.0), depth_km=(1.0, 15.0), fc=(0.5, 3.0), steps=(500, 1500),
                   DH=100.0, DT=0.008, nPML=12, precision=2, n_st=5, M0=1e15):
    """Seeded Green's-function-style request stream: Poisson arrivals at rate_hz, grid edges
    uniform in edge_xy (x, y) and edge_z (z), random layered media, random double couples at a
    random depth (kept inside the grid above the bottom PML), Ricker moment rate of random peak
    frequency, random step counts, an n_st x n_st surface station grid. M0 = 1e15 N m (Mw 4):
    with the reference's FP16 scales (Cv = 5e3, Cs = 1e-6) a point source of 1e17 N m in one
    100 m cell drives the velocity at that cell past the FP16 range (|V| < 13 m/s). Returns a list of
    (arrival_s, request) with request dicts as cgfdm.Server.submit takes them (desc is built
    by the caller from request['grid'] etc. to avoid importing the binding here)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for r in range(n):
        t += rng.exponential(1.0 / rate_hz)
        NX = int(rng.integers(edge_xy[0], edge_xy[1] + 1))
        NY = int(rng.integers(edge_xy[0], edge_xy[1] + 1))
        NZ = int(rng.integers(edge_z[0], edge_z[1] + 1))
        nl = int(rng.integers(layers[0], layers[1] + 1))
        lay = layered_profile(nl, rng, vs_range, max_depth_m=NZ * DH)
        NT = int(rng.integers(steps[0], steps[1] + 1))
        f = float(rng.uniform(*fc))
        strike, dip, rake = float(rng.uniform(0, 360)), float(rng.uniform(10, 90)), float(rng.uniform(-180, 180))
        dep = float(rng.uniform(*depth_km)) * 1e3
        kmin = nPML + 4  # stay above the bottom PML with the 7-point stencil
        Z = int(min(max(NZ - 1 - round(dep / DH), kmin), NZ - 4))
        X = int(rng.integers(NX // 4, 3 * NX // 4))
        Y = int(rng.integers(NY // 4, 3 * NY // 4))
        mu = layer_mu_at((NZ - 1 - Z) * DH, lay)
        table = double_couple_table(M0, strike, dip, rake, f, NT + 2, DT, mu)
        st = surface_stations(NX, NY, NZ, n=n_st)
        out.append((t, {"grid": (NX, NY, NZ), "DH": DH, "DT": DT, "NT": NT, "nPML": nPML,
                        "precision": precision, "layers": lay,
                        "table": (np.array([[X, Y, Z]], np.int32), table, DT), "stations": st}))
    return out


# ------------------------------------------------------------------ kinematic finite-fault rupture
WGS84_A = 6378137.0
WGS84_F = 1.0 / 298.257223563


def fractal_terrain(NX, NY, DH, hurst=0.8, relief=2000.0, min_wavelength=1000.0, seed=0):
    """Seeded self-affine (fractional Brownian) surface on the padded (NY+6, NX+6) grid: spectral
    synthesis with amplitude |k|^-(hurst+1) (power |k|^-(2 hurst + 2)) and random phases, a
    Gaussian roll-off exp(-(|k| min_wavelength)^2) that removes wavelengths below min_wavelength
    (so the terrain-following grid stays smooth on the cell scale), shifted and scaled to heights
    in [0, relief] metres."""
    rng = np.random.default_rng(seed)
    n1, n0 = NY + 6, NX + 6
    kx = np.fft.fftfreq(n0, d=DH)
    ky = np.fft.fftfreq(n1, d=DH)
    K = np.hypot(*np.meshgrid(kx, ky))
    amp = np.zeros_like(K)
    nz = K > 0
    amp[nz] = K[nz] ** -(hurst + 1.0) * np.exp(-(K[nz] * min_wavelength) ** 2)
    spec = amp * (rng.standard_normal(K.shape) + 1j * rng.standard_normal(K.shape))
    h = np.real(np.fft.ifft2(spec))
    h -= h.min()
    return h * (relief / max(h.max(), 1e-30))


def von_karman_slip(n_strike, n_dip, d_strike, d_dip, corr_strike, corr_dip, hurst, mean, seed, taper=0.15):
    """Seeded random slip (m) on an n_dip x n_strike subfault grid with the von Karman spectrum
    P(k) = 1 / (1 + (a_s k_s)^2 + (a_d k_d)^2)^(hurst + 1) (Mai and Beroza 2002, k in rad/m, a the
    correlation lengths), made non-negative, tapered to zero over `taper` of the fault at all four
    edges (a buried rupture) with a cosine, and scaled to the given mean."""
    rng = np.random.default_rng(seed)
    ks = 2 * math.pi * np.fft.fftfreq(n_strike, d=d_strike)
    kd = 2 * math.pi * np.fft.fftfreq(n_dip, d=d_dip)
    KS, KD = np.meshgrid(ks, kd)
    P = 1.0 / (1.0 + (corr_strike * KS) ** 2 + (corr_dip * KD) ** 2) ** (hurst + 1.0)
    phase = np.exp(2j * math.pi * rng.random(P.shape))
    f = np.real(np.fft.ifft2(np.sqrt(P) * phase))
    f -= f.min()

    def edge(n):
        w = np.ones(n)
        m = max(1, int(round(taper * n)))
        ramp = 0.5 * (1 - np.cos(math.pi * (np.arange(m) + 0.5) / m))
        w[:m] = ramp
        w[n - m:] = ramp[::-1]
        return w

    f *= edge(n_dip)[:, None] * edge(n_strike)[None, :]
    return f * (mean / f.mean())


def triangle_slip_rate(t, t0, rise, slip):
    """Slip rate (m/s) of a triangle source-time function starting at t0 with rise time `rise`
    whose integral is `slip`; arrays broadcast (t along the last axis)."""
    u = (t - t0) / rise
    return np.where((u >= 0) & (u <= 1), 2.0 * slip / rise * (1.0 - np.abs(2.0 * u - 1.0)), 0.0)


def geodesic_direct(lon0, lat0, azimuth, distance):
    """End point (degrees) of the WGS84 geodesic from (lon0, lat0) (degrees) along azimuth
    (radians, clockwise from north) for `distance` metres: Vincenty's direct solution (Survey
    Review 23(176), 1975). The inverse of the azimuthal equidistant projection centred on
    (lon0, lat0): x = s sin(az), y = s cos(az)."""
    a, f = WGS84_A, WGS84_F
    b = a * (1 - f)
    az = np.asarray(azimuth, np.float64)
    s = np.asarray(distance, np.float64)
    phi1 = math.radians(lat0)
    tanU1 = (1 - f) * math.tan(phi1)
    cosU1 = 1 / math.sqrt(1 + tanU1 * tanU1)
    sinU1 = tanU1 * cosU1
    sinA1, cosA1 = np.sin(az), np.cos(az)
    sig1 = np.arctan2(tanU1, cosA1)
    sinA = cosU1 * sinA1
    cos2A = 1 - sinA * sinA
    u2 = cos2A * (a * a - b * b) / (b * b)
    A = 1 + u2 / 16384 * (4096 + u2 * (-768 + u2 * (320 - 175 * u2)))
    B = u2 / 1024 * (256 + u2 * (-128 + u2 * (74 - 47 * u2)))
    sig = s / (b * A)
    for _ in range(100):
        c2sm = np.cos(2 * sig1 + sig)
        sinS, cosS = np.sin(sig), np.cos(sig)
        dsig = B * sinS * (c2sm + B / 4 * (cosS * (-1 + 2 * c2sm * c2sm)
                                             - B / 6 * c2sm * (-3 + 4 * sinS * sinS) * (-3 + 4 * c2sm * c2sm)))
        new = s / (b * A) + dsig
        done = np.max(np.abs(new - sig)) < 1e-14
        sig = new
        if done:
            break
    c2sm = np.cos(2 * sig1 + sig)
    sinS, cosS = np.sin(sig), np.cos(sig)
    t = sinU1 * sinS - cosU1 * cosS * cosA1
    phi2 = np.arctan2(sinU1 * cosS + cosU1 * sinS * cosA1, (1 - f) * np.sqrt(sinA * sinA + t * t))
    lam = np.arctan2(sinS * sinA1, cosU1 * cosS - sinU1 * sinS * cosA1)
    C = f / 16 * cos2A * (4 + f * (4 - 3 * cos2A))
    L = lam - (1 - C) * f * sinA * (sig + C * sinS * (c2sm + C * cosS * (-1 + 2 * c2sm * c2sm)))
    return lon0 + np.degrees(L), np.degrees(phi2)


def aeqd_inverse(lon0, lat0, x, y):
    """(lon, lat) in degrees of the point at (x east, y north) metres in the azimuthal
    equidistant projection on WGS84 centred on (lon0, lat0)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return geodesic_direct(lon0, lat0, np.arctan2(x, y), np.hypot(x, y))


def write_fault_file(path, lon, lat, z, area, strike, dip, rake, rate, dt):
    """A finite-fault source file in the reference's layout (the output of
    constructFiniteFaultModel/readUSGS.py, read by multiSource.cpp:14-37 and
    cgfdm_finite_fault_map): int32 npts, int32 nt, float32 dt, then per point float32 lon, lat
    (degrees), z (m), area (m^2), strike, dip (degrees), rake[nt] (degrees), rate[nt] (m/s)."""
    rake = np.asarray(rake, np.float32)
    rate = np.asarray(rate, np.float32)
    npts, nt = rate.shape
    head = np.stack([lon, lat, z, area, strike, dip], axis=1).astype(np.float32)
    rec = np.concatenate([head, np.broadcast_to(rake, (npts, nt)), rate], axis=1).astype(np.float32)
    with open(path, "wb") as f:
        np.array([npts, nt], np.int32).tofile(f)
        np.array([dt], np.float32).tofile(f)
        rec.tofile(f)


def finite_fault_rupture(lon0=100.0, lat0=30.0, n_strike=80, n_dip=20, length=60e3, width=20e3, strike=90.0,
                         dip=60.0, rake=90.0, mean_slip=2.0, hurst=0.75, vr=2800.0, rise=(1.0, 3.0), dt=0.1,
                         y_top=5e3, z_top=-500.0, seed=11):
    """The kinematic rupture of the finite-fault workload: an n_dip x n_strike grid of subfaults on
    a length x width plane (strike 90: along x, east; dipping `dip` degrees to the south, top edge
    at y = y_top and depth -z_top), von Karman slip (correlation lengths 2 km + L/3 along strike
    and 1 km + W/3 down dip, Mai and Beroza 2002; Hurst `hurst`; mean `mean_slip` m), rupture
    times from a front at vr m/s out of a seeded random hypocentre, triangle slip-rate functions
    whose rise time grows with the square root of the slip from rise[0] to rise[1] s, constant
    rake. Returns the fields of write_fault_file plus the subfault x, y (m) and the hypocentre."""
    rng = np.random.default_rng(seed)
    ds, dd = length / n_strike, width / n_dip
    slip = von_karman_slip(n_strike, n_dip, ds, dd, 2e3 + length / 3, 1e3 + width / 3, hurst, mean_slip,
                           seed=seed + 1)
    s = (np.arange(n_strike) + 0.5) * ds                    # along strike from the west end
    w = (np.arange(n_dip) + 0.5) * dd                       # down dip from the top edge
    S, Wd = np.meshgrid(s, w)
    d = math.radians(dip)
    x = S - length / 2
    y = y_top - Wd * math.cos(d)
    z = z_top - Wd * math.sin(d)
    hs, hw = rng.uniform(0.15, 0.85) * length, rng.uniform(0.5, 0.85) * width
    t0 = np.hypot(S - hs, Wd - hw) / vr
    tr = rise[0] + (rise[1] - rise[0]) * np.sqrt(slip / slip.max())
    nt = int(math.ceil((t0.max() + tr.max()) / dt)) + 3
    t = np.arange(nt) * dt
    rate = triangle_slip_rate(t[None, :], t0.reshape(-1, 1), tr.reshape(-1, 1), slip.reshape(-1, 1))
    lon, lat = aeqd_inverse(lon0, lat0, x.ravel(), y.ravel())
    npts = x.size
    return {"lon": lon, "lat": lat, "z": z.ravel(), "area": np.full(npts, ds * dd), "strike": np.full(npts, strike),
            "dip": np.full(npts, dip), "rake": np.full((npts, nt), rake), "rate": rate, "dt": dt,
            "x": x.ravel(), "y": y.ravel(), "slip": slip.ravel(), "t0": t0.ravel(), "rise": tr.ravel(),
            "hypocentre": (hs - length / 2, y_top - hw * math.cos(d), z_top - hw * math.sin(d))}


def finite_fault_workload(DH=100.0, seed=2031, lon0=100.0, lat0=30.0, build=True):
    """The kinematic finite-fault workload on a 120 km x 120 km x 40 km box: 1201 x 1201 x 401
    points at DH = 100 m (241 x 241 x 81 at 500 m), DT = 8 ms * DH / 100 m (inside the
    reference's CFL limit 1.34 Hmin / Vpmax on this model) and 4000 steps at 100 m, i.e. 32 s of
    simulated time, fractal terrain (Hurst 0.8, 2 km relief, wavelengths
    below 10 DH removed), the 3-layer medium with a 5 % random perturbation of 1 km correlation,
    the rupture of finite_fault_rupture (80 x 20 subfaults, 60 km x 20 km, strike 90, dip 60, von
    Karman slip of mean 2 m, 2.8 km/s rupture front, triangle slip rates of 1-3 s rise time) with
    the projection centred on (lon0, lat0) at the grid centre, and 20 x 20 = 400 surface stations.
    With build=False the terrain and medium are left out (None)."""
    NX = NY = int(round(120e3 / DH)) + 1
    NZ = int(round(40e3 / DH)) + 1
    DT = 0.008 * DH / 100.0
    NT = int(round(32.0 / DT))
    depth = NZ * DH
    terrain = medium = None
    if build:
        terrain = fractal_terrain(NX, NY, DH, hurst=0.8, relief=2000.0, min_wavelength=10 * DH, seed=seed)
        medium = build_medium(NX, NY, NZ, DH, terrain=terrain, depth=depth, pert=0.05, corr_m=1000.0, seed=seed + 1)
    return {"grid": (NX, NY, NZ), "DH": DH, "DT": DT, "NT": NT, "depth": depth, "nPML": 12,
            "lon0": lon0, "lat0": lat0, "terrain": terrain, "medium": medium,
            "rupture": finite_fault_rupture(lon0, lat0, seed=seed + 2),
            "stations": surface_stations(NX, NY, NZ, n=20)}


def write_rupture(path, rupture):
    """Writes the rupture of finite_fault_rupture as a reference-layout source file."""
    r = rupture
    write_fault_file(path, r["lon"], r["lat"], r["z"], r["area"], r["strike"], r["dip"], r["rake"], r["rate"],
                     r["dt"])
