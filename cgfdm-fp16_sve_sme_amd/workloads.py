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
