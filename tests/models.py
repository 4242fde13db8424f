"""Synthetic models shared by the tests and the bench (seeded, reproducible)."""
import numpy as np


def homogeneous(NX, NY, NZ, vs=3464.0, vp=6000.0, rho=2670.0):
    """Reference default medium (medium.cpp:75-77)."""
    shape = (NZ, NY, NX)
    return (np.full(shape, vs, np.float32), np.full(shape, vp, np.float32),
            np.full(shape, rho, np.float32))


def layered_on_depth(z, tops_km, vs_list, ratio=1.73):
    """Layered medium by depth below the surface; Vp = ratio*Vs, rho = 1740*(Vp/1000)^0.25."""
    vs = np.full(z.shape, vs_list[0], np.float64)
    for top, v in zip(tops_km[1:], vs_list[1:]):
        vs = np.where(z <= -top * 1e3, v, vs)
    vp = ratio * vs
    rho = 1740.0 * (vp / 1000.0) ** 0.25
    return vs.astype(np.float32), vp.astype(np.float32), rho.astype(np.float32)


def gauss_hill(NX, NY, DH, h0=1500.0, width=6000.0, cx=None, cy=None):
    """Surface height z = h0*exp(-r^2/width^2) on the padded (NY+6, NX+6) grid."""
    cx = NX // 2 if cx is None else cx
    cy = NY // 2 if cy is None else cy
    x = (np.arange(NX + 6) - 3 - cx) * DH
    y = (np.arange(NY + 6) - 3 - cy) * DH
    X, Y = np.meshgrid(x, y)
    return h0 * np.exp(-(X ** 2 + Y ** 2) / width ** 2)


def smooth_random_terrain(NX, NY, DH, rms=1000.0, corr=5000.0, seed=0):
    """Seeded Gaussian-correlated random height map on the padded grid, scaled to `rms`."""
    rng = np.random.default_rng(seed)
    n1, n0 = NY + 6, NX + 6
    noise = rng.standard_normal((n1, n0))
    kx = np.fft.fftfreq(n0, d=DH)
    ky = np.fft.fftfreq(n1, d=DH)
    KX, KY = np.meshgrid(kx, ky)
    filt = np.exp(-(np.pi * corr) ** 2 * (KX ** 2 + KY ** 2))
    h = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))
    h -= h.mean()
    h *= rms / max(h.std(), 1e-30)
    return h


def terrain_z(terrain, NZ, depth):
    """z of the terrain-following grid at the interior points, (NZ, NY, NX)."""
    h = terrain[3:-3, 3:-3]
    K = np.arange(NZ, dtype=np.float64)[:, None, None]
    D = abs(depth)
    return -D + (h[None] + D) / (NZ - 1) * K


def flat_z(NX, NY, NZ, DH):
    K = np.arange(NZ, dtype=np.float64)
    z = (K + 1) * DH - NZ * DH
    return np.broadcast_to(z[:, None, None], (NZ, NY, NX))


def smooth_state(NX, NY, NZ, seed=0, vamp=1e-2, tamp=1e5, Cv=1.0, Cs=1.0, nmodes=4):
    """A smooth random wavefield (9, NZ, NY, NX) in storage units: sums of a few long waves."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    out = np.zeros((9, NZ, NY, NX), np.float64)
    for v in range(9):
        acc = np.zeros((NZ, NY, NX))
        for _ in range(nmodes):
            kx, ky, kz = rng.uniform(0.05, 0.35, 3)
            ph = rng.uniform(0, 2 * np.pi, 3)
            acc += rng.standard_normal() * np.sin(kx * x + ph[0]) * np.sin(ky * y + ph[1]) * np.sin(kz * z + ph[2])
        out[v] = acc * ((vamp * Cv) if v < 3 else (tamp * Cs))
    return out


def surface_stations(NX, NY, NZ, n=5, margin=None):
    """An n x n grid of stations on the free surface (Z = NZ-1)."""
    margin = max(2, NX // 8) if margin is None else margin
    xs = np.linspace(margin, NX - 1 - margin, n).astype(int)
    ys = np.linspace(margin, NY - 1 - margin, n).astype(int)
    return np.array([[x, y, NZ - 1] for y in ys for x in xs], np.int32)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
