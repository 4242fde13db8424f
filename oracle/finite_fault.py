"""Finite-fault source file -> moment-rate table on the grid: a numpy restatement of the
reference's multi-source set-up (multiSource.cpp, init_MultiSource :824-922 and the functions it
calls), used to check cgfdm_finite_fault_map.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker. The product never imports it.

Pins (tests/test_finite_fault.py): the projection against closed forms (equator arc, meridian arc
by quadrature, the direct geodesic problem integrated as an ODE, symmetry); the moment tensor
against Aki & Richards (trace 0, double-couple eigenvalues, strike/dip/rake special cases); the
node mapping against cell-centre and cell-edge cases; the merge of points sharing a node against
conservation of the summed rates.
"""
import numpy as np

WGS84_A = 6378137.0
WGS84_F = 1.0 / 298.257223563
REF_D2R = float(np.float32(3.141592657)) / 180.0  # the reference's DEGREE2RADIAN (macro.h:60-63)
HALO = 3


def vincenty_inverse(phi1, lam1, phi2, lam2):
    """Geodesic distance (m) and azimuth at the first point (radians) on WGS84, Vincenty's
    inverse solution (Survey Review 23(176), 1975). Angles in radians."""
    a, f = WGS84_A, WGS84_F
    b = a * (1.0 - f)
    L = lam2 - lam1
    U1 = np.arctan((1.0 - f) * np.tan(phi1))
    U2 = np.arctan((1.0 - f) * np.tan(phi2))
    sU1, cU1, sU2, cU2 = np.sin(U1), np.cos(U1), np.sin(U2), np.cos(U2)
    lam = L
    for _ in range(200):
        sl, cl = np.sin(lam), np.cos(lam)
        t1, t2 = cU2 * sl, cU1 * sU2 - sU1 * cU2 * cl
        sinS = np.sqrt(t1 * t1 + t2 * t2)
        if sinS == 0.0:
            return 0.0, 0.0
        cosS = sU1 * sU2 + cU1 * cU2 * cl
        sig = np.arctan2(sinS, cosS)
        sina = cU1 * cU2 * sl / sinS
        cos2a = 1.0 - sina * sina
        cos2sm = cosS - 2.0 * sU1 * sU2 / cos2a if cos2a != 0.0 else 0.0
        C = f / 16.0 * cos2a * (4.0 + f * (4.0 - 3.0 * cos2a))
        prev = lam
        lam = L + (1.0 - C) * f * sina * (sig + C * sinS * (cos2sm + C * cosS * (-1.0 + 2.0 * cos2sm * cos2sm)))
        if abs(lam - prev) < 1e-13:
            u2 = cos2a * (a * a - b * b) / (b * b)
            A = 1.0 + u2 / 16384.0 * (4096.0 + u2 * (-768.0 + u2 * (320.0 - 175.0 * u2)))
            B = u2 / 1024.0 * (256.0 + u2 * (-128.0 + u2 * (74.0 - 47.0 * u2)))
            dsig = B * sinS * (cos2sm + B / 4.0 * (cosS * (-1.0 + 2.0 * cos2sm * cos2sm)
                                                   - B / 6.0 * cos2sm * (-3.0 + 4.0 * sinS * sinS)
                                                   * (-3.0 + 4.0 * cos2sm * cos2sm)))
            s = b * A * (sig - dsig)
            sl, cl = np.sin(lam), np.cos(lam)
            return s, np.arctan2(cU2 * sl, cU1 * sU2 - sU1 * cU2 * cl)
    raise ValueError("geodesic did not converge")


def aeqd(lon0, lat0, lon_rad, lat_rad):
    """Azimuthal equidistant projection (PROJ "+proj=aeqd +ellps=WGS84", multiSource.cpp:149-170):
    centre in degrees, point in radians -> (x east, y north) metres."""
    s, az = vincenty_inverse(np.radians(lat0), np.radians(lon0), lat_rad, lon_rad)
    return s * np.sin(az), s * np.cos(az)


def write_fault(path, lon, lat, z, area, strike, dip, rake, rate, dt):
    """Writes a finite-fault file in the reference's layout (read by readSourceInfo,
    readSourceLocation, readSourceParams: int32 npts, int32 nt, float32 dt, then per point
    lon, lat, z, area, strike, dip, rake[nt], rate[nt] as float32)."""
    npts, nt = rake.shape
    with open(path, "wb") as f:
        np.array([npts, nt], np.int32).tofile(f)
        np.array([dt], np.float32).tofile(f)
        for p in range(npts):
            np.array([lon[p], lat[p], z[p], area[p], strike[p], dip[p]], np.float32).tofile(f)
            rake[p].astype(np.float32).tofile(f)
            rate[p].astype(np.float32).tofile(f)


def read_fault(path):
    with open(path, "rb") as f:
        npts, nt = np.fromfile(f, np.int32, 2)
        dt = float(np.fromfile(f, np.float32, 1)[0])
        rec = np.fromfile(f, np.float32, int(npts) * (6 + 2 * int(nt))).reshape(int(npts), 6 + 2 * int(nt))
    return rec, int(nt), dt


def coord_z(desc, terrain, i, j, k):
    """Grid z of padded node (i, j, k) as float32, as construct_flat_coord /
    construct_terrain_coord store it (coord.cpp:106-145, 258-290)."""
    f32 = np.float32
    if desc.use_terrain:
        depth = f32(abs(f32(desc.depth)))
        height = f32(terrain[j, i])
        DZ = f32(float(f32(height + depth)) / float(desc.NZ - 1))
        return f32(-depth + f32(DZ * f32(k - HALO)))
    DH = f32(desc.DH)
    return f32(f32(f32(k - HALO + 1) * DH) - f32(f32(desc.NZ) * DH))


def moment_tensor(strike_deg, dip_deg, rake_deg):
    """solveMomentRate (multiSource.cpp:478-545) for one point and time: (xx, yy, zz, xy, xz, yz)
    per unit area and rate, with the reference's float/double mix (angles in float radians,
    sin/cos of an angle in float, of twice an angle in double)."""
    f32 = np.float32
    s, d, r = (f32(v * REF_D2R) for v in (strike_deg, dip_deg, rake_deg))
    sd, cd, sr, cr = f32(np.sin(d)), f32(np.cos(d)), f32(np.sin(r)), f32(np.cos(r))
    ss, cs = f32(np.sin(s)), f32(np.cos(s))
    s2s, c2s = np.sin(2.0 * float(s)), np.cos(2.0 * float(s))
    s2d, c2d = np.sin(2.0 * float(d)), np.cos(2.0 * float(d))
    M11 = f32(-(float(f32(sd * cr)) * s2s + s2d * float(sr) * float(ss) * float(ss)))
    M22 = f32(float(f32(sd * cr)) * s2s - s2d * float(sr) * float(cs) * float(cs))
    M33 = f32(-f32(M11 + M22))
    M12 = f32(float(f32(sd * cr)) * c2s + 0.5 * s2d * float(sr) * s2s)
    M13 = f32(-(float(f32(f32(cd * cr) * cs)) + c2d * float(sr) * float(ss)))
    M23 = f32(-(float(f32(f32(cd * cr) * ss)) - c2d * float(sr) * float(cs)))
    return M22, M11, M33, M12, f32(-M23), f32(-M13)


def finite_fault_map(path, desc, terrain, lon0, lat0):
    """(xyz [n, 3] interior indices, rate [nt, n, 6], dt) of the file on the grid of desc."""
    f32 = np.float32
    rec, nt, dt = read_fault(path)
    npts = rec.shape[0]
    nx_, ny_ = desc.NX + 2 * HALO, desc.NY + 2 * HALO
    DH = f32(desc.DH)
    # projTrans (multiSource.cpp:109-173)
    x = np.zeros(npts, np.float32)
    y = np.zeros(npts, np.float32)
    for p in range(npts):
        xx, yy = aeqd(lon0, lat0, float(rec[p, 0]) * REF_D2R, float(rec[p, 1]) * REF_D2R)
        x[p], y[p] = f32(xx), f32(yy)
    # srcCoord2Index (:177-320)
    z = rec[:, 2].astype(np.float32)
    zmax = max(f32(-f32(desc.depth)), z.max()) if npts else f32(0)
    z = (z - zmax).astype(np.float32)
    point_node = {}
    # the NO_SOURCE_SMOOTH branch (macro.h:76 defines it; multiSource.cpp:243-264): interior
    # columns [HALO, nx+HALO) and interior layers [HALO, nz+HALO) only, other points dropped
    for p in range(npts):
        sx = int(float(f32(x[p] / DH)) + 0.5) + desc.centerX + HALO
        sy = int(float(f32(y[p] / DH)) + 0.5) + desc.centerY + HALO
        if not (HALO <= sx < desc.NX + HALO and HALO <= sy < desc.NY + HALO):
            continue
        for k in range(HALO, desc.NZ + HALO):
            zk = coord_z(desc, terrain, sx, sy, k)
            dz1 = float(f32(zk - coord_z(desc, terrain, sx, sy, k - 1))) * 0.5
            dz2 = float(f32(coord_z(desc, terrain, sx, sy, k + 1) - zk)) * 0.5
            if float(zk) - dz1 <= float(z[p]) < dz2 + float(zk):
                point_node[p] = (k * ny_ + sy) * nx_ + sx
    # readSourceParams + solveMomentRate (:365-545), points in increasing order
    pts = sorted(point_node)
    mr = np.zeros((len(pts), nt, 6), np.float32)
    for q, p in enumerate(pts):
        area, strike, dip = (f32(v) for v in rec[p, 3:6])
        for it in range(nt):
            M = moment_tensor(strike, dip, rec[p, 6 + it])
            rt = f32(rec[p, 6 + nt + it])
            mr[q, it] = [f32(f32(m * area) * rt) for m in M]
    # dealDuplicateIndex (:551-607): a node keeps its last point; earlier ones added in order
    last = {}
    for q, p in enumerate(pts):
        last[point_node[p]] = q
    for q, p in enumerate(pts):
        L = last[point_node[p]]
        if L > q:
            mr[L] = (mr[L] + mr[q]).astype(np.float32)
    # changeStorageOrder (:752-796): nodes in increasing order, rates [nt][node][6]
    nodes = sorted(last)
    xyz = np.zeros((len(nodes), 3), np.int32)
    rate = np.zeros((nt, len(nodes), 6), np.float32)
    for J, node in enumerate(nodes):
        k, rem = divmod(node, nx_ * ny_)
        xyz[J] = (rem % nx_ - HALO, rem // nx_ - HALO, k - HALO)
        rate[:, J] = mr[last[node]]
    return xyz, rate, dt
