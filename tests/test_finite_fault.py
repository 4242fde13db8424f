"""Finite-fault source files (reference useMultiSource: multiSource.cpp): the oracle's projection,
moment tensor and node mapping pinned against closed forms, and the library's host-side
cgfdm_finite_fault_map / cgfdm_aeqd_forward checked against the oracle. CPU only."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import finite_fault as ff  # noqa: E402
import models  # noqa: E402

A, F = ff.WGS84_A, ff.WGS84_F
E2 = F * (2 - F)


@pytest.fixture(scope="module")
def cg():
    """The library's host-only entry points (no device call)."""
    import cgfdm
    cgfdm.lib()
    return cgfdm


def _meridian_arc(phi1, phi2):
    from scipy.integrate import quad
    return quad(lambda p: A * (1 - E2) / (1 - E2 * np.sin(p) ** 2) ** 1.5, phi1, phi2, epsabs=1e-9, epsrel=1e-13, limit=200)[0]


def _direct(phi, lam, az, s):
    """End point of the geodesic of length s leaving (phi, lam) with azimuth az, by integrating
    the geodesic equations of the ellipsoid of revolution (Clairaut form) in arc length."""
    from scipy.integrate import solve_ivp

    def rhs(_, u):
        p, l, a = u
        w = np.sqrt(1 - E2 * np.sin(p) ** 2)
        M, N = A * (1 - E2) / w ** 3, A / w
        return [np.cos(a) / M, np.sin(a) / (N * np.cos(p)), np.sin(a) * np.tan(p) / N]
    sol = solve_ivp(rhs, (0, s), [phi, lam, az], rtol=1e-12, atol=1e-14, method="DOP853")
    return sol.y[0, -1], sol.y[1, -1]


# ---------------------------------------------------------------- projection pins (oracle)
def test_aeqd_equator_and_meridian_closed_forms():
    for dlon in (0.1, 1.0, 3.0):
        x, y = ff.aeqd(10.0, 0.0, np.radians(10.0 + dlon), 0.0)
        assert x == pytest.approx(A * np.radians(dlon), abs=1e-6) and abs(y) < 1e-6
    for lat0, dlat in ((0.0, 1.0), (30.0, 0.5), (-45.0, 2.0)):
        x, y = ff.aeqd(20.0, lat0, np.radians(20.0), np.radians(lat0 + dlat))
        assert abs(x) < 1e-6
        assert y == pytest.approx(_meridian_arc(np.radians(lat0), np.radians(lat0 + dlat)), abs=1e-4)
    # WGS84 degree of latitude at the equator
    assert ff.aeqd(0.0, 0.0, 0.0, np.radians(1.0))[1] == pytest.approx(110574.3886, abs=1e-3)


@pytest.mark.parametrize("lat0,az_deg,s", [(30.0, 40.0, 120e3), (-12.0, 200.0, 80e3), (55.0, 300.0, 250e3)])
def test_aeqd_inverts_the_direct_geodesic(lat0, az_deg, s):
    """The point reached along azimuth az after distance s projects to (s sin az, s cos az)."""
    phi, lam = _direct(np.radians(lat0), np.radians(100.0), np.radians(az_deg), s)
    x, y = ff.aeqd(100.0, lat0, lam, phi)
    assert x == pytest.approx(s * np.sin(np.radians(az_deg)), abs=1e-3)
    assert y == pytest.approx(s * np.cos(np.radians(az_deg)), abs=1e-3)


def test_aeqd_symmetry_about_the_central_meridian():
    xe, ye = ff.aeqd(90.0, 35.0, np.radians(90.7), np.radians(35.2))
    xw, yw = ff.aeqd(90.0, 35.0, np.radians(89.3), np.radians(35.2))
    assert xe == pytest.approx(-xw, abs=1e-7) and ye == pytest.approx(yw, abs=1e-7)


# ---------------------------------------------------------------- moment tensor pins (oracle)
@pytest.mark.parametrize("sdr", [(0, 90, 0), (30, 60, 90), (217, 33, -40), (123, 81, 170)])
def test_moment_tensor_is_a_unit_double_couple(sdr):
    xx, yy, zz, xy, xz, yz = (float(v) for v in ff.moment_tensor(*sdr))
    M = np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])
    assert abs(np.trace(M)) < 1e-6
    assert np.allclose(np.sort(np.linalg.eigvalsh(M)), [-1, 0, 1], atol=1e-6)


def test_moment_tensor_special_cases():
    # vertical strike-slip on a north-striking fault: only the horizontal shear (xy) in the
    # east-north-up frame
    xx, yy, zz, xy, xz, yz = (float(v) for v in ff.moment_tensor(0.0, 90.0, 0.0))
    assert abs(xx) < 1e-7 and abs(yy) < 1e-7 and abs(zz) < 1e-7 and abs(xz) < 1e-7 and abs(yz) < 1e-7
    assert abs(abs(xy) - 1.0) < 1e-6
    # 45-degree thrust striking north: vertical-east dip-slip, zz = -xx, no y terms
    xx, yy, zz, xy, xz, yz = (float(v) for v in ff.moment_tensor(0.0, 45.0, 90.0))
    assert abs(yy) < 1e-6 and abs(xy) < 1e-6 and abs(yz) < 1e-6 and abs(xz) < 1e-6
    assert abs(zz + xx) < 1e-6 and abs(abs(zz) - 1.0) < 1e-6


# ---------------------------------------------------------------- node mapping (oracle + library)
def _desc(cg, NX=40, NY=36, NZ=30, use_terrain=0):
    return cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.01, NT=10, nPML=6, use_terrain=use_terrain, depth=NZ * 100.0)


def _fault(rng, lon0, lat0, npts, nt, spread_m=1500.0, zmin=-2200.0, zmax=-300.0):
    """Random points around the centre (in degrees via a local tangent plane), random geometry."""
    R = 6371e3
    dx, dy = rng.uniform(-spread_m, spread_m, (2, npts))
    lon = lon0 + np.degrees(dx / (R * np.cos(np.radians(lat0))))
    lat = lat0 + np.degrees(dy / R)
    z = rng.uniform(zmin, zmax, npts)
    area = rng.uniform(1e4, 4e4, npts)
    strike, dip = rng.uniform(0, 360, npts), rng.uniform(10, 90, npts)
    rake = rng.uniform(-180, 180, (npts, nt))
    rate = rng.uniform(0, 2, (npts, nt))
    return lon, lat, z, area, strike, dip, rake, rate


def test_cell_centre_and_edge_mapping(tmp_path, cg):
    """Points at a node's depth map to it; the shallowest point defines depth 0; a point outside
    the grid columns is dropped; points sharing a node are summed into one row."""
    d = _desc(cg)
    lon0, lat0 = 100.0, 30.0
    # three points on the centre column: depths 0 (the reference), -100 m, -100 m (duplicate)
    lon = np.array([lon0, lon0, lon0, lon0 + 1.0])
    lat = np.array([lat0, lat0, lat0, lat0])
    z = np.array([-500.0, -600.0, -600.0, -600.0])
    nt = 3
    rake = np.zeros((4, nt))
    rate = np.array([[1, 2, 3], [1, 1, 1], [2, 2, 2], [5, 5, 5]], float)
    path = tmp_path / "fault.bin"
    ff.write_fault(path, lon, lat, z, np.full(4, 1e4), np.zeros(4), np.full(4, 90.0), rake, rate, 0.1)
    xyz, r, dt = ff.finite_fault_map(path, d, None, lon0, lat0)
    # flat grid: top node (k = NZ-1) at z = 0; z = -100 m is one layer down
    assert [tuple(v) for v in xyz] == [(d.centerX, d.centerY, d.NZ - 2), (d.centerX, d.centerY, d.NZ - 1)]
    assert dt == pytest.approx(0.1)
    M = np.array([float(v) for v in ff.moment_tensor(0.0, 90.0, 0.0)])
    np.testing.assert_allclose(r[:, 1], np.outer(rate[0], M) * 1e4, rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(r[:, 0], np.outer(rate[1] + rate[2], M) * 1e4, rtol=1e-6, atol=1e-3)


@pytest.mark.parametrize("use_terrain", [0, 1])
def test_merge_conserves_the_summed_rates(tmp_path, cg, use_terrain):
    """With every point inside the grid, the merged table sums to the sum over all points of
    moment tensor x area x slip rate, and holds one row per node."""
    rng = np.random.default_rng(11 + use_terrain)
    d = _desc(cg, use_terrain=use_terrain)
    terrain = models.gauss_hill(d.NX, d.NY, d.DH, h0=200.0, width=900.0) if use_terrain else None
    lon0, lat0 = 102.5, 31.0
    lon, lat, z, area, strike, dip, rake, rate = _fault(rng, lon0, lat0, 60, 5, spread_m=250.0, zmin=-600.0)
    path = tmp_path / "fault.bin"
    ff.write_fault(path, lon, lat, z, area, strike, dip, rake, rate, 0.05)
    xyz, table, _ = ff.finite_fault_map(path, d, terrain, lon0, lat0)
    assert len({tuple(v) for v in xyz}) == len(xyz) < 60  # merged, one row per node
    want = np.zeros((5, 6))
    for p in range(60):
        for it in range(5):
            M = np.array([float(v) for v in ff.moment_tensor(np.float32(strike[p]), np.float32(dip[p]),
                                                            np.float32(rake[p, it]))])
            want[it] += M * float(np.float32(area[p])) * float(np.float32(rate[p, it]))
    np.testing.assert_allclose(table.astype(np.float64).sum(axis=1), want, rtol=1e-5, atol=1e-3 * np.abs(want).max())


@pytest.mark.parametrize("use_terrain", [0, 1])
def test_library_finite_fault_map_matches_oracle(tmp_path, cg, use_terrain):
    rng = np.random.default_rng(5 + use_terrain)
    d = _desc(cg, use_terrain=use_terrain)
    terrain = models.gauss_hill(d.NX, d.NY, d.DH, h0=250.0, width=1000.0) if use_terrain else None
    lon0, lat0 = 95.3, 28.7
    fault = _fault(rng, lon0, lat0, 80, 6, spread_m=2500.0)
    path = tmp_path / "fault.bin"
    ff.write_fault(path, *fault, 0.02)
    xo, ro, dto = ff.finite_fault_map(path, d, terrain, lon0, lat0)
    xg, rg, dtg = cg.finite_fault_map(path, d, terrain, lon0, lat0)
    assert dtg == pytest.approx(dto)
    np.testing.assert_array_equal(xg, xo)
    np.testing.assert_allclose(rg, ro, rtol=2e-6, atol=1e-6 * np.abs(ro).max())


def test_library_aeqd_matches_oracle(cg):
    rng = np.random.default_rng(3)
    lon = 120.0 + rng.uniform(-2, 2, 50)
    lat = -20.0 + rng.uniform(-2, 2, 50)
    x, y = cg.aeqd_forward(120.0, -20.0, lon, lat)
    for q in range(50):
        xo, yo = ff.aeqd(120.0, -20.0, np.radians(lon[q]), np.radians(lat[q]))
        assert x[q] == pytest.approx(xo, abs=1e-6) and y[q] == pytest.approx(yo, abs=1e-6)


def test_missing_file_is_an_error(cg, tmp_path):
    with pytest.raises(cg.CgfdmError):
        cg.finite_fault_map(tmp_path / "none.bin", _desc(cg), None, 0.0, 0.0)


@pytest.mark.parametrize("use_terrain", [0, 1])
def test_points_outside_the_interior_are_dropped(tmp_path, cg, use_terrain):
    """The reference compiles srcCoord2Index with NO_SOURCE_SMOOTH (macro.h:76,
    multiSource.cpp:243-264): only interior columns [HALO, N+HALO) and interior layers are kept.
    Points in the halo ring of columns, below the bottom interior layer or above the top one are
    dropped, not mapped onto halo nodes (which cgfdm_set_source_table would reject)."""
    d = _desc(cg, NX=20, NY=18, NZ=16, use_terrain=use_terrain)
    DH = d.DH
    # terrain: a basin 250 m deep at the centre column, so the shallowest point (depth 0 by
    # definition) lies above the surface there
    terrain = models.gauss_hill(d.NX, d.NY, DH, h0=-250.0, width=300.0) if use_terrain else None
    lon0, lat0 = 0.0, 0.0
    R = 6378137.0

    def lon_of(X):  # column X (interior index) on the equator
        return np.degrees((X - d.centerX) * DH / R)

    zc = -400.0  # the inside points, relative depth 0 once the shallowest point is subtracted
    # int(x/DH + 0.5) rounds toward zero, so west of the centre a point at column X lands on
    # column X+1: -2 and -4 land on the halo columns -1 and -3
    cols = [d.centerX, -2, -4, d.NX, d.NX + 2, d.centerX, d.centerX + 2]
    depth = [zc, zc - 100, zc - 100, zc - 100, zc - 100,
             zc - (d.NZ + 1) * DH,   # one layer below the bottom interior layer
             zc - (d.NZ + 2) * DH]   # two layers below
    if use_terrain:
        cols.append(d.centerX + 1)
        depth.append(zc - 50.0)
    cols.append(d.centerX)
    depth.append(zc - 600.0)  # inside on either grid
    n = len(cols)
    lon = np.array([lon_of(X) for X in cols])
    lat = np.zeros(n)
    nt = 2
    path = tmp_path / "edge.bin"
    ff.write_fault(path, lon, lat, np.array(depth), np.full(n, 1e4), np.zeros(n), np.full(n, 90.0),
                   np.zeros((n, nt)), np.ones((n, nt)), 0.1)
    xo, ro, _ = ff.finite_fault_map(path, d, terrain, lon0, lat0)
    xg, rg, _ = cg.finite_fault_map(path, d, terrain, lon0, lat0)
    np.testing.assert_array_equal(xg, xo)
    np.testing.assert_allclose(rg, ro, rtol=2e-6)
    lim = np.array([d.NX, d.NY, d.NZ])
    assert ((xo >= 0) & (xo < lim)).all(), xo
    got = [tuple(int(c) for c in v) for v in xo]
    if use_terrain:
        # the basin: the shallowest point and its neighbour 50 m down sit above the surface there
        # (surface at about -250 m, layers 90 m apart) and are dropped; the point 600 m down
        # lands 350 m below the surface, four layers under the top
        assert got == [(d.centerX, d.centerY, d.NZ - 5)]
    else:
        # the top layer of the centre column and the point six layers below it
        assert got == [(d.centerX, d.centerY, d.NZ - 7), (d.centerX, d.centerY, d.NZ - 1)]
