"""Pins of the CPU oracle against what the mathematics fixes, independent of its own formulas.

The reference ships no numeric golden outputs and cannot be compiled here (every translation unit
includes <proj.h>, which the image lacks), so each oracle function is checked against closed
forms, invariants and textbook special cases chosen so that a dropped term, a wrong sign, a wrong
stencil offset or a transposed operand fails:

  operator L (axis.h:129-133)          polynomial moments of the stencil on lines
  RK4 (wave_rk.cpp:13-99)              amplification polynomial of y' = z y
  wave_deriv + metric (a4, a17, a18)   linear fields on flat and tilted-terrain grids give the
                                       textbook elastodynamic RHS at a point where they vanish
  free surface (a5)                    traction rate sigma_dot . grad(zeta) = 0 at the surface
                                       (lambda = mu, where the reference's chi = 2 lambda + mu is
                                       the physical lambda + 2 mu); odd stress imaging is exact
  PML profiles (a10)                   special cases nPML = 10, 20 where R is a power of ten
  Ricker source (a11)                  time derivative of the Ricker wavelet; weights sum to 1
  moment tensor (a13)                  M = n s^T + s n^T from fault geometry (Aki & Richards)
  table source (a12)                   injection = -rate*mu*DT/(J*DH^3) at t = it*DT, before stage 0
  stations / PGV / slices (a14-a16)    equal to the state sampled after each step
"""
import math

import numpy as np
import pytest

import models

F64 = 0


def _desc(cg, NX, NY, NZ, **kw):
    kw.setdefault("DH", 100.0)
    kw.setdefault("DT", 0.008)
    kw.setdefault("NT", 10)
    kw.setdefault("nPML", 6)
    return cg.make_desc(NX, NY, NZ, **kw)


@pytest.fixture(scope="module")
def cg():
    import cgfdm
    return cgfdm


# macro.h:126-130; the reference writes them as float literals, so they enter as float32 values
COEF = {m: float(np.float32(a)) for m, a in {-1: -0.30874, 0: -0.6326, 1: 1.2330, 2: -0.3334, 3: 0.04168}.items()}


# ---------------------------------------------------------------- operator L
@pytest.mark.parametrize("FB", [1, -1])
def test_L_polynomial_moments(oracle_lib, FB):
    po = oracle_lib
    DH = 50.0
    x = np.arange(40) * DH
    i0 = 17
    x0 = x[i0]
    # m-th moment of the stencil about the evaluation point: sum_m a_m (FB m)^p
    mom = [sum(a * (FB * m) ** p for m, a in COEF.items()) for p in range(3)]
    # constant: FB * sum(a)/DH, which is not zero (-6e-5)
    out = po.apply_L_line(np.ones_like(x), FB, 1.0 / DH)
    assert out[i0] == pytest.approx(FB * mom[0] / DH, rel=1e-12)
    assert abs(mom[0] + 6e-5) < 1e-7
    # a line through zero at x0: sum(a m) = 0.99998, for either direction
    out = po.apply_L_line(3.0 * (x - x0), FB, 1.0 / DH)
    assert out[i0] == pytest.approx(3.0 * FB * mom[1], rel=1e-12)
    assert FB * mom[1] == pytest.approx(0.99998, abs=1e-7)
    # a parabola: second moment, sign follows FB
    out = po.apply_L_line((x - x0) ** 2 / DH, FB, 1.0 / DH)
    assert out[i0] == pytest.approx(FB * mom[2], rel=1e-12)
    # forward/backward average kills the second moment (metric construction)
    fp = po.apply_L_line((x - x0) ** 2, 1, 1.0 / DH)[i0]
    fm = po.apply_L_line((x - x0) ** 2, -1, 1.0 / DH)[i0]
    assert abs(0.5 * (fp + fm)) < 1e-9


# ---------------------------------------------------------------- RK4
@pytest.mark.parametrize("z", [-0.3, 0.2, -1.1, 0.05])
def test_rk4_amplification(oracle_lib, z):
    po = oracle_lib
    n = 5
    W = np.linspace(1.0, 2.0, n)
    W0 = W.copy()
    t = np.zeros(n)
    m = np.zeros(n)
    for irk in range(4):
        h = z * W.copy()          # h = DT*lambda*W at the stage value
        po.wave_rk(irk, h, W, t, m)
    amp = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    # 0.16666667f and 0.33333333f differ from 1/6, 1/3 by ~1e-8 relative
    np.testing.assert_allclose(W, amp * W0, rtol=2e-7)


# ---------------------------------------------------------------- interior RHS on linear fields
def _linear_fields(rng, X, Y, Z, P):
    """Velocity V = G (x-P) and symmetric stress T = H.(x-P); returns (9,...) fields, G, H."""
    G = rng.standard_normal((3, 3))
    H = rng.standard_normal((3, 3, 3))
    H = 0.5 * (H + H.transpose(1, 0, 2))
    d = [X - P[0], Y - P[1], Z - P[2]]
    V = [sum(G[a, b] * d[b] for b in range(3)) for a in range(3)]
    T = {(a, b): sum(H[a, b, c] * d[c] for c in range(3)) for a in range(3) for b in range(3)}
    f = np.stack([V[0], V[1], V[2], T[0, 0], T[1, 1], T[2, 2], T[0, 1], T[0, 2], T[1, 2]])
    return f, G, H


def _textbook_rhs(G, H, lam, mu, b):
    """v_dot = b div(sigma), sigma_dot = lam tr(grad v) I + mu (grad v + grad v^T)."""
    vdot = b * np.array([sum(H[a, c, c] for c in range(3)) for a in range(3)])
    S = lam * np.trace(G) * np.eye(3) + mu * (G + G.T)
    return np.array([vdot[0], vdot[1], vdot[2], S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])


def _coords(desc, terrain):
    NX, NY, NZ = desc.NX, desc.NY, desc.NZ
    x = (np.arange(NX) - desc.centerX) * desc.DH
    y = (np.arange(NY) - desc.centerY) * desc.DH
    Z3, Y3, X3 = np.meshgrid(np.arange(NZ), y, x, indexing="ij")
    if terrain is None:
        z = models.flat_z(NX, NY, NZ, desc.DH)
    else:
        z = models.terrain_z(terrain, NZ, desc.depth)
    return X3, Y3, np.array(z)


@pytest.mark.parametrize("terrain_kind", ["flat", "tilted"])
@pytest.mark.parametrize("FB", [(1, 1, 1), (-1, 1, -1), (1, -1, -1)])
def test_interior_rhs_is_textbook_elastodynamics(cg, oracle_lib, terrain_kind, FB):
    po = oracle_lib
    NX, NY, NZ = 24, 22, 20
    DH = 100.0
    tilted = terrain_kind == "tilted"
    d = _desc(cg, NX, NY, NZ, pml=0, free_surface=0, use_terrain=int(tilted), depth=2500.0)
    terrain = None
    if tilted:
        xs = (np.arange(NX + 6) - 3 - d.centerX) * DH
        ys = (np.arange(NY + 6) - 3 - d.centerY) * DH
        XX, YY = np.meshgrid(xs, ys)
        terrain = 0.15 * XX - 0.08 * YY + 40.0
    vs, vp, rho = 2800.0, 5200.0, 2500.0
    o = po.Oracle(d, F64)
    o.set_model(terrain, *models.homogeneous(NX, NY, NZ, vs, vp, rho))
    X, Y, Z = _coords(d, terrain)
    rng = np.random.default_rng(7)
    for (pi, pj, pk) in [(12, 11, 10), (8, 14, 6)]:
        P = (X[pk, pj, pi], Y[pk, pj, pi], Z[pk, pj, pi])
        f, G, H = _linear_fields(rng, X, Y, Z, P)
        o.set_state(f)
        h = o.rhs(*FB)[:, pk, pj, pi] / d.DT
        mu = rho * vs ** 2
        lam = rho * (vp ** 2 - 2 * vs ** 2)
        ref = _textbook_rhs(G, H, lam, mu, 1.0 / rho)
        scale = np.abs(ref).max()
        np.testing.assert_allclose(h, ref, rtol=0, atol=1e-9 * scale)


@pytest.mark.parametrize("terrain_kind", ["flat", "hill", "rough"])
def test_metric_off_diagonals_vanish_on_terrain_following_grids(cg, oracle_lib, terrain_kind):
    """x depends on i only and y on j only, so the averaged forward/backward derivatives
    x_eta, x_zeta, y_xi, y_zeta cancel bit for bit and xi_y, xi_z, eta_x, eta_z (reference
    components 1, 2, 3, 5, built from products with them) are exactly zero; xi_x, eta_y,
    zeta_z are not. The stage tile kernels leave the zero terms out."""
    po = oracle_lib
    NX, NY, NZ = 22, 20, 14
    d = _desc(cg, NX, NY, NZ, pml=0, free_surface=1, use_terrain=int(terrain_kind != "flat"), depth=1400.0)
    terrain = None
    if terrain_kind == "hill":
        terrain = models.gauss_hill(NX, NY, 100.0, h0=300.0, width=700.0)
    elif terrain_kind == "rough":
        terrain = np.random.default_rng(5).normal(0.0, 80.0, (NY + 6, NX + 6))
    o = po.Oracle(d, F64)
    o.set_model(terrain, *models.homogeneous(NX, NY, NZ))
    c = o.cjm64()
    for comp in (1, 2, 3, 5):
        assert not c[comp].any(), comp
    for comp in (0, 4, 8):
        assert np.abs(c[comp][3:-3, 3:-3, 3:-3]).min() > 0.5, comp
    if terrain is not None:
        assert np.abs(c[6]).max() > 1e-3 and np.abs(c[7]).max() > 1e-3


# ---------------------------------------------------------------- free surface
@pytest.mark.parametrize("terrain_kind", ["flat", "tilted"])
def test_free_surface_traction_rate_vanishes(cg, oracle_lib, terrain_kind):
    po = oracle_lib
    NX, NY, NZ = 26, 24, 16
    DH = 100.0
    tilted = terrain_kind == "tilted"
    d = _desc(cg, NX, NY, NZ, pml=0, free_surface=1, use_terrain=int(tilted), depth=1500.0)
    terrain = None
    if tilted:
        xs = (np.arange(NX + 6) - 3 - d.centerX) * DH
        ys = (np.arange(NY + 6) - 3 - d.centerY) * DH
        XX, YY = np.meshgrid(xs, ys)
        terrain = 0.2 * XX + 0.1 * YY
    vs = 3000.0
    vp = math.sqrt(3.0) * vs  # lambda = mu
    o = po.Oracle(d, F64)
    o.set_model(terrain, *models.homogeneous(NX, NY, NZ, vs, vp, 2600.0))
    w = models.smooth_state(NX, NY, NZ, seed=11)
    o.set_state(w)
    cjm = o.cjm64()
    for FB in [(1, 1, 1), (-1, 1, -1), (1, -1, 1)]:
        h = o.rhs(*FB) / d.DT
        k = NZ - 1
        js, is_ = slice(8, NY - 8), slice(8, NX - 8)
        S = {(0, 0): h[3], (1, 1): h[4], (2, 2): h[5], (0, 1): h[6], (0, 2): h[7], (1, 2): h[8]}
        zt = [cjm[6 + c, k + 3, 3:-3, 3:-3] for c in range(3)]
        for a in range(3):
            tr = sum(S[tuple(sorted((a, b)))][k] * zt[b] for b in range(3))
            mag = np.abs(h[3:6, k]).max()
            # vp = sqrt(3) vs is rounded to float32, so lambda = mu only to ~1e-7
            assert np.abs(tr[js, is_]).max() <= 1e-7 * mag, (FB, a)


def test_free_surface_odd_stress_imaging_is_exact(cg, oracle_lib):
    po = oracle_lib
    NX, NY, NZ = 20, 20, 14
    d = _desc(cg, NX, NY, NZ, pml=0, free_surface=1)
    vs, vp, rho = 3000.0, 5400.0, 2400.0
    o = po.Oracle(d, F64)
    o.set_model(None, *models.homogeneous(NX, NY, NZ, vs, vp, rho))
    X, Y, Z = _coords(d, None)
    # stresses linear in z and zero on the surface (z = 0): odd about it, so the
    # antisymmetric image above the surface continues them exactly
    rng = np.random.default_rng(3)
    c = rng.standard_normal(6)
    w = np.zeros((9, NZ, NY, NX))
    for q in range(6):
        w[3 + q] = c[q] * Z
    o.set_state(w)
    H = np.zeros((3, 3, 3))
    for q, (a, b) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]):
        H[a, b, 2] = H[b, a, 2] = c[q]
    ref = _textbook_rhs(np.zeros((3, 3)), H, 0.0, 0.0, 1.0 / rho)[:3]
    for FB in [(1, 1, 1), (-1, -1, -1), (1, -1, 1)]:
        h = o.rhs(*FB) / d.DT
        got = h[:3, NZ - 1, NY // 2, NX // 2]
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


# ---------------------------------------------------------------- PML profiles
@pytest.mark.parametrize("nPML,logR", [(10, -3.0), (20, -4.0)])
def test_pml_profile_special_cases(cg, oracle_lib, nPML, logR):
    po = oracle_lib
    NX, NY, NZ = 50, 46, 44
    DH = 200.0
    d = _desc(cg, NX, NY, NZ, DH=DH, nPML=nPML, free_surface=1)
    o = po.Oracle(d, F64)
    o.set_model(None, *models.homogeneous(NX, NY, NZ))
    p = o.pml()
    L = DH * (nPML - 1)
    d0 = -3.0 * 6000.0 / (2.0 * L) * (logR * math.log(10.0))
    alpha0 = 3.141592657 * 0.5
    ax, bx, dx = p["alpha_x"], p["beta_x"], p["d_x"]
    # outermost PML cell: l = L -> alpha 0, beta 3 (stored 1/3), d = d0 (stored d0/3)
    assert ax[3] == pytest.approx(0.0, abs=1e-12)
    assert bx[3] == pytest.approx(1.0 / 3.0, rel=1e-12)
    assert dx[3] == pytest.approx(d0 / 3.0, rel=1e-9)
    # innermost PML cell: l = 0 -> alpha0, beta 1, d 0
    i1 = 3 + nPML - 1
    assert ax[i1] == pytest.approx(alpha0, rel=1e-7)
    assert bx[i1] == pytest.approx(1.0) and dx[i1] == pytest.approx(0.0, abs=1e-12)
    # mirror symmetry of the two x faces, and no damping inside
    n = NX + 6
    np.testing.assert_allclose(bx[3:3 + nPML], bx[n - 4:n - 4 - nPML:-1], rtol=1e-12)
    assert np.all(bx[3 + nPML:n - 3 - nPML] == 1.0) and np.all(dx[3 + nPML:n - 3 - nPML] == 0.0)
    # free surface: z is damped at the bottom only
    bz = p["beta_z"]
    assert bz[3] == pytest.approx(1.0 / 3.0, rel=1e-12) and bz[NZ + 2] == 1.0


# ---------------------------------------------------------------- sources
@pytest.mark.parametrize("fc", [0.8, 2.0])
def test_ricker_source_is_ricker_derivative(oracle_lib, fc):
    po = oracle_lib
    DT = 1e-3
    its = np.arange(0, int(3.0 / fc / DT), 37)
    s = np.array([po.source_function(fc, int(it), DT) for it in its]) / 1e16

    def ricker(t):
        r = math.pi * fc * (t - 1.2 / fc)
        return (1 - 2 * r * r) * math.exp(-r * r)

    eps = 1e-6
    ref = np.array([(ricker(it * DT + eps) - ricker(it * DT - eps)) / (2 * eps) for it in its])
    np.testing.assert_allclose(s, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    assert abs(po.source_function(fc, int(round(1.2 / fc / DT)), DT)) < 1e-3 * np.abs(s).max() * 1e16


def test_gauss_weights_partition_of_unity(oracle_lib):
    w = oracle_lib.gauss_weights()
    assert w.sum() == pytest.approx(1.0, abs=2e-9)
    assert np.allclose(w, w[::-1, :, :]) and np.allclose(w, w.transpose(2, 1, 0))
    assert w.argmax() == np.ravel_multi_index((3, 3, 3), w.shape)


@pytest.mark.parametrize("sdr", [(0, 90, 0), (30, 70, 20), (164, 46, 122), (90, 60, -90), (275, 35, 45)])
def test_moment_tensor_from_fault_geometry(oracle_lib, sdr):
    strike, dip, rake = sdr
    m = oracle_lib.moment_tensor(strike, dip, rake)
    M = np.array([[m[0], m[3], m[4]], [m[3], m[1], m[5]], [m[4], m[5], m[2]]])  # x=E, y=N, z=Up
    phi, dl, lam = np.radians([strike, dip, rake])
    # Aki & Richards: fault normal and slip in (North, East, Down)
    n_ned = np.array([-np.sin(dl) * np.sin(phi), np.sin(dl) * np.cos(phi), -np.cos(dl)])
    s_ned = np.array([np.cos(lam) * np.cos(phi) + np.cos(dl) * np.sin(lam) * np.sin(phi),
                      np.cos(lam) * np.sin(phi) - np.cos(dl) * np.sin(lam) * np.cos(phi),
                      -np.sin(lam) * np.sin(dl)])
    to_enu = np.array([[0, 1, 0], [1, 0, 0], [0, 0, -1]])
    n, s = to_enu @ n_ned, to_enu @ s_ned
    ref = np.outer(n, s) + np.outer(s, n)
    # the reference converts degrees with its float PI = 3.14159274: ~1e-7 absolute
    np.testing.assert_allclose(M, ref, atol=5e-7)
    assert abs(np.trace(M)) < 1e-12
    np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(M)), [-1, 0, 1], atol=5e-7)


def test_table_source_injection_scaling_and_timing(cg, oracle_lib):
    po = oracle_lib
    NX, NY, NZ = 16, 16, 16
    DH, DT = 100.0, 0.008
    d = _desc(cg, NX, NY, NZ, DH=DH, DT=DT, pml=0, free_surface=0)
    vs, vp, rho = 3000.0, 5000.0, 2500.0
    o = po.Oracle(d, F64)
    o.set_model(None, *models.homogeneous(NX, NY, NZ, vs, vp, rho))
    nt, dt = 50, 0.002
    t = np.arange(nt) * dt
    rate = np.zeros((nt, 1, 6), np.float32)
    comp = np.array([1.0, -2.0, 0.5, 0.7, -0.3, 0.9])
    rate[:, 0, :] = (1.0 + 10.0 * t)[:, None] * comp[None, :]   # linear in time
    P = (8, 7, 9)
    o.set_source_table(np.array([P], np.int32), rate, dt)
    it = 3
    for _ in range(it):           # advance the counter without fields: zero state stays zero
        o.step(1)
    o.set_state(np.zeros((9, NZ, NY, NX)))
    o.inject_source()
    w = o.state()
    J = 0.99998 ** 3              # flat grid: x_xi = first moment of the stencil
    mu = rho * vs * vs
    tnow = it * DT
    expect = -(1.0 + 10.0 * tnow) * comp * mu / (J * DH ** 3) * DT
    np.testing.assert_allclose(w[3:, P[2], P[1], P[0]], expect.astype(np.float64), rtol=2e-6)
    w[3:, P[2], P[1], P[0]] = 0
    assert np.all(w == 0)


# ---------------------------------------------------------------- stations / PGV / slices
def test_stations_and_pgv_sample_the_state(cg, oracle_lib):
    po = oracle_lib
    NX, NY, NZ = 30, 28, 22
    d = _desc(cg, NX, NY, NZ, NT=12, nPML=5)
    o = po.Oracle(d, F64)
    o.set_model(None, *models.homogeneous(NX, NY, NZ))
    o.set_source_point(15, 14, 12, 3.0)
    st = np.array([[3, 4, NZ - 1], [15, 14, 12], [27, 20, 5]], np.int32)
    o.set_stations(st)
    states = []
    for _ in range(12):
        o.step(1)
        states.append(o.state())
    S = o.stations()
    for n, w in enumerate(states):
        for s, (x, y, z) in enumerate(st):
            np.testing.assert_allclose(S[s, n], w[:, z, y, x].astype(np.float32), rtol=1e-6)
    top = np.stack([w[:3, NZ - 1] for w in states])  # (it, 3, NY, NX)
    pgvh = np.sqrt(top[:, 0] ** 2 + top[:, 1] ** 2).max(0)
    pgv = np.sqrt((top ** 2).sum(1)).max(0)
    got = o.pgv()
    np.testing.assert_allclose(got[:, :, 0], pgvh, rtol=1e-6, atol=1e-30)
    np.testing.assert_allclose(got[:, :, 1], pgv, rtol=1e-6, atol=1e-30)


# ---------------------------------------------------------------- storage emulation
def test_fp16_bf16_rounding_matches_numpy(oracle_lib):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-9, 5, 4000),
                        np.array([65504.0, 65520.0, 1e-8, 6e-8, -3e-5, 0.0], np.float64)]).astype(np.float32)
    with np.errstate(over="ignore"):
        ref16 = x.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(oracle_lib.round_f16(x), ref16)
    # bf16: round to nearest even on the upper 16 bits
    b = x.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    np.testing.assert_array_equal(oracle_lib.round_bf16(x), r.astype(np.uint32).view(np.float32))
