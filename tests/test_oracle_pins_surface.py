"""Pins of the oracle's free-surface, PML free-surface, rDZ and point-source amplitude functions
against closed forms (CPU only).

  L2 / L3 (freeSurface.cpp:210-218, axis.h:92-93)
      A velocity quadratic in z, V = G (x - P) + Q (z - z_P)^2 with zero stress, evaluated at a
      point P of the layers k = NZ-2 and NZ-3: the z derivative is the 2-point forward difference
      (exact slope plus FB3*DH/2 f'') at NZ-2 and the 3-point one-sided difference of the
      reference's Cf weights at NZ-3, both Taylor expansions of a quadratic; the stress rate is then
      lambda tr(grad V) I + mu (grad V + grad V^T) of that gradient. The full operator L, a wrong
      layer or a wrong direction gives a different second-order term and fails.
  rDZ (contravariant.cpp:257-351)
      On a flat grid, -rDZ.DX = [[0,0,-1],[0,0,0],[-lambda/chi,0,0]] and -rDZ.DY likewise in y, with
      the reference's chi = 2 lambda + mu (or lambda + 2 mu on request), at Vp/Vs != sqrt 3. On a
      tilted surface with chi = lambda + 2 mu the traction rate sigma_dot . grad(zeta) vanishes for
      lambda != mu; with the reference's chi it does not (the quirk the switch fixes).
  PML free surface (pml_freeSurface.cpp:81-116)
      The auxiliary increment of an x or y slab at the surface layer, for a linear velocity field,
      equals the single-direction elastic rate scaled by d (pml_deriv.cpp:176-207) plus the stress
      rate of the vertical gradient the free surface implies (V_x,z = -V_z,x and V_z,z =
      -(lambda/chi) V_x,x for the x slab): closed forms per component; with chi = lambda + 2 mu
      its zz and xz components (the surface traction) vanish. One layer down the surface term is
      absent.
  point source (singleSource.cpp:77-127)
      The injection into Txx, Tyy, Tzz at the source cell is -s(t) w000 Cs DT / (J DH^3) with s the
      time derivative of the Ricker wavelet times M0 at t = it*DT, w000 the centre value of the
      normalised 7^3 Gaussian (computed here from its definition), J = 0.99998^3 on a flat grid;
      the neighbours follow the Gaussian's shape and no other variable changes.
"""
import math

import numpy as np
import pytest

import models

F64 = 0
# macro.h:126-130 and axis.h:93 as the float literals the reference writes
COEF = {m: float(np.float32(a)) for m, a in {-1: -0.30874, 0: -0.6326, 1: 1.2330, 2: -0.3334, 3: 0.04168}.items()}
CF = [float(np.float32(c)) for c in (-1.16666667, 1.33333333, -0.16666667)]
M1 = sum(a * m for m, a in COEF.items())  # first moment of L, 0.99998: flat-grid x_xi


@pytest.fixture(scope="module")
def cg():
    import cgfdm
    return cgfdm


def _desc(cg, NX, NY, NZ, **kw):
    kw.setdefault("DH", 100.0)
    kw.setdefault("DT", 0.008)
    kw.setdefault("NT", 10)
    kw.setdefault("nPML", 6)
    return cg.make_desc(NX, NY, NZ, **kw)


def _xyz(d):
    x = (np.arange(d.NX) - d.centerX) * d.DH
    y = (np.arange(d.NY) - d.centerY) * d.DH
    z = models.flat_z(d.NX, d.NY, d.NZ, d.DH)
    Z3, Y3, X3 = np.meshgrid(np.arange(d.NZ), y, x, indexing="ij")
    return X3, Y3, np.array(z)


def _stress_rate(Geff, lam, mu):
    S = lam * np.trace(Geff) * np.eye(3) + mu * (Geff + Geff.T)
    return np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])


# ---------------------------------------------------------------- L2 / L3 free-surface layers
@pytest.mark.parametrize("FB", [(1, 1, 1), (-1, 1, -1), (1, -1, 1), (-1, -1, -1)])
def test_free_surface_L2_L3_layers(cg, oracle_lib, FB):
    po = oracle_lib
    NX, NY, NZ = 22, 20, 14
    d = _desc(cg, NX, NY, NZ, pml=0, free_surface=1)
    vs, vp, rho = 2600.0, 5100.0, 2450.0  # lambda != mu
    mu, lam = rho * vs * vs, rho * (vp * vp - 2 * vs * vs)
    o = po.Oracle(d, F64)
    o.set_model(None, *models.homogeneous(NX, NY, NZ, vs, vp, rho))
    X, Y, Z = _xyz(d)
    rng = np.random.default_rng(21)
    DH = d.DH
    FB3 = FB[2]
    for k, (c1, c2) in ((NZ - 2, (1.0, 1.0)),
                        (NZ - 3, (CF[1] + 2 * CF[2], CF[1] + 4 * CF[2]))):
        pi, pj = 10, 9
        P = (X[k, pj, pi], Y[k, pj, pi], Z[k, pj, pi])
        G = rng.standard_normal((3, 3)) * 1e-3
        Q = rng.standard_normal(3) * 1e-6
        dz = Z - P[2]
        w = np.zeros((9, NZ, NY, NX))
        for a in range(3):
            w[a] = G[a, 0] * (X - P[0]) + G[a, 1] * (Y - P[1]) + G[a, 2] * dz + Q[a] * dz * dz
        o.set_state(w)
        h = o.rhs(*FB)[:, k, pj, pi] / d.DT
        # one-sided z difference of f = G2 s + Q s^2 at s = 0 with step FB3*DH:
        # c1 * f'(0) + FB3 * c2 * (DH/2) * f''(0), f'' = 2Q; the flat grid's zeta_z = 1/0.99998
        dvdz = (c1 * G[:, 2] + FB3 * c2 * DH * Q) / M1
        Geff = np.column_stack([G[:, 0], G[:, 1], dvdz])
        ref = _stress_rate(Geff, lam, mu)
        np.testing.assert_allclose(h[3:], ref, rtol=0, atol=1e-9 * np.abs(ref).max())
        # zero stress, zero image: no velocity rate
        assert np.abs(h[:3]).max() <= 1e-12 * np.abs(ref).max()


# ---------------------------------------------------------------- rDZ
@pytest.mark.parametrize("chi_mode", [0, 1])
def test_rdz_flat_closed_form_lambda_ne_mu(cg, oracle_lib, chi_mode):
    po = oracle_lib
    NX, NY, NZ = 16, 14, 10
    d = _desc(cg, NX, NY, NZ, pml=0, free_surface=1, chi_mode=chi_mode)
    vs, vp, rho = 2000.0, 4300.0, 2300.0  # Vp/Vs = 2.15: lambda = 2.62 mu
    o = po.Oracle(d, F64)
    o.set_model(None, *models.homogeneous(NX, NY, NZ, vs, vp, rho))
    mu = float(np.float32(rho)) * vs * vs
    lam = float(np.float32(rho)) * (vp * vp - 2 * vs * vs)
    chi = lam + 2 * mu if chi_mode else 2 * lam + mu
    r = o.rdz()[:, 3:-3, 3:-3].reshape(2, -1, 3, 3)
    want_x = np.array([[0, 0, -1], [0, 0, 0], [-lam / chi, 0, 0]])
    want_y = np.array([[0, 0, 0], [0, 0, -1], [0, -lam / chi, 0]])
    np.testing.assert_allclose(r[0], np.broadcast_to(want_x, r[0].shape), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(r[1], np.broadcast_to(want_y, r[1].shape), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("chi_mode", [0, 1])
def test_free_surface_traction_rate_lambda_ne_mu(cg, oracle_lib, chi_mode):
    """Tilted surface, lambda = 2.6 mu: with chi = lambda + 2 mu the traction rate
    sigma_dot . grad(zeta) vanishes at the surface; with the reference's 2 lambda + mu it does not."""
    po = oracle_lib
    NX, NY, NZ = 26, 24, 16
    DH = 100.0
    d = _desc(cg, NX, NY, NZ, pml=0, free_surface=1, use_terrain=1, depth=1500.0, chi_mode=chi_mode)
    xs = (np.arange(NX + 6) - 3 - d.centerX) * DH
    ys = (np.arange(NY + 6) - 3 - d.centerY) * DH
    XX, YY = np.meshgrid(xs, ys)
    terrain = 0.2 * XX + 0.1 * YY
    o = po.Oracle(d, F64)
    o.set_model(terrain, *models.homogeneous(NX, NY, NZ, 2000.0, 4300.0, 2300.0))
    o.set_state(models.smooth_state(NX, NY, NZ, seed=11))
    cjm = o.cjm64()
    k = NZ - 1
    js, is_ = slice(8, NY - 8), slice(8, NX - 8)
    zt = [cjm[6 + c, k + 3, 3:-3, 3:-3] for c in range(3)]
    worst = 0.0
    for FB in [(1, 1, 1), (-1, 1, -1), (1, -1, 1)]:
        h = o.rhs(*FB) / d.DT
        S = {(0, 0): h[3], (1, 1): h[4], (2, 2): h[5], (0, 1): h[6], (0, 2): h[7], (1, 2): h[8]}
        mag = np.abs(h[3:6, k]).max()
        for a in range(3):
            tr = sum(S[tuple(sorted((a, b)))][k] * zt[b] for b in range(3))
            worst = max(worst, float(np.abs(tr[js, is_]).max() / mag))
    if chi_mode:
        assert worst <= 1e-9, worst
    else:
        assert worst >= 1e-3, worst


# ---------------------------------------------------------------- PML free surface
@pytest.mark.parametrize("chi_mode", [0, 1])
@pytest.mark.parametrize("slab", [0, 1, 3, 4])
def test_pml_free_surface_aux_term(cg, oracle_lib, slab, chi_mode):
    po = oracle_lib
    NX, NY, NZ = 24, 22, 14
    nP = 6
    d = _desc(cg, NX, NY, NZ, nPML=nP, pml=1, free_surface=1, chi_mode=chi_mode)
    vs, vp, rho = 2000.0, 4300.0, 2300.0
    o = po.Oracle(d, F64)
    o.set_model(None, *models.homogeneous(NX, NY, NZ, vs, vp, rho))
    mu = float(np.float32(rho)) * vs * vs
    lam = float(np.float32(rho)) * (vp * vp - 2 * vs * vs)
    chi = lam + 2 * mu if chi_mode else 2 * lam + mu
    a = slab % 3                      # slab normal: 0 x, 1 y
    hi = slab >= 3                    # second face of the axis
    prof = o.pml()["d_" + "xy"[a]]    # stored d/beta per padded index
    X, Y, Z = _xyz(d)
    rng = np.random.default_rng(40 + slab)
    G = rng.standard_normal((3, 3)) * 1e-3
    # a point inside the slab whose stencils (x: i-1..i+3 for FB1 = 1, y: j-3..j+1 for FB2 = -1)
    # stay inside the grid
    l0 = 2 if hi else 3
    if a == 0:
        pi, pj = ((NX - nP + l0) if hi else l0), NY // 2
        pidx = pi + 3
    else:
        pi, pj = NX // 2, ((NY - nP + l0) if hi else l0)
        pidx = pj + 3
    dd = prof[pidx]
    assert dd > 0
    for k in (NZ - 1, NZ - 2):
        P = (X[k, pj, pi], Y[k, pj, pi], Z[k, pj, pi])
        w = np.zeros((9, NZ, NY, NX))
        for c in range(3):
            w[c] = G[c, 0] * (X - P[0]) + G[c, 1] * (Y - P[1]) + G[c, 2] * (Z - P[2])
        o.set_state(w)
        o.rhs(1, -1, 1)
        aux = o.aux_rhs(slab) / d.DT
        i0 = pi - (NX - nP if (hi and a == 0) else 0) if a == 0 else pi
        j0 = pj - (NY - nP if (hi and a == 1) else 0) if a == 1 else pj
        got = aux[k, j0, i0]
        g = G[:, a] * dd                       # d times the gradient along the slab normal
        # single-direction elastic rate along the normal (pml_deriv.cpp:176-207)
        Gs = np.zeros((3, 3))
        Gs[:, a] = g
        rate = _stress_rate(Gs, lam, mu)
        vel = np.zeros(3)                      # zero stress, zero auxiliaries
        if k == NZ - 1:
            # the vertical gradient the free surface implies for this tangential gradient:
            # V_n,z = -V_z,n and V_z,z = -(lambda/chi) V_n,n (flat-grid -rDZ.DX / -rDZ.DY)
            Gz = np.zeros((3, 3))
            Gz[a, 2] = -g[2]
            Gz[2, 2] = -(lam / chi) * g[a]
            rate = rate + _stress_rate(Gz, lam, mu)
            if chi_mode:
                # surface traction of the auxiliary rate: zz and the normal shear vanish
                assert abs(rate[2]) <= 1e-9 * np.abs(rate).max()
                assert abs(rate[4 if a == 0 else 5]) <= 1e-9 * np.abs(rate).max()
        np.testing.assert_allclose(got[3:], rate, rtol=0, atol=1e-9 * np.abs(rate).max())
        np.testing.assert_allclose(got[:3], vel, atol=1e-12 * np.abs(rate).max())


# ---------------------------------------------------------------- point source amplitude
@pytest.mark.parametrize("Cs", [1.0, 1e-6])
def test_point_source_amplitude(cg, oracle_lib, Cs):
    po = oracle_lib
    NX, NY, NZ = 17, 17, 17
    DH, DT = 100.0, 0.008
    d = _desc(cg, NX, NY, NZ, DH=DH, DT=DT, NT=60, pml=0, free_surface=0, Cs=Cs, Cv=1.0, Crho=1.0)
    o = po.Oracle(d, F64)
    o.set_model(None, *models.homogeneous(NX, NY, NZ))
    fc = 2.0
    P = (8, 8, 8)
    o.set_source_point(*P, fc)
    it = 37
    for _ in range(it):
        o.step(1)
    o.set_state(np.zeros((9, NZ, NY, NX)))
    o.inject_source()
    w = o.state()
    # s(t) = M0 d/dt[(1 - 2 r^2) exp(-r^2)], r = pi fc (t - 1.2/fc), t = it*DT
    t = it * DT
    r = math.pi * fc * (t - 1.2 / fc)
    s = 1e16 * math.pi * fc * (-4 * r - 2 * r * (1 - 2 * r * r)) * math.exp(-r * r)
    gs = {q: math.exp(-(q / 1.5) ** 2) / (1.5 * math.sqrt(math.pi)) for q in range(-3, 4)}
    norm = sum(gs[a] * gs[b] * gs[c] for a in gs for b in gs for c in gs)
    assert norm == pytest.approx(0.998125703461425, rel=1e-13)
    J = M1 ** 3
    amp0 = -s * gs[0] ** 3 / norm / (J * DH ** 3) * Cs * DT
    for v in (3, 4, 5):
        assert w[v, P[2], P[1], P[0]] == pytest.approx(amp0, rel=1e-6)
    # Gaussian shape around the cell
    for off in [(1, 0, 0), (0, 2, 0), (1, 2, 3), (-3, -1, 2)]:
        got = w[3, P[2] + off[2], P[1] + off[1], P[0] + off[0]]
        want = amp0 * gs[off[0]] * gs[off[1]] * gs[off[2]] / gs[0] ** 3
        assert got == pytest.approx(want, rel=1e-6)
    # outside the 7^3 stencil and in the other variables: nothing
    mask = np.zeros_like(w, bool)
    mask[3:6, P[2] - 3:P[2] + 4, P[1] - 3:P[1] + 4, P[0] - 3:P[0] + 4] = True
    assert not w[~mask].any()
    np.testing.assert_array_equal(w[3], w[4])
    np.testing.assert_array_equal(w[3], w[5])
