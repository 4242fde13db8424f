"""Scale factors of 16-bit storage chosen from the source (cgfdm_source_scales); a host-only entry
point, no GPU.

Pins: the point source keeps the reference's calibrated scales (5e3, 1e-6) at 100 m
(macro.h:65-73, singleSource.cpp:29-56) and scales with DH^2; every choice keeps the reference's
ratio Cv/Cs = 5e9 (the stored medium rho*Cs/Cv does not change); for a table, the scaled
estimates are what the rule says they are: ten times the moment rate gives a tenth of the scales
(the equations are linear, so the stored numbers do not change), the velocity estimate of a table
that holds slip rate * area is independent of the medium, a planar patch of nodes slipping
together is bounded by half its slip rate (its far-reaching coherent field) rather than by one
node's near field, and the stress bound takes over when the accumulated moment is large and its
rate small; a slow double couple at one node runs finite in the FP16 oracle with the chosen
scales, and not with 6.25x larger ones. The GPU tests (tests/test_gpu_configs.py)
check that the chosen scales keep M0 = 1e17 N m sources and the finite-fault rupture finite and
within the FP16 tolerance of the FP64 oracle."""
import math

import numpy as np
import pytest

import workloads as wl


def test_point_source_keeps_the_reference_scales(gpu):
    cg = gpu
    assert cg.source_scales(100.0) == pytest.approx((5e3, 1e-6), rel=1e-12)
    assert cg.source_scales(50.0) == pytest.approx((1250.0, 2.5e-7), rel=1e-12)
    assert cg.source_scales(400.0) == pytest.approx((8e4, 1.6e-5), rel=1e-12)


def test_table_scales_follow_the_peak_moment_rate(gpu):
    cg = gpu
    mu, dt, fc, nt = 3.0e10, 0.005, 2.0, 400
    base = wl.double_couple_table(1e16, 0.0, 90.0, 0.0, fc, nt, dt, mu)
    one = np.array([[50, 50, 40]], np.int32)
    cv1, cs1 = cg.source_scales(100.0, one, base, dt, mu)
    assert cv1 / cs1 == pytest.approx(5e9, rel=1e-12)
    # linear in the source: ten times the moment rate, a tenth of the scales
    cv10, _ = cg.source_scales(100.0, one, 10 * base, dt, mu)
    assert cv10 == pytest.approx(cv1 / 10, rel=1e-6)
    # velocity-bound here: Cv * 2.5 * max|rate| / (4 pi DH^2) = 4e3 (one node's field one cell
    # away, with the discrete scheme's excess), and DH doubled gives Cv x 4
    assert cv1 * 2.5 * np.abs(base).max() / (4 * math.pi * 1e4) == pytest.approx(4e3, rel=1e-6)
    assert cg.source_scales(200.0, one, base, dt, mu)[0] == pytest.approx(4 * cv1, rel=1e-6)
    # the velocity bound does not depend on the medium (the table is rate / mu at the source)
    assert cg.source_scales(100.0, one, base, dt, 3 * mu)[0] == pytest.approx(cv1, rel=1e-6)
    # a slow ramp in moment (200 s at a constant small rate): the stress bound
    # (mu * moment / (4 pi DH^3)) decides instead of the velocity bound
    step = np.zeros((nt, 1, 6), np.float32)
    step[:, 0, 3] = 1e-3
    cv_s, _ = cg.source_scales(100.0, one, step, 0.5, mu)
    assert cv_s < 4e3 / (2.5 * 1e-3 / (4 * math.pi * 1e4))
    M = mu * 1e-3 * 0.5 * nt
    assert cv_s == pytest.approx(5e9 * 1e4 / (M / (4 * math.pi * 1e6)), rel=1e-5)


def test_planar_patch_is_bounded_by_half_its_slip_rate(gpu):
    """A square patch of nodes 5 cells apart (500 m at DH = 100 m, each with area 500^2 m^2), all
    slipping at 2 m/s: one node's near field, 2.5 * area * 2 / (4 pi DH^2) = 99.5 m/s, dominates
    at 100 m spacing; at DH = 500 m (nodes one cell apart) the patch's field decides instead,
    within 10 % of half the slip rate (1 m/s)."""
    cg = gpu
    n = 41
    X, Y = np.meshgrid(np.arange(n), np.arange(n))
    slip_rate, dt, nt = 2.0, 0.01, 50
    for dh, step in ((100.0, 5), (500.0, 1)):
        xyz = np.stack([X.ravel() * step, Y.ravel() * step, np.full(n * n, 60)], 1).astype(np.int32)
        area = (step * dh) ** 2
        tab = np.zeros((nt, n * n, 6), np.float32)
        tab[:, :, 3] = area * slip_rate  # the xy component of a vertical strike-slip, unit tensor
        cv, _ = cg.source_scales(dh, xyz, tab, dt, 3e10)
        v = 4e3 / cv
        if dh == 100.0:
            assert v == pytest.approx(2.5 * area * slip_rate / (4 * math.pi * dh * dh), rel=1e-6)
        else:
            # one node's near field would say 2.5 * area * 2 / (4 pi DH^2) = 0.4 m/s, below (with
            # its four neighbours one cell away, (2.5 + 4) * 0.16 = 1.03 m/s, within the 10 %)
            assert v > 5 * area * slip_rate / (4 * math.pi * dh * dh)
            assert 0.9 * slip_rate / 2 < v < 1.1 * slip_rate / 2


def test_bad_arguments_are_rejected(gpu):
    import ctypes
    cg = gpu
    cv, cs = ctypes.c_double(), ctypes.c_double()
    assert cg.lib().cgfdm_source_scales(0.0, None, None, 0, 0, 0.0, 0.0, ctypes.byref(cv), ctypes.byref(cs)) == -1
    t = np.ones((4, 1, 6), np.float32)
    x = np.zeros((1, 3), np.int32)
    for xp, npts, nt, dt, mu in ((x, 1, 4, 0.0, 1e10), (x, 1, 4, 0.01, 0.0), (x, 0, 4, 0.01, 1e10),
                                 (None, 1, 4, 0.01, 1e10)):
        xa = None if xp is None else xp.ctypes.data
        assert cg.lib().cgfdm_source_scales(100.0, xa, t.ctypes.data, npts, nt, dt, mu, ctypes.byref(cv),
                                            ctypes.byref(cs)) == -1
    zero = np.zeros((4, 1, 6), np.float32)  # no source at all: the reference scales
    assert cg.source_scales(100.0, x, zero, 0.01, 1e10) == pytest.approx((5e3, 1e-6))


def test_point_source_scales_follow_the_wavelet_and_the_medium(gpu):
    """cgfdm_point_source_scales: the table rule applied to the point source's equivalent table
    (the Ricker moment rate of peak frequency fc over the 7^3 Gaussian weights, divided by the
    shear modulus at the source). Velocity-bound for these inputs, so the scale is proportional to
    mu_src and DH^2 and inversely to fc (the peak moment rate is proportional to fc), with the
    reference's ratio Cv/Cs = 5e9."""
    cg = gpu
    mu = 2670.0 * 3464.0 ** 2
    cv, cs = cg.point_source_scales(100.0, 2.0, mu, mu)
    assert cv / cs == pytest.approx(5e9, rel=1e-12)
    assert cg.point_source_scales(100.0, 2.0, mu / 2, mu)[0] == pytest.approx(cv / 2, rel=1e-5)
    assert cg.point_source_scales(200.0, 2.0, mu, mu)[0] == pytest.approx(4 * cv, rel=1e-5)
    assert cg.point_source_scales(100.0, 4.0, mu, mu)[0] == pytest.approx(cv / 2, rel=1e-3)
    with pytest.raises(cg.CgfdmError):
        cg.point_source_scales(100.0, 0.0, mu, mu)


def test_point_source_scales_bound_the_field_of_a_soft_medium(gpu, oracle_lib):
    """In a soft medium (Vs 1800 m/s), the FP32 oracle's peak particle velocity times the chosen
    Cv stays below the rule's target 4e3 (the estimate is an upper bound; 1.75e3 here), while the
    reference's fixed Cv = 5e3 would put it past the FP16 maximum (65504: 14.1 m/s x 5e3)."""
    cg, po = gpu, oracle_lib
    NX, NY, NZ = 40, 40, 36
    vs = np.full((NZ, NY, NX), 1800.0, np.float32)
    vp = (1.73 * vs).astype(np.float32)
    rho = (1740.0 * (vp / 1000.0) ** 0.25).astype(np.float32)
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=160, nPML=6)
    o = po.Oracle(d, po.F32)
    o.set_model(None, vs, vp, rho)
    o.set_source_point(NX // 2, NY // 2, NZ // 2, 2.5)
    peak = 0.0
    for _ in range(16):
        o.step(10)
        peak = max(peak, float(np.abs(o.state()[:3]).max()))
    o.close()
    mu = float(rho[0, 0, 0]) * 1800.0 ** 2
    cv, _ = cg.point_source_scales(100.0, 2.5, mu, mu)
    assert peak * cv < 4e3
    assert peak * 5e3 > 65504.0


def test_single_node_double_couple_stays_inside_fp16(gpu, oracle_lib):
    """A slow double couple at one node (Ricker fc 0.5 Hz, M0 1e15 N m, Vs 4000 m/s, DT 8 ms:
    the hardest corner of the serving request stream). The FP32 oracle's peak particle velocity
    times the chosen Cv stays below the rule's target 4e3, and the FP16 oracle (16-bit state and
    RK stage values, as the reference's FP16 build stores them) runs finite with the chosen
    scales; with 6.25x larger scales (the earlier rule: node factor 1, target 1e4) its RK stage
    values overflow although the stored peak (~1.5e4) is far below 65504."""
    cg, po = gpu, oracle_lib
    N, DH, DT, fc, steps = 40, 100.0, 0.008, 0.5, 340
    vs = 4000.0
    vp, rho = wl.vp_rho_from_vs(np.array([vs]))
    vp, rho = float(vp[0]), float(rho[0])
    mu = rho * vs * vs
    tab = wl.double_couple_table(1e15, 244.17, 17.89, 105.51, fc, steps + 2, DT, mu)
    c = N // 2
    xyz = np.array([[c, c, c]], np.int32)
    cv, cs = cg.source_scales(DH, xyz, tab, DT, mu)
    one = np.ones((N, N, N), np.float32)

    def run(prec, Cv):
        d = cg.make_desc(N, N, N, DH=DH, DT=DT, NT=steps, nPML=6, precision=prec, Cv=Cv, Cs=Cv / 5e9,
                         free_surface=0)
        o = po.Oracle(d, po.F32 if prec == cg.FP32 else po.F16)
        o.set_model(None, vs * one, vp * one, rho * one)
        o.set_source_table(xyz, tab, DT)
        peak = 0.0
        for n in range(0, steps, 20):
            o.step(20)
            s = o.state()
            if not np.isfinite(s).all():
                o.close()
                return None
            peak = max(peak, float(np.abs(s[:3]).max()))
        o.close()
        return peak

    peak = run(cg.FP32, 0.0)
    assert 1e3 < peak * cv < 4e3
    assert run(cg.FP16, cv) is not None
    assert run(cg.FP16, 6.25 * cv) is None
