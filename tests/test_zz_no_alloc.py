"""cgfdm_get_slice_noalloc, the snapshot entry point that allocates no device memory, against
cgfdm_get_slice through the C ABI: bit-identical planes on two calls in a row, the second a
smaller plane, in FP32 and FP16 storage; planes above its fixed staging size are rejected."""
import numpy as np
import pytest

import models

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_slice_noalloc_matches_slice(gpu, prec):
    cg = gpu
    NX, NY, NZ = 64, 56, 40
    scales = {"Cv": 5e3, "Cs": 1e-6, "Crho": 1e6} if prec == "fp16" else {}
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=30, nPML=8, precision=prec, **scales)
    g = cg.Domain(d)
    g.set_model(None, *models.homogeneous(NX, NY, NZ))
    g.set_source_point(NX // 2, NY // 2, NZ - 10, 2.0)
    g.step(30)
    print(f"{prec}: 30 steps done", flush=True)
    # first call: the surface plane (NY x NX = 3584 points), second: an x plane (NZ x NY = 2240)
    for axis, index, var in ((2, NZ - 1, 2), (0, NX // 2, 5)):
        ref = g.slice(axis, index, var)
        got = g.slice_noalloc(axis, index, var)
        assert got.shape == ref.shape
        assert np.abs(ref).max() > 0
        np.testing.assert_array_equal(got, ref)
    # the staged plane of a later call replaces the earlier one: a y plane after the x plane
    ref = g.slice(1, NY // 2, 0)
    np.testing.assert_array_equal(g.slice_noalloc(1, NY // 2, 0), ref)
    g.close()


def test_slice_noalloc_rejects_planes_beyond_its_staging_size(gpu):
    cg = gpu
    # a surface plane of 2049 x 2049 = 4198401 points exceeds the 4194304-point staging plane;
    # the check happens before any device work, so no model is needed
    d = cg.make_desc(2049, 2049, 6, DH=100.0, DT=0.008, NT=1, nPML=2, precision="fp16",
                     Cv=5e3, Cs=1e-6, Crho=1e6)
    g = cg.Domain(d)
    out = np.zeros((2049, 2049), np.float32)
    rc = cg.lib().cgfdm_get_slice_noalloc(g.h, 2, 0, 0, out.ctypes.data)
    assert rc == -1 and "staging plane" in cg.lib().cgfdm_last_error().decode()
    # an x plane of the same domain (6 x 2049 points) fits
    assert cg.lib().cgfdm_get_slice_noalloc(g.h, 0, 0, 0, np.zeros((6, 2049), np.float32).ctypes.data) == 0
    g.close()
