"""The structure the stage tile kernels rely on for zeta, pinned on the oracle's FP64 metric (a
restatement of the reference's solve_con_jac, contravariant.cpp:71-171, on the coordinates of
coord.cpp:106-307): on the flat and terrain-following grids, zeta_x and zeta_y are (k - 3) times
a value per column and zeta_z is constant along a column, at every interior point. The library's
set-up checks the same on the device (k_metric against k_zcol) and refuses a grid without it."""
import numpy as np
import pytest

import models


@pytest.mark.parametrize("terrain_kind", ["flat", "hill", "random"])
def test_zeta_is_linear_along_columns(gpu, oracle_lib, terrain_kind):
    cg, po = gpu, oracle_lib
    NX, NY, NZ, DH = 30, 26, 22, 100.0
    if terrain_kind == "flat":
        terrain, use = None, 0
    elif terrain_kind == "hill":
        terrain, use = models.gauss_hill(NX, NY, DH, h0=400.0, width=900.0), 1
    else:
        terrain, use = models.smooth_random_terrain(NX, NY, DH, rms=300.0, corr=600.0, seed=4), 1
    d = cg.make_desc(NX, NY, NZ, DH=DH, DT=0.008, NT=1, nPML=5, use_terrain=use, depth=NZ * DH)
    o = po.Oracle(d, po.F64)
    o.set_model(terrain, *models.homogeneous(NX, NY, NZ))
    c = o.cjm64()  # (13, nz_, ny_, nx_), reference component order
    o.close()
    k = np.arange(3, 3 + NZ)
    zx, zy, zz = (c[q][3:3 + NZ, 3:3 + NY, 3:3 + NX] for q in (6, 7, 8))
    kk = (k - 3).astype(np.float64)[:, None, None]
    top = NZ - 1
    ax, ay = zx[top] / top, zy[top] / top
    scale = np.abs(zz).max()
    assert np.abs(zx - ax[None] * kk).max() <= 1e-12 * scale
    assert np.abs(zy - ay[None] * kk).max() <= 1e-12 * scale
    assert np.abs(zz - zz[top][None]).max() <= 1e-12 * scale
    if terrain_kind != "flat":
        assert np.abs(ax).max() > 1e-6 * scale  # the slopes make zeta_x non-zero here
    else:
        assert np.abs(ax).max() == 0.0 and np.abs(ay).max() == 0.0
