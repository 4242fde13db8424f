"""PML pin (a6, a7, a9): a small domain closed by the auxiliary-variable PML must reproduce, at
receivers near its edges, the seismograms of a domain large enough that no boundary is reached
in the window. A wrong sign or a dropped term in the auxiliary equations reflects or amplifies the
outgoing wave and fails the bound.
"""
import numpy as np
import pytest

import models
from models import rel_l2


def _run(cg, po, N, src, sts, fs, nsteps, nPML=10):
    NX, NY, NZ = N
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=nsteps, nPML=nPML, free_surface=fs)
    o = po.Oracle(d, po.F32)
    o.set_model(None, *models.homogeneous(NX, NY, NZ))
    o.set_source_point(*src, 2.5)
    o.set_stations(np.array(sts, np.int32))
    o.step(nsteps)
    return o.stations()


@pytest.mark.parametrize("fs", [0, 1])
def test_pml_absorbs(cg_mod, oracle_lib, fs):
    cg, po = cg_mod, oracle_lib
    nsteps = 110
    small = (36, 36, 36)
    off = 28  # the large domain embeds the small one with this margin on every damped side
    big = (36 + 2 * off, 36 + 2 * off, 36 + (0 if fs else 2 * off) + off)
    zoff = off
    src = (18, 18, 22 if fs else 18)
    # receivers 3 cells inside the PML's inner edge, on every damped face
    sts = [(13, 18, src[2]), (22, 18, src[2]), (18, 13, src[2]), (18, 22, src[2]), (18, 18, 13)]
    if fs:
        sts.append((18, 18, 35))  # on the free surface
    else:
        sts.append((18, 18, 22))
    a = _run(cg, po, small, src, sts, fs, nsteps)
    shift = lambda p: (p[0] + off, p[1] + off, p[2] + zoff)
    b = _run(cg, po, big, shift(src), [shift(p) for p in sts], fs, nsteps)
    assert np.abs(b[:, :, :3]).max() > 0
    err = rel_l2(a[:, :, :3], b[:, :, :3])
    assert err < 0.05, err
    # and without the PML the same small domain reflects strongly
    NX, NY, NZ = small
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=nsteps, nPML=10, free_surface=fs, pml=0)
    o = po.Oracle(d, po.F32)
    o.set_model(None, *models.homogeneous(NX, NY, NZ))
    o.set_source_point(*src, 2.5)
    o.set_stations(np.array(sts, np.int32))
    o.step(nsteps)
    err_nopml = rel_l2(o.stations()[:, :, :3], b[:, :, :3])
    assert err_nopml > 4 * err, (err_nopml, err)


@pytest.fixture(scope="module")
def cg_mod():
    import cgfdm
    return cgfdm
