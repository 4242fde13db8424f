"""The non-finite guard (cgfdm_set_check_interval / cgfdm_nonfinite_step): a wavefield value that
is not finite is reported by the first scan after it appears, and a finite run is never flagged.
The reference has no such guard (propagate.cpp:142-198 checks nothing), so there is no reference
behaviour to match; the checks below are the guard's own contract."""
import numpy as np
import pytest

import models

pytestmark = pytest.mark.gpu


def _domain(cg, prec, NT=12):
    NX, NY, NZ = 40, 36, 30
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=NT, nPML=8, precision=prec)
    g = cg.Domain(d)
    g.set_model(None, *models.homogeneous(NX, NY, NZ))
    return g, (NX, NY, NZ)


@pytest.mark.parametrize("prec", ["fp16", "bf16", "fp32"])
def test_guard_reports_a_non_finite_value_at_the_next_scan(gpu, prec):
    cg = gpu
    g, (NX, NY, NZ) = _domain(cg, prec)
    g.set_check_interval(4)
    w = np.zeros((9, NZ, NY, NX), np.float32)
    # one stored value beyond every storage type's range (the state is given in stored units):
    # infinite in FP16, BF16 and FP32 storage alike
    w[3, NZ // 2, NY // 2, NX // 2] = np.float32(np.inf)
    g.set_state(w)
    g.step(3)  # scans after step counts 4, 8, ...: none yet
    assert g.nonfinite_step() == -1
    g.sync()
    g.step(2)  # the scan after step count 4 sees it
    assert g.nonfinite_step() == 4
    with pytest.raises(cg.CgfdmError, match="ENONFINITE"):
        g.sync()
    with pytest.raises(cg.CgfdmError, match="ENONFINITE"):
        g.pgv()
    g.step(4)  # the first step found is kept
    assert g.nonfinite_step() == 4
    g.close()


def test_guard_quiet_on_a_finite_run_and_off_when_disabled(gpu):
    cg = gpu
    g, (NX, NY, NZ) = _domain(cg, "fp16")
    g.set_check_interval(2)
    g.set_source_point(NX // 2, NY // 2, NZ - 10, 2.0)
    g.step(12)
    g.sync()
    assert g.nonfinite_step() == -1
    assert np.isfinite(g.pgv()).all()
    g.close()
    g, (NX, NY, NZ) = _domain(cg, "fp16")
    g.set_check_interval(0)
    w = np.zeros((9, NZ, NY, NX), np.float32)
    w[0, 5, 5, 5] = np.float32(np.nan)
    g.set_state(w)
    g.step(10)
    assert g.nonfinite_step() == -1  # no scan ran
    g.close()
