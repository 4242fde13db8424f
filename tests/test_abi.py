"""The C-ABI library: it loads, exports every entry point include/cgfdm.h declares, its descriptor
layout matches the header as a C compiler sees it, and argument errors are reported before any
device call (these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cgfdm.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cgfdm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(gpu):
    lib = gpu.lib()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("cname,pyname", [("cgfdm_desc", "Desc"), ("cgfdm_server_config", "ServerConfig"),
                                          ("cgfdm_request", "Request"), ("cgfdm_request_info", "RequestInfo"),
                                          ("cgfdm_server_stats", "ServerStats")])
def test_struct_layout_matches_header(gpu, tmp_path, cname, pyname):
    """The binding's ctypes mirror of each struct has the header's size and field offsets."""
    cls = getattr(gpu, pyname)
    fields = [f[0] for f in cls._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){",
             f'printf("size %zu\\n", sizeof({cname}));']
    lines += [f'printf("{f} %zu\\n", offsetof({cname}, {f}));' for f in fields]
    lines += ["return 0;}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-o", str(exe), str(src)])
    out = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    assert int(out["size"]) == ctypes.sizeof(cls)
    for f in fields:
        assert int(out[f]) == getattr(cls, f).offset, f


def _create_rc(gpu, desc):
    h = ctypes.c_void_p()
    rc = gpu.lib().cgfdm_domain_create(ctypes.byref(desc), 0, ctypes.byref(h))
    return rc, gpu.lib().cgfdm_last_error().decode()


def test_pml_wider_than_grid_is_rejected(gpu):
    d = gpu.make_desc(10, 40, 40, nPML=12)
    rc, msg = _create_rc(gpu, d)
    assert rc == -5 and "nPML" in msg


def test_bad_descriptors_are_rejected(gpu):
    d = gpu.make_desc(20, 20, 20, nPML=6)
    d.precision = 7
    assert _create_rc(gpu, d)[0] == -1
    d = gpu.make_desc(20, 20, 20, nPML=6, DT=0.0)
    assert _create_rc(gpu, d)[0] == -1
    d = gpu.make_desc(2000, 2000, 600, nPML=12)  # > 2^31 padded points
    rc, msg = _create_rc(gpu, d)
    assert rc == -1 and "2^31" in msg
    h = ctypes.c_void_p()
    assert gpu.lib().cgfdm_domain_create(None, 0, ctypes.byref(h)) == -1
    # queries on a null domain are argument errors, reported before any device call
    flag = ctypes.c_int(7)
    assert gpu.lib().cgfdm_uniform_metric_pair(None, ctypes.byref(flag)) == -1 and flag.value == 7
    assert gpu.lib().cgfdm_partition(None, None) == -1
    # a server configuration with an unknown admission order is refused before any device call
    c = gpu.ServerConfig(0, gpu.FP16, 0, 0, 4, 8, 1, 2)
    assert gpu.lib().cgfdm_server_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert "admit_order" in gpu.lib().cgfdm_last_error().decode()
    # and one with more rescaled runs than the server allows
    c = gpu.ServerConfig(0, gpu.FP16, 0, 0, 4, 8, 1, 0, 9)
    assert gpu.lib().cgfdm_server_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert "rescale_retries" in gpu.lib().cgfdm_last_error().decode()


def test_estimate_bytes_scales_with_precision(gpu):
    d32 = gpu.make_desc(201, 201, 101, precision="fp32")
    d16 = gpu.make_desc(201, 201, 101, precision="fp16")
    b32, b16 = gpu.estimate_bytes(d32), gpu.estimate_bytes(d16)
    num = 207 * 207 * 107
    assert b32 >= num * 4 * 49 and b16 >= num * 2 * 49
    assert 1.6 < b32 / b16 < 2.0  # PML auxiliaries stay FP32


def test_no_fallback_when_library_missing(gpu, monkeypatch):
    import importlib
    import cgfdm
    monkeypatch.setattr(cgfdm, "LIB_PATH", "/nonexistent/libcgfdm.so")
    monkeypatch.setattr(cgfdm, "_lib", None)
    with pytest.raises(RuntimeError):
        cgfdm.lib()


@pytest.mark.parametrize("fs,pml,npml,shape", [(1, 1, 12, (60, 50, 40)), (0, 1, 6, (30, 26, 20)),
                                               (1, 0, 6, (20, 18, 12)), (0, 0, 5, (16, 16, 16)),
                                               (1, 1, 8, (18, 40, 30)), (1, 1, 4, (9, 9, 10)),
                                               (1, 1, 12, (150, 40, 30)), (1, 1, 12, (101, 30, 20)),
                                               (0, 1, 6, (270, 30, 24)), (1, 1, 12, (1001, 30, 30)),
                                               (1, 1, 12, (200, 30, 30))])
def test_stage_decomposition_covers_interior_once(gpu, fs, pml, npml, shape):
    """Bulk box, PML face/edge/corner boxes, the free-surface box and general boxes tile the
    interior exactly once; a face box lies inside exactly one slab (its normal axis), an edge or
    corner box inside exactly the slabs of its mask, all outside the free-surface layers; the
    free-surface box is the top three layers of the bulk columns."""
    import numpy as np
    NX, NY, NZ = shape
    d = gpu.make_desc(NX, NY, NZ, nPML=npml, free_surface=fs, pml=pml)
    boxes = gpu.decompose(d)
    cover = np.zeros((NZ + 6, NY + 6, NX + 6), np.int32)
    for i0, i1, j0, j1, k0, k1, kind, axis in boxes:
        assert i0 < i1 and j0 < j1 and k0 < k1
        cover[k0:k1, j0:j1, i0:i1] += 1
        if kind == 1:
            ranges = [(i0, i1, NX), (j0, j1, NY), (k0, k1, NZ)]
            for a, (b0, b1, n) in enumerate(ranges):
                in_low = b1 <= 3 + npml
                in_high = b0 >= 3 + n - npml
                top = 3 + n - (3 if (a == 2 and fs) else npml)  # free-surface layers or top slab
                if a == axis:
                    assert in_low or (in_high and not (a == 2 and fs))
                else:
                    assert b0 >= 3 + npml and b1 <= top
            if fs:
                assert k1 <= 3 + NZ - 3
        if kind == 4:  # edge or corner: inside exactly the slabs of the normals in the mask
            assert axis in (3, 5, 6, 7)
            ranges = [(i0, i1, NX), (j0, j1, NY), (k0, k1, NZ)]
            for a, (b0, b1, n) in enumerate(ranges):
                in_low = b1 <= 3 + npml
                in_high = b0 >= 3 + n - npml and not (a == 2 and fs)
                top = 3 + n - (3 if (a == 2 and fs) else npml)
                if (axis >> a) & 1:
                    assert in_low or in_high
                else:
                    assert b0 >= 3 + npml and b1 <= top
        if kind == 3:  # free-surface layers above the bulk columns
            assert fs and (k0, k1) == (3 + NZ - 3, 3 + NZ)
            np_ = npml if pml else 0
            assert (i0, i1, j0, j1) == (3 + np_, 3 + NX - np_, 3 + np_, 3 + NY - np_)
    inner = cover[3:NZ + 3, 3:NY + 3, 3:NX + 3]
    assert inner.min() == 1 and inner.max() == 1
    assert cover.sum() == NX * NY * NZ
    kinds = [b[6] for b in boxes]
    if pml and min(NX, NY, NZ) > 2 * npml + 3:
        assert kinds.count(0) == 1 and kinds.count(1) == (5 if fs else 6)
        assert kinds.count(3) == (1 if fs else 0)
        assert kinds.count(4) == (12 if fs else 20)  # edges and corners below the surface layers
        assert kinds.count(2) == (4 if fs else 0)    # the surface layers of the PML columns
