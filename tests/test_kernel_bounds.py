"""Every memory access of the fused RK-stage kernels stays inside the device array it addresses.

tests/kernel_bounds/bounds_check.cpp compiles the library's kernel headers for the CPU (with the
stand-in device vocabulary of tests/kernel_bounds/hip/hip_runtime.h), lays out a domain's arrays
the way the library does and runs every workgroup and thread of the bulk, PML, free-surface and
general kernels for all four RK stages, every tile shape and both ways of reading the first
metric pair; buffer accesses are checked against their descriptors and the arrays, plain pointer
accesses run against host memory with guard regions. The grids below cover ragged tiles in x, y
and z, PML with and without a free surface, and the three storage types. No GPU is used.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "kernel_bounds", "bounds_check.cpp")
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(CLANG):
        pytest.skip("clang++ of the ROCm toolchain not found")
    exe = str(tmp_path_factory.mktemp("bounds") / "bounds_check")
    subprocess.check_call([CLANG, "-O2", "-std=c++17", "-fopenmp", "-Wno-unknown-attributes",
                           "-I", os.path.join(HERE, "kernel_bounds"),
                           "-I", os.path.join(ROOT, "cgfdm-fp16_sve_sme_amd", "csrc"), SRC, "-o", exe])
    return exe


@pytest.mark.parametrize("grid,npml,prec,fs,steps", [
    ((80, 70, 40), 8, "fp16", 1, 2),     # ragged 64/128-wide and 4/8/16-tall tiles
    ((150, 37, 29), 6, "fp32", 0, 1),    # no free surface: top PML slab, general boxes absent
    ((41, 130, 33), 12, "bf16", 1, 1),   # bulk box narrower than one tile in x
    ((270, 30, 24), 6, "fp16", 1, 8),    # three 128-wide bulk tiles; every MacCormack pattern
    ((40, 36, 30), 8, "fp32", 1, 1),     # the command-line driver's test grid
    ((200, 40, 36), 12, "fp16", 1, 1),   # right x slab across a layout-block boundary
])
def test_stage_kernels_stay_in_bounds(checker, grid, npml, prec, fs, steps):
    """Every access inside its array, and (--coverage) every interior point of the stage output
    and every point of the PML registers written, nothing else."""
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    r = subprocess.run([checker, *map(str, grid), str(npml), prec, str(fs), "-1", "0", "--steps", str(steps),
                        "--coverage"], capture_output=True, text=True, env=env, timeout=600)
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert "past a descriptor 0, outside every array 0" in r.stdout
    assert "coverage: 0 failed checks" in r.stdout


def test_coverage_check_catches_a_planted_gap(checker):
    """An interior output element reset to the sentinel after each stage must be reported."""
    r = subprocess.run([checker, "80", "40", "20", "8", "fp16", "1", "3", "0", "--coverage", "--plant-coverage"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode != 0
    assert "not written" in r.stderr


def test_checker_catches_a_planted_overrun(checker, tmp_path):
    """A domain whose arrays are registered one variable plane short must be reported (the
    checker is not vacuous): the kernels then address past the end of the shortened arrays."""
    r = subprocess.run([checker, "40", "36", "20", "6", "fp16", "1", "0", "0", "--shrink-arrays"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode != 0
    assert "past a descriptor" in r.stderr or "outside every device array" in r.stderr or "FAULT" in r.stderr
