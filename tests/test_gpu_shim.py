"""The drop-in boundary under the reference's own code: ps::GpuFeaSolver (parts-sim_amd/host/) over
libpsx.so, compiled by oracle/Makefile (target shim) with the reference's VoxelFEA sources.

- oracle/_ref/test_gpu is the reference's native/tests/test_gpu.cpp, unchanged: on its 40x4x4 and
  100x10x10 cantilevers the GPU solve through the shim must match VoxelFEA::solve to
  max |du| / max |u| < 1e-6 (test_gpu.cpp:50-51).
- oracle/_ref/shim_driver adds solve() in every precision, precondition() against
  VoxelFEA::precondition, warm start, cancellation, the error paths, and the reference's own
  solveStatic and breakTest (static_study.cpp) taking their GPU engine through the shim.
Both run as child processes (each initialises the GPU itself); they are built here and travel with
the tree, so a missing binary means the reference sources were absent at build time."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

REF = os.path.join(ROOT, "oracle", "_ref")


def _run(name):
    exe = os.path.join(REF, name)
    if not os.path.exists(exe):
        pytest.skip(exe + " not built (reference sources absent at build time)")
    print("running", exe, flush=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(p.stdout, p.stderr, flush=True)
    return p


def test_reference_test_gpu_through_the_shim():
    p = _run("test_gpu")
    assert p.returncode == 0
    assert "skipped" not in p.stdout
    assert "max relative difference" in p.stdout and "FAIL" not in p.stdout


def test_shim_driver_and_reference_callers():
    p = _run("shim_driver")
    assert p.returncode == 0 and "SHIM OK" in p.stdout
    assert "solveStatic: engine GPU" in p.stdout and "breakTest: engine GPU" in p.stdout
