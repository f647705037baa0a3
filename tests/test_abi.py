"""The C-ABI library loads and exports every entry point include/psx.h declares; without a GPU the
device calls fail loudly (no CPU fallback)."""
import os
import re

import pytest

from conftest import ROOT


def declared():
    src = open(os.path.join(ROOT, "include", "psx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(psx_[a-z0-9_]+)\s*\(", src)) - {"psx_progress_fn"})


def test_every_declared_symbol_is_exported(P):
    L = P.lib()
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(P.SIGNATURES), set(names) ^ set(P.SIGNATURES)


def test_no_device_fails_loudly(P, W):
    if P.available():
        pytest.skip("a GPU is present")
    w = W.cantilever(2)
    with pytest.raises(P.PsxError) as e:
        P.Model(w["dims"], w["density"], w["nu"], w["bc"])
    assert e.value.code == P.ERR_NO_DEVICE


def test_missing_library_raises(P, monkeypatch):
    monkeypatch.setattr(P, "_lib", None)
    monkeypatch.setattr(P, "LIB_PATH", "/nonexistent/libpsx.so")
    with pytest.raises(RuntimeError):
        P.lib()


def test_server_without_device_raises(P):
    if P.available():
        pytest.skip("a GPU is present")
    with pytest.raises(P.PsxError) as e:
        P.Server(max_inflight=2)
    assert e.value.code == P.ERR_NO_DEVICE


def test_grid_for_box_matches_oracle(P, cpu):
    import numpy as np

    import meshes
    for pos in (meshes.uv_sphere()[0], meshes.torus()[0], meshes.box((-3, -2, -1), (5, 2.5, 1.7))[0]):
        mn, mx = meshes.bounds(pos)
        for n, pad in ((10, 0), (37, 2), (128, 1)):
            o, h, d = P.grid_for_box(mn, mx, n, pad)
            oo, ho, do = cpu.grid_for_box(mn, mx, n, pad)
            assert np.array_equal(o, oo) and h == ho and d == do
