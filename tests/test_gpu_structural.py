"""StructuralModel on the device's voxelizer (include/psx.h psx_structural_model): voxel densities
(fill fraction rule and the thin-wall layer), active and surface nodes, voxel counts and the
wall thickness must equal the reference's StructuralModel(part, resolution) bit for bit, on parts
its own buildPart makes (oracle/_ref)."""
import numpy as np
import pytest

import meshes
from conftest import ref_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not ref_available(), reason="oracle/_ref (the reference compiled from its sources) is not built")]


@pytest.mark.parametrize("name,res", [("thin", 48), ("thick", 40), ("torus", 56), ("sphere", 32)])
def test_structural_model_equals_reference(P, cpu, name, res):
    pos, idx = {"thin": meshes.hollow_box((0, 0, 0), (40, 20, 30), 1.2), "thick": meshes.hollow_box((0, 0, 0), (30, 24, 20), 4.0),
                "torus": meshes.torus(), "sphere": meshes.uv_sphere()}[name]
    hnd, V, T, N, A = cpu.ref_build_part(pos, idx)
    mh, o, h, d, dens, act, surf, wall, cnt = cpu.ref_structural_model(hnd, res)
    # the grid: gridForBox of the part's bounding box
    p = V.reshape(-1, 3)
    go, gh, gd = P.grid_for_box(p.min(0).astype(np.float64), p.max(0).astype(np.float64), res)
    assert np.array_equal(go, o) and gh == h and tuple(gd) == tuple(d)
    print("structural model:", name, tuple(d), "wall", wall, "counts", cnt, flush=True)
    dp, ap, sp, info = P.structural_model(V, T, o, h, d)
    assert np.array_equal(dp.view(np.uint32), dens.view(np.uint32)), int(np.sum(dp != dens))
    assert np.array_equal(ap, act) and np.array_equal(sp, surf)
    assert (info["voxel_count"], info["thin_voxels"]) == cnt
    assert info["wall_thickness"] == wall or (np.isinf(wall) and np.isinf(info["wall_thickness"]))
    assert info["used_shell"] == (1 if wall < 2.5 * h else 0)
