"""The device voxelizer (include/psx.h psx_voxelize) against the oracle's restatement of
voxelize.cpp:25-128, bit for bit: the oracle is pinned to the reference's own compiled voxelizer,
to closed forms (box, octahedron) and to the reference's sphere-volume test (tests/test_oracle.py)."""
import numpy as np
import pytest

import meshes

pytestmark = pytest.mark.gpu

MESHES = {
    "sphere": meshes.uv_sphere,
    "torus": meshes.torus,
    "soup": meshes.soup,
    "box": lambda: meshes.box((-3, -2, -1), (5, 2.5, 1.7)),
    "octahedron": lambda: meshes.octahedron(7.3, (0.21, -0.37, 0.13)),
}


@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_voxelize_matches_oracle(P, cpu, mesh):
    pos, idx = MESHES[mesh]()
    for n, sub, pad in ((24, 2, 0), (17, 3, 1), (40, 4, 0), (30, 5, 2)):
        g = P.grid_for_box(*meshes.bounds(pos), n, pad)
        f = P.voxelize(pos, idx, *g, sub)
        fo = cpu.voxelize(pos, idx, *g, sub)
        print(mesh, n, sub, pad, g[2], int((f != fo).sum()), flush=True)
        assert np.array_equal(f, fo)


def test_voxelize_empty_and_errors(P):
    pos, idx = meshes.box((0, 0, 0), (1, 1, 1))
    g = P.grid_for_box(*meshes.bounds(pos), 4)
    assert not P.voxelize(pos, idx[:0], *g, 2).any()  # no triangles: nothing filled
    with pytest.raises(P.PsxError):
        P.voxelize(pos, np.array([0, 1, 99], np.uint32), *g, 2)  # index past the vertices
