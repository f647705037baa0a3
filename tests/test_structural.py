"""Input contract of a CAD part on the host (include/psx.h psx_wall_thickness, psx_thin_wall_density,
psx_assemble): the typical wall thickness (voxelize.cpp:190-217), the thin-wall density
(voxelize.cpp:219-276) and the assembly of fixtures and loads onto grid nodes (structural.cpp:107-209)
must equal the reference's bit for bit — fixture masks, nodal forces, load totals, warnings — on
parts its own buildPart makes (oracle/_ref, compiled from the reference's sources). Closed forms pin
what the reference leaves open: the wall of a hollow box is its thickness, every load and gravity
total is what the nodal forces sum to, and the reference's own error messages come back."""
import numpy as np
import pytest

import meshes
from conftest import ref_available

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref (the reference compiled from its sources) is not built")

MAT = {"E": 200.0, "nu": 0.3, "yield": 250.0, "uts": 400.0, "density": 7850.0, "brittle": False}


@pytest.fixture(scope="module")
def parts(cpu):
    out = {}
    for name, (pos, idx) in {"thin": meshes.hollow_box((0, 0, 0), (40, 20, 30), 1.2), "thick": meshes.hollow_box((0, 0, 0), (30, 24, 20), 4.0),
                             "torus": meshes.torus()}.items():
        out[name] = cpu.ref_build_part(pos, idx)
    return out


def _eq(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name", ["thin", "thick", "torus"])
def test_wall_thickness_equals_reference(P, cpu, parts, name):
    hnd, V, T, N, A = parts[name]
    for samples in (1500, 200):
        r, p = cpu.ref_wall_thickness(hnd, samples), P.wall_thickness(V, T, samples)
        assert r == p or (np.isinf(r) and np.isinf(p)), (name, samples, r, p)
    if name == "thin":
        # axis-aligned walls: the ray from just inside a face meets the opposite face at the wall thickness
        assert abs(P.wall_thickness(V, T) - 1.2) < 1e-6
    if name == "thick":
        assert abs(P.wall_thickness(V, T) - 4.0) < 1e-6


@pytest.mark.parametrize("name", ["thin", "thick"])
def test_thin_wall_density_equals_reference(P, cpu, parts, name):
    hnd, V, T, N, A = parts[name]
    mh, o, h, d, dens, act, surf, wall, cnt = cpu.ref_structural_model(hnd, 40)
    for maxt in (1.5 * h, 5.0 * h):
        r = cpu.ref_thin_wall_density(hnd, o, h, d, maxt)
        p = P.thin_wall_density(V, T, o, h, d, maxt)
        assert _eq(r, p), (name, maxt, int(np.sum(r != p)))
        assert (p >= 0).all() and (p <= 1).all()
    if name == "thin":
        assert cnt[1] > 0  # the thin-wall layer was used


def _loads(nt, seed):
    rng = np.random.default_rng(seed)
    return [{"type": "force", "magnitude": 100.0, "dir": (0.0, -1.0, 0.2), "patches": [{"tris": np.arange(nt // 2, nt // 2 + nt // 8)}]},
            {"type": "pressure", "magnitude": 2.0, "patches": [{"tris": np.arange(nt // 3, nt // 3 + nt // 10), "clip": ((20.0, 10.0, 15.0), 8.0)},
                                                               {"tris": np.arange(nt // 3, nt // 3 + nt // 20)}]},
            {"type": "wind", "forces": (rng.standard_normal(3 * nt) * (rng.random(3 * nt) < 0.1)).astype(np.float32)}]


@pytest.mark.parametrize("name", ["thin", "thick", "torus"])
def test_assemble_equals_reference(P, cpu, parts, name):
    hnd, V, T, N, A = parts[name]
    nt = T.size // 3
    mh, o, h, d, dens, act, surf, wall, cnt = cpu.ref_structural_model(hnd, 36)
    nn = int(np.prod(d + 1))
    cases = [([[{"tris": np.arange(0, nt // 10)}]], _loads(nt, 3), True),
             ([[{"tris": np.arange(0, nt // 20)}, {"tris": np.arange(nt // 40, nt // 15), "clip": ((0.0, 0.0, 0.0), 9.0)}]], _loads(nt, 4)[:2], False),
             ([[{"tris": [5]}], [{"tris": np.arange(nt - 50, nt)}]], [], True)]
    for fix, loads, grav in cases:
        br, fr, ir, wr = cpu.ref_assemble(mh, nn, fix, loads, grav, MAT, 1e-3)
        bp, fp, ip, wp = P.assemble(V, T, o, h, d, dens, fix, loads, grav, MAT, 1e-3)
        assert _eq(br, bp), (name, int(np.sum(br != bp)))
        assert _eq(fr, fp), (name, np.abs(fr - fp).max())
        assert ir["fixed_nodes"] == ip["fixed_nodes"] and ir["total"] == ip["total"] and wr == wp, (ir, ip, wr, wp)
        # fixtures hold whole nodes; the nodal forces add up to the reported total
        assert (bp.reshape(-1, 3).min(1) == bp.reshape(-1, 3).max(1)).all()
        assert np.allclose(fp.reshape(-1, 3).sum(0), ip["total"], rtol=1e-9, atol=1e-9 * max(1.0, np.abs(fp).max()))
    # a force load puts exactly its magnitude along its direction
    fx = [{"type": "force", "magnitude": 37.5, "dir": (3.0, 0.0, 4.0), "patches": [{"tris": np.arange(0, nt // 4)}]}]
    bp, fp, ip, wp = P.assemble(V, T, o, h, d, dens, [], fx, False, MAT, 1e-3)
    assert np.allclose(ip["total"], [22.5, 0.0, 30.0], rtol=1e-12) and not bp.any()
    # gravity: total weight = material density x solid volume x g, downward along y
    bp, fp, ip, wp = P.assemble(V, T, o, h, d, dens, [], [], True, MAT, 1e-3)
    mass = MAT["density"] * float(np.sum(dens.astype(np.float64))) * (h * 1e-3) ** 3
    assert ip["total"][0] == 0 and ip["total"][2] == 0 and abs(ip["total"][1] + mass * 9.81) < 1e-9 * mass * 9.81


def test_assemble_errors_and_warnings(P, cpu, parts):
    hnd, V, T, N, A = parts["thin"]
    nt = T.size // 3
    mh, o, h, d, dens, act, surf, wall, cnt = cpu.ref_structural_model(hnd, 24)
    nn = int(np.prod(d + 1))
    bad = [({"E": -1.0}, [], 1e-3), ({"nu": 0.5}, [], 1e-3), ({"uts": 100.0}, [], 1e-3), ({}, [], 0.0),
           ({}, [{"type": "force", "magnitude": -1.0, "patches": []}], 1e-3), ({}, [{"type": "force", "magnitude": 1.0, "dir": (0, 0, 0), "patches": []}], 1e-3),
           ({}, [{"type": "wind", "forces": np.zeros(6, np.float32)}], 1e-3)]
    for mat_over, loads, units in bad:
        mat = dict(MAT, **mat_over)
        with pytest.raises(ValueError) as er:
            cpu.ref_assemble(mh, nn, [], loads, False, mat, units)
        with pytest.raises(P.PsxError) as ep:
            P.assemble(V, T, o, h, d, dens, [], loads, False, mat, units)
        assert ep.value.code == P.ERR_ARGUMENT and str(er.value) in str(ep.value), (str(er.value), str(ep.value))
    # warnings: a fixture under no voxel (a clip far away), a load with no samples, a tiny fixture
    fix = [[{"tris": np.arange(0, 10), "clip": ((1e3, 1e3, 1e3), 1.0)}]]
    loads = [{"type": "force", "magnitude": 1.0, "patches": [{"tris": np.arange(0, 10), "clip": ((1e3, 1e3, 1e3), 1.0)}]}]
    br, fr, ir, wr = cpu.ref_assemble(mh, nn, fix, loads, False, MAT, 1e-3)
    bp, fp, ip, wp = P.assemble(V, T, o, h, d, dens, fix, loads, False, MAT, 1e-3)
    assert wr == wp and len(wp) == 2 and ip["n_warnings"] == 2
