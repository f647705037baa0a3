"""The bend-test pipeline end to end on the product alone (test_model.cpp:79-110): a 200 x 20 x 10 mm
box, StructuralModel on the grid of each resolution (psx_structural_model: device voxelizer and the
density rule), fixtures and loads assembled onto grid nodes (psx_assemble: the x = min face clamped,
an 800 N tip load on the x = max face), the hierarchy built on the device and the static solve. The
tip deflection must be within 7 % of Timoshenko beam theory at every resolution, as the reference's
own test asserts; the reaction at the clamp balances the load."""
import numpy as np
import pytest

import meshes

pytestmark = pytest.mark.gpu

STEEL = {"E": 200.0, "nu": 0.29, "yield": 350.0, "uts": 420.0, "density": 7870.0, "brittle": False}


def _select(V, T, test):
    p = V.reshape(-1, 3).astype(np.float64)
    t = T.reshape(-1, 3)
    c = (p[t[:, 0]] + p[t[:, 1]] + p[t[:, 2]]) / 3.0
    n = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return np.nonzero(test(c, n))[0].astype(np.int32)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_beam_deflection_matches_theory(P, precision):
    V, T = meshes.box((-100.0, -10.0, -5.0), (100.0, 10.0, 5.0))
    eps = 0.02
    clamp = [[{"tris": _select(V, T, lambda c, n: (c[:, 0] < -100 + eps) & (n[:, 0] < -0.9))}]]
    tip = [{"type": "force", "magnitude": 800.0, "dir": (0.0, -1.0, 0.0), "patches": [{"tris": _select(V, T, lambda c, n: (c[:, 0] > 100 - eps) & (n[:, 0] > 0.9))}]}]
    assert len(clamp[0][0]["tris"]) == 2 and len(tip[0]["patches"][0]["tris"]) == 2
    E, nu = 200e9, STEEL["nu"]
    I = 0.01 * 0.02 ** 3 / 12
    theory = 800 * 0.008 / (3 * E * I) + 800 * 0.2 / ((5.0 / 6) * (E / (2 * (1 + nu))) * 0.02 * 0.01)
    p = V.reshape(-1, 3)
    for res in (40, 64, 90, 104, 128):
        o, h, d = P.grid_for_box(p.min(0).astype(np.float64), p.max(0).astype(np.float64), res)
        dens, act, surf, info = P.structural_model(V, T, o, h, d)
        bc, f, ai, warn = P.assemble(V, T, o, h, d, dens, clamp, tip, False, STEEL, 1e-3)
        assert not warn and abs(ai["total"][1] + 800.0) < 1e-9 and ai["fixed_nodes"] > 0
        m = P.Model(d, dens, nu, bc, 1100, device_build=True)
        s = m.solve(f, tol=1e-7, max_iter=1000, precision=precision)
        assert s["converged"]
        tip_u = np.abs(s["u"][1::3]).max() / (E * h * 1e-3)
        ratio = tip_u / theory
        # equilibrium: the clamp's reaction balances the 800 N load
        reac = m.reactions(s["u"], f)
        print("%s %d voxels %s: tip %.4f mm vs %.4f mm (x%.3f), %d its, reaction %s" % (precision, res, d, tip_u * 1e3, theory * 1e3, ratio,
                                                                                       s["iterations"], np.round(reac, 3)), flush=True)
        assert 0.93 < ratio < 1.07
        assert abs(reac[1] - 800.0) < 1.0 and abs(reac[0]) < 1.0 and abs(reac[2]) < 1.0
        m.close()
