"""The benchmark configurations at their full sizes, where the oracle is too slow to run, checked through
properties that hold at any size: symmetry of the operator (x.Ky = y.Kx), the null space of an
unsupported part's operator (rigid translations), equilibrium of the solution (the support reactions
balance the applied load, test_fea.cpp:98), agreement of the mixed-precision and FP64 solutions and
of the served multi-instance and lone solutions.

Sizes: the 256x64x64 void block (3.26 M DOF, FP64 MG-PCG to 1e-8), the 512x128x128 void block
(25.6 M DOF, mixed precision to 1e-6), the 128x64x64 SIMP block with 16 load cases (1.64 M DOF each)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_void_block_256_fp64(P, W):
    w = W.void_block((256, 64, 64), seed=1, E=70e9, nu=0.33, pressure=5e6)
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
    n = m.n_dof
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    Kx, Ky = m.apply(x, 0, True), m.apply(y, 0, True)
    assert abs(x @ Ky - y @ Kx) <= 1e-12 * np.sqrt((x @ Kx) * (y @ Ky)) * 10
    f = w["f"]
    s = m.solve(f, tol=1e-8, precision="fp64")
    print("C2 fp64", s["iterations"], s["residual"], flush=True)
    assert s["converged"] and s["residual"] <= 1e-7
    total = f.reshape(-1, 3).sum(axis=0)
    R = m.reactions(s["u"], f)
    assert np.allclose(R, -total, rtol=1e-6, atol=1e-6 * np.abs(total).max())
    mx = m.solve(f, tol=1e-8, precision="mixed")
    assert mx["converged"] and _rel(mx["u"], s["u"]) < 1e-6
    m.close()


def test_void_block_512_mixed(P, W):
    w = W.void_block((512, 128, 128), seed=1)
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
    f = w["f"]
    s = m.solve(f, tol=1e-6, precision="mixed")
    print("C3 mixed", s["iterations"], s["residual"], s["folds"], flush=True)
    assert s["converged"] and s["residual"] <= 1e-5
    total = f.reshape(-1, 3).sum(axis=0)
    R = m.reactions(s["u"], f)
    assert np.allclose(R, -total, rtol=1e-3, atol=1e-3 * np.abs(total).max())
    m.close()


def test_unsupported_part_null_space(P, W):
    # no held DOF: K annihilates rigid translations at every node (FP64 and FP32 operators)
    w = W.void_block((128, 64, 64), seed=2)
    bc = np.zeros_like(w["bc"])
    dims = w["dims"]
    nodes = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)
    bc[:3] = 1  # one node held so the hierarchy has a direct coarse solve
    m = P.Model(dims, w["density"], w["nu"], bc, 1200)
    for d in range(3):
        t = np.zeros((nodes, 3))
        t[:, d] = 1.0
        y = m.apply(t.reshape(-1), 0, True)
        scale = np.abs(m.apply(np.random.default_rng(d).standard_normal(3 * nodes), 0, True)).max()
        assert np.abs(y).max() <= 1e-12 * scale
        y32 = m.apply32(t.reshape(-1).astype(np.float32), 0, True)
        assert np.abs(y32).max() <= 1e-6 * scale
    m.close()


def test_simp_block_16_cases(P, W):
    w = W.simp_block((128, 64, 64), seed=4, nrhs=16)
    F = w["f"]
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
    U, infos = m.solve_multi(F, tol=1e-6, precision="mixed")
    assert all(i["converged"] for i in infos)
    for k in (0, 7, 15):
        lone = m.solve(F[k], tol=1e-6, precision="mixed")
        assert np.array_equal(U[k], lone["u"])
        R = m.reactions(U[k], F[k])
        total = F[k].reshape(-1, 3).sum(axis=0)
        assert np.allclose(R, -total, rtol=1e-3, atol=1e-3 * np.abs(total).max())
    m.close()
