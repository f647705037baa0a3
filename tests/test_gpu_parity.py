"""HIP path against the oracle (the CPU restatement of VoxelFEA) on the same seeded inputs.

Tolerances: FP64 operator 1e-13 relative (different summation order only); FP32 operator 1e-5;
V-cycle 1e-4 (FP32 rounding in a different order); solutions 1e-7 relative at tol 1e-10 — far
below the 1e-6 displacement contract (test_gpu.cpp:50-51).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases(W):
    return [W.cantilever(4), W.c1_cantilever(), W.void_block((48, 16, 16), seed=3), W.simp_block((32, 16, 16), nrhs=1)]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def models(P, W, cpu):
    out = []
    for w in _cases(W):
        print("model", w["name"], flush=True)
        g = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
        o = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
        out.append((w, g, o))
    return out


def test_levels_match(models):
    for w, g, o in models:
        assert g.num_levels == o.num_levels
        for l in range(g.num_levels):
            assert g.level(l)["freeDof"] == o.level(l)["freeDof"]
            assert g.level(l)["nOwn"] == o.level(l)["nOwn"]


def test_apply64_every_level(models):
    for w, g, o in models:
        for l in range(g.num_levels):
            n = g.level(l)["nDof"]
            x = np.random.default_rng(l).standard_normal(n)
            for raw in (False, True):
                y = g.apply(x, l, raw)
                yo = o.apply(x, l, raw)
                print(w["name"], l, raw, _rel(y, yo), flush=True)
                assert _rel(y, yo) < 1e-13


def test_apply32_every_level(models):
    for w, g, o in models:
        for l in range(g.num_levels):
            n = g.level(l)["nDof"]
            x = np.random.default_rng(l).standard_normal(n).astype(np.float32)
            y = g.apply32(x, l)
            yo = o.apply32(x, l)
            assert _rel(y.astype(np.float64), yo.astype(np.float64)) < 1e-5


def test_lanczos_lmax(models):
    for w, g, o in models:
        lg = g.estimate_smoothers()
        lo = o.prepare_smoothers()
        for l in range(g.num_levels - 1):
            assert abs(lg[l] / lo[l] - 1) < 1e-9, (w["name"], l, lg[l], lo[l])


def test_precondition(models):
    for w, g, o in models:
        r = np.random.default_rng(7).standard_normal(g.n_dof)
        z = g.precondition(r)
        zo = o.precondition(r)
        print(w["name"], "precondition", _rel(z, zo), flush=True)
        assert _rel(z, zo) < 1e-4


@pytest.mark.parametrize("precision", ["fp64", "ref"])
def test_solve_matches_oracle(models, precision):
    for w, g, o in models:
        s = g.solve(w["f"] if w["f"].ndim == 1 else w["f"][0], tol=1e-10, max_iter=500, precision=precision, check_every=4)
        so = o.solve(w["f"] if w["f"].ndim == 1 else w["f"][0], tol=1e-10, max_iter=500)
        err = _rel(s["u"], so["u"])
        print(w["name"], precision, "its", s["iterations"], so["iterations"], "res", s["residual"], "err", err, flush=True)
        assert s["converged"] and so["converged"]
        assert err < 1e-7
        assert abs(s["iterations"] - so["iterations"]) <= 2
