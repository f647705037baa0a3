"""The serving loop (include/psx.h psx_server_*): requests admitted as slots and device memory free
up, run concurrently on their own streams. Every result must equal the same request solved alone,
bit for bit (per-request deterministic reductions, nothing shared on the device)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _parts(W):
    return [(W.void_block((24, 16, 12), seed=11), 2), (W.void_block((40, 12, 10), seed=12), 1), (W.cantilever(3), 3),
            (W.void_block((20, 20, 20), seed=13, void_density=0.0, void_fraction=0.15), 2), (W.void_block((32, 8, 8), seed=14), 1)]


def _loads(w, nrhs, seed):
    f = w["f"] if w["f"].ndim == 1 else w["f"][0]
    rng = np.random.default_rng(seed)
    return np.stack([f * (1 + k) + 0.01 * np.abs(f).max() * rng.standard_normal(f.size) * (f != 0) for k in range(nrhs)])


# pack_max 0: the small parts run in the shared-launch pack; 1: each on its own stream; mfma 1: the
# smoothing of levels 0 and 1 as element products on the matrix cores at degree 3 (the bench's setting)
@pytest.mark.parametrize("precision,device_build,pack_max,mfma", [("fp64", True, 0, 0), ("mixed", True, 0, 0), ("mixed", False, 0, 0),
                                                                  ("mixed", True, 1, 0), ("mixed", True, 0, 1)])
def test_served_results_equal_lone_solves(P, W, precision, device_build, pack_max, mfma):
    parts = _parts(W)
    kw = dict(mfma_smoother=1, cheb_degree=3) if mfma else {}
    alone = []
    for q, (w, nrhs) in enumerate(parts):
        m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300, device_build=device_build)
        F = _loads(w, nrhs, q)
        alone.append(np.stack([m.solve(F[k], tol=1e-9, precision=precision, **kw)["u"] for k in range(nrhs)]))
        m.close()
    srv = P.Server(max_inflight=3, pack_max=pack_max)
    tickets = [srv.submit(w["dims"], w["density"], w["nu"], w["bc"], _loads(w, nrhs, q), coarsest_max_dof=300, tol=1e-9, precision=precision,
                          device_build=device_build, **kw)
               for q, (w, nrhs) in enumerate(parts)]
    for q, t in enumerate(tickets):
        u, res = srv.wait(t)
        print(parts[q][0]["name"], res, flush=True)
        assert res["status"] == 0 and res["converged"]
        assert res["packed"] == (1 if pack_max != 1 else 0)
        assert np.array_equal(u, alone[q])
    assert srv.pending() == 0
    srv.close()


def test_resident_part_and_admission(P, W):
    w, _ = _parts(W)[0]
    f = _loads(w, 2, 5)
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300, device_build=True)
    lone = [m.solve(f[k], tol=1e-9, precision="mixed")["u"] for k in range(2)]
    # a budget smaller than any request: requests run one at a time, in order
    srv = P.Server(max_inflight=4, hbm_budget=1)
    t1 = srv.submit(w["dims"], w["density"], w["nu"], w["bc"], f[0], part_id=7, coarsest_max_dof=300, tol=1e-9, precision="mixed")
    t2 = srv.submit(w["dims"], w["density"], w["nu"], w["bc"], f[1], part_id=7, coarsest_max_dof=300, tol=1e-9, precision="mixed")
    u1, r1 = srv.wait(t1)
    u2, r2 = srv.wait(t2)
    assert r1["reused_part"] == 0 and r2["reused_part"] == 1 and r2["build_ms"] == 0
    assert r2["queue_ms"] >= r1["solve_ms"] * 0.5  # admitted after the first finished
    assert np.array_equal(u1[0], lone[0]) and np.array_equal(u2[0], lone[1])
    # warm start from the first solution converges at once
    t3 = srv.submit(w["dims"], w["density"], w["nu"], w["bc"], f[0], x0=u1[0], part_id=7, coarsest_max_dof=300, tol=1e-6, precision="mixed")
    u3, r3 = srv.wait(t3)
    assert r3["iterations"] == 0 and r3["converged"]
    srv.close()


def test_failed_request_leaves_others_running(P, W):
    w, _ = _parts(W)[1]
    srv = P.Server(max_inflight=2)
    bad = srv.submit(w["dims"], w["density"], 0.7, w["bc"], w["f"], coarsest_max_dof=300)  # nu outside (-1, 0.5)
    good = srv.submit(w["dims"], w["density"], w["nu"], w["bc"], w["f"], coarsest_max_dof=300, tol=1e-8)
    with pytest.raises(P.PsxError):
        srv.wait(bad)
    u, res = srv.wait(good)
    assert res["converged"]
    srv.close()
