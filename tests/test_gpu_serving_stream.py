"""The serving stream (Poisson arrivals of void-block parts of uneven sizes, 1-8 load cases each, 30 %
warm re-solves of an earlier part with 1 % of its solid voxels removed and the floating voxels
pruned, warm-started from the earlier solution, as breakTest re-solves, static_study.cpp:219, 235)
through the request loop, on a bounded prefix of the stream with edges capped at 128 voxels.

Parts of at most 3 * 2^20 DOFs run in the server's shared-launch pack here (pack_max_dof), the
larger ones on their own streams. Every served result must equal the same request solved alone on its own model, bit for bit,
and the three smallest parts' served solutions must match the oracle's FP64 VoxelFEA::solve to 1e-6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_REQUESTS, EDGE_MAX, TOL, COARSEST = 24, 128, 1e-6, 1200


def _inputs(P, W, reqs, parts, served):
    """(dims, density, nu, bc, F, x0) of request r; warm requests need their parent's served u."""
    out = {}
    for r in reqs:
        if r["warm"]:
            base = parts[r["parent"]]
            dens = W.remove_voxels(base["density"], r["remove_fraction"], r["remove_seed"])
            dens, removed = P.prune_floating(base["dims"], dens, W.held_nodes(base["bc"]))
            out[r["id"]] = (base["dims"], dens, base["nu"], base["bc"], W.serving_loads(base, r["nrhs"], r["remove_seed"]), served[r["parent"]])
        else:
            w = parts[r["id"]]
            out[r["id"]] = (w["dims"], w["density"], w["nu"], w["bc"], W.serving_loads(w, r["nrhs"], r["part_seed"]), None)
    return out


def test_serving_stream_prefix(P, W, cpu):
    reqs = W.serving_stream(N_REQUESTS, seed=5, edge_max=EDGE_MAX)
    warm = [r for r in reqs if r["warm"]]
    cold = [r for r in reqs if not r["warm"]]
    assert len(warm) >= 4 and len(cold) >= 8
    parts = {r["id"]: W.void_block(r["dims"], seed=r["part_seed"], nu=r["nu"], E=r["E"]) for r in cold}
    srv = P.Server(max_inflight=4, pack_max_dof=3 << 20)
    served, results = {}, {}
    # the cold requests together, then the warm re-solves together (each needs its parent's result)
    for phase in (cold, warm):
        args = _inputs(P, W, phase, parts, served)
        tickets = []
        for r in phase:
            dims, dens, nu, bc, F, x0 = args[r["id"]]
            tickets.append((r, srv.submit(dims, dens, nu, bc, F, x0=x0, coarsest_max_dof=COARSEST, tol=TOL, precision="mixed")))
        for r, t in tickets:
            u, res = srv.wait(t)
            print("request", r["id"], "warm" if r["warm"] else "cold", r["dims"], r["nrhs"], res["iterations"], "%.1f ms" % res["total_ms"], flush=True)
            assert res["status"] == 0 and res["converged"]
            served[r["id"]] = u
            results[r["id"]] = (args[r["id"]], res)
    srv.close()
    packed = sum(res["packed"] for _, res in results.values())
    print("packed", packed, "of", len(results), flush=True)
    assert packed >= 8
    # each alone: its own model, the same multi-instance solve
    for r in reqs:
        (dims, dens, nu, bc, F, x0), res = results[r["id"]]
        m = P.Model(dims, dens, nu, bc, COARSEST, device_build=True)
        U, infos = m.solve_multi(F, tol=TOL, X0=None if x0 is None else x0, precision="mixed")
        m.close()
        assert all(i["converged"] for i in infos) and sum(i["iterations"] for i in infos) == res["iterations"]
        assert np.array_equal(U, served[r["id"]]), r["id"]
    # the three smallest parts against the oracle's FP64 solve
    smallest = sorted(cold, key=lambda r: int(np.prod(r["dims"])))[:3]
    for r in smallest:
        (dims, dens, nu, bc, F, x0), res = results[r["id"]]
        o = cpu.Model(dims, dens, nu, bc, COARSEST)
        for k in range(F.shape[0]):
            uo = o.solve(F[k], tol=1e-10, max_iter=1000)["u"]
            err = np.linalg.norm(served[r["id"]][k] - uo) / np.linalg.norm(uo)
            print("oracle", r["id"], r["dims"], k, "%.2e" % err, flush=True)
            assert err < 1e-6
