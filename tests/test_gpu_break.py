"""The break test on the device (include/psx.h psx_break_test, psx_prune_floating_device) against
the oracle's restatement of breakTest (static_study.cpp:197-279), whose steps are checked against
the oracle's own solve, stresses and prune in the CPU suite.

Both sides solve to a relative residual of 1e-10, so the nodal stresses agree to about a float
rounding. The crack is the worst voxels by their worst corner stress; neighbours share corners, so
equal measures are common, and with tie_order 0 the device orders them as the reference's std::sort
does (the same sort over the same measures), so the sequences agree voxel for voxel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _break_case(W, dims=(20, 6, 6), seed=3):
    nx, ny, nz = dims
    bc = W.clamp_x0(dims)
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    f = np.zeros((NZ, NY, NX, 3))
    f[:, :, nx, 2] = -2000.0 / (NY * NZ)  # 2 kN down over the free end
    density = np.random.default_rng(seed).uniform(0.6, 1.0, nx * ny * nz).astype(np.float32)
    return dims, density, bc, f.reshape(-1)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_prune_device_matches_oracle(P, W, cpu, seed):
    rng = np.random.default_rng(seed)
    dims = (37, 23, 11)
    density = (rng.uniform(0, 1, 37 * 23 * 11) > 0.45).astype(np.float32) * rng.uniform(0.1, 1, 37 * 23 * 11).astype(np.float32)
    held = np.zeros(38 * 24 * 12, np.uint8)
    held.reshape(12, 24, 38)[:, :, 0] = 1
    out, removed = P.prune_floating_device(dims, density, held)
    ref, rref = cpu.prune_floating(dims, density, held)
    print(seed, removed, rref, flush=True)
    assert removed == rref and removed > 0
    assert np.array_equal(out, ref)


def test_prune_device_long_path(P, cpu):
    # a serpentine path: one BFS level per voxel along it (many batches of levels)
    nx, ny, nz = 40, 40, 1
    d = np.zeros((nz, ny, nx), np.float32)
    for j in range(0, ny, 2):
        d[0, j, :] = 1
        d[0, j + 1 if j + 1 < ny else j, (nx - 1) if (j // 2) % 2 == 0 else 0] = 1
    d[0, 31, 0] = 0  # cut the path: rows 32.. float
    held = np.zeros((nz + 1) * (ny + 1) * (nx + 1), np.uint8)
    held[0] = 1
    out, removed = P.prune_floating_device((nx, ny, nz), d.reshape(-1), held)
    ref, rref = cpu.prune_floating((nx, ny, nz), d.reshape(-1), held)
    assert removed == rref == 164 and np.array_equal(out, ref)


def _compare(dev, ref, exact_order=True):
    assert dev["reason"] == ref["reason"]
    assert dev["removed"] == ref["removed"]
    assert len(dev["steps"]) == len(ref["steps"])
    assert dev["peak"] == pytest.approx(ref["peak"], rel=1e-6)
    for a, b in zip(dev["steps"], ref["steps"]):
        assert a["voxels"] == b["voxels"]
        assert a["lambda"] == pytest.approx(b["lambda"], rel=1e-6)
        assert a["max_disp"] == pytest.approx(b["max_disp"], rel=1e-6)
        if exact_order:
            assert list(a["cracked"]) == list(b["cracked"])
        else:
            assert sorted(a["cracked"]) == sorted(b["cracked"])
        assert list(a["detached"]) == list(b["detached"])


@pytest.mark.parametrize("principal", [False, True])
def test_break_sequence_matches_oracle(P, W, cpu, principal):
    dims, density, bc, f = _break_case(W)
    kw = dict(principal=principal, max_steps=12, tol=1e-10)
    dev = P.break_test(dims, density, 0.3, bc, f, 200e9, 1e-3, 150e6, precision="fp64", **kw)
    ref = cpu.break_test(dims, density, 0.3, bc, f, 200e9, 1e-3, 150e6, **kw)
    print(principal, dev["reason"], [len(s["cracked"]) for s in dev["steps"]], [len(s["detached"]) for s in dev["steps"]], flush=True)
    _compare(dev, ref)


def test_break_to_separation_mixed_and_index_ties(P, W, cpu):
    dims, density, bc, f = _break_case(W, (16, 4, 4), seed=5)
    ref = cpu.break_test(dims, density, 0.3, bc, f, 200e9, 1e-3, 150e6, max_steps=40, tol=1e-10)
    assert ref["reason"] == "separated"
    dev = P.break_test(dims, density, 0.3, bc, f, 200e9, 1e-3, 150e6, max_steps=40, tol=1e-10, precision="mixed")
    _compare(dev, ref)
    # equal measures by voxel index: the same steps up to the first tie that straddles the cap (two
    # voxels sharing the worst corner with room for one), where the lower voxel index is cracked
    idx = P.break_test(dims, density, 0.3, bc, f, 200e9, 1e-3, 150e6, max_steps=40, tol=1e-10, precision="mixed", tie_order=1)
    assert idx["reason"] == "separated"
    first = next(s for s, (a, b) in enumerate(zip(idx["steps"], ref["steps"])) if sorted(a["cracked"]) != sorted(b["cracked"]))
    print("index ties diverge at step", first, flush=True)
    _compare({**idx, "steps": idx["steps"][:first], "peak": 0}, {**ref, "steps": ref["steps"][:first], "peak": 0}, exact_order=False)
    a, b = idx["steps"][first], ref["steps"][first]
    assert a["lambda"] == pytest.approx(b["lambda"], rel=1e-6) and len(a["cracked"]) == len(b["cracked"])
    assert sorted(a["cracked"]) < sorted(b["cracked"])


def test_break_fields_and_stop(P, W, cpu):
    dims, density, bc, f = _break_case(W, (12, 4, 4), seed=7)
    density.reshape(4, 4, 12)[1:3, 1:3, 9] = 0  # a hole near the free end
    seen = []
    dev = P.break_test(dims, density, 0.3, bc, f, 70e9, 2e-3, 80e6, max_steps=10, tol=1e-10, precision="fp64", fields=True,
                       on_step=lambda st: seen.append(st["step"]) or st["step"] == 2)
    assert dev["reason"] == "stopped" and seen == [0, 1, 2] and len(dev["steps"]) == 3
    ref = cpu.break_test(dims, density, 0.3, bc, f, 70e9, 2e-3, 80e6, max_steps=3, tol=1e-10, fields=True)
    _compare({**dev, "reason": ref["reason"]}, ref)
    for a, b in zip(dev["steps"], ref["steps"]):
        scale = np.abs(b["vm"]).max()
        assert np.abs(a["vm"] - b["vm"]).max() <= 1e-5 * scale
        umax = np.abs(a["u"]).max()
        # the largest displacement is the step's max_disp (float rounding of u lambda)
        assert np.sqrt((a["u"].reshape(-1, 3).astype(np.float64) ** 2).sum(axis=1)).max() == pytest.approx(a["max_disp"], rel=1e-6)
        assert umax > 0
        assert np.all(a["vm"][a["active"] == 0] == 0)


def test_break_rejects_bad_input(P, W):
    dims, density, bc, f = _break_case(W, (6, 3, 3))
    with pytest.raises(P.PsxError):
        P.break_test(dims, density, 0.3, bc, f, 200e9, 1e-3, 0.0)
    # nothing held: the load is unsupported before any step
    r = P.break_test(dims, density, 0.3, np.zeros_like(bc), f, 200e9, 1e-3, 100e6)
    assert r["reason"] == "unsupported load" and r["steps"] == [] and r["removed"] == density.size
