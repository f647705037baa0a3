"""The allocation-free prune (include/psx.h psx_prune_floating_device_noalloc): its static scratch is
shared by every call on a GPU and keeps the last call's contents, so consecutive calls, each smaller
than the one before, must each equal the allocating entry point's result (psx_prune_floating_device)
bit for bit; the oracle's pruneFloating checks both."""
import ctypes as C

import numpy as np
import pytest

LIMIT = 1 << 23


def _part(dims, seed, solid=0.55):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    density = (rng.uniform(0, 1, n) < solid).astype(np.float32) * rng.uniform(0.1, 1, n).astype(np.float32)
    nx, ny, nz = dims
    held = np.zeros((nz + 1, ny + 1, nx + 1), np.uint8)
    held[:, :, 0] = 1
    return density, held.reshape(-1)


def _serpentine():
    nx, ny, nz = 40, 40, 1
    d = np.zeros((nz, ny, nx), np.float32)
    for j in range(0, ny, 2):
        d[0, j, :] = 1
        d[0, j + 1 if j + 1 < ny else j, (nx - 1) if (j // 2) % 2 == 0 else 0] = 1
    d[0, 31, 0] = 0
    held = np.zeros((nz + 1) * (ny + 1) * (nx + 1), np.uint8)
    held[0] = 1
    return (nx, ny, nz), d.reshape(-1), held


def test_oversized_grid_is_rejected_before_any_device_work(P):
    # non-null stand-in pointers: the size check comes first and touches nothing
    dummy = C.c_void_p(256)
    removed = C.c_int(-1)
    code = P.lib().psx_prune_floating_device_noalloc(LIMIT + 1, 1, 1, dummy, dummy, dummy, C.byref(removed))
    assert code == P.ERR_ARGUMENT
    assert "limit" in P.lib().psx_last_error().decode()
    assert removed.value == -1


@pytest.mark.gpu
def test_noalloc_equals_allocating_prune_on_shrinking_calls(P, W, cpu):
    cases = []
    w = W.void_block((512, 128, 128), seed=1, void_density=0.0)  # the largest accepted part: 2^23 voxels
    cases.append((w["dims"], w["density"], W.held_nodes(w["bc"])))
    d, h = _part((96, 64, 48), 7)
    cases.append(((96, 64, 48), d, h))
    d, h = _part((37, 23, 11), 3)
    cases.append(((37, 23, 11), d, h))
    cases.append(_serpentine())  # about 800 search levels: the ring of level marks wraps many times
    for dims, density, held in cases:
        print("noalloc prune", dims, flush=True)
        a, ra = P.prune_floating_device(dims, density, held)
        b, rb = P.prune_floating_device(dims, density, held, no_alloc=True)
        assert ra == rb
        assert np.array_equal(a, b)
        if int(np.prod(dims)) < 200000:
            ref, rref = cpu.prune_floating(dims, density, held)
            assert rb == rref and np.array_equal(b, ref)
