"""Host-side input rules of the product (no device): pruneFloating against the compiled reference's
golden fixture and the oracle on random parts."""
import os

import numpy as np

from conftest import ROOT


def test_prune_floating_matches_golden_and_oracle(P, cpu):
    g = np.load(os.path.join(ROOT, "tests", "golden", "prune_16x10x8.npz"), allow_pickle=False)
    out, removed = P.prune_floating(tuple(g["dims"]), g["density"], g["held"])
    assert np.array_equal(out, g["out"]) and removed == int(g["removed"])
    rng = np.random.default_rng(3)
    for dims in [(9, 7, 5), (20, 6, 11), (4, 4, 4)]:
        nx, ny, nz = dims
        d = rng.uniform(0.1, 1, nx * ny * nz).astype(np.float32)
        d[rng.uniform(size=d.size) < 0.45] = 0
        held = np.zeros((nz + 1) * (ny + 1) * (nx + 1), np.uint8)
        held.reshape(nz + 1, ny + 1, nx + 1)[:, :, 0] = 1
        a, ra = P.prune_floating(dims, d, held)
        b, rb = cpu.prune_floating(dims, d, held)
        assert np.array_equal(a, b) and ra == rb


def test_grids_beyond_the_kernels_index_widths_are_rejected(P):
    # the device kernels index node planes with 32-bit offsets (3 NX NY components) and nodes and
    # voxels with signed 32-bit values: both build paths reject a grid beyond them before reading
    # any array (stand-in pointers), with the argument error and no device work
    import ctypes as C
    dummy = C.c_void_p(256)
    for dims in [(40000, 20000, 1), (50000, 1, 50000), (2000, 2000, 600), (1 << 16, 1 << 15, 1), (10000, 9000, 2)]:
        for device_build in (0, 1):
            d = P.ModelDesc(dims[0], dims[1], dims[2], dummy, 0.3, dummy, None, 1100, 0, device_build)
            h = C.c_void_p()
            code = P.lib().psx_create(C.byref(d), C.byref(h))
            assert code == P.ERR_ARGUMENT, (dims, code)
            assert "32-bit" in P.lib().psx_last_error().decode()
        code = P.lib().psx_hierarchy_create(C.byref(d), C.byref(h))
        assert code == P.ERR_ARGUMENT and "32-bit" in P.lib().psx_last_error().decode()
    # the largest configuration the bench solves stays accepted by the rule
    assert 3 * 513 * 129 < 2 ** 31 and 513 * 129 * 129 < 2 ** 31 - 2
