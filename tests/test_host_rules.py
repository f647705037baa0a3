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
