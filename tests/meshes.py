"""Triangle meshes for the voxelizer tests: float32 positions (3 per vertex) and uint32 indices."""
import numpy as np


def uv_sphere(r=10.0, nu=64, nv=32):
    """The reference's test sphere (test_model.cpp:52-66): quads of a UV sphere, two triangles each."""
    def P(i, j):
        th, ph = np.pi * j / nv, 2 * np.pi * i / nu
        return [r * np.sin(th) * np.cos(ph), r * np.cos(th), r * np.sin(th) * np.sin(ph)]
    pos = []
    for j in range(nv):
        for i in range(nu):
            a, b, c, d = P(i, j), P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)
            for q in (a, c, b, a, d, c):
                pos += q
    pos = np.array(pos, np.float32)
    return pos, np.arange(pos.size // 3, dtype=np.uint32)


def box(lo, hi):
    """Axis-aligned box as 12 outward triangles (test_model.cpp:32-40)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0), (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    idx = [0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4, 2, 3, 7, 2, 7, 6, 1, 2, 6, 1, 6, 5, 0, 4, 7, 0, 7, 3]
    return np.array(v, np.float32).reshape(-1), np.array(idx, np.uint32)


def octahedron(R, c=(0.0, 0.0, 0.0)):
    """|x - c|_1 <= R as 8 triangles."""
    cx, cy, cz = c
    v = [(cx + R, cy, cz), (cx - R, cy, cz), (cx, cy + R, cz), (cx, cy - R, cz), (cx, cy, cz + R), (cx, cy, cz - R)]
    idx = [0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0, 4, 2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3, 5]
    return np.array(v, np.float32).reshape(-1), np.array(idx, np.uint32)


def torus(R=8.0, r=3.0, nu=48, nv=24, rot=0.3):
    pos = []
    ca, sa = np.cos(rot), np.sin(rot)

    def P(i, j):
        u, w = 2 * np.pi * i / nu, 2 * np.pi * j / nv
        x, y, z = (R + r * np.cos(w)) * np.cos(u), (R + r * np.cos(w)) * np.sin(u), r * np.sin(w)
        return [x, ca * y - sa * z, sa * y + ca * z]
    for j in range(nv):
        for i in range(nu):
            a, b, c, d = P(i, j), P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)
            for q in (a, b, c, a, c, d):
                pos += q
    pos = np.array(pos, np.float32)
    return pos, np.arange(pos.size // 3, dtype=np.uint32)


def soup(n=300, seed=1):
    """Random triangles (open, overlapping, some degenerate): every code path of the rasterizer."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    pos = (c + rng.normal(0, 2, (n, 3, 3))).astype(np.float32).reshape(-1)
    pos[:9] = pos[9:18]  # a duplicated triangle
    pos[18:27] = np.repeat(pos[18:21].reshape(1, 3), 3, axis=0).reshape(-1)  # a point triangle
    return pos, np.arange(pos.size // 3, dtype=np.uint32)


def bounds(pos):
    p = np.asarray(pos, np.float32).reshape(-1, 3)
    return p.min(0).astype(np.float64), p.max(0).astype(np.float64)


def hollow_box(lo, hi, t):
    """A box with walls of thickness t: the outer box outward, the inner cavity inward (its triangles
    reversed), so every wall has an opposite face facing the other way."""
    v0, i0 = box(lo, hi)
    v1, i1 = box(tuple(a + t for a in lo), tuple(b - t for b in hi))
    i1 = i1.reshape(-1, 3)[:, ::-1].reshape(-1) + v0.size // 3
    return np.concatenate([v0, v1]), np.concatenate([i0, i1]).astype(np.uint32)
