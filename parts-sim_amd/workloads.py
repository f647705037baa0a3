"""Synthetic voxel parts for the linear-static solve (numpy only, seeded, no I/O).

Every generator returns a dict with the solver's input contract (native/src/fea/voxel_fea.hpp:132-137,
native/src/fea/static_study.hpp:14-26): ``dims`` (nx, ny, nz) voxel counts, ``density`` float32
[nx*ny*nz] with voxel e = i + nx*(j + ny*k), ``bc`` uint8 [3*nodes] with DOF 3n+d and node
n = i + NX*(j + NY*k), ``f`` float64 [3*nodes] nodal forces in N (or [nrhs, 3*nodes]), ``nu``,
``E`` [Pa] and ``h`` [m]. The solver works in the normalised system (E = 1, h = 1), so
u_physical = u / (E * h).
"""
import math

import numpy as np


def nodes(dims):
    nx, ny, nz = dims
    return (nx + 1) * (ny + 1) * (nz + 1)


def clamp_x0(dims):
    """All three DOFs of every node on the x = 0 face held (native/tests/test_fea.cpp:65-68)."""
    nx, ny, nz = dims
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    bc = np.zeros((NZ, NY, NX, 3), np.uint8)
    bc[:, :, 0, :] = 1
    return bc.reshape(-1)


def cantilever(n, length=0.1, b=0.01, F=100.0, E=200e9, nu=0.3, component=1, consistent=True):
    """native/tests/test_fea.cpp:59-88: length x b x b beam of n voxels through the thickness,
    clamped at x = 0, tip load F along -y spread over the x = L face with edge weights 1/2."""
    h = b / n
    nx = int(round(length / h))
    dims = (nx, n, n)
    NX, NY, NZ = nx + 1, n + 1, n + 1
    f = np.zeros((NZ, NY, NX, 3))
    wj = np.ones(NY)
    wk = np.ones(NZ)
    if consistent:
        wj[[0, -1]] = 0.5
        wk[[0, -1]] = 0.5
        f[:, :, nx, component] = -F * np.outer(wk, wj) / (n * n)
    else:  # native/tests/test_gpu.cpp:171: equal shares per face node
        f[:, :, nx, component] = -F / (NY * NZ)
    return {"dims": dims, "density": np.ones(nx * n * n, np.float32), "bc": clamp_x0(dims), "f": f.reshape(-1),
            "nu": nu, "E": E, "h": h, "name": "cantilever-%dx%dx%d" % dims}


def c1_cantilever():
    """64x16x16 unit voxels (h = 1 mm), E = 200 GPa, nu = 0.3, x = 0 clamped, uniform -z traction of
    1 MPa on the x = 64 face (256 N, consistent nodal weights: edges 1/2)."""
    nx, ny, nz = 64, 16, 16
    dims = (nx, ny, nz)
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    total = 1e6 * (ny * 1e-3) * (nz * 1e-3)
    wj = np.ones(NY)
    wk = np.ones(NZ)
    wj[[0, -1]] = 0.5
    wk[[0, -1]] = 0.5
    f = np.zeros((NZ, NY, NX, 3))
    f[:, :, nx, 2] = -total * np.outer(wk, wj) / (ny * nz)
    return {"dims": dims, "density": np.ones(nx * ny * nz, np.float32), "bc": clamp_x0(dims), "f": f.reshape(-1),
            "nu": 0.3, "E": 200e9, "h": 1e-3, "name": "C1 cantilever 64x16x16"}


def spherical_voids(dims, seed, void_fraction=0.2, rmin=2.0, rmax=8.0):
    """Boolean [nz, ny, nx] void mask: spheres with centres uniform in the box and radii uniform in
    [rmin, rmax] voxels (voxel centres inside a sphere are void), added until the void fraction is
    reached. numpy PCG64 with the given seed."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    void = np.zeros((nz, ny, nx), bool)
    target = void_fraction * void.size
    count = 0
    while count < target:
        cx, cy, cz = rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz)
        r = rng.uniform(rmin, rmax)
        i0, i1 = max(0, int(cx - r)), min(nx, int(cx + r) + 1)
        j0, j1 = max(0, int(cy - r)), min(ny, int(cy + r) + 1)
        k0, k1 = max(0, int(cz - r)), min(nz, int(cz + r) + 1)
        if i0 >= i1 or j0 >= j1 or k0 >= k1:
            continue
        zi = (np.arange(k0, k1) + 0.5 - cz)[:, None, None]
        yi = (np.arange(j0, j1) + 0.5 - cy)[None, :, None]
        xi = (np.arange(i0, i1) + 0.5 - cx)[None, None, :]
        inside = xi * xi + yi * yi + zi * zi <= r * r
        sub = void[k0:k1, j0:j1, i0:i1]
        count += int(np.count_nonzero(inside & ~sub))
        sub |= inside
    return void


def void_block(dims=(256, 64, 64), seed=1, E=70e9, nu=0.33, pressure=5e6, void_density=1e-3, load_from=200 / 256,
               void_fraction=0.2, h=1e-3):
    """Solid block with seeded spherical voids of stiffness ``void_density`` (still active), x = 0 face
    clamped, -z pressure on the top face (k = nz) over voxel columns i >= load_from * nx, applied as
    -p h^2 / 4 per corner of every loaded face quad."""
    nx, ny, nz = dims
    void = spherical_voids(dims, seed, void_fraction)
    dens = np.where(void, np.float32(void_density), np.float32(1.0)).astype(np.float32)
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    f = np.zeros((NZ, NY, NX, 3))
    i0 = int(round(load_from * nx))
    quad = -pressure * h * h / 4
    top = np.zeros((NY, NX))
    for dj in (0, 1):
        for di in (0, 1):
            top[dj:dj + ny, i0 + di:nx + di] += quad
    f[nz, :, :, 2] = top
    return {"dims": dims, "density": dens.reshape(-1), "bc": clamp_x0(dims), "f": f.reshape(-1), "nu": nu, "E": E,
            "h": h, "seed": seed, "void_fraction": float(void.mean()), "name": "void block %dx%dx%d seed %d" % (nx, ny, nz, seed)}


def simp_block(dims=(128, 64, 64), seed=4, nrhs=16, E0=200e9, nu=0.3, force=1e3, penalty=3.0, h=1e-3):
    """SIMP-style densities: x ~ U[0.001, 1] per voxel, stiffness scale x^penalty; x = 0 clamped;
    ``nrhs`` load cases, each a ``force`` N point load at a uniformly random node of the x = nx face
    in a direction uniform on the unit sphere."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.001, 1.0, nx * ny * nz)
    dens = (x ** penalty).astype(np.float32)
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    F = np.zeros((nrhs, NZ, NY, NX, 3))
    for c in range(nrhs):
        j, k = rng.integers(0, NY), rng.integers(0, NZ)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        F[c, k, j, nx, :] = force * d
    return {"dims": dims, "density": dens, "bc": clamp_x0(dims), "f": F.reshape(nrhs, -1), "nu": nu, "E": E0, "h": h,
            "seed": seed, "name": "SIMP block %dx%dx%d x%d RHS" % (nx, ny, nz, nrhs)}


def serving_stream(n_requests, seed=5, rate=200.0, edge_min=32, edge_max=256, warm_fraction=0.3, remove_fraction=0.01):
    """Poisson stream of solve requests on seeded void-block parts: longest edge log-uniform in
    [edge_min, edge_max] voxels, the other two edges = longest / U[1, 4]; E ~ U[2, 210] GPa,
    nu ~ U[0.28, 0.36]; 1-8 load cases; ``warm_fraction`` of requests re-solve an earlier part with
    ``remove_fraction`` of its solid voxels removed (warm start from the earlier solution).
    Returns a list of request dicts with an arrival time ``t`` [s]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(n_requests):
        t += rng.exponential(1.0 / rate)
        if out and rng.uniform() < warm_fraction:
            base = out[int(rng.integers(0, len(out)))]
            root = base.get("parent", base["id"])
            req = {"id": q, "t": t, "parent": root, "warm": True, "remove_seed": int(rng.integers(1 << 31)),
                   "remove_fraction": remove_fraction, "nrhs": base["nrhs"], "dims": base["dims"], "E": base["E"],
                   "nu": base["nu"], "part_seed": base["part_seed"]}
        else:
            L = math.exp(rng.uniform(math.log(edge_min), math.log(edge_max)))
            a = max(4, int(round(L)))
            b = max(4, int(round(L / rng.uniform(1, 4))))
            c = max(4, int(round(L / rng.uniform(1, 4))))
            req = {"id": q, "t": t, "warm": False, "dims": (a, b, c), "E": rng.uniform(2e9, 210e9), "nu": rng.uniform(0.28, 0.36),
                   "nrhs": int(rng.integers(1, 9)), "part_seed": int(rng.integers(1 << 31))}
        out.append(req)
    return out


def serving_loads(w, nrhs, seed):
    """The load cases of a served request on part ``w``: the part's pressure load scaled by 1, 1.5, 2, ...
    plus seeded noise of 5 % of its largest nodal force on the loaded DOFs (nrhs x nDof)."""
    rng = np.random.default_rng(seed)
    f = w["f"]
    return np.stack([f * (1 + 0.5 * k) + 0.05 * np.abs(f).max() * rng.standard_normal(f.size) * (f != 0) for k in range(nrhs)])


def remove_voxels(density, fraction, seed):
    """Sets ``fraction`` of the solid voxels (density > 0) to 0, chosen with the given seed."""
    rng = np.random.default_rng(seed)
    d = density.copy()
    solid = np.flatnonzero(d > 0)
    k = int(round(fraction * len(solid)))
    if k:
        d[rng.choice(solid, k, replace=False)] = 0
    return d


def held_nodes(bc):
    """native/src/fea/static_study.cpp:17-21: a node is held when any of its DOFs is."""
    b = np.asarray(bc, np.uint8).reshape(-1, 3)
    return (b[:, 0] | b[:, 1] | b[:, 2]).astype(np.uint8)
