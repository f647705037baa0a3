"""Pins the oracle (oracle/oracle.cpp, the CPU restatement of the reference's VoxelFEA) before any
GPU result is compared with it: closed forms and invariants, brute-force assembled matrices (scipy),
the reference's own known-answer tests (native/tests/test_fea.cpp), the golden fixtures written
from the compiled reference (tests/golden), and — where oracle/_ref is built — bit-for-bit equality
with the reference's own code.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import ROOT, ref_available

GOLD = os.path.join(ROOT, "tests", "golden")
CX = [0, 1, 1, 0, 0, 1, 1, 0]
CY = [0, 0, 1, 1, 0, 0, 1, 1]
CZ = [0, 0, 0, 0, 1, 1, 1, 1]


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def assemble(dims, density, K0):
    """Global stiffness sum_e rho_e K0 (unmasked), DOF 3n+d."""
    nx, ny, nz = dims
    NX, NY = nx + 1, ny + 1
    rows, cols, vals = [], [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                rho = density[i + nx * (j + ny * k)]
                if rho <= 0:
                    continue
                dofs = []
                for a in range(8):
                    n = (i + CX[a]) + NX * ((j + CY[a]) + NY * (k + CZ[a]))
                    dofs += [3 * n, 3 * n + 1, 3 * n + 2]
                dofs = np.array(dofs)
                rows.append(np.repeat(dofs, 24))
                cols.append(np.tile(dofs, 24))
                vals.append((rho * K0).reshape(-1))
    N = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def prolongation(fdims):
    """P (fine DOFs x coarse DOFs): trilinear, coarse node I at fine 2I (voxel_fea.cpp:124-134)."""
    def axis(nf):
        nc = (nf + 1) // 2
        A = np.zeros((nf + 1, nc + 1))
        for i in range(nf + 1):
            if i % 2 == 0:
                A[i, i // 2] = 1
            else:
                A[i, (i - 1) // 2] = A[i, (i + 1) // 2] = 0.5
        return sp.csr_matrix(A)
    Px, Py, Pz = (axis(n) for n in fdims)
    P = sp.kron(sp.kron(Pz, Py), Px)
    return sp.kron(P, sp.identity(3)).tocsr()


def random_part(dims, seed, p_void=0.25, soft=False):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.2, 1.0, int(np.prod(dims))).astype(np.float32)
    d[rng.uniform(size=d.size) < p_void] = 1e-3 if soft else 0
    return d


# ---------------- element (hex8.cpp) ----------------

def test_element_symmetric_null_space(cpu):
    """test_fea.cpp:12-22."""
    K, _, _ = cpu.hex_element(0.3)
    assert np.array_equal(K, K.T)
    for d in range(3):
        assert np.abs(K[:, d::3].sum(axis=1)).max() < 1e-12
    # rigid rotations too (closed form: K annihilates every rigid-body mode of the cube)
    xyz = np.array([[CX[a], CY[a], CZ[a]] for a in range(8)], float)
    for axis in range(3):
        w = np.zeros(3)
        w[axis] = 1
        u = np.cross(w, xyz).reshape(-1)
        assert np.abs(K @ u).max() < 1e-12


def test_element_gauss_reproduces_K(cpu):
    """test_fea.cpp:24-38: K = sum_g Bbar_g^T D Bbar_g / 8."""
    K, _, gb = cpu.hex_element(0.3)
    nu = 0.3
    lam, mu = nu / ((1 + nu) * (1 - 2 * nu)), 1 / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[range(3), range(3)] = lam + 2 * mu
    D[range(3, 6), range(3, 6)] = mu
    Kg = sum(B.T @ D @ B for B in gb) / 8
    assert np.abs(Kg - K).max() / np.abs(K).max() < 1e-12


def test_element_golden(cpu):
    g = gold("hex8")
    K, cs, gb = cpu.hex_element(0.3)
    assert np.array_equal(K, g["K_03"]) and np.array_equal(cs, g["corner_03"]) and np.array_equal(gb, g["gauss_03"])
    assert np.array_equal(cpu.hex_element(0.33)[0], g["K_033"])


def test_element_rejects_bad_nu(cpu):
    with pytest.raises(ValueError):
        cpu.hex_element(0.5)


def test_principal_known_tensor(cpu):
    """test_fea.cpp:40-47."""
    out = cpu.principal([50, -20, 10, 30, 0, 0])
    r = np.hypot(35.0, 30.0)
    assert abs(out[0] - (15 + r)) < 1e-9 and abs(out[1] - 10) < 1e-9 and abs(out[2] - (15 - r)) < 1e-9
    t = np.random.default_rng(3).standard_normal(6)
    S = np.array([[t[0], t[3], t[5]], [t[3], t[1], t[4]], [t[5], t[4], t[2]]])
    assert np.allclose(cpu.principal(t), np.sort(np.linalg.eigvalsh(S))[::-1], atol=1e-12)


# ---------------- operator, transfers, coarse solve ----------------

@pytest.mark.parametrize("seed", [0, 1])
def test_apply_matches_assembled(cpu, seed):
    dims = (5, 4, 3)
    dens = random_part(dims, seed)
    bc = np.zeros(3 * 6 * 5 * 4, np.uint8)
    bc[:9] = 1
    m = cpu.Model(dims, dens, 0.3, bc, coarsest_max_dof=10 ** 9)
    K = assemble(dims, dens, cpu.hex_element(0.3)[0])
    fixed, _, _, _ = m.level_arrays(0)
    x = np.random.default_rng(seed).standard_normal(m.n_dof)
    x[fixed == 1] = 0
    y = K @ x
    assert np.abs(m.apply(x, raw=True) - y).max() < 1e-13 * np.abs(y).max()
    ym = np.where(fixed == 1, 0, y)
    assert np.abs(m.apply(x) - ym).max() < 1e-13 * np.abs(y).max()
    # symmetry and the rigid translation null space of the raw operator on interior rows
    z = np.random.default_rng(seed + 9).standard_normal(m.n_dof)
    assert abs(z @ m.apply(x, raw=True) - x @ m.apply(z, raw=True)) < 1e-11 * np.linalg.norm(x) * np.linalg.norm(z)


def test_coarse_levels_are_galerkin(cpu):
    """Each coarse operator is P^T K P of the finer one (voxel_fea.cpp:268-386), own matrices in FP32."""
    dims = (8, 6, 4)
    dens = random_part(dims, 5, soft=True)
    bc = np.zeros(3 * 9 * 7 * 5, np.uint8)
    m = cpu.Model(dims, dens, 0.3, bc, coarsest_max_dof=0, diag_add=None)
    assert m.num_levels >= 3
    K = assemble(dims, dens, cpu.hex_element(0.3)[0])
    fd = dims
    for l in range(1, m.num_levels):
        P = prolongation(fd)
        K = (P.T @ K @ P).tocsr()
        x = np.random.default_rng(l).standard_normal(K.shape[0])
        y = K @ x
        assert np.abs(m.apply(x, level=l, raw=True) - y).max() < 2e-6 * np.abs(y).max()
        fd = tuple((n + 1) // 2 for n in fd)


def test_restrict_prolong_are_P(cpu):
    dims = (7, 5, 4)
    dens = np.ones(int(np.prod(dims)), np.float32)
    bc = np.zeros(3 * 8 * 6 * 5, np.uint8)
    m = cpu.Model(dims, dens, 0.3, bc, coarsest_max_dof=0)
    P = prolongation(dims)
    rf = np.random.default_rng(1).standard_normal(P.shape[0])
    assert np.allclose(m.restrict(rf, 0, zero_fixed=False), P.T @ rf, atol=1e-14)
    zc = np.random.default_rng(2).standard_normal(P.shape[1])
    assert np.allclose(m.prolong_add(zc, np.zeros(P.shape[0]), 0), P @ zc, atol=1e-14)


def test_coarsest_cholesky(cpu, W):
    w_dims = (4, 2, 2)
    dens = random_part(w_dims, 2, p_void=0)
    bc = W.clamp_x0(w_dims)
    m = cpu.Model(w_dims, dens, 0.3, bc, coarsest_max_dof=10 ** 9)
    fixed, _, _, _ = m.level_arrays(0)
    K = assemble(w_dims, dens, cpu.hex_element(0.3)[0]).toarray()
    free = np.flatnonzero(fixed == 0)
    r = np.random.default_rng(0).standard_normal(m.n_dof)
    r[fixed == 1] = 0
    z = m.coarse_solve64(r)
    zr = np.zeros_like(r)
    zr[free] = np.linalg.solve(K[np.ix_(free, free)], r[free])
    assert np.abs(z - zr).max() < 1e-10 * np.abs(zr).max()


def test_lanczos_lmax_bounds(cpu):
    dims = (6, 4, 4)
    dens = random_part(dims, 3, p_void=0.0)
    bc = np.zeros(3 * 7 * 5 * 5, np.uint8)
    m = cpu.Model(dims, dens, 0.3, bc, coarsest_max_dof=100)
    lm = m.prepare_smoothers()[0]
    fixed, inv, _, _ = m.level_arrays(0)
    K = assemble(dims, dens, cpu.hex_element(0.3)[0]).toarray()
    free = fixed == 0
    S = np.sqrt(inv[free])
    true = np.linalg.eigvalsh(S[:, None] * K[np.ix_(free, free)] * S[None, :]).max()
    assert 0.8 * true < lm <= true * (1 + 1e-12)
    c = cpu.cheb_coefficients(lm)
    a, b = 0.1 * lm, 1.15 * lm
    theta, delta = (a + b) / 2, (b - a) / 2
    assert abs(c[0] - 1 / theta) < 1e-15 * c[0]
    sigma = theta / delta
    rho1 = 1 / (2 * sigma - 1 / sigma)
    assert abs(c[1] - rho1 / sigma) < 1e-14 and abs(c[2] - 2 * rho1 / delta) < 1e-14 * abs(c[2])


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_chebyshev_smoother_is_the_chebyshev_polynomial(cpu, degree):
    """From z = 0 with r = K e, the smoother leaves the error p(A) e with A = D^-1 K and
    p(l) = T_k((theta - l) / delta) / T_k(theta / delta) on [0.1, 1.15] lmax: the closed form of the
    degree-k Chebyshev iteration (degree 2 is the reference's smoother, voxel_fea.cpp:692-725)."""
    dims = (5, 4, 3)
    dens = random_part(dims, 11, p_void=0.2, soft=True)
    bc = np.zeros(3 * 6 * 5 * 4, np.uint8)
    bc[: 3 * 6 * 5] = 1  # bottom face held
    m = cpu.Model(dims, dens, 0.3, bc, coarsest_max_dof=50)
    m.set_cheb_degree(degree)
    lm = m.prepare_smoothers()[0]
    fixed, inv, _, _ = m.level_arrays(0)
    free = np.flatnonzero(fixed == 0)
    K = assemble(dims, dens, cpu.hex_element(0.3)[0]).toarray()[np.ix_(free, free)]
    e = np.zeros(m.n_dof)
    e[free] = np.random.default_rng(degree).standard_normal(free.size)
    r = np.zeros(m.n_dof)
    r[free] = K @ e[free]
    z = m.smooth(r.astype(np.float32)).astype(np.float64)
    a, b = 0.1 * lm, 1.15 * lm
    theta, delta = (a + b) / 2, (b - a) / 2
    Dh = np.sqrt(inv[free])
    lam, V = np.linalg.eigh(Dh[:, None] * K * Dh[None, :])
    T = np.polynomial.chebyshev.Chebyshev.basis(degree)
    p = T((theta - lam) / delta) / T(theta / delta)
    want = (V @ (p * (V.T @ (e[free] / Dh)))) * Dh
    assert np.linalg.norm((e - z)[free] - want) < 2e-5 * np.linalg.norm(e)
    assert np.abs(z[fixed == 1]).max() == 0


def test_tridiagonal_max(cpu):
    a = np.random.default_rng(4).standard_normal(8)
    b = np.random.default_rng(5).standard_normal(7)
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    assert abs(cpu.tridiagonal_max(a, b) - np.linalg.eigvalsh(T).max()) < 1e-10


# ---------------- solve ----------------

def test_solve_matches_direct(cpu, W):
    w = W.void_block((12, 6, 6), seed=2)
    m = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], coarsest_max_dof=100)
    s = m.solve(w["f"], tol=1e-12, max_iter=500)
    fixed, _, _, _ = m.level_arrays(0)
    free = np.flatnonzero(fixed == 0)
    K = assemble(w["dims"], w["density"], cpu.hex_element(w["nu"])[0])[free][:, free]
    u = np.zeros(m.n_dof)
    u[free] = spla.spsolve(K.tocsc(), w["f"][free])
    assert s["converged"]
    assert np.linalg.norm(s["u"] - u) < 1e-9 * np.linalg.norm(u)


@pytest.mark.parametrize("n", [2, 4, 8])
def test_cantilever_timoshenko(cpu, W, n):
    """test_fea.cpp:59-100: tip deflection within 6 % of Timoshenko, reaction 100 N +- 0.1 %."""
    w = W.cantilever(n)
    m = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"])
    s = m.solve(w["f"], tol=1e-7)
    E, h, nx = w["E"], w["h"], w["dims"][0]
    u = s["u"] / (E * h)
    NX, NY, NZ = nx + 1, n + 1, n + 1
    tip = u.reshape(NZ, NY, NX, 3)[:, :, nx, 1].mean()
    b, L, F, nu = 0.01, 0.1, 100.0, 0.3
    I, G = b ** 4 / 12, E / (2 * (1 + nu))
    expected = -F * L ** 3 / (3 * E * I) - F * L / ((5.0 / 6) * G * b * b)
    assert s["converged"]
    assert abs(tip / expected - 1) < 0.06
    assert abs(m.reactions(s["u"])[1] - 100) < 0.1
    if n == 8:  # test_fea.cpp:102-113: mid-span surface von Mises vs M c / I
        vm, _, _ = m.stresses(u, E, h, True)
        mid = nx // 2 + NX * (n + NY * (n // 2))
        theory = 100 * 0.05 * 0.005 / (0.01 * 1e-6 / 12)
        assert abs(vm[mid] / theory - 1) < 0.05
        assert 0.95 * 60e6 < vm.max() < 1.4 * 60e6


def test_reaction_with_load_on_fixed_node(cpu):
    """test_fea.cpp:124-132."""
    bc = np.ones(24, np.uint8)
    m = cpu.Model((1, 1, 1), np.ones(1, np.float32), 0.3, bc)
    f = np.zeros(24)
    f[1] = -100
    s = m.solve(f)
    R = m.reactions(s["u"], f)
    assert R[0] == 0 and R[1] == 100 and R[2] == 0


def test_patch_constant_stress(cpu):
    """Linear displacement -> constant stress D eps at every node (exact for the element)."""
    dims = (3, 3, 3)
    m = cpu.Model(dims, np.ones(27, np.float32), 0.25, np.zeros(3 * 64, np.uint8), coarsest_max_dof=10 ** 9)
    G = np.array([[1e-3, 2e-4, 0], [0, -5e-4, 3e-4], [1e-4, 0, 2e-4]])
    X = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"), -1)[..., ::-1].reshape(-1, 3)  # (x, y, z), x fastest
    u = (X @ G.T).reshape(-1)
    E, h = 1e9, 1.0
    T = m.stress_tensors(u, E, h)
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[2, 0] + G[0, 2]])
    nu = 0.25
    lam, mu = nu / ((1 + nu) * (1 - 2 * nu)), 1 / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[range(3), range(3)] = lam + 2 * mu
    D[range(3, 6), range(3, 6)] = mu
    sig = E * D @ eps
    assert np.abs(T - sig[None, :]).max() < 1e-6 * np.abs(sig).max()
    vm, _, _ = m.stresses(u, E, h, True)
    s = sig
    vmt = np.sqrt(0.5 * ((s[0] - s[1]) ** 2 + (s[1] - s[2]) ** 2 + (s[2] - s[0]) ** 2) + 3 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2))
    assert np.abs(vm - vmt).max() < 1e-5 * vmt


def test_prune_floating_known_answer(cpu):
    """test_fea.cpp:115-122."""
    held = np.zeros(5 * 2 * 2, np.uint8)
    held[0] = 1
    out, removed = cpu.prune_floating((4, 1, 1), np.array([1, 1, 0, 1], np.float32), held)
    assert list(out) == [1, 1, 0, 0] and removed == 1


def test_prune_floating_golden_and_bfs(cpu):
    g = gold("prune_16x10x8")
    out, removed = cpu.prune_floating(tuple(g["dims"]), g["density"], g["held"])
    assert np.array_equal(out, g["out"]) and removed == int(g["removed"])
    # brute force: scipy connected components of the face-adjacency graph
    nx, ny, nz = (int(v) for v in g["dims"])
    solid = (g["density"] > 0).reshape(nz, ny, nx)
    from scipy import ndimage
    lab, _ = ndimage.label(solid)
    held = g["held"].reshape(nz + 1, ny + 1, nx + 1)
    touch = np.zeros_like(solid)
    for a in range(8):
        touch |= held[CZ[a]:CZ[a] + nz, CY[a]:CY[a] + ny, CX[a]:CX[a] + nx].astype(bool)
    keep = np.isin(lab, np.unique(lab[solid & touch]))
    assert np.array_equal((out > 0).reshape(nz, ny, nx), keep & solid)


@pytest.mark.parametrize("name", ["cantilever4", "c1_cantilever", "void_block_48x16x16", "simp_block_24x12x12"])
def test_golden_fixtures(cpu, name):
    """The restatement reproduces the compiled reference's outputs bit for bit at 8 threads."""
    g = gold(name)
    assert int(g["threads"]) == 8
    m = cpu.Model(tuple(g["dims"]), g["density"], float(g["nu"]), g["bc"], int(g["coarsest"]))
    s = m.solve(g["f"], tol=float(g["tol"]), max_iter=500)
    assert m.num_levels == int(g["levels"])
    assert s["iterations"] == int(g["iterations"])
    assert np.array_equal(s["u"], g["u"])
    assert s["residual"] == float(g["residual"])
    assert np.array_equal(np.array([m.level(l)["lmax"] for l in range(m.num_levels)]), g["lmax"])
    E, h = float(g["E"]), float(g["h"])
    vm, p1, p3 = m.stresses(s["u"] / (E * h), E, h, True)
    assert np.array_equal(vm, g["vm"]) and np.array_equal(p1, g["p1"]) and np.array_equal(p3, g["p3"])
    assert np.array_equal(m.reactions(s["u"], g["f"]), g["reactions"])
    if "apply_y" in g:
        assert np.array_equal(m.apply(np.random.default_rng(11).standard_normal(m.n_dof)), g["apply_y"])
        m2 = cpu.Model(tuple(g["dims"]), g["density"], float(g["nu"]), g["bc"], int(g["coarsest"]))
        assert np.array_equal(m2.precondition(np.random.default_rng(12).standard_normal(m.n_dof)), g["precond_z"])


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_bitwise_against_compiled_reference(cpu, W):
    for w in [W.void_block((20, 10, 8), seed=7), W.simp_block((16, 8, 8), nrhs=1)]:
        f = w["f"] if w["f"].ndim == 1 else w["f"][0]
        out = {}
        for kind in ("orc", "ref"):
            m = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], 300, kind=kind)
            s = m.solve(f, tol=1e-9)
            ys = [m.apply(np.random.default_rng(l).standard_normal(m.level(l)["nDof"]), level=l) for l in range(m.num_levels)]
            y32 = m.apply32(np.random.default_rng(3).standard_normal(m.n_dof).astype(np.float32))
            out[kind] = (s, ys, y32, [m.level(l)["lmax"] for l in range(m.num_levels)])
        a, b = out["orc"], out["ref"]
        assert np.array_equal(a[0]["u"], b[0]["u"]) and a[0]["iterations"] == b[0]["iterations"]
        assert all(np.array_equal(p, q) for p, q in zip(a[1], b[1]))
        assert np.array_equal(a[2], b[2]) and a[3] == b[3]


# ---- voxelize (voxelize.cpp:14-128) ----

import meshes


def _octahedron_fractions(R, c, origin, h, dims, sub):
    """Closed form: a sample is solid iff |p - c|_1 < R; None when a sample sits within 1e-6 of the
    surface (the rays' jitter then decides)."""
    nx, ny, nz = dims
    ds = h / sub
    ax = [origin[d] + (np.arange(dims[d] * sub) + 0.5) * ds - c[d] for d in range(3)]
    X, Y, Z = np.meshgrid(np.abs(ax[0]), np.abs(ax[1]), np.abs(ax[2]), indexing="ij")
    L1 = X + Y + Z
    if np.any(np.abs(L1 - R) < 1e-6):
        return None
    solid = (L1 < R).astype(np.int32)
    f = solid.reshape(nx, sub, ny, sub, nz, sub).sum(axis=(1, 3, 5)) / sub ** 3
    return f.transpose(2, 1, 0).reshape(-1).astype(np.float32)


def test_voxelize_sphere_volume(cpu):
    # test_model.cpp:52-77: the voxel volume of the UV sphere at 40 voxels, sub 4, within 1 % of the
    # mesh volume
    pos, idx = meshes.uv_sphere()
    o, h, dims = cpu.grid_for_box(*meshes.bounds(pos), 40)
    f = cpu.voxelize(pos, idx, o, h, dims, 4)
    p = pos.reshape(-1, 3, 3).astype(np.float64)
    vol = abs(np.einsum("ij,ij->i", p[:, 0], np.cross(p[:, 1], p[:, 2])).sum()) / 6
    assert abs(f.sum() * h ** 3 / vol - 1) < 0.01


def test_voxelize_box_closed_form(cpu):
    # a box on voxel boundaries fills its voxels exactly; shifted by half a voxel, boundary voxels
    # are half, quarter or eighth full
    pos, idx = meshes.box((0, 0, 0), (10, 6, 4))
    o, h, dims = cpu.grid_for_box(*meshes.bounds(pos), 10)
    assert h == 1 and dims == (10, 6, 4) and list(o) == [0, 0, 0]
    for sub in (2, 3, 4):
        assert np.all(cpu.voxelize(pos, idx, o, h, dims, sub) == 1.0)
    o2 = np.array([-0.5, -0.5, -0.5])
    f = cpu.voxelize(pos, idx, o2, 1.0, (11, 7, 5), 2).reshape(5, 7, 11)
    expect = np.ones((5, 7, 11))
    for ax in range(3):
        sl = [slice(None)] * 3
        for e in (0, -1):
            sl[ax] = e
            expect[tuple(sl)] *= 0.5
    assert np.array_equal(f, expect.astype(np.float32))


@pytest.mark.parametrize("sub", [2, 3, 5])
def test_voxelize_octahedron_closed_form(cpu, sub):
    R, c = 7.3, (0.21, -0.37, 0.13)
    pos, idx = meshes.octahedron(R, c)
    o, h, dims = cpu.grid_for_box(*meshes.bounds(pos), 12, pad=1)
    expect = _octahedron_fractions(R, c, o, h, dims, sub)
    assert expect is not None
    assert np.array_equal(cpu.voxelize(pos, idx, o, h, dims, sub), expect)


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("mesh", ["sphere", "torus", "soup", "box"])
def test_voxelize_bitwise_vs_reference(cpu, mesh):
    pos, idx = {"sphere": meshes.uv_sphere, "torus": meshes.torus, "soup": meshes.soup, "box": lambda: meshes.box((-3, -2, -1), (5, 2.5, 1.7))}[mesh]()
    for n, sub, pad in ((24, 2, 0), (17, 3, 1), (30, 5, 2)):
        mn, mx = meshes.bounds(pos)
        g = cpu.grid_for_box(mn, mx, n, pad)
        gr = cpu.grid_for_box(mn, mx, n, pad, kind="ref")
        assert np.array_equal(g[0], gr[0]) and g[1] == gr[1] and g[2] == gr[2]
        assert np.array_equal(cpu.voxelize(pos, idx, *g, sub), cpu.voxelize(pos, idx, *g, sub, kind="ref"))


# ---- topology optimization (topology.cpp:21-183) ----

def _topo_case(W, dims=(24, 8, 6)):
    nx, ny, nz = dims
    bc = W.clamp_x0(dims)
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    f = np.zeros((NZ, NY, NX, 3))
    f[:, :, nx, 2] = -1.0 / (NY * NZ)  # a downward load over the free end
    fill = np.ones(nx * ny * nz, np.float32)
    keep = np.zeros(nx * ny * nz, np.uint8)
    keep.reshape(nz, ny, nx)[:, :, 0] = 1  # the clamped layer stays solid
    keep.reshape(nz, ny, nx)[:, :, -1] = 1  # and the loaded one
    return dims, fill, keep, bc, f.reshape(-1)


def test_element_energies_sum_to_compliance(cpu, W):
    dims, fill, keep, bc, f = _topo_case(W)
    rho = np.random.default_rng(2).uniform(0.2, 1.0, fill.size).astype(np.float32)
    m = cpu.Model(dims, rho, 0.3, bc, 300)
    u = m.solve(f, tol=1e-12)["u"]
    e = m.element_energies(u)
    assert abs((rho.astype(np.float64) * e).sum() - f @ u) <= 1e-9 * abs(f @ u)
    if ref_available():
        r = cpu.Model(dims, rho, 0.3, bc, 300, kind="ref")
        assert np.array_equal(e, r.element_energies(u))


def test_density_filter_constant_and_adjoint(cpu):
    dims = (9, 7, 5)
    fill = (np.random.default_rng(1).uniform(0, 1, 9 * 7 * 5) > 0.3).astype(np.float32)
    m = int((fill > 0).sum())
    for rmin in (1.0, 1.5, 2.4):
        assert np.allclose(cpu.topo_filter(dims, fill, rmin, np.full(m, 0.37)), 0.37, rtol=1e-14)
        x, y = np.random.default_rng(3).standard_normal((2, m))
        Fx = cpu.topo_filter(dims, fill, rmin, x)
        Fty = cpu.topo_filter(dims, fill, rmin, y, transpose=True)
        assert abs(Fx @ y - x @ Fty) <= 1e-12 * np.abs(x).sum() * np.abs(y).max()
    # rmin <= 1: the ball holds only the voxel itself
    x = np.random.default_rng(4).standard_normal(m)
    assert np.array_equal(cpu.topo_filter(dims, fill, 1.0, x), x)


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
def test_optimize_topology_bitwise_vs_reference(cpu, W):
    dims, fill, keep, bc, f = _topo_case(W)
    kw = dict(nu=0.3, vol_frac=0.4, penal=3.0, rmin=1.5, move=0.2, tol=0.01, max_iter=6, coarsest_max_dof=300)
    d, h = cpu.optimize_topology(dims, fill, keep, bc, f, **kw)
    dr, hr = cpu.optimize_topology(dims, fill, keep, bc, f, kind="ref", **kw)
    print(h, flush=True)
    assert np.array_equal(d, dr) and np.array_equal(h, hr)
    assert h[-1, 2] < 0.45 and h[-1, 1] < h[1, 1] * 5


def _break_case(W, dims=(20, 6, 6), seed=3):
    nx, ny, nz = dims
    bc = W.clamp_x0(dims)
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    f = np.zeros((NZ, NY, NX, 3))
    f[:, :, nx, 2] = -2000.0 / (NY * NZ)  # 2 kN down over the free end
    # stiffness varies per voxel, so no two corner stresses tie
    density = np.random.default_rng(seed).uniform(0.6, 1.0, nx * ny * nz).astype(np.float32)
    return dims, density, bc, f.reshape(-1)


def test_break_first_step_restated(cpu, W):
    # one step of breakTest (static_study.cpp:227-266) recomputed from the oracle's own pieces:
    # solve, corner stresses, per-voxel worst corner, the 0.95 threshold and the cap, the prune
    dims, density, bc, f = _break_case(W)
    E, h, strength = 200e9, 1e-3, 150e6
    for principal in (False, True):
        r = cpu.break_test(dims, density, 0.3, bc, f, E, h, strength, principal=principal, max_steps=1, tol=1e-10)
        assert r["removed"] == 0 and len(r["steps"]) == 1
        st = r["steps"][0]
        m = cpu.Model(dims, density, 0.3, bc, 1100)
        u = m.solve(f, tol=1e-10, max_iter=400)["u"] * (1 / (E * h))
        vm, p1, _ = m.stresses(u, E, h)
        nodal = p1 if principal else vm
        nx, ny, nz = dims
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        n0 = (i + (nx + 1) * (j + (ny + 1) * k)).reshape(-1)
        offs = [0, 1, 1 + (nx + 1), nx + 1]
        offs = offs + [o + (nx + 1) * (ny + 1) for o in offs]
        measure = np.maximum(0, np.max(np.stack([nodal[n0 + o] for o in offs]), axis=0))
        smax = float(measure.max())
        assert st["lambda"] == strength / smax
        cap = max(1, int(np.floor(0.02 * measure.size ** (2 / 3) + 0.5)))
        cand = np.flatnonzero(measure >= 0.95 * smax)
        order = cand[np.argsort(-measure[cand], kind="stable")]
        got = np.asarray(st["cracked"])
        # the worst voxels in non-increasing order of their measure; equal measures (neighbours share
        # their worst corner) come in std::sort's order, which the standard leaves unspecified
        assert len(got) == min(cap, cand.size) >= 1
        assert np.all(np.diff(measure[got]) <= 0)
        assert np.array_equal(np.sort(measure[got]), np.sort(measure[order[:len(got)]]))
        if cand.size <= cap or measure[order[cap - 1]] > measure[order[cap]]:
            assert set(got.tolist()) == set(order[:len(got)].tolist())
        want = got.tolist()
        assert st["max_disp"] == np.sqrt(np.max((u.reshape(-1, 3) ** 2).sum(axis=1))) * st["lambda"]
        nxt = density.copy()
        nxt[want] = 0
        pruned, _ = cpu.prune_floating(dims, nxt, W.held_nodes(bc))
        assert list(st["detached"]) == [int(e) for e in np.flatnonzero((nxt > 0) & ~(pruned > 0))]


def _measures(dims, density, nodal):
    """Each voxel's worst corner value (static_study.cpp:239-247), 0 for empty voxels."""
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (i + (nx + 1) * (j + (ny + 1) * k)).reshape(-1)
    offs = [0, 1, 1 + (nx + 1), nx + 1]
    offs = offs + [o + (nx + 1) * (ny + 1) for o in offs]
    m = np.maximum(0, np.max(np.stack([nodal[n0 + o] for o in offs]), axis=0))
    return np.where(density > 0, m, -1.0)


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("case", ["uniform", "random"])
def test_break_sequence_bitwise_vs_reference_breaktest(cpu, W, case):
    # the reference's own breakTest (static_study.cpp:197-279, CPU engine: coarsest 1100, tol 1e-6,
    # 400 iterations) against the restatement: every step's lambda, displacement, crack list in order,
    # detached voxels and von Mises field, bit for bit. Neighbouring voxels share their worst corner,
    # so equal measures straddle the crack size; std::sort's order among them must be the reference's.
    dims, density, bc, f = _break_case(W, (20, 6, 6), seed=3)
    if case == "uniform":
        density = np.ones_like(density)
    kw = dict(principal=False, max_steps=6, tol=1e-6, max_iter=400, coarsest_max_dof=1100, fields=True)
    E, h, strength = 200e9, 1e-3, 150e6
    a = cpu.break_test(dims, density, 0.3, bc, f, E, h, strength, **kw)
    b = cpu.break_test(dims, density, 0.3, bc, f, E, h, strength, kind="ref", **kw)
    print(case, a["reason"], len(a["steps"]), [list(s["cracked"]) for s in a["steps"]], flush=True)
    assert a["reason"] == b["reason"] and a["removed"] == b["removed"] and a["peak"] == b["peak"]
    assert len(a["steps"]) == len(b["steps"]) >= 3
    straddled = 0
    dens = density.copy()
    for sa, sb in zip(a["steps"], b["steps"]):
        assert sa["lambda"] == sb["lambda"] and sa["max_disp"] == sb["max_disp"] and sa["voxels"] == sb["voxels"]
        assert np.array_equal(sa["cracked"], sb["cracked"]) and np.array_equal(sa["detached"], sb["detached"])
        assert np.array_equal(sa["vm"], sb["vm"])
        # equal measures on both sides of the crack size: the tie order decided which were cracked
        m = _measures(dims, dens, sa["vm"])
        last = m[sa["cracked"][-1]]
        straddled += int(np.count_nonzero(m == last) > np.count_nonzero(m[sa["cracked"]] == last))
        dens[sa["cracked"]] = 0
        dens[sa["detached"]] = 0
    if case == "uniform":
        assert straddled >= 1


def test_break_sequence_invariants(cpu, W):
    dims, density, bc, f = _break_case(W, (16, 4, 4), seed=5)
    r = cpu.break_test(dims, density, 0.3, bc, f, 200e9, 1e-3, 150e6, max_steps=40, tol=1e-8)
    steps = r["steps"]
    print(r["reason"], len(steps), [len(s["cracked"]) for s in steps], flush=True)
    # every step removes its cracked and detached voxels from the next model
    for a, b in zip(steps, steps[1:]):
        assert b["voxels"] == a["voxels"] - len(a["cracked"]) - len(a["detached"])
    assert r["peak"] == max(s["lambda"] for s in steps)
    # a clamped cantilever ends separated from its load within the step budget
    assert r["reason"] == "separated"
