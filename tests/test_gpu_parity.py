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


@pytest.mark.parametrize("precision", ["fp64", "ref", "mixed"])
def test_solve_matches_oracle(models, precision):
    for w, g, o in models:
        f = w["f"] if w["f"].ndim == 1 else w["f"][0]
        s = g.solve(f, tol=1e-10, max_iter=1000, precision=precision, check_every=4)
        so = o.solve(f, tol=1e-10, max_iter=500)
        err = _rel(s["u"], so["u"])
        print(w["name"], precision, "its", s["iterations"], so["iterations"], "folds", s["folds"], "res", s["residual"], "err", err, flush=True)
        assert s["converged"] and so["converged"]
        assert s["residual"] <= 1e-9
        assert err < 1e-7
        if precision != "mixed":
            assert abs(s["iterations"] - so["iterations"]) <= 2


def test_solve_warm_start_and_cancel(models):
    w, g, o = models[2]
    s = g.solve(w["f"], tol=1e-8, precision="fp64")
    # warm start from the solution: converged at once (static_study.cpp:219, 235 reuse u as x0)
    s2 = g.solve(w["f"], tol=1e-6, x0=s["u"], precision="fp64")
    assert s2["iterations"] == 0 and s2["converged"]
    calls = []
    s3 = g.solve(w["f"], tol=1e-12, precision="mixed", check_every=2, on_progress=lambda it, rel: calls.append(it) or len(calls) >= 2)
    assert s3["cancelled"] and not s3["converged"] and len(calls) == 2


def test_reactions_equal_load(models):
    """Equilibrium: the support reactions balance the applied load (test_fea.cpp:98)."""
    for w, g, o in models:
        f = w["f"] if w["f"].ndim == 1 else w["f"][0]
        s = g.solve(f, tol=1e-10, precision="fp64")
        R = g.reactions(s["u"], f)
        Ro = o.reactions(s["u"], f)
        total = f.reshape(-1, 3).sum(axis=0)
        assert np.allclose(R, Ro, rtol=1e-9, atol=1e-9 * np.abs(total).max())
        assert np.allclose(R, -total, rtol=1e-6, atol=1e-6 * np.abs(total).max())


def test_fine_kernel_matches_generic_kernel(P, W):
    """The z-marching fine-level kernel against the one-thread-per-node kernel, every epilogue mode
    exercised through apply (FP64 and FP32) on parts with empty voxels, soft voids and box edges."""
    import os
    cases = [W.void_block((37, 19, 11), seed=8), W.void_block((20, 9, 17), seed=2, void_density=0.0, void_fraction=0.25), W.cantilever(3)]
    for w in cases:
        a = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300)
        os.environ["PSX_FINE"] = "0"
        try:
            b = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300)
        finally:
            del os.environ["PSX_FINE"]
        for l in range(a.num_levels):
            x = np.random.default_rng(l).standard_normal(a.level(l)["nDof"])
            for raw in (False, True):
                assert _rel(a.apply(x, l, raw), b.apply(x, l, raw)) < 1e-14
        x = np.random.default_rng(1).standard_normal(a.n_dof)
        x32 = x.astype(np.float32)
        assert _rel(a.apply32(x32).astype(float), b.apply32(x32).astype(float)) < 1e-6
        r = np.random.default_rng(2).standard_normal(a.n_dof)
        assert _rel(a.precondition(r), b.precondition(r)) < 1e-5


def test_coarsest_buffer_reuse_does_not_leak(models):
    """A raw product on the coarsest level leaves nonzeros at its held DOFs in a work buffer; the
    V-cycle's coarsest solve must overwrite them (regression)."""
    w, g, o = models[1]
    last = g.num_levels - 1
    x = np.random.default_rng(3).standard_normal(g.level(last)["nDof"])
    g.apply(x, last, raw=True)
    s = g.solve(w["f"], tol=1e-10, precision="fp64")
    assert s["converged"]


@pytest.mark.parametrize("pieces", [2, 3])
def test_element_matrix_core_operator(P, W, pieces):
    """The element-by-element matrix-core operator against the FP64 operator. Each voxel's product
    Kb (x_e - c) is a sum of BF16 piece products with FP32 accumulation (c: one node's value per tile):
    two pieces (the mixed V-cycle's smoothing) carry x - c to about 2^-17, three to FP32 rounding, so
    the product is within 2e-5 (two) or 3e-7 (three) of K x for rough vectors and within 2e-4 (two)
    or 3e-6 (three) for smooth ones, relative in the 2-norm (measured: 6e-6 / 7e-8 and 6e-5 / 8e-7;
    the FP32 VALU operator: 1e-7 and 1.3e-6). Parts: voids at 1e-3 stiffness, an odd grid with empty
    voxels (tiles with one node column and one node row), per-voxel SIMP densities."""
    parts = [W.void_block((40, 16, 12), seed=5), W.void_block((21, 13, 11), seed=2, void_density=0.0, void_fraction=0.3),
             W.simp_block((32, 16, 16), seed=1, nrhs=1)]
    for w in parts:
        g = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300)
        n = g.n_dof
        nx, ny, nz = w["dims"]
        rough = np.random.default_rng(4).standard_normal(n)
        Z, Y, X = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
        smooth = np.stack([np.sin(X / 9.0) + 0.1 * Y, np.cos(Y / 5.0) * Z / 7.0, 0.02 * X * Z], -1).reshape(-1)
        for x, tol in ((rough, 2e-5 if pieces == 2 else 3e-7), (smooth, 2e-4 if pieces == 2 else 3e-6)):
            y64 = g.apply(x)
            ym = g.apply_mfma(x.astype(np.float32), 0, pieces).astype(np.float64)
            err = np.linalg.norm(ym - y64) / np.linalg.norm(y64)
            print(w["dims"], pieces, "pieces: rel err", err, flush=True)
            assert err < tol
        # a rigid translation is the reference value itself in every tile: exactly zero everywhere,
        # void surfaces and the part's faces included
        t = np.tile([1.0, -2.0, 0.5], n // 3)
        assert np.all(g.apply_mfma(t.astype(np.float32), 0, pieces) == 0)
        # the child-scale level: uniform elements as above, own elements sum_c s_c M_c (x_e - x_e0) as
        # two-piece products over the 8 child matrices in a pass of their own (both forms), against the
        # FP64 child form; a translation is again exactly zero
        if g.num_levels > 1 and g.level(1)["nOwn"] > 0:
            info = g.level(1)
            n1 = info["nDof"]
            mx, my, mz = (d + 1 for d in info["dims"])
            Z1, Y1, X1 = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
            smooth1 = np.stack([np.sin(X1 / 5.0) + 0.1 * Y1, np.cos(Y1 / 3.0) * Z1 / 4.0, 0.03 * X1 * Z1], -1).reshape(-1)
            rough1 = np.random.default_rng(6).standard_normal(n1)
            for x, tol in ((rough1, 2e-5), (smooth1, 2e-4)):
                y64 = g.apply_child(x, 1)
                ym = g.apply_mfma(x.astype(np.float32), 1, pieces).astype(np.float64)
                err = np.linalg.norm(ym - y64) / np.linalg.norm(y64)
                print(w["dims"], "level 1 nOwn", info["nOwn"], pieces, "pieces: rel err", err, flush=True)
                assert err < tol
            t1 = np.tile([1.0, -2.0, 0.5], n1 // 3)
            assert np.all(g.apply_mfma(t1.astype(np.float32), 1, pieces) == 0)
        g.close()


def test_element_matrix_core_operator_rejects(P, W):
    # pieces other than 2 or 3, and the levels with stored own matrices (below the child-scale one)
    w = W.void_block((32, 16, 16), seed=1)
    g = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 60)
    with pytest.raises(Exception):
        g.apply_mfma(np.zeros(g.n_dof, np.float32), 0, 4)
    assert g.num_levels > 2 and g.level(2)["nOwn"] > 0
    with pytest.raises(Exception):
        g.apply_mfma(np.zeros(g.level(2)["nDof"], np.float32), 2, 2)
    g.close()


def test_mixed_iterations_close_to_fp64(models):
    for w, g, o in models:
        f = w["f"] if w["f"].ndim == 1 else w["f"][0]
        a = g.solve(f, tol=1e-8, precision="fp64")
        b = g.solve(f, tol=1e-8, precision="mixed")
        print(w["name"], "fp64 its", a["iterations"], "mixed its", b["iterations"], "folds", b["folds"], flush=True)
        assert b["converged"] and b["iterations"] <= 1.5 * a["iterations"] + 5


def test_child_form_is_the_galerkin_product(P, cpu):
    """Own coarse elements in child-scale form (sum_c s_c P_c^T K0 P_c, as the FP64 and mixed solves
    apply them) equal P^T K P of the assembled finest operator to FP64 rounding; the stored FP32
    entries (the reference's form) differ by their rounding only."""
    from test_oracle import assemble, prolongation, random_part
    for dims, soft in (((16, 10, 8), True), ((14, 9, 7), False)):
        dens = random_part(dims, 7, p_void=0.3, soft=soft)
        bc = np.zeros(3 * int(np.prod([n + 1 for n in dims])), np.uint8)
        g = P.Model(dims, dens, 0.3, bc, 60)
        assert g.level(1)["nOwn"] > 0
        K1 = (lambda Pm: (Pm.T @ assemble(dims, dens, cpu.hex_element(0.3)[0]) @ Pm).tocsr())(prolongation(dims))
        x = np.random.default_rng(3).standard_normal(K1.shape[0])
        y = K1 @ x
        yc = g.apply_child(x, 1, raw=True)
        print(dims, "child", np.abs(yc - y).max() / np.abs(y).max(), flush=True)
        assert np.abs(yc - y).max() < 1e-12 * np.abs(y).max()
        assert np.abs(g.apply(x, 1, raw=True) - y).max() < 2e-6 * np.abs(y).max()
        y32 = g.apply_child(x.astype(np.float32), 1, raw=True).astype(np.float64)
        assert np.abs(y32 - y).max() < 2e-5 * np.abs(y).max()


def test_child_form_held_dofs_and_epilogues(P, W):
    """The child-scale operator with held DOFs against the stored operator (FP32 rounding apart), and
    a V-cycle built on it against the reference-form V-cycle."""
    for w in (W.void_block((40, 18, 14), seed=5), W.void_block((21, 13, 11), seed=2, void_density=0.0, void_fraction=0.3)):
        g = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300)
        for l in range(1, g.num_levels):
            x = np.random.default_rng(l).standard_normal(g.level(l)["nDof"])
            a, b = g.apply_child(x, l), g.apply(x, l)
            assert _rel(a, b) < 3e-6
        s = g.solve(w["f"], tol=1e-10, precision="fp64")
        r = g.solve(w["f"], tol=1e-10, precision="ref")
        assert s["converged"] and r["converged"] and _rel(s["u"], r["u"]) < 1e-8


@pytest.mark.parametrize("precision", ["ref", "fp64", "mixed"])
def test_degree3_smoother_matches_oracle(P, W, cpu, precision):
    """Degree-3 Chebyshev smoothing (the mixed-precision configuration) against the oracle's."""
    w = W.void_block((48, 16, 16), seed=3)
    o = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
    o.set_cheb_degree(3)
    so = o.solve(w["f"], tol=1e-10, max_iter=500)
    g = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
    s = g.solve(w["f"], tol=1e-10, max_iter=1000, precision=precision, cheb_degree=3)
    print(precision, "its", s["iterations"], so["iterations"], _rel(s["u"], so["u"]), flush=True)
    assert s["converged"] and so["converged"] and _rel(s["u"], so["u"]) < 1e-7
    if precision == "ref":
        assert abs(s["iterations"] - so["iterations"]) <= 2


@pytest.mark.parametrize("form", [1, 4, 5])
def test_mixed_with_matrix_core_operators(P, W, form):
    """Mixed solves with the matrix-core element product in the V-cycle (form 1), as the outer
    operator (form 4) and both (5): the FP64 solution to 1e-7 and iteration counts within 2 of the
    FP32 VALU form's (void block; odd grid with empty voxels; per-voxel densities)."""
    for w in (W.void_block((48, 16, 16), seed=3), W.void_block((21, 13, 11), seed=2, void_density=0.0, void_fraction=0.3),
              W.simp_block((32, 16, 16), seed=1, nrhs=1)):
        f = w["f"] if w["f"].ndim == 1 else w["f"][0]
        g = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
        a = g.solve(f, tol=1e-10, precision="mixed", mfma_smoother=form)
        v = g.solve(f, tol=1e-10, precision="mixed")
        b = g.solve(f, tol=1e-10, precision="fp64")
        print(w["dims"], "form", form, "its", a["iterations"], "valu", v["iterations"], "fp64", b["iterations"], flush=True)
        assert a["converged"] and _rel(a["u"], b["u"]) < 1e-7
        assert abs(a["iterations"] - v["iterations"]) <= 2
        g.close()


def test_matrix_core_forms_with_diagonal_shift(P, W):
    """A shifted operator K + diag(a) (a shifted eigen model's, restricted to the coarse levels as
    the FP64 path restricts it): the element products add a x and fold a into the inverse diagonal
    on level 0 (from the voxel scales) and level 1 (the stored inverse), to the tolerances of the
    unshifted forms; the mixed solve with them meets the FP64 solution to 1e-7."""
    w = W.void_block((48, 16, 16), seed=3)
    n0 = 3 * W.nodes(w["dims"])
    shift = 0.05 * np.random.default_rng(8).random(n0)
    g = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200, diag_add=shift)
    x = np.random.default_rng(4).standard_normal(n0)
    y64 = g.apply(x)
    ym = g.apply_mfma(x.astype(np.float32), 0, 2).astype(np.float64)
    err0 = _rel(ym, y64)
    info = g.level(1)
    x1 = np.random.default_rng(6).standard_normal(info["nDof"])
    err1 = _rel(g.apply_mfma(x1.astype(np.float32), 1, 2).astype(np.float64), g.apply_child(x1, 1))
    f = w["f"] if w["f"].ndim == 1 else w["f"][0]
    a = g.solve(f, tol=1e-10, precision="mixed", mfma_smoother=1, cheb_degree=3)
    v = g.solve(f, tol=1e-10, precision="mixed", cheb_degree=3)
    b = g.solve(f, tol=1e-10, precision="fp64", cheb_degree=3)
    print("shifted: level 0 %.2e level 1 (nOwn %d) %.2e its %d valu %d fp64 %d" % (err0, info["nOwn"], err1, a["iterations"], v["iterations"],
                                                                                 b["iterations"]), flush=True)
    assert info["nOwn"] > 0
    assert err0 < 2e-5 and err1 < 2e-5
    assert a["converged"] and _rel(a["u"], b["u"]) < 1e-7
    assert abs(a["iterations"] - v["iterations"]) <= 2
    g.close()


def test_stresses_match_oracle(models):
    """Nodal von Mises, principal stresses and tensors on the device against the oracle's
    (voxel_fea.cpp:968-1084): the 1e-5 relative von Mises bar, met with margin."""
    for w, g, o in models:
        f = w["f"] if w["f"].ndim == 1 else w["f"][0]
        u = o.solve(f, tol=1e-10, max_iter=500)["u"]
        E, h = w.get("E", 200e9), w.get("h", 1e-3)
        vm, p1, p3 = g.stresses(u, E, h)
        vo, q1, q3 = o.stresses(u, E, h)
        scale = float(np.abs(vo).max())
        print(w["name"], "vm", np.abs(vm - vo).max() / scale, "p1", np.abs(p1 - q1).max() / scale, flush=True)
        assert scale > 0
        assert np.abs(vm.astype(float) - vo).max() <= 1e-6 * scale
        assert np.abs(p1.astype(float) - q1).max() <= 1e-6 * scale and np.abs(p3.astype(float) - q3).max() <= 1e-6 * scale
        vm2, a, b = g.stresses(u, E, h, principals=False)
        assert a is None and np.array_equal(vm2, vm)
        t, to = g.stress_tensors(u, E, h), o.stress_tensors(u, E, h)
        assert np.abs(t.astype(float) - to).max() <= 1e-6 * float(np.abs(to).max())


@pytest.mark.parametrize("forced", ["1", "0"])
def test_dense_form_forced(P, W, cpu, forced, monkeypatch):
    """The per-voxel operator form (every node sums its 8 voxels), forced on for parts where most
    nodes are regular and off for the SIMP block, against the oracle: operator, V-cycle and solve."""
    monkeypatch.setenv("PSX_DENSE", forced)
    for w in [W.void_block((48, 16, 16), seed=3), W.cantilever(4), W.simp_block((32, 16, 16), nrhs=1)]:
        g = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
        o = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], 1200)
        n = g.level(0)["nDof"]
        x = np.random.default_rng(5).standard_normal(n)
        for raw in (False, True):
            assert _rel(g.apply(x, 0, raw), o.apply(x, 0, raw)) < 1e-13
        x32 = x.astype(np.float32)
        assert _rel(g.apply32(x32, 0).astype(np.float64), o.apply32(x32, 0).astype(np.float64)) < 1e-5
        r = np.random.default_rng(6).standard_normal(n)
        assert _rel(g.precondition(r), o.precondition(r)) < 1e-4
        f = w["f"] if w["f"].ndim == 1 else w["f"][0]
        u = g.solve(f, tol=1e-10, precision="mixed")["u"]
        uo = o.solve(f, tol=1e-12)["u"]
        print(w["name"], forced, _rel(u, uo), flush=True)
        assert _rel(u, uo) < 1e-7
        g.close()


@pytest.mark.parametrize("dims", [(21, 13, 11), (23, 9, 7), (37, 23, 11)])
def test_precondition_odd_grids_match_oracle(P, W, cpu, dims):
    # grids with odd cell counts on the levels they coarsen: the fine node 2I of a coarse node on the
    # last coarse plane lies outside the fine grid, so restriction (the 3 x 3 x 3 gather around 2I)
    # and prolongation (trilinear from 1, 2, 4 or 8 coarse nodes) take their one-sided branches
    # (voxel_fea.cpp:739-815). One V-cycle against the oracle's at every depth from 2 levels (the
    # transfer between the fine level and the coarsest) to the full hierarchy; a wrong transfer
    # weight on a boundary plane moves z by far more than the FP32 V-cycle's 1e-4.
    nx, ny, nz = dims
    rng = np.random.default_rng(sum(dims))
    density = rng.uniform(0.2, 1.0, nx * ny * nz).astype(np.float32)
    density[rng.uniform(0, 1, density.size) < 0.15] = 0.0  # empty voxels: mixed coarse elements
    density.reshape(nz, ny, nx)[:, :, 0] = 1.0  # the clamped face stays attached
    bc = W.clamp_x0(dims)
    for coarsest in (3000, 300, 60):
        g = P.Model(dims, density, 0.3, bc, coarsest)
        o = cpu.Model(dims, density, 0.3, bc, coarsest)
        assert g.num_levels == o.num_levels
        for k in range(2):
            r = np.random.default_rng(k).standard_normal(g.n_dof)
            z, zo = g.precondition(r), o.precondition(r)
            print(dims, coarsest, g.num_levels, [g.level(l)["dims"] for l in range(g.num_levels)], k, _rel(z, zo), flush=True)
            assert _rel(z, zo) < 1e-4
        g.close()
