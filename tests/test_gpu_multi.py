"""Multi-instance solves (include/psx.h psx_solve_multi): several load cases of one part in one set
of launches per V-cycle level. Each load case must give, bit for bit, the solution a lone solve gives
(the instance's tiles, reductions and stop test are those of a single launch), and the lone solves
are pinned to the oracle elsewhere; here the SIMP-block shape (every voxel its own stiffness,
point loads) is also checked against the oracle's FP64 solve directly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(P, w, coarsest=300):
    return P.Model(w["dims"], w["density"], w["nu"], w["bc"], coarsest)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "ref"])
def test_multi_equals_lone_solves(P, W, precision):
    w = W.simp_block((24, 12, 12), seed=3, nrhs=5)
    F = w["f"]
    m = _model(P, w)
    lone = [m.solve(F[k], tol=1e-9, precision=precision) for k in range(F.shape[0])]
    U, infos = m.solve_multi(F, tol=1e-9, precision=precision)
    for k in range(F.shape[0]):
        print(precision, k, infos[k], flush=True)
        assert infos[k]["converged"]
        assert infos[k]["iterations"] == lone[k]["iterations"]
        assert np.array_equal(U[k], lone[k]["u"])
    m.close()


def test_multi_zero_load_and_warm_start(P, W):
    w = W.simp_block((20, 10, 10), seed=5, nrhs=3)
    F = w["f"].copy()
    F[1] = 0.0  # no load: u = 0 at once
    m = _model(P, w)
    U, infos = m.solve_multi(F, tol=1e-8, precision="mixed")
    assert infos[1]["converged"] and infos[1]["iterations"] == 0 and not U[1].any()
    assert infos[0]["converged"] and infos[2]["converged"]
    # warm start from the solutions: every case is converged before the first iteration
    U2, infos2 = m.solve_multi(F, tol=1e-6, X0=U, precision="mixed")
    assert all(i["iterations"] == 0 and i["converged"] for i in infos2)
    # a smaller batch after a larger one reuses the grown work vectors
    U3, infos3 = m.solve_multi(F[2:3], tol=1e-8, precision="mixed")
    assert np.array_equal(U3[0], U[2])
    m.close()


def test_multi_matches_oracle(P, W, cpu):
    w = W.simp_block((16, 8, 8), seed=7, nrhs=3)
    m = _model(P, w, 200)
    U, infos = m.solve_multi(w["f"], tol=1e-10, precision="mixed")
    ref = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], 200)
    for k in range(3):
        r = ref.solve(w["f"][k], tol=1e-12)
        err = np.linalg.norm(U[k] - r["u"]) / np.linalg.norm(r["u"])
        print(k, infos[k]["iterations"], r["iterations"], err, flush=True)
        assert err < 1e-8
    m.close()


def test_multi_device_vectors(P, W):
    w = W.simp_block((20, 10, 10), seed=9, nrhs=4)
    F = w["f"]
    m = _model(P, w)
    U, _ = m.solve_multi(F, tol=1e-9, precision="mixed")
    d_f = P.DeviceArray(F.size, F.reshape(-1))
    d_u = P.DeviceArray(F.size)
    infos = m.solve_multi_device(4, d_f, d_u, tol=1e-9, precision="mixed")
    assert all(i["converged"] for i in infos)
    assert np.array_equal(d_u.download().reshape(F.shape), U)
    m.close()


def test_block_precondition_equals_lone(P, W):
    # the block V-cycle of a LOBPCG preconditioner: each vector's result is its lone V-cycle
    for w in (W.simp_block((24, 12, 12), seed=3, nrhs=1), W.void_block((32, 16, 16), seed=2)):
        m = _model(P, w)
        R = np.random.default_rng(11).standard_normal((6, m.n_dof))
        R[2] *= 1e-9  # the V-cycle scales each vector to order one on its own
        Z = m.precondition_multi(R)
        for k in range(R.shape[0]):
            assert np.array_equal(Z[k], m.precondition(R[k])), k
        # a smaller block after a larger one
        assert np.array_equal(m.precondition_multi(R[4:6]), Z[4:6])
        m.close()


def test_block_precondition_matches_oracle_per_vector(P, W, cpu):
    # LOBPCG's block preconditioner as the reference's eigen solver drives it (eigen.cpp:508-553):
    # a block of free-DOF vectors scattered to full length (held DOFs 0), one V-cycle per vector
    # (cpuPreconditioner), gathered back. The device's block V-cycle against the oracle's
    # VoxelFEA::precondition of each vector, to the FP32 V-cycle's bar; one part carries a diagonal
    # shift, as a shifted eigen model's does.
    can = W.cantilever(4)
    can["diag_add"] = 0.01 * np.random.default_rng(5).random(3 * W.nodes(can["dims"]))
    for w in (W.void_block((32, 16, 16), seed=2), can):
        da = w.get("diag_add")
        m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300, diag_add=da)
        o = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], 300, diag_add=da)
        fixed = o.level_arrays(0)[0]
        free = np.flatnonzero(fixed == 0)
        rng = np.random.default_rng(13)
        block = rng.standard_normal((5, free.size))
        block[3] *= 1e-7
        full = np.zeros((5, m.n_dof))
        full[:, free] = block
        Z = m.precondition_multi(full)
        for k in range(5):
            zo = o.precondition(full[k])
            err = np.linalg.norm(Z[k] - zo) / np.linalg.norm(zo)
            print(w["name"], k, err, flush=True)
            assert err < 1e-4
            assert not Z[k][fixed != 0].any()
        m.close()


@pytest.mark.parametrize("nrhs", [5, 16])
def test_matrix_core_own_elements(P, W, cpu, nrhs):
    # own-element levels on the matrix cores (mfma_smoother bit 2): each load case's column of the
    # contraction is computed the same way for any batch, so a batch equals its lone solves bit for
    # bit; the BF16 smoother changes only the preconditioner, so the solution still meets the FP64
    # oracle at a tight tolerance
    w = W.simp_block((24, 12, 12), seed=3, nrhs=nrhs)
    F = w["f"]
    m = _model(P, w)
    U, infos = m.solve_multi(F, tol=1e-10, precision="mixed", mfma_smoother=2)
    ref = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], 300)
    for k in (0, nrhs - 1):
        lone = m.solve(F[k], tol=1e-10, precision="mixed", mfma_smoother=2)
        print(nrhs, k, infos[k]["iterations"], lone["iterations"], flush=True)
        assert infos[k]["converged"] and infos[k]["iterations"] == lone["iterations"]
        assert np.array_equal(U[k], lone["u"])
        r = ref.solve(F[k], tol=1e-12)
        assert np.linalg.norm(U[k] - r["u"]) / np.linalg.norm(r["u"]) < 1e-8
    # the VALU smoother on the same batch (FP16 own matrices, FP32 arithmetic): the BF16 contraction
    # costs at most a few iterations (measured: 31-35 against 31 at tol 1e-10)
    V, infos32 = m.solve_multi(F, tol=1e-10, precision="mixed")
    assert max(abs(a["iterations"] - b["iterations"]) for a, b in zip(infos, infos32)) <= 6
    m.close()


def test_matrix_core_element_product_batch_equals_lone(P, W, cpu):
    """The element-by-element matrix-core operator in the V-cycle and as the outer operator
    (mfma_smoother 1 | 4) over a batch of load cases: each block works on one case's vectors with that
    case's reference value, so every case equals its lone solve bit for bit, and the FP64 oracle to
    1e-8."""
    w = W.simp_block((24, 12, 12), seed=5, nrhs=5)
    F = w["f"]
    m = _model(P, w)
    U, infos = m.solve_multi(F, tol=1e-10, precision="mixed", mfma_smoother=5)
    ref = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], 300)
    for k in (0, 2, 4):
        lone = m.solve(F[k], tol=1e-10, precision="mixed", mfma_smoother=5)
        assert infos[k]["converged"] and infos[k]["iterations"] == lone["iterations"]
        assert np.array_equal(U[k], lone["u"])
        r = ref.solve(F[k], tol=1e-12)
        assert np.linalg.norm(U[k] - r["u"]) / np.linalg.norm(r["u"]) < 1e-8
    m.close()
