"""Packed solves (include/psx.h psx_solve_packed): the load cases of several models advance together,
each step of a batch one shared launch for every model that takes it. Every model's solutions,
iteration counts and residuals must be, bit for bit, those of its own multi-instance solve
(psx_solve_multi_device), which the multi-instance and parity tests pin to the oracle; the parts
differ in size, level count and operator form (uniform, per-voxel densities, voids), so the merged
launch sequence interleaves kernels of different levels and forms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _parts(W):
    out = []
    w = W.simp_block((24, 12, 12), seed=3, nrhs=3)
    out.append((w, w["f"]))
    w = W.cantilever(4)
    out.append((w, w["f"][None, :]))
    w = W.void_block((48, 16, 16), seed=2)
    out.append((w, np.stack([w["f"], 0.5 * w["f"]])))
    w = W.simp_block((37, 23, 11), seed=11, nrhs=2)
    out.append((w, w["f"]))
    w = W.void_block((20, 9, 7), seed=5, void_fraction=0.3)
    out.append((w, w["f"][None, :]))
    return out


def _models(P, parts, coarsest=300):
    return [P.Model(w["dims"], w["density"], w["nu"], w["bc"], coarsest) for w, _ in parts]


def _lone(P, models, parts, x0s=None, **kw):
    outs = []
    for i, (m, (w, F)) in enumerate(zip(models, parts)):
        d_f = P.DeviceArray(F.size, F.reshape(-1))
        d_u = P.DeviceArray(F.size)
        d_x = None if x0s is None or x0s[i] is None else P.DeviceArray(F.size, x0s[i].reshape(-1))
        infos = m.solve_multi_device(F.shape[0], d_f, d_u, d_x0=d_x, **kw)
        outs.append((d_u.download().reshape(F.shape), infos))
    return outs


def _packed(P, models, parts, x0s=None, **kw):
    d_f = [P.DeviceArray(F.size, F.reshape(-1)) for _, F in parts]
    d_u = [P.DeviceArray(F.size) for _, F in parts]
    d_x = None
    if x0s is not None:
        d_x = [None if x is None else P.DeviceArray(F.size, x.reshape(-1)) for x, (_, F) in zip(x0s, parts)]
    infos = P.solve_packed(models, d_f, d_u, [F.shape[0] for _, F in parts], d_x0=d_x, **kw)
    return [(u.download().reshape(F.shape), inf) for u, inf, (_, F) in zip(d_u, infos, parts)]


def _same(lone, packed):
    for i, ((U, I), (Up, Ip)) in enumerate(zip(lone, packed)):
        for k in range(U.shape[0]):
            print(i, k, I[k]["iterations"], Ip[k]["iterations"], I[k]["residual"], Ip[k]["residual"], flush=True)
            assert I[k]["converged"] and Ip[k]["converged"]
            assert Ip[k]["iterations"] == I[k]["iterations"]
            assert Ip[k]["folds"] == I[k]["folds"]
            assert Ip[k]["residual"] == I[k]["residual"]
        assert np.array_equal(Up, U), i


@pytest.mark.parametrize("precision", ["mixed", "fp64", "ref"])
def test_packed_equals_own_solves(P, W, precision):
    parts = _parts(W)
    models = _models(P, parts)
    lone = _lone(P, models, parts, tol=1e-9, precision=precision)
    packed = _packed(P, models, parts, tol=1e-9, precision=precision)
    _same(lone, packed)
    # launched one kernel at a time instead of replaying the merged batch as a graph: the same
    _same(lone, _packed(P, models, parts, tol=1e-9, precision=precision, eager=True))
    for m in models:
        m.close()


def test_packed_matrix_core_smoothing_equals_own_solves(P, W):
    # the mixed V-cycle's smoothing as element products on the matrix cores (levels 0 and 1, the
    # child-scale level's own elements in their pass of their own), degree 3 as the bench line runs
    # it: the packed launches give every model its own multi-instance results bit for bit, replayed
    # from the merged graph and launched eagerly
    parts = _parts(W)
    models = _models(P, parts)
    kw = dict(tol=1e-9, precision="mixed", cheb_degree=3, mfma_smoother=1)
    lone = _lone(P, models, parts, **kw)
    _same(lone, _packed(P, models, parts, **kw))
    _same(lone, _packed(P, models, parts, eager=True, **kw))
    for m in models:
        m.close()


def test_packed_warm_start_and_degree3(P, W):
    parts = _parts(W)
    models = _models(P, parts)
    base = _lone(P, models, parts, tol=1e-4, precision="mixed")
    # warm starts for some models (the others from zero), Chebyshev degree 3, a short check interval
    x0s = [U if i % 2 == 0 else None for i, (U, _) in enumerate(base)]
    kw = dict(tol=1e-9, precision="mixed", cheb_degree=3, check_every=3)
    _same(_lone(P, models, parts, x0s=x0s, **kw), _packed(P, models, parts, x0s=x0s, **kw))
    for m in models:
        m.close()


def test_packed_single_and_errors(P, W):
    parts = _parts(W)[:1]
    models = _models(P, parts)
    _same(_lone(P, models, parts, tol=1e-8, precision="mixed"), _packed(P, models, parts, tol=1e-8, precision="mixed"))
    F = parts[0][1]
    d_f, d_u = P.DeviceArray(F.size, F.reshape(-1)), P.DeviceArray(F.size)
    with pytest.raises(P.PsxError) as e:
        P.solve_packed([models[0], models[0]], [d_f, d_f], [d_u, d_u], [F.shape[0]] * 2)
    assert e.value.code == P.ERR_ARGUMENT
    for m in models:
        m.close()


def test_packed_edge_members(P, W):
    """Members whose runs need no iteration (a zero load, a start at the solution) and a part with a
    diagonal shift, packed with ordinary members: each equals its own solve."""
    parts = _parts(W)[:3]
    w = W.cantilever(4)
    shift = np.full(3 * W.nodes(w["dims"]), 1e6)
    m_shift = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300, diag_add=shift)
    models = _models(P, parts) + [m_shift]
    parts = parts + [(w, w["f"][None, :])]
    F0 = parts[0][1].copy()
    F0[1] = 0.0  # a zero load case beside loaded ones
    parts[0] = (parts[0][0], F0)
    exact = _lone(P, models, parts, tol=1e-10, precision="mixed")
    # model 1 starts at its solution: no iteration; the others from zero
    x0s = [None, exact[1][0], None, None]
    lone = _lone(P, models, parts, x0s=x0s, tol=1e-8, precision="mixed")
    packed = _packed(P, models, parts, x0s=x0s, tol=1e-8, precision="mixed")
    assert lone[1][1][0]["iterations"] == 0 and lone[0][1][1]["iterations"] == 0
    _same(lone, packed)
    for m in models:
        m.close()


def test_packed_argument_errors(P, W):
    parts = _parts(W)[:1]
    models = _models(P, parts)
    F = parts[0][1]
    d_f, d_u = P.DeviceArray(F.size, F.reshape(-1)), P.DeviceArray(F.size)
    with pytest.raises(P.PsxError) as e:
        P.solve_packed(models, [d_f], [d_u], [0])  # no load cases
    assert e.value.code == P.ERR_ARGUMENT
    with pytest.raises(P.PsxError) as e:
        P.solve_packed(models * 65, [d_f] * 65, [d_u] * 65, [F.shape[0]] * 65)
    assert e.value.code == P.ERR_ARGUMENT and "64" in str(e.value)
    # the model still solves normally afterwards
    _same(_lone(P, models, parts, tol=1e-8, precision="mixed"), _packed(P, models, parts, tol=1e-8, precision="mixed"))
    for m in models:
        m.close()
