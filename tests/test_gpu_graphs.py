"""Solves replay one captured HIP graph per batch of conjugate-gradient iterations (the same kernels
with the same arguments as the eager launches, which read their scalars from the device-side state).
Every replay must give the eager result bit for bit, with eager work and host synchronisations of the
program between the replays, and a graph must follow the buffers it was captured for."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


# (mixed, 1): the smoothing of levels 0 and 1 as element products on the matrix cores, degree 3
@pytest.mark.parametrize("precision,mfma", [("fp64", 0), ("mixed", 0), ("ref", 0), ("mixed", 1)])
def test_graph_replays_equal_eager(P, W, precision, mfma):
    w = W.void_block((40, 16, 12), seed=21)
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300)
    f = w["f"]
    kw = dict(mfma_smoother=1, cheb_degree=3) if mfma else {}
    eager = m.solve(f, tol=1e-9, precision=precision, eager=True, **kw)
    for rep in range(3):
        # the second and third solves replay the graph the first one captured, with eager solves
        # and reads in between
        g = m.solve(f * (1 + rep), tol=1e-9, precision=precision, check_every=4, **kw)
        e = m.solve(f * (1 + rep), tol=1e-9, precision=precision, check_every=4, eager=True, **kw)
        print(precision, mfma, rep, g["iterations"], e["iterations"], flush=True)
        assert g["iterations"] == e["iterations"]
        assert np.array_equal(g["u"], e["u"])
    assert np.array_equal(m.solve(f, tol=1e-9, precision=precision, **kw)["u"], eager["u"])
    m.close()


def test_graph_follows_grown_buffers(P, W):
    w = W.simp_block((20, 10, 10), seed=2, nrhs=4)
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300)
    F = w["f"]
    one = m.solve(F[0], tol=1e-8, precision="mixed")
    U, _ = m.solve_multi(F, tol=1e-8, precision="mixed")  # grows the work vectors: a new graph
    again = m.solve(F[0], tol=1e-8, precision="mixed")
    eager = m.solve(F[0], tol=1e-8, precision="mixed", eager=True)
    assert np.array_equal(one["u"], eager["u"]) and np.array_equal(again["u"], eager["u"]) and np.array_equal(U[0], eager["u"])
    m.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_follows_new_smoother_estimates(P, W, precision):
    # a captured batch holds the Chebyshev coefficients of its levels: after psx_set_lmax the next
    # solve must use the new ones, exactly as an eager solve does
    w = W.void_block((40, 16, 12), seed=22)
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300, device_build=True)
    f = w["f"]
    before = m.solve(f, tol=1e-9, precision=precision)
    lm = m.estimate_smoothers()
    m.set_lmax(0, 1.3 * lm[0])
    after = m.solve(f, tol=1e-9, precision=precision)
    eager = m.solve(f, tol=1e-9, precision=precision, eager=True)
    print(precision, before["iterations"], after["iterations"], eager["iterations"], flush=True)
    assert after["iterations"] == eager["iterations"] and np.array_equal(after["u"], eager["u"])
    assert not np.array_equal(after["u"], before["u"])
    m.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_length_does_not_change_the_solve(P, W, precision):
    """The direction update p = z + beta p runs inside the outer operator, the new direction written
    beside the old (two buffers alternating), except in the last iteration of an odd-length batch,
    which updates in place so that every batch starts from the same buffer. Both forms compute the
    same expression, so the solve is bit for bit the same for any batch length, replayed or eager."""
    w = W.void_block((40, 16, 12), seed=21)
    m = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 300)
    f = w["f"]
    kw = dict(tol=1e-9, precision=precision, cheb_degree=3)
    if precision == "mixed":
        kw["mfma_smoother"] = 1
    base = m.solve(f, check_every=8, **kw)
    for ce, eager in ((1, False), (3, False), (5, True), (7, False)):
        o = m.solve(f, check_every=ce, eager=eager, **kw)
        print(precision, ce, eager, o["iterations"], base["iterations"], flush=True)
        assert o["iterations"] == base["iterations"] and o["folds"] == base["folds"]
        assert np.array_equal(o["u"], base["u"])
    m.close()
