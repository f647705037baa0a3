"""Hierarchy build on the device (psx_model_desc.build_on_device): every level array a model's kernels
read must equal the host build's bit for bit (voxel words with their own-matrix numbering, own
Galerkin matrices, inverse diagonals, node info and scale words, child-scale records, diagonal
shifts, level sizes, coarsest free-DOF map). The host build is itself pinned to the oracle
(test_hierarchy.py and the parity tests), so this carries that pin over. The coarsest dense inverse
comes from a different factorisation (Gauss-Jordan sweeps instead of Cholesky), so it is compared to
rounding, and so are the Lanczos estimates and the solutions that follow from it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _parts(P, W):
    rng = np.random.default_rng(11)
    # ragged grid: odd sizes, empty voxels (absent children), two densities; voxels not face-connected
    # to the clamp removed (no mechanisms)
    dims = (23, 9, 7)
    dens = np.where(rng.random(23 * 9 * 7) < 0.15, 0.0, np.where(rng.random(23 * 9 * 7) < 0.3, 0.25, 1.0)).astype(np.float32)
    dens, _ = P.prune_floating(dims, dens, W.held_nodes(W.clamp_x0(dims)))
    ragged = {"dims": dims, "density": dens, "bc": W.clamp_x0(dims), "nu": 0.31, "name": "ragged 23x9x7"}
    f = np.zeros(3 * W.nodes(dims))
    f[2::3][-(dims[1] + 1) * (dims[0] + 1):] = -1.0
    ragged["f"] = f
    can = W.cantilever(4)
    can["diag_add"] = 0.01 * rng.random(3 * W.nodes(can["dims"]))
    simp = W.simp_block((24, 12, 12), seed=3, nrhs=1)
    simp["f"] = simp["f"][0]
    return [("void", W.void_block((48, 16, 16), seed=2), 300), ("simp", simp, 300), ("ragged", ragged, 120), ("cantilever+shift", can, 100),
            ("void64", W.void_block((64, 32, 32), seed=7), 1200)]


PARTS = ["void", "simp", "ragged", "cantilever+shift", "void64"]


@pytest.fixture(scope="module")
def parts(P, W):
    return {name: (w, c) for name, w, c in _parts(P, W)}


def _pair(P, w, coarsest):
    args = (w["dims"], w["density"], w["nu"], w["bc"], coarsest, w.get("diag_add"))
    return P.Model(*args), P.Model(*args, device_build=True)


@pytest.mark.parametrize("name", PARTS)
def test_device_build_equals_host_build(P, parts, name):
    w, coarsest = parts[name]
    print("device build:", name, w["dims"], flush=True)
    mh, md = _pair(P, w, coarsest)
    assert md.num_levels == mh.num_levels
    for l in range(mh.num_levels):
        ih, id_ = mh.level(l), md.level(l)
        for key in ("dims", "nNodes", "nDof", "freeDof", "nElems", "nOwn"):
            assert ih[key] == id_[key], (l, key, ih[key], id_[key])
        ah, ad = mh.level_arrays(l), md.level_arrays(l)
        assert set(ah) == set(ad), (l, set(ah) ^ set(ad))
        for key in ah:
            a, b = ah[key], ad[key]
            assert a.dtype == b.dtype and a.shape == b.shape, (l, key)
            # bitwise (NaN-free arrays; the view compares the bit patterns of floats)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (l, key, int(np.sum(a != b)))
    ih, ah_inv = mh.coarse()
    idd, ad_inv = md.coarse()
    assert np.array_equal(ih, idd)
    assert np.abs(ad_inv - ah_inv).max() <= 1e-10 * np.abs(ah_inv).max()
    lh, ld = mh.estimate_smoothers(), md.estimate_smoothers()
    for a, b in zip(lh[:-1], ld[:-1]):
        assert abs(a - b) <= 1e-12 * a, (a, b)
    for prec in ("fp64", "mixed", "ref"):
        sh = mh.solve(w["f"], tol=1e-9, max_iter=500, precision=prec)
        sd = md.solve(w["f"], tol=1e-9, max_iter=500, precision=prec)
        err = np.linalg.norm(sd["u"] - sh["u"]) / np.linalg.norm(sh["u"])
        print(name, prec, sh["iterations"], sd["iterations"], "rel diff %.2e" % err, flush=True)
        assert sd["converged"] and abs(sd["iterations"] - sh["iterations"]) <= 1
        assert err <= 1e-7
    mh.close()
    md.close()


def test_device_build_reactions_and_errors(P, W):
    w = W.cantilever(4)
    mh = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 100)
    md = P.Model(w["dims"], w["density"], w["nu"], w["bc"], 100, device_build=True)
    s = mh.solve(w["f"], tol=1e-10, precision="fp64")
    assert np.array_equal(mh.reactions(s["u"], w["f"]), md.reactions(s["u"], w["f"]))
    bad = w["density"].copy()
    bad[3] = 1.5
    with pytest.raises(P.PsxError) as e:
        P.Model(w["dims"], bad, w["nu"], w["bc"], 100, device_build=True)
    assert e.value.code == 1
    with pytest.raises(P.PsxError) as e:
        P.Model(w["dims"], w["density"], 0.6, w["bc"], 100, device_build=True)
    assert e.value.code == 1
    # every DOF held, or a single level with more than 3000 free DOFs: no coarse direct solve, as
    # the host build reports it
    w5 = W.cantilever(5)  # 5400 free DOFs on its finest level
    for part, bc, levels in ((w, np.ones_like(w["bc"]), 0), (w5, w5["bc"], 1)):
        for dev in (False, True):
            with pytest.raises(P.PsxError) as e:
                P.Model(part["dims"], part["density"], part["nu"], bc, 100, None, levels, device_build=dev)
            assert e.value.code == 3, (levels, dev)
