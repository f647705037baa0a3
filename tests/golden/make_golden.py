"""Writes the golden fixtures in this folder from the reference's own solver, compiled from its
sources into oracle/_ref/libpsref.so (oracle/Makefile). Run where the reference sources exist:

    make -C oracle && PARTS_SIM_THREADS=8 python tests/golden/make_golden.py

The fixtures are data only (inputs and outputs as .npz); the reference's chunked reductions follow
PARTS_SIM_THREADS, recorded in every file.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "parts-sim_amd"))
os.environ.setdefault("PARTS_SIM_THREADS", "8")

import cpu  # noqa: E402
import workloads as W  # noqa: E402


def case(w, name, tol, coarsest=1100, x0=None, E=None, h=None, vectors=True):
    f = w["f"] if w["f"].ndim == 1 else w["f"][0]
    m = cpu.Model(w["dims"], w["density"], w["nu"], w["bc"], coarsest, kind="ref")
    s = m.solve(f, tol=tol, max_iter=500, x0=x0)
    lmax = np.array([m.level(l)["lmax"] for l in range(m.num_levels)])
    x = np.random.default_rng(11).standard_normal(m.n_dof)
    y = m.apply(x)
    r = np.random.default_rng(12).standard_normal(m.n_dof)
    z = m.precondition(r)
    E = E or w["E"]
    h = h or w["h"]
    u_phys = s["u"] / (E * h)
    vm, p1, p3 = m.stresses(u_phys, E, h, True)
    R = m.reactions(s["u"], f)
    # inputs: the generator is deterministic, but the fixture keeps them so it stands alone; the
    # apply / precondition inputs are numpy PCG64 normals with seeds 11 / 12
    extra = {"apply_y": y, "precond_z": z} if vectors else {}
    np.savez_compressed(os.path.join(HERE, name + ".npz"), dims=np.array(w["dims"]), density=w["density"], bc=w["bc"], f=f, nu=w["nu"],
                        E=E, h=h, tol=tol, coarsest=coarsest, threads=int(os.environ["PARTS_SIM_THREADS"]), u=s["u"],
                        iterations=s["iterations"], residual=s["residual"], converged=s["converged"], lmax=lmax,
                        vm=vm, p1=p1, p3=p3, reactions=R, levels=m.num_levels, **extra)
    print(name, w["dims"], "levels", m.num_levels, "its", s["iterations"], "res", s["residual"])


def main():
    K, cs, gb = cpu.hex_element(0.3, "ref")
    K33, _, _ = cpu.hex_element(0.33, "ref")
    np.savez_compressed(os.path.join(HERE, "hex8.npz"), nu=np.array([0.3, 0.33]), K_03=K, corner_03=cs, gauss_03=gb, K_033=K33)
    case(W.cantilever(4), "cantilever4", 1e-8)
    case(W.c1_cantilever(), "c1_cantilever", 1e-8, coarsest=1200, vectors=False)
    case(W.void_block((48, 16, 16), seed=3), "void_block_48x16x16", 1e-8, coarsest=1200, vectors=False)
    case(W.simp_block((24, 12, 12), nrhs=1), "simp_block_24x12x12", 1e-6, coarsest=600)
    # pruneFloating on a sparse random part (40 % solid voxels: many floating clusters)
    dims = (16, 10, 8)
    dens = (np.random.default_rng(9).uniform(size=16 * 10 * 8) < 0.4).astype(np.float32)
    held = W.held_nodes(W.clamp_x0(dims))
    out, removed = cpu.prune_floating(dims, dens, held, kind="ref")
    np.savez_compressed(os.path.join(HERE, "prune_16x10x8.npz"), dims=np.array(dims), density=dens, held=held, out=out, removed=removed)
    print("prune removed", removed)


if __name__ == "__main__":
    main()
