"""The topology loop on the device (include/psx.h psx_optimize_topology, psx_element_energies,
psx_topology_filter) against the oracle's restatement of optimizeTopology (topology.cpp:104-183),
which is pinned bit for bit to the compiled reference in the CPU suite.

The element energies are FP64 in the reference's order (equal to rounding of the element matrix),
the filter and its transpose gather in the reference's offset order (bit for bit). The whole loop
solves with this library's solver where the oracle solves with its own VoxelFEA, both to a relative
residual of 1e-8, and sums volumes in another order; after 8 iterations the densities differ by one
float rounding (6e-8 measured, 1e-5 allowed) and the compliances to 1e-6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cantilever(W, dims):
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


def test_element_energies_match_oracle(P, W, cpu):
    dims, fill, keep, bc, f = _cantilever(W, (20, 8, 6))
    rho = np.random.default_rng(2).uniform(0.2, 1.0, fill.size).astype(np.float32)
    rho[::7] = 0  # empty voxels have no element
    m = P.Model(dims, rho, 0.3, bc, 300)
    u = np.random.default_rng(5).standard_normal(m.n_dof)
    e = m.element_energies(u)
    ref = cpu.Model(dims, rho, 0.3, bc, 300).element_energies(u)
    print("energies", e.size, np.abs(e - ref).max() / np.abs(ref).max(), flush=True)
    assert e.shape == ref.shape
    assert np.abs(e - ref).max() <= 1e-13 * np.abs(ref).max()
    m.close()


@pytest.mark.parametrize("rmin", [1.0, 1.5, 2.4])
def test_filter_bitwise(P, cpu, rmin):
    dims = (13, 9, 7)
    fill = (np.random.default_rng(1).uniform(0, 1, 13 * 9 * 7) > 0.3).astype(np.float32)
    fill[fill > 0] *= np.random.default_rng(6).uniform(0.5, 1, int((fill > 0).sum())).astype(np.float32)
    x = np.random.default_rng(3).uniform(0, 1, int((fill > 0).sum()))
    for t in (False, True):
        out = P.topology_filter(dims, fill, rmin, x, transpose=t)
        ref = cpu.topo_filter(dims, fill, rmin, x, transpose=t)
        assert np.array_equal(out, ref), (rmin, t, np.abs(out - ref).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_optimize_topology_matches_oracle(P, W, cpu, precision):
    dims, fill, keep, bc, f = _cantilever(W, (24, 8, 6))
    kw = dict(vol_frac=0.4, penal=3.0, rmin=1.5, move=0.2, tol=0.01, max_iter=8, coarsest_max_dof=300)
    d, h, kept = P.optimize_topology(dims, fill, keep, bc, f, 0.3, precision=precision, **kw)
    dr, hr = cpu.optimize_topology(dims, fill, keep, bc, f, nu=0.3, **kw)
    print(precision, h, hr, np.abs(d - dr).max(), flush=True)
    assert h.shape == hr.shape
    assert np.array_equal(h[:, 0], hr[:, 0])
    assert np.allclose(h[:, 1], hr[:, 1], rtol=1e-6, atol=0)
    assert np.allclose(h[:, 2], hr[:, 2], rtol=0, atol=1e-7)
    assert np.allclose(h[:, 3], hr[:, 3], rtol=0, atol=1e-5)
    assert np.abs(d - dr).max() <= 1e-5
    assert kept == pytest.approx(2 / 24)
    # the budget holds and the kept voxels stay solid
    assert np.all(d[keep > 0] == 1)
    assert h[-1, 2] < 0.45


def test_optimize_topology_void_and_stop(P, W, cpu):
    # voxels with fill 0 stay out of the design; on_iter stops the loop
    dims, fill, keep, bc, f = _cantilever(W, (16, 8, 6))
    fill.reshape(6, 8, 16)[2:4, 3:5, 4:12] = 0
    fill.reshape(6, 8, 16)[:, :, 6] *= 0.5
    seen, fields = [], []

    def on_iter(it, c, v, ch, dens):
        seen.append(it)
        fields.append(dens)
        return it == 3

    d, h, _ = P.optimize_topology(dims, fill, keep, bc, f, 0.3, max_iter=10, precision="fp64", coarsest_max_dof=300, on_iter=on_iter)
    assert seen == [0, 1, 2, 3] and h.shape == (4, 4)
    assert np.all(d[fill == 0] == 0)
    # the loop stopped after iteration 3's solve: its final densities are the ones shown then
    assert np.array_equal(fields[-1], d)
    for dens, row in zip(fields, h):
        assert np.all(dens[fill == 0] == 0) and np.all(dens[keep > 0] == 1)
        # the shown densities carry the iteration's volume fraction
        assert abs((dens.astype(np.float64) * fill).sum() / fill.astype(np.float64).sum() - row[2]) < 1e-6
    dr, hr = cpu.optimize_topology(dims, fill, keep, bc, f, nu=0.3, max_iter=4, coarsest_max_dof=300)
    # the oracle at max_iter 4 runs the same 4 solves, then one more density update
    assert np.allclose(h[:, 1], hr[:, 1], rtol=1e-5, atol=0)
    with pytest.raises(P.PsxError):
        P.optimize_topology(dims, np.zeros_like(fill), keep, bc, f, 0.3)
