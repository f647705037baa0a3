// ORACLE — test infrastructure only, never the product.
//
// C ABI over the reference's own VoxelFEA, compiled from the reference sources where they lie
// (oracle/Makefile: native/src/fea/hex8.cpp, native/src/fea/voxel_fea.cpp, native/src/util/parallel.cpp,
// and for the voxelizer native/src/core/voxelize.cpp with the mesh and BVH sources it links against;
// for the input contract of a CAD part native/src/fea/structural.cpp and native/src/core/materials.cpp;
// for the break test native/src/fea/static_study.cpp with the no-GPU builds of native/src/gpu/gpu.cpp
// and native/src/gpu/gpu_fea.cpp, gpu_fea.cpp:654-663)
// into oracle/_ref/libpsref.so. The tests use it to pin the restatement in oracle/oracle.cpp bit for
// bit, and bench.py may time it as the CPU baseline. The entry points have the same shapes as
// oracle.cpp's orc_* functions.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

// the member templates the tests call (applyT<float>, restrict, prolongAdd) are private: open them
// for this test harness only, and instantiate them in this translation unit
#define private public
#include "fea/voxel_fea.hpp"
#undef private
#include "fea/voxel_fea.cpp"
#include "core/voxelize.hpp"
#include "fea/structural.hpp"
#include "fea/topology.hpp"
#include "fea/static_study.hpp"

static std::string g_err;

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

int ref_hex_element(double nu, double* K, double* corner, double* gauss) {
    try {
        const auto h = ps::hexElement(nu);
        std::memcpy(K, h.K.data(), 576 * sizeof(double));
        for (int a = 0; a < 8; a++) {
            if (corner) std::memcpy(corner + 144 * a, h.cornerStress[a].data(), 144 * sizeof(double));
            if (gauss) std::memcpy(gauss + 144 * a, h.gaussB[a].data(), 144 * sizeof(double));
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void* ref_create(int nx, int ny, int nz, const float* density, double nu, const uint8_t* bc, int64_t coarsestMaxDof, const double* diagAdd) {
    try {
        const int64_t nE = int64_t(nx) * ny * nz, nD = 3 * int64_t(nx + 1) * (ny + 1) * (nz + 1);
        std::vector<float> dens(density, density + nE);
        std::vector<uint8_t> b(bc, bc + nD);
        std::vector<double> da;
        if (diagAdd) da.assign(diagAdd, diagAdd + nD);
        return new ps::VoxelFEA({nx, ny, nz}, dens, nu, b, coarsestMaxDof, diagAdd ? &da : nullptr);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void ref_destroy(void* h) { delete static_cast<ps::VoxelFEA*>(h); }

int ref_num_levels(void* h) { return int(static_cast<ps::VoxelFEA*>(h)->levels.size()); }

void ref_level_info(void* h, int l, int* dims, int64_t* info, double* lmax) {
    const ps::Level& L = static_cast<ps::VoxelFEA*>(h)->levels[l];
    const int d[6] = {L.nx, L.ny, L.nz, L.NX, L.NY, L.NZ};
    std::memcpy(dims, d, sizeof d);
    info[0] = L.nNodes;
    info[1] = L.nDof;
    info[2] = L.freeDof;
    info[3] = int64_t(L.K.size() / 576);
    *lmax = L.lmax;
}

void ref_level_arrays(void* h, int l, uint8_t* fixed, double* invDiag, float* nodeScale, float* voxelScale) {
    const ps::Level& L = static_cast<ps::VoxelFEA*>(h)->levels[l];
    if (fixed) std::copy(L.fixed.begin(), L.fixed.end(), fixed);
    if (invDiag) std::copy(L.invDiag.begin(), L.invDiag.end(), invDiag);
    if (nodeScale) std::copy(L.nodeScale.begin(), L.nodeScale.end(), nodeScale);
    if (voxelScale) std::copy(L.voxelScale.begin(), L.voxelScale.end(), voxelScale);
}

void ref_level_kb(void* h, int l, double* Kb576, double* Sb324) {
    const ps::Level& L = static_cast<ps::VoxelFEA*>(h)->levels[l];
    if (Kb576) std::memcpy(Kb576, L.Kb.data(), 576 * sizeof(double));
    if (Sb324) std::memcpy(Sb324, L.Sb.data(), 324 * sizeof(double));
}

void ref_apply(void* h, int l, const double* x, double* y, int raw) { static_cast<ps::VoxelFEA*>(h)->apply(l, x, y, raw != 0); }

void ref_apply32(void* h, int l, const float* x, float* y, int raw) { static_cast<ps::VoxelFEA*>(h)->applyT<float>(l, x, y, raw != 0); }

void ref_restrict(void* h, int l, const double* rf, double* rc, int zeroFixed) {
    auto* M = static_cast<ps::VoxelFEA*>(h);
    M->restrict<double>(M->levels[l], M->levels[l + 1], rf, rc, zeroFixed != 0);
}

void ref_prolong_add(void* h, int l, const double* zc, double* zf) {
    auto* M = static_cast<ps::VoxelFEA*>(h);
    M->prolongAdd<double>(M->levels[l], M->levels[l + 1], zc, zf);
}

void ref_prepare_smoothers(void* h) { static_cast<ps::VoxelFEA*>(h)->prepareSmoothers(); }

int64_t ref_coarse_m(void* h) { return static_cast<ps::VoxelFEA*>(h)->coarse.m; }

void ref_coarse_solve64(void* h, const double* r, double* z) { static_cast<ps::VoxelFEA*>(h)->coarseSolve64(r, z); }

void ref_precondition(void* h, const double* r, double* z) { static_cast<ps::VoxelFEA*>(h)->precondition(r, z); }

int ref_solve(void* h, const double* f, const double* x0, double tol, int maxIter, double* u, int* iterations, double* residual) {
    try {
        auto* M = static_cast<ps::VoxelFEA*>(h);
        const int64_t n = M->nDof();
        std::vector<double> fv(f, f + n), xv;
        ps::SolveOptions o;
        o.tol = tol;
        o.maxIter = maxIter;
        if (x0) {
            xv.assign(x0, x0 + n);
            o.x0 = &xv;
        }
        const auto r = M->solve(fv, o);
        std::copy(r.u.begin(), r.u.end(), u);
        *iterations = r.iterations;
        *residual = r.residual;
        return r.converged ? 1 : 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void ref_reactions(void* h, const double* u, const double* f, double* R) {
    auto* M = static_cast<ps::VoxelFEA*>(h);
    const int64_t n = M->nDof();
    std::vector<double> uv(u, u + n), fv;
    if (f) fv.assign(f, f + n);
    const auto r = M->reactions(uv, f ? &fv : nullptr);
    R[0] = r[0]; R[1] = r[1]; R[2] = r[2];
}

void ref_stresses(void* h, const double* u, double E, double hvox, int principals, float* vm, float* p1, float* p3) {
    auto* M = static_cast<ps::VoxelFEA*>(h);
    std::vector<double> uv(u, u + M->nDof());
    const auto s = M->stresses(uv, E, hvox, principals != 0);
    std::copy(s.nodeVM.begin(), s.nodeVM.end(), vm);
    if (principals) {
        std::copy(s.nodeP1.begin(), s.nodeP1.end(), p1);
        std::copy(s.nodeP3.begin(), s.nodeP3.end(), p3);
    }
}

void ref_stress_tensors(void* h, const double* u, double E, double hvox, float* out) {
    auto* M = static_cast<ps::VoxelFEA*>(h);
    std::vector<double> uv(u, u + M->nDof());
    const auto s = M->stressTensors(uv, E, hvox);
    std::copy(s.begin(), s.end(), out);
}

void ref_principal(double sx, double sy, double sz, double txy, double tyz, double tzx, double* out) { ps::principalStresses(sx, sy, sz, txy, tyz, tzx, out); }

double ref_von_mises(const double* s) { return ps::vonMises(s); }

int ref_prune_floating(int nx, int ny, int nz, const float* density, const uint8_t* held, float* out) {
    const int64_t nE = int64_t(nx) * ny * nz, nN = int64_t(nx + 1) * (ny + 1) * (nz + 1);
    const auto r = ps::pruneFloating({nx, ny, nz}, std::vector<float>(density, density + nE), std::vector<uint8_t>(held, held + nN));
    std::copy(r.density.begin(), r.density.end(), out);
    return r.removed;
}

double ref_tridiagonal_max(const double* a, int k, const double* b) {
    return ps::tridiagonalMax(std::vector<double>(a, a + k), std::vector<double>(b, b + std::max(0, k - 1)));
}

void ref_cheb_coefficients(double lmax, double* out3) {
    const auto c = ps::chebyshevCoefficients(lmax);
    out3[0] = c.first;
    out3[1] = c.c1;
    out3[2] = c.c2;
}

// wall time of the reference's own multithreaded solve (bench.py's CPU baseline): build + solve
double ref_timed_solve(int nx, int ny, int nz, const float* density, double nu, const uint8_t* bc, int64_t coarsestMaxDof, const double* f, double tol,
                       int maxIter, double* u, int* iterations, double* residual, double* buildSeconds) {
    try {
        const auto t0 = std::chrono::steady_clock::now();
        void* h = ref_create(nx, ny, nz, density, nu, bc, coarsestMaxDof, nullptr);
        if (!h) return -1;
        const auto t1 = std::chrono::steady_clock::now();
        const int ok = ref_solve(h, f, nullptr, tol, maxIter, u, iterations, residual);
        const auto t2 = std::chrono::steady_clock::now();
        ref_destroy(h);
        if (ok < 0) return -1;
        *buildSeconds = std::chrono::duration<double>(t1 - t0).count();
        return std::chrono::duration<double>(t2 - t1).count();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// the same in phases: build, smoother eigenvalue estimates, then maxIter PCG iterations (a bounded
// sample of a long solve); out[0..2] = seconds of each phase
int ref_timed_phases(int nx, int ny, int nz, const float* density, double nu, const uint8_t* bc, int64_t coarsestMaxDof, const double* f,
                     double tol, int maxIter, double* u, int* iterations, double* residual, double* out) {
    try {
        const auto t0 = std::chrono::steady_clock::now();
        void* h = ref_create(nx, ny, nz, density, nu, bc, coarsestMaxDof, nullptr);
        if (!h) return -1;
        const auto t1 = std::chrono::steady_clock::now();
        ref_prepare_smoothers(h);
        const auto t2 = std::chrono::steady_clock::now();
        const int ok = ref_solve(h, f, nullptr, tol, maxIter, u, iterations, residual);
        const auto t3 = std::chrono::steady_clock::now();
        ref_destroy(h);
        if (ok < 0) return -1;
        out[0] = std::chrono::duration<double>(t1 - t0).count();
        out[1] = std::chrono::duration<double>(t2 - t1).count();
        out[2] = std::chrono::duration<double>(t3 - t2).count();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

unsigned ref_threads() { return ps::ThreadPool::instance().size(); }

void ref_element_energies(void* h, const double* u, double* out) {
    auto* M = static_cast<ps::VoxelFEA*>(h);
    const std::vector<double> uv(u, u + M->nDof());
    const auto e = ps::elementEnergies(*M, uv);
    std::copy(e.begin(), e.end(), out);
}

// the reference's optimizeTopology with its own VoxelFEA as the solve (tol 1e-8, 1000 iterations,
// warm-started as the loop asks); history: it, compliance, volume, change per iteration
int ref_optimize_topology(const int* dims, const float* fill, const uint8_t* keep, const uint8_t* bc, const double* f, double nu, double volFrac,
                          double penal, double rmin, double move, double tol, int maxIter, int64_t coarsestMaxDof, float* density, double* history) {
    try {
        const int64_t nE = int64_t(dims[0]) * dims[1] * dims[2], nD = 3 * int64_t(dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1);
        ps::TopologyOptions o;
        o.dims = {dims[0], dims[1], dims[2]};
        o.fill.assign(fill, fill + nE);
        o.keep.assign(keep, keep + nE);
        o.volFrac = volFrac;
        o.penal = penal;
        o.rmin = rmin;
        o.move = move;
        o.tol = tol;
        o.maxIter = maxIter;
        const std::vector<uint8_t> b(bc, bc + nD);
        const std::vector<double> fv(f, f + nD);
        std::unique_ptr<ps::VoxelFEA> fea;
        o.solve = [&](const std::vector<float>& dens, const std::vector<double>* x0) {
            fea = std::make_unique<ps::VoxelFEA>(o.dims, dens, nu, b, coarsestMaxDof, nullptr);
            ps::SolveOptions so;
            so.tol = 1e-8;
            so.maxIter = 1000;
            so.x0 = x0;
            auto r = fea->solve(fv, so);
            return ps::TopologySolve{std::move(r.u), fv, fea.get()};
        };
        const auto res = ps::optimizeTopology(o);
        std::copy(res.density.begin(), res.density.end(), density);
        for (size_t q = 0; q < res.history.size(); q++) {
            history[4 * q] = res.history[q].it;
            history[4 * q + 1] = res.history[q].compliance;
            history[4 * q + 2] = res.history[q].volume;
            history[4 * q + 3] = res.history[q].change;
        }
        return int(res.history.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void ref_grid_for_box(const double* mn, const double* mx, int n, int pad, double* origin, double* h, int* dims) {
    const ps::Grid g = ps::gridForBox({mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}, n, pad);
    for (int d = 0; d < 3; d++) {
        origin[d] = g.origin[d];
        dims[d] = g.dims[d];
    }
    *h = g.h;
}

int ref_voxelize(const float* pos, int64_t nVert, const uint32_t* index, int64_t nTri, const double* origin, double h, const int* dims, int sub, float* frac) {
    try {
        ps::Grid g;
        for (int d = 0; d < 3; d++) {
            g.origin[d] = origin[d];
            g.dims[d] = dims[d];
        }
        g.h = h;
        const std::vector<float> p(pos, pos + 3 * nVert);
        const std::vector<uint32_t> ix(index, index + 3 * nTri);
        const auto f = ps::voxelize(p, ix, g, sub);
        std::memcpy(frac, f.data(), f.size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// ---- input contract of a CAD part: Parts built by the reference's buildPart (weld, orient, refine,
// triangle data), whose final vertices and triangles the tests hand to the product as well ----
static std::vector<std::shared_ptr<ps::Part>> g_parts;

int ref_build_part(const float* pos, int64_t nFloats, const uint32_t* index, int64_t nIndex, int64_t* nVert, int64_t* nTri) {
    try {
        ps::MeshSource src;
        src.positions.assign(pos, pos + nFloats);
        src.index.assign(index, index + nIndex);
        g_parts.push_back(ps::buildPart(src));
        *nVert = g_parts.back()->nVert;
        *nTri = g_parts.back()->nTri;
        return int(g_parts.size()) - 1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void ref_part_mesh(int part, float* vertices, uint32_t* tris, float* normal, float* area) {
    const ps::Part& p = *g_parts.at(size_t(part));
    std::copy(p.vertices.begin(), p.vertices.end(), vertices);
    std::copy(p.tris.begin(), p.tris.end(), tris);
    if (normal) std::copy(p.triNormal.begin(), p.triNormal.end(), normal);
    if (area) std::copy(p.triArea.begin(), p.triArea.end(), area);
}

double ref_wall_thickness(int part, int samples) {
    ps::Part& p = *g_parts.at(size_t(part));
    p.wallThickness.reset();
    return ps::typicalWallThickness(p, samples);
}

void ref_thin_wall_density(int part, const double* origin, double h, const int* dims, double maxThickness, float* out) {
    ps::Grid g;
    for (int d = 0; d < 3; d++) g.origin[d] = origin[d], g.dims[d] = dims[d];
    g.h = h;
    const auto v = ps::thinWallDensity(*g_parts.at(size_t(part)), g, maxThickness);
    std::copy(v.begin(), v.end(), out);
}

static std::vector<std::unique_ptr<ps::StructuralModel>> g_models;

// StructuralModel(part, resolution): the grid first (density etc. NULL), then the arrays
int ref_structural_model(int part, int resolution, double* origin, double* h, int* dims, float* density, uint8_t* active, uint8_t* surface, double* wall,
                         int* counts) {
    try {
        g_parts.at(size_t(part))->wallThickness.reset();
        g_models.push_back(std::make_unique<ps::StructuralModel>(g_parts.at(size_t(part)), resolution));
        const ps::StructuralModel& m = *g_models.back();
        for (int d = 0; d < 3; d++) origin[d] = m.grid.origin[d], dims[d] = m.grid.dims[d];
        *h = m.grid.h;
        if (density) std::copy(m.density.begin(), m.density.end(), density);
        if (active) std::copy(m.activeNode.begin(), m.activeNode.end(), active);
        if (surface) std::copy(m.surfaceNode.begin(), m.surfaceNode.end(), surface);
        if (wall) *wall = m.wallThickness;
        if (counts) counts[0] = m.voxelCount, counts[1] = m.thinVoxels;
        return int(g_models.size()) - 1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// assemble on a model from ref_structural_model. Patches flat: patch p has ntris[p] triangles (from
// tris, in order), clip[4 p .. 4 p + 3] = centre, radius when hasClip[p]; fixture q takes fixNP[q]
// patches in order, then load q takes loadNP[q]; wind load q takes windN[q] forces from windForces
int ref_assemble(int model, const int64_t* ntris, const int32_t* tris, const int* hasClip, const double* clip, int nFix, const int* fixNP, int nLoad,
                 const int* loadType, const double* loadMag, const double* loadDir, const int* loadNP, const float* windForces, const int64_t* windN, int gravity,
                 const double* mat, double toMeters, uint8_t* bc, double* f, int* fixedNodes, double* total, char* warnings, size_t wlen) {
    try {
        const ps::StructuralModel& m = *g_models.at(size_t(model));
        int pi = 0;
        int64_t at = 0;
        auto takePatches = [&](int n) {
            std::vector<ps::Patch> out;
            for (int q = 0; q < n; q++, pi++) {
                ps::Patch p;
                p.tris.assign(tris + at, tris + at + ntris[pi]);
                at += ntris[pi];
                if (hasClip[pi]) p.clip = ps::Clip{{clip[4 * pi], clip[4 * pi + 1], clip[4 * pi + 2]}, clip[4 * pi + 3]};
                out.push_back(std::move(p));
            }
            return out;
        };
        std::vector<ps::Fixture> fx;
        for (int q = 0; q < nFix; q++) fx.push_back({"F" + std::to_string(q), takePatches(fixNP[q])});
        std::vector<ps::Load> lds;
        int64_t wat = 0;
        for (int q = 0; q < nLoad; q++) {
            ps::Load l;
            l.type = ps::Load::Type(loadType[q]);
            l.name = "L" + std::to_string(q);
            l.magnitude = loadMag[q];
            l.dir = {loadDir[3 * q], loadDir[3 * q + 1], loadDir[3 * q + 2]};
            l.patches = takePatches(loadNP[q]);
            if (l.type == ps::Load::Wind) {
                l.forces.assign(windForces + wat, windForces + wat + windN[q]);
                wat += windN[q];
            }
            lds.push_back(std::move(l));
        }
        ps::Material mt;
        mt.E = mat[0], mt.nu = mat[1], mt.yield = mat[2], mt.uts = mat[3], mt.density = mat[4], mt.brittle = mat[5] != 0;
        const ps::Assembly a = m.assemble(fx, lds, gravity != 0, mt, toMeters);
        std::copy(a.bc.begin(), a.bc.end(), bc);
        std::copy(a.f.begin(), a.f.end(), f);
        *fixedNodes = a.fixedNodes;
        for (int d = 0; d < 3; d++) total[d] = a.total[d];
        std::string all;
        for (const auto& w : a.warnings) all += w + "\n";
        std::snprintf(warnings, wlen, "%s", all.c_str());
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}


// breakTest (static_study.cpp:197-279) as the reference runs it without a GPU (useGPU = false): each
// step's VoxelFEA with coarsestMaxDof 1100 solved to 1e-6 in at most 400 iterations
// (static_study.cpp:67-68, 83, 117-120). Same output shape as orc_break_test; tol, maxIter and
// coarsestMaxDof must be those fixed values (the reference takes no options for them).
int ref_break_test(const int* dims, const float* density, double nu, const uint8_t* bc, const double* f, double E, double h, double strength,
                   int principal, int maxSteps, double tol, int maxIter, int64_t coarsestMaxDof, double* scalars, int* counts, int* cracked,
                   int crackedCap, int* detached, int detachedCap, float* vm, double* peak, int* removed, int* reason) {
    try {
        if (tol != 1e-6 || maxIter != 400 || coarsestMaxDof != 1100)
            throw std::invalid_argument("the reference's CPU break test solves to 1e-6 in 400 iterations with coarsestMaxDof 1100");
        ps::StructuralInput in;
        in.dims = {dims[0], dims[1], dims[2]};
        const int64_t nV = int64_t(dims[0]) * dims[1] * dims[2], nNodes = int64_t(dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1);
        in.density.assign(density, density + nV);
        in.bc.assign(bc, bc + 3 * nNodes);
        in.f.assign(f, f + 3 * nNodes);
        in.nu = nu;
        in.E = E;
        in.h = h;
        in.useGPU = false;
        std::vector<ps::BreakStep> steps;
        const ps::BreakResult r = ps::breakTest(in, strength, principal != 0, maxSteps, [&](ps::BreakStep&& b) { steps.push_back(std::move(b)); }, nullptr);
        static const char* names[] = {"step limit", "separated", "unsupported load", "solver did not converge", "no stress"};
        *reason = -1;
        for (int k = 0; k < 5; k++)
            if (r.reason == names[k]) *reason = k;
        *peak = r.peak;
        *removed = r.removed;
        int nc = 0, nd = 0;
        for (size_t s = 0; s < steps.size(); s++) {
            const auto& b = steps[s];
            scalars[3 * s] = b.lambda;
            scalars[3 * s + 1] = b.maxDisp;
            scalars[3 * s + 2] = b.voxels;
            counts[2 * s] = int(b.cracked.size());
            counts[2 * s + 1] = int(b.detached.size());
            for (int e : b.cracked)
                if (nc < crackedCap) cracked[nc++] = e;
            for (int e : b.detached)
                if (nd < detachedCap) detached[nd++] = e;
            if (vm) std::memcpy(vm + s * size_t(nNodes), b.nodeVM.data(), size_t(nNodes) * sizeof(float));
        }
        return int(steps.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
