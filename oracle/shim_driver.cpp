// Test driver of the drop-in boundary — test infrastructure only, never the product.
//
// Drives ps::GpuFeaSolver through the C++ shim (parts-sim_amd/host/gpu_fea.{hpp,cpp} over libpsx.so)
// the way the reference's callers do: a VoxelFEA built on the CPU with GPU_COARSEST_DOF, the solver
// constructed from it, then solve() and precondition() (static_study.cpp:96-101,
// study_runner.cpp:56-88). oracle/Makefile compiles it with the reference's own VoxelFEA sources and
// places the shim's gpu_fea.hpp first on the include path, where the reference's native/src/gpu/
// header would be. Checks, on native/tests/test_gpu.cpp's two cantilevers (40x4x4 and 100x10x10):
//   solve() in every precision against VoxelFEA::solve: max |du| / max |u| < 1e-6 (test_gpu.cpp:51);
//   precondition() against VoxelFEA::precondition of the same model: relative L2 < 1e-4 (FP32 V-cycle);
//   warm start, cancellation through onProgress, the size-mismatch invalid_argument, profile();
// the constructor's runtime_error for a model without a coarsest direct solve (gpu_fea.cpp:254-255);
// and the reference's own callers linked against the shim: solveStatic and breakTest
// (static_study.cpp:134-279) with useGPU = true take the GPU engine through it and agree with their CPU
// engine (both solve to 1e-6, static_study.cpp:82).
// Prints one line per check and "SHIM OK" at the end; the exit status is 0 only when all pass.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "fea/static_study.hpp"
#include "fea/voxel_fea.hpp"
#include "gpu/gpu.hpp"
#include "gpu/gpu_fea.hpp"

using namespace ps;

static int g_failures = 0;
#define EXPECT(cond)                                                                  \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::printf("FAIL line %d: %s\n", __LINE__, #cond);                       \
            g_failures++;                                                             \
        }                                                                             \
    } while (0)

struct Part {
    std::array<int, 3> dims;
    std::vector<float> density;
    std::vector<uint8_t> bc;
    std::vector<double> f;
};

// test_gpu.cpp:20-34: 0.1 m x b x b steel beam, n voxels through the thickness, clamped at x = 0,
// 100 N along -y spread over the free end's nodes
static Part cantilever(int n) {
    const double b = 0.01, Lm = 0.1, h = b / n;
    const int nx = int(std::lround(Lm / h)), ny = n, nz = n, NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    Part p;
    p.dims = {nx, ny, nz};
    p.density.assign(size_t(nx) * ny * nz, 1.f);
    p.bc.assign(3 * size_t(NX) * NY * NZ, 0);
    p.f.assign(p.bc.size(), 0.0);
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++) {
            const size_t n0 = size_t(NX) * (j + size_t(NY) * k);
            p.bc[3 * n0] = p.bc[3 * n0 + 1] = p.bc[3 * n0 + 2] = 1;
            p.f[3 * (nx + n0) + 1] = -100.0 / (NY * NZ);
        }
    return p;
}

static double maxRelDiff(const std::vector<double>& a, const std::vector<double>& b) {
    double d = 0, m = 0;
    for (size_t i = 0; i < a.size(); i++) {
        d = std::max(d, std::abs(a[i] - b[i]));
        m = std::max(m, std::abs(a[i]));
    }
    return m > 0 ? d / m : d;
}

static double l2Rel(const std::vector<double>& a, const std::vector<double>& b) {
    double d = 0, m = 0;
    for (size_t i = 0; i < a.size(); i++) {
        d += (a[i] - b[i]) * (a[i] - b[i]);
        m += b[i] * b[i];
    }
    return std::sqrt(d / std::max(m, 1e-300));
}

int main() {
    if (!gpuAvailable()) {
        std::printf("no GPU: %s\n", gpuUnavailableReason().c_str());
        return 2;
    }
    std::printf("device: %s\n", gpuName().c_str());
    EXPECT(!gpuShaderF16());
    EXPECT(std::string(wgslSource("fea")).empty());
    for (int n : {4, 10}) {
        const Part p = cantilever(n);
        VoxelFEA cpu(p.dims, p.density, 0.3, p.bc);
        SolveOptions o;
        o.tol = 1e-8;
        const SolveResult a = cpu.solve(p.f, o);
        VoxelFEA forGpu(p.dims, p.density, 0.3, p.bc, GPU_COARSEST_DOF);
        GpuFeaSolver g(forGpu);
        o.maxIter = 3000;
        for (const char* prec : {"mixed", "fp64", "ref"}) {
            setenv("PSX_PRECISION", prec, 1);
            const SolveResult r = g.solve(p.f, o);
            const double d = maxRelDiff(a.u, r.u);
            std::printf("solve %dx%dx%d %s: GPU %d its (converged %d, residual %.2e), CPU %d its, max |du| / max |u| = %.2e\n", p.dims[0], p.dims[1],
                        p.dims[2], prec, r.iterations, int(r.converged), r.residual, a.iterations, d);
            EXPECT(r.converged);
            EXPECT(d < 1e-6);
        }
        unsetenv("PSX_PRECISION");
        // one V-cycle: the shim's against the same VoxelFEA's (its smoother estimates came from the GPU)
        std::vector<double> r(size_t(forGpu.nDof())), zg(r.size()), zc(r.size());
        uint64_t s = 12345;
        for (double& v : r) {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            v = double(int64_t(s >> 11) % 2000001 - 1000000) * 1e-6;
        }
        g.precondition(r.data(), zg.data());
        forGpu.precondition(r.data(), zc.data());
        const double e = l2Rel(zg, zc);
        std::printf("precondition %dx%dx%d: relative L2 difference %.2e\n", p.dims[0], p.dims[1], p.dims[2], e);
        EXPECT(e < 1e-4);
        // warm start from the CPU solution: at most a few iterations
        SolveOptions w = o;
        w.x0 = &a.u;
        const SolveResult rw = g.solve(p.f, w);
        std::printf("warm start: %d its, converged %d\n", rw.iterations, int(rw.converged));
        EXPECT(rw.converged && rw.iterations <= 5);
        // cancellation: onProgress returning true stops the solve (study_runner.cpp:81-88 checks it)
        SolveOptions c = o;
        c.onProgress = [](int, double) { return true; };
        const SolveResult rc = g.solve(p.f, c);
        std::printf("cancelled: %d, converged %d\n", int(rc.cancelled), int(rc.converged));
        EXPECT(rc.cancelled && !rc.converged);
        // a load vector of the wrong size is an invalid_argument (gpu_fea.cpp:462)
        bool threw = false;
        try {
            (void)g.solve(std::vector<double>(5, 0.0), o);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        EXPECT(threw);
        const std::string prof = g.profile(2);
        EXPECT(prof.find("L0") != std::string::npos);
    }
    // a single level of 36 k free DOFs has no coarsest direct solve: the constructor throws
    // runtime_error (gpu_fea.cpp:254-255), as the callers expect before they fall back to the CPU
    {
        const Part p = cantilever(10);
        VoxelFEA one(p.dims, p.density, 0.3, p.bc, int64_t(1) << 40);
        bool threw = false;
        try {
            GpuFeaSolver g(one);
        } catch (const std::runtime_error& e) {
            threw = true;
            std::printf("no coarse solve: %s\n", e.what());
        }
        EXPECT(threw && one.levels.size() == 1);
    }
    // the reference's bend and break studies with the GPU engine (the shim) against the CPU engine
    {
        const Part p = cantilever(6);
        StructuralInput in;
        in.dims = p.dims;
        in.density = p.density;
        in.bc = p.bc;
        in.f = p.f;
        in.nu = 0.3;
        in.E = 200e9;
        in.h = 0.01 / 6;
        in.useGPU = true;
        const StaticResult gs = solveStatic(in, nullptr);
        dropKeptModel();
        in.useGPU = false;
        const StaticResult cs = solveStatic(in, nullptr);
        dropKeptModel();
        const std::vector<double> ug(gs.u.begin(), gs.u.end()), uc(cs.u.begin(), cs.u.end());
        const std::vector<double> vg(gs.nodeVM.begin(), gs.nodeVM.end()), vc(cs.nodeVM.begin(), cs.nodeVM.end());
        const double du = maxRelDiff(uc, ug), dv = maxRelDiff(vc, vg);
        std::printf("solveStatic: engine %s (%s) %d its, CPU engine %s %d its, max |du| / max |u| = %.2e, von Mises %.2e\n", gs.engine.c_str(),
                    gs.gpuNote.c_str(), gs.iterations, cs.engine.c_str(), cs.iterations, du, dv);
        EXPECT(gs.engine == "GPU" && cs.engine == "CPU" && gs.converged);
        EXPECT(du < 1e-4 && dv < 1e-4);
        std::vector<BreakStep> bg, bcpu;
        in.useGPU = true;
        const BreakResult rg = breakTest(in, 150e6, false, 3, [&](BreakStep&& b) { bg.push_back(std::move(b)); }, nullptr);
        in.useGPU = false;
        const BreakResult rc = breakTest(in, 150e6, false, 3, [&](BreakStep&& b) { bcpu.push_back(std::move(b)); }, nullptr);
        std::printf("breakTest: engine %s %zu steps (%s), CPU engine %zu steps (%s)\n", rg.engine.c_str(), bg.size(), rg.reason.c_str(), bcpu.size(),
                    rc.reason.c_str());
        EXPECT(rg.engine == "GPU" && bg.size() == bcpu.size() && !bg.empty());
        for (size_t k = 0; k < std::min(bg.size(), bcpu.size()); k++) {
            const double dl = std::abs(bg[k].lambda / bcpu[k].lambda - 1);
            std::printf("  step %zu: lambda %.6e vs %.6e (%.1e), cracked %zu vs %zu\n", k, bg[k].lambda, bcpu[k].lambda, dl, bg[k].cracked.size(),
                        bcpu[k].cracked.size());
            EXPECT(dl < 1e-4 && bg[k].voxels == bcpu[k].voxels);
        }
    }
    std::printf(g_failures ? "SHIM FAIL (%d)\n" : "SHIM OK\n", g_failures);
    return g_failures ? 1 : 0;
}
