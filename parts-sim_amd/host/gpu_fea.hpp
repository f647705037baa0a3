// Drop-in for the reference's native/src/gpu/gpu_fea.hpp: the same class, constructor, methods and
// constant (gpu_fea.hpp:15-35), implemented over the C ABI of include/psx.h (HIP on MI355X) instead
// of WebGPU. Placed at native/src/gpu/gpu_fea.hpp, it lets static_study.cpp, study_runner.cpp,
// nonlinear / topology drivers, bench.cpp and test_gpu.cpp compile and run unchanged.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../fea/voxel_fea.hpp"

namespace ps {

constexpr int64_t GPU_COARSEST_DOF = 1100;  // same hierarchy as the CPU solver

class GpuFeaSolver {
public:
    /** Throws when the GPU is unavailable or the upload fails. fea must use coarsestMaxDof <= GPU_COARSEST_DOF. */
    explicit GpuFeaSolver(VoxelFEA& fea);
    ~GpuFeaSolver();

    /** One multigrid V-cycle on the GPU (float32): z ~ K^-1 r, like VoxelFEA::precondition. */
    void precondition(const double* r, double* z);
    /** K u = f to the requested tolerance (FP32 CG with FP64 dots and FP64 true-residual refinement on the GPU). */
    SolveResult solve(const std::vector<double>& f, const SolveOptions& opts);
    /** Diagnostics: GPU time of each kernel and of the whole V-cycle, one line each. */
    std::string profile(int reps = 20);


private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
    VoxelFEA& fea_;
};

}  // namespace ps
