// ps::GpuFeaSolver over the psx C ABI (see gpu_fea.hpp). The model the reference's VoxelFEA built on
// the CPU is rebuilt on the device side from the same inputs (finest-level densities, held DOFs,
// diagonal shift, Poisson's ratio) with the same number of levels, so both hierarchies agree; the
// smoother eigenvalue estimates flow both ways like gpu_fea.cpp:410-421 (the CPU model's values
// are used when it has them, and the GPU's are written back when it has none).
#include "gpu_fea.hpp"

#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "psx.h"

namespace ps {

namespace {
[[noreturn]] void raise(int code) {
    const std::string msg = psx_last_error();
    if (code == PSX_ERR_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
void check(int code) {
    if (code != PSX_OK) raise(code);
}
// PSX_PRECISION=fp64|ref|mixed selects the arithmetic of solve(); mixed by default (the reference's
// GPU solver is FP32 with FP64 reliable updates, gpu_fea.cpp:452-457)
int precisionFromEnv() {
    const char* p = std::getenv("PSX_PRECISION");
    if (!p) return PSX_PREC_MIXED;
    if (!std::strcmp(p, "fp64")) return PSX_PREC_FP64;
    if (!std::strcmp(p, "ref")) return PSX_PREC_REF;
    return PSX_PREC_MIXED;
}
}  // namespace

struct GpuFeaSolver::Impl {
    psx_model m = nullptr;
    ~Impl() { psx_destroy(m); }
};

GpuFeaSolver::GpuFeaSolver(VoxelFEA& fea) : fea_(fea) {
    if (!psx_available()) throw std::runtime_error("no usable GPU adapter was found");
    if (fea.coarse.A.empty() || fea.coarse.m == 0) throw std::runtime_error("the coarsest level has no direct solver");
    const Level& L0 = fea.levels[0];
    std::vector<float> density(size_t(L0.nx) * L0.ny * L0.nz, 0.f);
    for (size_t q = 0; q < L0.elems.size(); q++) density[L0.elems[q]] = float(L0.rho[q]);
    psx_model_desc d{};
    d.nx = L0.nx;
    d.ny = L0.ny;
    d.nz = L0.nz;
    d.density = density.data();
    d.nu = fea.nu;
    d.bc = L0.bc.data();
    d.diag_add = L0.diagAdd.empty() ? nullptr : L0.diagAdd.data();
    // stop at the same level as the CPU model: its coarsest free-DOF count is below every finer one
    d.coarsest_max_dof = fea.levels.back().freeDof;
    d.max_levels = int(fea.levels.size());
    // the CPU model is built already; the device builds its own copy of the levels (the same data
    // bit for bit) instead of a second host build
    d.build_on_device = 1;
    impl_ = std::make_unique<Impl>();
    check(psx_create(&d, &impl_->m));
    if (psx_num_levels(impl_->m) != int(fea.levels.size())) throw std::runtime_error("GPU hierarchy does not match the CPU model");
    for (size_t l = 0; l + 1 < fea.levels.size(); l++)
        if (fea.levels[l].lmax > 0) check(psx_set_lmax(impl_->m, int(l), fea.levels[l].lmax));
    check(psx_estimate_smoothers(impl_->m));
    for (size_t l = 0; l + 1 < fea.levels.size(); l++) {
        Level& lv = fea.levels[l];
        if (lv.lmax > 0) continue;
        psx_level_info li{};
        check(psx_get_level_info(impl_->m, int(l), &li));
        lv.lmax = li.lmax;
        lv.omega = 1.2 / (lv.lmax * 1.05);
    }
}

GpuFeaSolver::~GpuFeaSolver() = default;

void GpuFeaSolver::precondition(const double* r, double* z) { check(psx_precondition(impl_->m, r, z)); }

SolveResult GpuFeaSolver::solve(const std::vector<double>& f, const SolveOptions& opts) {
    const int64_t n = fea_.nDof();
    if (int64_t(f.size()) != n || (opts.x0 && int64_t(opts.x0->size()) != n)) throw std::invalid_argument("Force or displacement vector has the wrong size.");
    SolveResult res;
    res.u.assign(size_t(n), 0.0);
    psx_solve_opts o{};
    o.tol = opts.tol;
    o.max_iter = opts.maxIter;
    o.precision = precisionFromEnv();
    o.check_every = 8;
    struct Bridge {
        const std::function<bool(int, double)>* fn;
    } bridge{&opts.onProgress};
    if (opts.onProgress) {
        o.on_progress = [](int it, double rel, void* user) -> int { return (*static_cast<Bridge*>(user)->fn)(it, rel) ? 1 : 0; };
        o.user = &bridge;
    }
    psx_solve_info info{};
    check(psx_solve(impl_->m, f.data(), opts.x0 ? opts.x0->data() : nullptr, &o, res.u.data(), &info));
    res.iterations = info.iterations;
    res.residual = info.residual;
    res.converged = info.converged != 0;
    res.cancelled = info.cancelled != 0;
    return res;
}

std::string GpuFeaSolver::profile(int reps) {
    std::string buf(1 << 16, '\0');
    check(psx_profile(impl_->m, reps, buf.data(), buf.size()));
    buf.resize(std::strlen(buf.c_str()));
    return buf;
}

}  // namespace ps
