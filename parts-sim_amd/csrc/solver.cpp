// The C ABI of include/psx.h over the device-resident model (model.hpp).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/psx.h"
#include "device.hpp"
#include "abi.hpp"
#include "packed.hpp"
#include "model.hpp"
#include "psx_core.hpp"


// =================================================================================================
// C ABI

using namespace psx::abi;

extern "C" {

const char* psx_last_error(void) { return g_error.c_str(); }

int psx_available(void) {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0 ? 1 : 0;
}

int psx_device_name(char* buf, size_t len) {
    return guarded([&] {
        if (!buf || len == 0) throw std::invalid_argument("bad buffer");
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
        return PSX_OK;
    });
}

int psx_create(const psx_model_desc* desc, psx_model* out) {
    return guarded([&] {
        if (!desc || !out) throw std::invalid_argument("desc and out are required");
        *out = nullptr;
        psx::checkGridLimits(desc->nx, desc->ny, desc->nz);  // before the device is touched
        auto h = std::make_unique<psx_model_s>();
        h->m = std::make_unique<psx::Model>(*desc);
        *out = h.release();
        return PSX_OK;
    });
}

void psx_destroy(psx_model m) { delete m; }

int psx_num_levels(psx_model m) { return m ? int(m->m->H.levels.size()) : 0; }

int psx_get_level_info(psx_model m, int level, psx_level_info* out) {
    return guarded([&] {
        if (!m || !out || level < 0 || level >= int(m->m->H.levels.size())) throw std::invalid_argument("bad level");
        const psx::Level& L = m->m->H.levels[level];
        out->nx = L.nx; out->ny = L.ny; out->nz = L.nz;
        out->n_nodes = L.nNodes;
        out->n_dof = L.nDof;
        out->free_dof = L.freeDof;
        out->n_elems = L.nElem;
        out->n_own = L.nOwn;
        out->lmax = L.lmax;
        return PSX_OK;
    });
}

int psx_model_level(psx_model m, int level, uint8_t* fixed, double* inv_diag, float* node_scale, int32_t* words, double* kb, float* own,
                    uint16_t* node_info) {
    return guarded([&] {
        if (!m || level < 0 || level >= int(m->m->H.levels.size())) throw std::invalid_argument("bad level");
        psx::Model& M = *m->m;
        const psx::Level& L = M.H.levels[size_t(level)];
        const psx::DevLevel& D = M.lv[size_t(level)].dl;
        auto get = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, M.stream));
        };
        std::vector<double> inv(fixed && !inv_diag ? size_t(L.nDof) : 0);
        double* invOut = inv_diag ? inv_diag : (fixed ? inv.data() : nullptr);
        get(invOut, D.invD64, size_t(L.nDof) * 8);
        get(node_scale, D.nodeScale, size_t(L.nNodes) * 4);
        get(words, D.words, size_t(L.nVox) * 4);
        get(kb, D.Kb64, 576 * 8);
        get(own, D.Kown, size_t(L.nOwn) * 576 * 4);
        get(node_info, D.nodeInfo, size_t(L.nNodes) * 2);
        M.sync();
        if (fixed)
            for (int64_t q = 0; q < L.nDof; q++) fixed[q] = invOut[q] == 0 ? 1 : 0;
        return PSX_OK;
    });
}

int psx_model_level_child(psx_model m, int level, float* child_s, int32_t* child_two, double* diag_add) {
    return guarded([&] {
        if (!m || level < 0 || level >= int(m->m->H.levels.size())) throw std::invalid_argument("bad level");
        psx::Model& M = *m->m;
        const psx::Level& L = M.H.levels[size_t(level)];
        const psx::DevLevel& D = M.lv[size_t(level)].dl;
        if ((child_s || child_two) && !D.childForm) throw std::invalid_argument("the level has no child-scale form");
        if (diag_add && !D.diagAdd) throw std::invalid_argument("the level has no diagonal shift");
        if (child_s) HIPCHK(hipMemcpyAsync(child_s, D.childS, size_t(L.nOwn) * 8 * 4, hipMemcpyDeviceToHost, M.stream));
        if (child_two) HIPCHK(hipMemcpyAsync(child_two, D.childTwo, size_t(L.nOwn) * 4 * 4, hipMemcpyDeviceToHost, M.stream));
        if (diag_add) HIPCHK(hipMemcpyAsync(diag_add, D.diagAdd, size_t(L.nDof) * 8, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        return PSX_OK;
    });
}

int psx_model_coarse(psx_model m, int64_t* mdense, int32_t* free_idx, double* ainv) {
    return guarded([&] {
        if (!m) throw std::invalid_argument("null model");
        psx::Model& M = *m->m;
        if (mdense) *mdense = M.H.m;
        if (free_idx) std::copy(M.H.freeIdx.begin(), M.H.freeIdx.end(), free_idx);
        if (ainv) {
            HIPCHK(hipMemcpyAsync(ainv, M.Ainv, size_t(M.H.m * M.H.m) * 8, hipMemcpyDeviceToHost, M.stream));
            M.sync();
        }
        return PSX_OK;
    });
}

int psx_estimate_smoothers(psx_model m) {
    return guarded([&] {
        if (!m) throw std::invalid_argument("null model");
        m->m->estimateSmoothers();
        return PSX_OK;
    });
}

int psx_set_lmax(psx_model m, int level, double lmax) {
    return guarded([&] {
        if (!m || level < 0 || level >= int(m->m->H.levels.size()) || !(lmax > 0)) throw std::invalid_argument("bad level or lmax");
        m->m->H.levels[level].lmax = lmax;
        m->m->dropGraph();  // a captured iteration batch holds the old Chebyshev coefficients
        return PSX_OK;
    });
}

int psx_apply(psx_model m, int level, const double* x, double* y, int raw) {
    return guarded([&] {
        if (!m || !x || !y || level < 0 || level >= int(m->m->H.levels.size())) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        auto& V = M.lv[level];
        const int64_t n = V.dl.nDof;
        HIPCHK(hipMemcpyAsync(V.a64, x, size_t(n) * 8, hipMemcpyHostToDevice, M.stream));
        M.op<double>(size_t(level), raw ? psx::AM_RAW : psx::AM_APPLY, V.a64, V.b64, nullptr, nullptr, 0, 0, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y, V.b64, size_t(n) * 8, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        return PSX_OK;
    });
}

int psx_apply32(psx_model m, int level, const float* x, float* y, int raw) {
    return guarded([&] {
        if (!m || !x || !y || level < 0 || level >= int(m->m->H.levels.size())) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        auto& V = M.lv[level];
        const int64_t n = V.dl.nDof;
        HIPCHK(hipMemcpyAsync(V.a32, x, size_t(n) * 4, hipMemcpyHostToDevice, M.stream));
        M.op<float>(size_t(level), raw ? psx::AM_RAW : psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y, V.b32, size_t(n) * 4, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        return PSX_OK;
    });
}

int psx_apply_child(psx_model m, int level, int fp32, const void* x, void* y, int raw) {
    return guarded([&] {
        if (!m || !x || !y || level < 0 || level >= int(m->m->H.levels.size())) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        auto& V = M.lv[level];
        const int64_t n = V.dl.nDof;
        const int mode = raw ? psx::AM_RAW : psx::AM_APPLY;
        if (fp32) {
            HIPCHK(hipMemcpyAsync(V.a32, x, size_t(n) * 4, hipMemcpyHostToDevice, M.stream));
            M.op<float>(size_t(level), mode, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr, psx::OF_CHILD);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(y, V.b32, size_t(n) * 4, hipMemcpyDeviceToHost, M.stream));
        } else {
            HIPCHK(hipMemcpyAsync(V.a64, x, size_t(n) * 8, hipMemcpyHostToDevice, M.stream));
            M.op<double>(size_t(level), mode, V.a64, V.b64, nullptr, nullptr, 0, 0, nullptr, psx::OF_CHILD);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(y, V.b64, size_t(n) * 8, hipMemcpyDeviceToHost, M.stream));
        }
        M.sync();
        return PSX_OK;
    });
}

int psx_apply_mfma(psx_model m, int level, int pieces, const float* x, float* y) {
    return guarded([&] {
        if (!m || !x || !y || level < 0 || level >= int(m->m->H.levels.size())) throw std::invalid_argument("bad arguments");
        if (pieces != 2 && pieces != 3) throw std::invalid_argument("pieces must be 2 or 3");
        psx::Model& M = *m->m;
        auto& V = M.lv[level];
        if (M.H.levels[size_t(level)].nOwn > 0 && !V.dl.ownVox)
            throw std::invalid_argument("the level has stored own element matrices: no matrix-core operator");
        const int64_t n = V.dl.nDof;
        HIPCHK(hipMemcpyAsync(V.a32, x, size_t(n) * 4, hipMemcpyHostToDevice, M.stream));
        M.op<float>(size_t(level), psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr, (pieces == 3 ? psx::OF_EL3 : psx::OF_MFMA) | psx::OF_CHILD);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y, V.b32, size_t(n) * 4, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        return PSX_OK;
    });
}

int psx_precondition_device(psx_model m, const double* d_r, double* d_z) {
    return guarded([&] {
        if (!m || !d_r || !d_z) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        if (!M.smoothersReady) M.estimateSmoothers();
        M.precondition32(d_r, d_z, nullptr);
        HIPCHK(hipGetLastError());
        M.sync();
        return PSX_OK;
    });
}

int psx_precondition(psx_model m, const double* r, double* z) {
    return guarded([&] {
        if (!m || !r || !z) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        const int64_t n = M.nDof();
        if (!M.smoothersReady) M.estimateSmoothers();
        HIPCHK(hipMemcpyAsync(M.f, r, size_t(n) * 8, hipMemcpyHostToDevice, M.stream));
        M.precondition32(M.f, M.u, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(z, M.u, size_t(n) * 8, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        return PSX_OK;
    });
}

int psx_precondition_multi_device(psx_model m, int nvec, const double* d_r, double* d_z) {
    return guarded([&] {
        if (!m || !d_r || !d_z || nvec < 1) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        if (!M.smoothersReady) M.estimateSmoothers();
        M.preconditionMulti(nvec, d_r, d_z);
        HIPCHK(hipGetLastError());
        M.sync();
        return PSX_OK;
    });
}

int psx_precondition_multi(psx_model m, int nvec, const double* r, double* z) {
    return guarded([&] {
        if (!m || !r || !z || nvec < 1) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        const int64_t n = M.nDof();
        if (!M.smoothersReady) M.estimateSmoothers();
        M.reserve(nvec);
        const size_t bytes = size_t(n) * size_t(nvec) * 8;
        HIPCHK(hipMemcpyAsync(M.f, r, bytes, hipMemcpyHostToDevice, M.stream));
        M.preconditionMulti(nvec, M.f, M.u);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(z, M.u, bytes, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        return PSX_OK;
    });
}

int psx_solve_device(psx_model m, const double* d_f, const double* d_x0, const psx_solve_opts* opts, double* d_u, psx_solve_info* info) {
    return guarded([&] {
        if (!m || !d_f || !d_u) throw std::invalid_argument("bad arguments");
        const psx_solve_opts o = defaults(opts);
        psx_solve_info tmp{};
        m->m->solveDevice(d_f, d_x0, o, d_u, info ? *info : tmp);
        HIPCHK(hipGetLastError());
        return PSX_OK;
    });
}

int psx_solve(psx_model m, const double* f, const double* x0, const psx_solve_opts* opts, double* u, psx_solve_info* info) {
    return guarded([&] {
        if (!m || !f || !u) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        const int64_t n = M.nDof();
        checkFinite(f, n, "Forces");
        if (x0) checkFinite(x0, n, "Displacements");
        const psx_solve_opts o = defaults(opts);
        HIPCHK(hipMemcpyAsync(M.f, f, size_t(n) * 8, hipMemcpyHostToDevice, M.stream));
        if (x0) HIPCHK(hipMemcpyAsync(M.x, x0, size_t(n) * 8, hipMemcpyHostToDevice, M.stream));
        psx_solve_info tmp{};
        M.solveDevice(M.f, x0 ? M.x : nullptr, o, M.u, info ? *info : tmp);
        HIPCHK(hipMemcpyAsync(u, M.u, size_t(n) * 8, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        return PSX_OK;
    });
}

int psx_solve_multi_device(psx_model m, int nrhs, const double* d_f, const double* d_x0, const psx_solve_opts* opts, double* d_u,
                           psx_solve_info* info) {
    return guarded([&] {
        if (!m || nrhs < 1 || !d_f || !d_u || !info) throw std::invalid_argument("bad arguments");
        const psx_solve_opts o = defaults(opts);
        m->m->solveMulti(nrhs, d_f, d_x0, o, d_u, info);
        HIPCHK(hipGetLastError());
        return PSX_OK;
    });
}

int psx_solve_packed(int n_models, const psx_model* models, const int* nrhs, const double* const* d_f, const double* const* d_x0,
                     const psx_solve_opts* opts, double* const* d_u, psx_solve_info* info) {
    return guarded([&] {
        if (n_models < 1 || !models || !nrhs || !d_f || !d_u || !info) throw std::invalid_argument("bad arguments");
        if (n_models > psx::PACK_MAX) throw std::invalid_argument("a pack holds at most 64 models");
        std::vector<psx::Model*> ms(size_t(n_models), nullptr);
        for (int i = 0; i < n_models; i++) {
            if (!models[i] || nrhs[i] < 1 || !d_f[i] || !d_u[i]) throw std::invalid_argument("bad arguments");
            ms[size_t(i)] = models[i]->m.get();
            for (int j = 0; j < i; j++)
                if (ms[size_t(j)] == ms[size_t(i)]) throw std::invalid_argument("a model can be in a pack once");
        }
        psx_solve_opts o = defaults(opts);
        o.on_progress = nullptr;
        // every model's run settings are reset however this ends
        struct Reset {
            std::vector<psx::Model*>& ms;
            ~Reset() {
                for (psx::Model* m : ms) m->resetRun();
            }
        } reset{ms};
        psx::PackedSolve pk;
        size_t at = 0;
        for (int i = 0; i < n_models; i++) {
            auto mb = std::make_unique<psx::PackedSolve::Member>();
            mb->m = ms[size_t(i)];
            if (mb->m->begin(mb->run, nrhs[i], d_f[i], d_x0 ? d_x0[i] : nullptr, o, d_u[i], info + at)) pk.add(std::move(mb));
            else mb->m->resetRun();
            at += size_t(nrhs[i]);
        }
        const bool graphs = !o.eager && !psx::Model::graphsOff();
        while (pk.size() > 0) pk.step(graphs);
        return PSX_OK;
    });
}

int psx_solve_multi(psx_model m, int nrhs, const double* f, const double* x0, const psx_solve_opts* opts, double* u, psx_solve_info* info) {
    return guarded([&] {
        if (!m || nrhs < 1 || !f || !u || !info) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        const int64_t n = M.nDof();
        const size_t nk = size_t(n) * size_t(nrhs);
        checkFinite(f, int64_t(nk), "Forces");
        if (x0) checkFinite(x0, int64_t(nk), "Displacements");
        const psx_solve_opts o = defaults(opts);
        M.reserve(nrhs);
        HIPCHK(hipMemcpyAsync(M.f, f, nk * 8, hipMemcpyHostToDevice, M.stream));
        if (x0) HIPCHK(hipMemcpyAsync(M.x, x0, nk * 8, hipMemcpyHostToDevice, M.stream));
        M.solveMulti(nrhs, M.f, x0 ? M.x : nullptr, o, M.u, info);
        HIPCHK(hipMemcpyAsync(u, M.u, nk * 8, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        return PSX_OK;
    });
}

int psx_reactions(psx_model m, const double* u, const double* f, double out[3]) {
    return guarded([&] {
        if (!m || !u || !out) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        const psx::Level& L = M.hostLevel0();
        const int64_t n = L.nDof;
        std::vector<double> y(n);
        HIPCHK(hipMemcpyAsync(M.x, u, size_t(n) * 8, hipMemcpyHostToDevice, M.stream));
        M.op<double>(0, psx::AM_RAW, M.x, M.q, nullptr, nullptr, 0, 0, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y.data(), M.q, size_t(n) * 8, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        out[0] = out[1] = out[2] = 0;
        for (int64_t node = 0; node < L.nNodes; node++)
            for (int d = 0; d < 3; d++) {
                const int64_t i = 3 * node + d;
                if (L.bc[i] && L.active[node]) out[d] += y[i] - (f ? f[i] : 0.0);
            }
        return PSX_OK;
    });
}

int psx_stresses_device(psx_model m, const double* d_u, double E, double h, int principals, float* d_vm, float* d_p1, float* d_p3) {
    return guarded([&] {
        if (!m || !d_u || !d_vm || (principals && (!d_p1 || !d_p3))) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        M.stressMatrices(E, h);
        psx::launchStresses(M.lv[0].dl, M.stressS, d_u, principals, d_vm, d_p1, d_p3, M.stream);
        HIPCHK(hipGetLastError());
        M.sync();
        return PSX_OK;
    });
}

int psx_stresses(psx_model m, const double* u, double E, double h, int principals, float* vm, float* p1, float* p3) {
    return guarded([&] {
        if (!m || !u || !vm || (principals && (!p1 || !p3))) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        const psx::DevLevel& D = M.lv[0].dl;
        const int64_t nn = D.nNodes;
        // the level-0 work vectors hold u and the outputs (3 nodal float arrays fit in a64)
        double* du = M.lv[0].b64;
        float* out = reinterpret_cast<float*>(M.lv[0].a64);
        HIPCHK(hipMemcpyAsync(du, u, size_t(D.nDof) * 8, hipMemcpyHostToDevice, M.stream));
        M.stressMatrices(E, h);
        psx::launchStresses(D, M.stressS, du, principals, out, out + nn, out + 2 * nn, M.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(vm, out, size_t(nn) * 4, hipMemcpyDeviceToHost, M.stream));
        if (principals) {
            HIPCHK(hipMemcpyAsync(p1, out + nn, size_t(nn) * 4, hipMemcpyDeviceToHost, M.stream));
            HIPCHK(hipMemcpyAsync(p3, out + 2 * nn, size_t(nn) * 4, hipMemcpyDeviceToHost, M.stream));
        }
        M.sync();
        return PSX_OK;
    });
}

int psx_stress_tensors(psx_model m, const double* u, double E, double h, float* out) {
    return guarded([&] {
        if (!m || !u || !out) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        const psx::DevLevel& D = M.lv[0].dl;
        double* du = M.lv[0].b64;
        // 6 floats per node = 2 doubles per DOF: the a64 and d64 work vectors side by side are not
        // contiguous, so the tensor goes through the CG vector p (nDof doubles = 6 nNodes floats)
        float* dout = reinterpret_cast<float*>(M.p);
        HIPCHK(hipMemcpyAsync(du, u, size_t(D.nDof) * 8, hipMemcpyHostToDevice, M.stream));
        M.stressMatrices(E, h);
        psx::launchStressTensors(D, M.stressS, du, dout, M.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dout, size_t(D.nNodes) * 6 * 4, hipMemcpyDeviceToHost, M.stream));
        M.sync();
        return PSX_OK;
    });
}

int psx_profile(psx_model m, int reps, char* buf, size_t len) {
    return guarded([&] {
        if (!m || !buf || len == 0) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        if (!M.smoothersReady) M.estimateSmoothers();
        reps = std::max(1, reps);
        std::string out;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        auto timeIt = [&](auto&& fn) {
            fn();
            HIPCHK(hipEventRecord(e0, M.stream));
            for (int k = 0; k < reps; k++) fn();
            HIPCHK(hipEventRecord(e1, M.stream));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            return double(t) / reps;
        };
        char line[512];
        for (size_t l = 0; l < M.lv.size(); l++) {
            auto& V = M.lv[l];
            const psx::Level& L = M.H.levels[l];
            const double t64 = timeIt([&] { M.op<double>(l, psx::AM_APPLY, V.a64, V.b64, nullptr, nullptr, 0, 0, nullptr); });
            const double t32 = timeIt([&] { M.op<float>(l, psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr); });
            double tmf = 0, tmf3 = 0, s32 = 0, smf = 0, c64 = 0, c32 = 0;
            if (L.nOwn == 0 || V.dl.ownVox) {
                // (on the child-scale level with its own elements' corner-row pass)
                const int ch = L.nOwn ? psx::OF_CHILD : 0;
                tmf = timeIt([&] { M.op<float>(l, psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr, psx::OF_MFMA | ch); });
                tmf3 = timeIt([&] { M.op<float>(l, psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr, psx::OF_EL3 | ch); });
                // a later Chebyshev step y = x + c1 (x - d) + c2 D^-1 (r - K x), the V-cycle's most frequent
                // launch, in the FP32 VALU form and as the element product on the matrix cores
                s32 = timeIt([&] { M.op<float>(l, psx::AM_CHEB2, V.a32, V.b32, V.r32, V.d32, 0.5, 0.5, nullptr, ch ? psx::OF_AUTO : 0); });
                smf = timeIt([&] { M.op<float>(l, psx::AM_CHEB2, V.a32, V.b32, V.r32, V.d32, 0.5, 0.5, nullptr, psx::OF_MFMA | (ch ? psx::OF_AUTO : 0)); });
            }
            if (V.dl.childForm) {
                c64 = timeIt([&] { M.op<double>(l, psx::AM_APPLY, V.a64, V.b64, nullptr, nullptr, 0, 0, nullptr, psx::OF_CHILD); });
                c32 = timeIt([&] { M.op<float>(l, psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr, psx::OF_CHILD); });
            }
            std::snprintf(line, sizeof line,
                          "L%zu nodes=%lld elems=%zu own matrices=%lld: matvec64 %.4f ms  matvec32 %.4f ms  matvec-mfma2 %.4f ms  matvec-mfma3 %.4f ms  "
                          "smooth32 %.4f ms  smooth-mfma2 %.4f ms  child-form64 %.4f ms  child-form32 %.4f ms\n",
                          l, (long long)L.nNodes, size_t(L.nElem), (long long)L.nOwn, t64, t32, tmf, tmf3, s32, smf, c64, c32);
            out += line;
        }
        const double tv64 = timeIt([&] { M.vcycle<double>(0, M.lv[0].r64, nullptr, psx::OF_AUTO); });
        const double tv32 = timeIt([&] { M.vcycle<float>(0, M.lv[0].r32, nullptr, psx::OF_AUTO); });
        const double tv3 = timeIt([&] { M.vcycle<float>(0, M.lv[0].r32, nullptr, psx::OF_AUTO, 3); });
        const double tvbf = timeIt([&] { M.vcycle<float>(0, M.lv[0].r32, nullptr, psx::OF_AUTO | psx::OF_MFMA); });
        const double tvref = timeIt([&] { M.vcycle<float>(0, M.lv[0].r32, nullptr, 0); });
        std::snprintf(line, sizeof line,
                      "V-cycle FP64 %.4f ms  FP32 %.4f ms  FP32 degree-3 %.4f ms  FP32 with the matrix-core element product %.4f ms  FP32 stored own matrices %.4f ms\n",
                      tv64, tv32, tv3, tvbf, tvref);
        out += line;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::snprintf(buf, len, "%s", out.c_str());
        return PSX_OK;
    });
}

int64_t psx_device_bytes(psx_model m) { return m ? m->m->totalBytes() : 0; }

}  // extern "C"

// ---- host-side inspection ----

struct psx_hierarchy_s {
    psx::Hierarchy H;
};

extern "C" {

int psx_hex_element(double nu, double* K, double* corner_stress) {
    return guarded([&] {
        if (!K) throw std::invalid_argument("K is required");
        const psx::Element el = psx::hexElement(nu);
        std::memcpy(K, el.K, sizeof el.K);
        if (corner_stress) std::memcpy(corner_stress, el.corner, sizeof el.corner);
        return PSX_OK;
    });
}

int psx_prune_floating(int nx, int ny, int nz, const float* density, const uint8_t* held, float* out) {
    int removed = 0;
    const int code = guarded([&] {
        if (nx <= 0 || ny <= 0 || nz <= 0 || !density || !held || !out) throw std::invalid_argument("bad arguments");
        removed = psx::pruneFloating(nx, ny, nz, density, held, out);
        return PSX_OK;
    });
    return code == PSX_OK ? removed : -code;
}

int psx_hierarchy_create(const psx_model_desc* d, psx_hierarchy* out) {
    return guarded([&] {
        if (!d || !out) throw std::invalid_argument("desc and out are required");
        *out = nullptr;
        auto h = std::make_unique<psx_hierarchy_s>();
        psx::buildHierarchy(h->H, d->nx, d->ny, d->nz, d->density, d->nu, d->bc, d->coarsest_max_dof > 0 ? d->coarsest_max_dof : 1100, d->diag_add,
                            d->max_levels);
        *out = h.release();
        return PSX_OK;
    });
}

void psx_hierarchy_destroy(psx_hierarchy h) { delete h; }

int psx_hierarchy_num_levels(psx_hierarchy h) { return h ? int(h->H.levels.size()) : 0; }

int psx_hierarchy_level_info(psx_hierarchy h, int level, psx_level_info* out) {
    return guarded([&] {
        if (!h || !out || level < 0 || level >= int(h->H.levels.size())) throw std::invalid_argument("bad level");
        const psx::Level& L = h->H.levels[level];
        out->nx = L.nx; out->ny = L.ny; out->nz = L.nz;
        out->n_nodes = L.nNodes;
        out->n_dof = L.nDof;
        out->free_dof = L.freeDof;
        out->n_elems = L.nElem;
        out->n_own = L.nOwn;
        out->lmax = L.lmax;
        return PSX_OK;
    });
}

int psx_hierarchy_level(psx_hierarchy h, int level, uint8_t* fixed, double* inv_diag, float* node_scale, int32_t* words, double* kb, float* own) {
    return guarded([&] {
        if (!h || level < 0 || level >= int(h->H.levels.size())) throw std::invalid_argument("bad level");
        const psx::Level& L = h->H.levels[level];
        if (fixed) std::copy(L.fixed.begin(), L.fixed.end(), fixed);
        if (inv_diag) std::copy(L.invDiag.begin(), L.invDiag.end(), inv_diag);
        if (node_scale) std::copy(L.nodeScale.begin(), L.nodeScale.end(), node_scale);
        if (words) std::copy(L.words.begin(), L.words.end(), words);
        if (kb) std::copy(L.Kb, L.Kb + 576, kb);
        if (own) std::copy(L.K.begin(), L.K.end(), own);
        return PSX_OK;
    });
}

int psx_hierarchy_coarse(psx_hierarchy h, int64_t* m, int32_t* free_idx, double* ainv) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null hierarchy");
        if (m) *m = h->H.m;
        if (free_idx) std::copy(h->H.freeIdx.begin(), h->H.freeIdx.end(), free_idx);
        if (ainv) std::copy(h->H.Ainv.begin(), h->H.Ainv.end(), ainv);
        return PSX_OK;
    });
}

}  // extern "C"
