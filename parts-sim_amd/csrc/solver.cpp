// Host orchestration of the device solve and the C ABI (include/psx.h).
//
// A model uploads its hierarchy once and keeps every vector in HBM. A conjugate-gradient run
// queues iterations on the model's stream without waiting: every kernel of an iteration returns at
// once when the device-side state says stop (converged, broken, out of iterations), so the host
// reads the state back only every `check_every` iterations and never steers a single iteration.
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
#include "psx_core.hpp"

namespace psx {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NoDevice : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(expr)                                                                                          \
    do {                                                                                                      \
        const hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) throw ::psx::DeviceError(std::string(#expr) + " failed: " + hipGetErrorString(e_));     \
    } while (0)

// Chebyshev smoother on [0.1, 1.15] lmax (voxel_fea.cpp:692-725): first step d = D^-1 r / theta,
// then d = c1[j] d + c2[j] D^-1 (r - K z) for j = 1 .. degree - 1 with the three-term recurrence
// rho_0 = 1 / sigma, rho_j = 1 / (2 sigma - rho_(j-1)), c1 = rho_j rho_(j-1), c2 = 2 rho_j / delta
// (degree 2 is the reference's)
constexpr int CHEB_MAX = 6;
struct Cheb {
    double c0, c1[CHEB_MAX], c2[CHEB_MAX];
};
static Cheb chebyshev(double lmax, int degree) {
    const double b = 1.15 * lmax, a = 0.1 * lmax;
    const double theta = (b + a) / 2, delta = (b - a) / 2, sigma = theta / delta;
    Cheb c{};
    c.c0 = 1 / theta;
    double rho = 1 / sigma;
    for (int j = 1; j < degree; j++) {
        const double next = 1 / (2 * sigma - rho);
        c.c1[j] = next * rho;
        c.c2[j] = 2 * next / delta;
        rho = next;
    }
    return c;
}

class Model {
public:
    Hierarchy H;
    hipStream_t stream = nullptr;
    int device = 0;
    int64_t bytes = 0;

    struct Lvl {
        DevLevel dl{};
        void* aop = nullptr;  // BF16 matrix-core operands of Kb (levels without own matrices)
        double *r64 = nullptr, *a64 = nullptr, *b64 = nullptr, *d64 = nullptr;
        float *r32 = nullptr, *a32 = nullptr, *b32 = nullptr, *d32 = nullptr;
    };
    std::vector<Lvl> lv;
    int32_t* freeIdx = nullptr;
    int32_t* denseIdx = nullptr;  // coarsest DOF -> dense index, -1 when held
    double* Ainv = nullptr;
    // level-0 conjugate-gradient vectors
    double *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *f = nullptr, *u = nullptr;
    float *p32 = nullptr, *q32 = nullptr;
    CgState* st = nullptr;
    CgState* stHost = nullptr;  // pinned mirror
    Reduce red{};
    bool smoothersReady = false;

    explicit Model(const psx_model_desc& d) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw NoDevice("no usable GPU was found");
        HIPCHK(hipGetDevice(&device));
        buildHierarchy(H, d.nx, d.ny, d.nz, d.density, d.nu, d.bc, d.coarsest_max_dof > 0 ? d.coarsest_max_dof : 1100, d.diag_add, d.max_levels);
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        upload();
    }
    ~Model() {
        if (stream) hipStreamSynchronize(stream);
        for (void* ptr : allocs) hipFree(ptr);
        if (stHost) hipHostFree(stHost);
        if (stream) hipStreamDestroy(stream);
    }

    template <class T> T* alloc(int64_t count) {
        void* ptr = nullptr;
        const size_t b = size_t(std::max<int64_t>(count, 1)) * sizeof(T);
        HIPCHK(hipMalloc(&ptr, b));
        HIPCHK(hipMemsetAsync(ptr, 0, b, stream));
        allocs.push_back(ptr);
        bytes += int64_t(b);
        return static_cast<T*>(ptr);
    }
    template <class T> T* upload(const T* src, int64_t count) {
        T* ptr = alloc<T>(count);
        if (count > 0) HIPCHK(hipMemcpyAsync(ptr, src, size_t(count) * sizeof(T), hipMemcpyHostToDevice, stream));
        return ptr;
    }
    void sync() { HIPCHK(hipStreamSynchronize(stream)); }
    int64_t nDof() const { return H.levels[0].nDof; }

    void upload() {
        lv.resize(H.levels.size());
        for (size_t l = 0; l < H.levels.size(); l++) {
            const Level& L = H.levels[l];
            DevLevel& D = lv[l].dl;
            D.nx = L.nx; D.ny = L.ny; D.nz = L.nz; D.NX = L.NX; D.NY = L.NY; D.NZ = L.NZ;
            D.nNodes = L.nNodes;
            D.nDof = L.nDof;
            D.words = upload(L.words.data(), L.nVox);
            D.Kown = upload(L.K.data(), int64_t(L.K.size()));
            float kb32[576];
            for (int u = 0; u < 576; u++) kb32[u] = float(L.Kb[u]);
            D.Kb64 = upload(L.Kb, 576);
            D.Kb32 = upload(kb32, 576);
            double S[243];
            float S32[243];
            stencil(L.Kb, S);
            for (int u = 0; u < 243; u++) S32[u] = float(S[u]);
            D.St64 = upload(S, 243);
            D.St32 = upload(S32, 243);
            D.nodeScale = upload(L.nodeScale.data(), L.nNodes);
            {
                std::vector<uint16_t> info(L.nNodes);
                for (int64_t n = 0; n < L.nNodes; n++) {
                    uint16_t w = L.diffMask[n];
                    for (int d = 0; d < 3; d++)
                        if (L.invDiag[3 * n + d] == 0) w |= uint16_t(1u << (8 + d));
                    info[n] = w;
                }
                D.nodeInfo = upload(info.data(), L.nNodes);
            }
            D.childForm = L.childForm && L.nOwn > 0;
            if (D.childForm) {
                D.childS = upload(L.childS.data(), int64_t(L.childS.size()));
                double rows[8 * 72];
                float rows32[8 * 72];
                for (int c = 0; c < 8; c++)
                    for (int u = 0; u < 72; u++) rows[c * 72 + u] = L.childM[size_t(c) * 576 + u];  // rows 0..2 of M_c
                for (int u = 0; u < 8 * 72; u++) rows32[u] = float(rows[u]);
                D.crow64 = upload(rows, 8 * 72);
                D.crow32 = upload(rows32, 8 * 72);
            }
            std::vector<float> inv32(L.invDiag.begin(), L.invDiag.end());
            D.invD64 = upload(L.invDiag.data(), L.nDof);
            D.invD32 = upload(inv32.data(), L.nDof);
            D.diagAdd = L.diagAdd.empty() ? nullptr : upload(L.diagAdd.data(), L.nDof);
            {
                const double dev = symmetryClasses(L.Kb, D.sym64);
                for (int q = 0; q < 22; q++) D.sym32[q] = float(D.sym64[q]);
                D.symOk = dev < 1e-12;
            }
            if (L.nOwn == 0) {
                std::vector<uint16_t> a(mfmaOperandBytes() / 2);
                mfmaOperands(L.Kb, a.data());
                lv[l].aop = upload(a.data(), int64_t(a.size()));
            }
            lv[l].r64 = alloc<double>(L.nDof);
            lv[l].a64 = alloc<double>(L.nDof);
            lv[l].b64 = alloc<double>(L.nDof);
            lv[l].d64 = alloc<double>(L.nDof);
            lv[l].r32 = alloc<float>(L.nDof);
            lv[l].a32 = alloc<float>(L.nDof);
            lv[l].b32 = alloc<float>(L.nDof);
            lv[l].d32 = alloc<float>(L.nDof);
        }
        freeIdx = upload(H.freeIdx.data(), H.m);
        {
            std::vector<int32_t> dense(H.levels.back().nDof, -1);
            for (int64_t q = 0; q < H.m; q++) dense[H.freeIdx[q]] = int32_t(q);
            denseIdx = upload(dense.data(), int64_t(dense.size()));
        }
        Ainv = upload(H.Ainv.data(), H.m * H.m);
        const int64_t n = nDof();
        x = alloc<double>(n);
        r = alloc<double>(n);
        p = alloc<double>(n);
        q = alloc<double>(n);
        f = alloc<double>(n);
        u = alloc<double>(n);
        p32 = alloc<float>(n);
        q32 = alloc<float>(n);
        st = alloc<CgState>(1);
        red.partials = alloc<double>(4 * (n / RED_TILE + 2));
        red.counter = alloc<unsigned>(1);
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&stHost), sizeof(CgState), hipHostMallocDefault));
        std::memset(stHost, 0, sizeof(CgState));
        sync();
    }

    // ----- operator dispatch: the z-marching kernel on levels without own matrices -----
    bool useFine = [] {
        const char* e = std::getenv("PSX_FINE");
        return !(e && e[0] == '0');
    }();
    // form: 0 = the reference's operator (own matrices as stored FP32 entries); OF_CHILD = own
    // elements in child-scale form where the level has it (exact Galerkin products; the FP64 and
    // mixed solves); OF_MFMA = the BF16 matrix-core operator (FP32 vectors) where the level allows it
    template <class T>
    void op(size_t l, int mode, const T* xx, T* yy, const T* rr, T* dd, double c1, double c2, const int* stop, int form = 0) {
        const DevLevel& D = lv[l].dl;
        if constexpr (sizeof(T) == 4) {
            if ((form & OF_MFMA) && lv[l].aop) {
                launchFineMfma(D, lv[l].aop, mode, xx, yy, rr, dd, c1, c2, stop, stream);
                return;
            }
        }
        if (!useFine) launchApply<T>(D, mode, xx, yy, rr, dd, c1, c2, stop, stream);
        else if (D.symOk && (H.levels[l].nOwn == 0 || ((form & OF_CHILD) && D.childForm))) launchFine<T>(D, mode, xx, yy, rr, dd, c1, c2, stop, stream);
        else launchApplyOwn<T>(D, mode, xx, yy, rr, dd, c1, c2, stop, stream);
    }

    // ----- multigrid -----

    // largest eigenvalue of D^-1 K per level: 10 Lanczos steps on the device from the reference's
    // LCG start vector (voxel_fea.cpp:579-610)
    void estimateSmoothers() {
        for (size_t l = 0; l + 1 < lv.size(); l++) {
            Level& L = H.levels[l];
            if (L.lmax > 0) continue;
            const DevLevel& D = lv[l].dl;
            const int64_t n = L.nDof;
            const auto v0 = lanczosStart(L);
            double *v = lv[l].a64, *vPrev = lv[l].d64, *w = lv[l].b64, *xx = lv[l].r64;
            HIPCHK(hipMemcpyAsync(v, v0.data(), size_t(n) * 8, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemsetAsync(vPrev, 0, size_t(n) * 8, stream));
            for (int j = 0; j < 10; j++) {
                launchLzX(v, D.invD64, xx, n, stream);
                op<double>(l, AM_APPLY, xx, w, nullptr, nullptr, 0, 0, nullptr);
                launchLzW(w, D.invD64, vPrev, v, n, st, j, red, stream);
                launchLzAxpy(w, v, n, st, j, red, stream);
                launchLzNext(w, v, vPrev, n, st, j, stream);
            }
            HIPCHK(hipMemcpyAsync(stHost, st, sizeof(CgState), hipMemcpyDeviceToHost, stream));
            sync();
            std::vector<double> alpha, beta;
            for (int j = 0; j < 10; j++) {
                const double a = stHost->lz[j], b = stHost->lz[16 + j];
                alpha.push_back(a);
                if (!(b > 1e-12 * std::abs(a))) break;
                beta.push_back(b);
            }
            beta.resize(alpha.size() - 1);
            L.lmax = tridiagonalMax(alpha, beta);
            HIPCHK(hipMemsetAsync(v, 0, size_t(n) * 8, stream));
            HIPCHK(hipMemsetAsync(w, 0, size_t(n) * 8, stream));
            HIPCHK(hipMemsetAsync(xx, 0, size_t(n) * 8, stream));
            HIPCHK(hipMemsetAsync(vPrev, 0, size_t(n) * 8, stream));
        }
        smoothersReady = true;
    }

    template <class T> struct Vecs {
        T *r, *a, *b, *d;
    };
    template <class T> Vecs<T> vecs(size_t l) {
        if constexpr (sizeof(T) == 8) return {lv[l].r64, lv[l].a64, lv[l].b64, lv[l].d64};
        else return {lv[l].r32, lv[l].a32, lv[l].b32, lv[l].d32};
    }

    // one V-cycle from rIn on level l (voxel_fea.cpp:817-828); returns the buffer holding z
    template <class T> T* vcycle(size_t l, const T* rIn, const int* stop, int form = 0, int degree = 2) {
        Vecs<T> V = vecs<T>(l);
        const DevLevel& D = lv[l].dl;
        if (l + 1 == lv.size()) {
            launchCoarse<T>(freeIdx, denseIdx, H.m, D.nDof, Ainv, rIn, V.b, stop, stream);
            return V.b;
        }
        const Cheb c = chebyshev(H.levels[l].lmax, degree);
        const DevLevel& C = lv[l + 1].dl;
        T *z = V.a, *t = V.b;
        launchChebFirst<T>(D, rIn, z, V.d, c.c0, stop, stream);
        for (int j = 1; j < degree; j++) {
            op<T>(l, AM_CHEB2, z, t, rIn, V.d, c.c1[j], c.c2[j], stop, form);
            std::swap(z, t);
        }
        op<T>(l, AM_RESID, z, t, rIn, nullptr, 0, 0, stop, form);
        Vecs<T> W = vecs<T>(l + 1);
        launchRestrict<T>(D, C, t, W.r, stop, stream);
        const T* zc = vcycle<T>(l + 1, W.r, stop, form, degree);
        launchProlong<T>(D, C, zc, z, stop, stream);
        op<T>(l, AM_CHEB1, z, t, rIn, V.d, c.c0, 0, stop, form);
        std::swap(z, t);
        for (int j = 1; j < degree; j++) {
            op<T>(l, AM_CHEB2, z, t, rIn, V.d, c.c1[j], c.c2[j], stop, form);
            std::swap(z, t);
        }
        return z;
    }

    // z = M r on level 0 with the FP32 cycle and r scaled to order one (voxel_fea.cpp:835-861)
    void precondition32(const double* dr, double* dz, const int* stop, int degree = 2) {
        const int64_t n = nDof();
        launchMaxAbs64(dr, n, &st->rmax, red, stream);
        launchToF32(dr, lv[0].r32, n, &st->rmax, stream);
        const float* z32 = vcycle<float>(0, lv[0].r32, stop, 0, degree);
        launchFromF32(z32, dz, n, &st->rmax, stream);
    }

    // ----- conjugate gradients -----

    void readState() {
        HIPCHK(hipMemcpyAsync(stHost, st, sizeof(CgState), hipMemcpyDeviceToHost, stream));
        sync();
    }

    void solveDevice(const double* df, const double* dx0, const psx_solve_opts& o, double* du, psx_solve_info& info) {
        const auto t0 = std::chrono::steady_clock::now();
        if (!smoothersReady) estimateSmoothers();
        const int64_t n = nDof();
        const DevLevel& D0 = lv[0].dl;
        info = psx_solve_info{};
        // x = x0 on free DOFs; bnorm = |f_free|; r = f - K x (FP64)
        launchMaskCopy64(dx0, D0.invD64, du, n, stream);
        launchNormFree64(df, D0.invD64, n, &st->bnorm, red, stream);
        op<double>(0, AM_APPLY, du, q, nullptr, nullptr, 0, 0, nullptr);
        launchResidual64(df, q, D0.invD64, r, n, &st->rr, red, stream);
        readState();
        const double bnorm = std::sqrt(stHost->bnorm);
        if (bnorm == 0) {
            HIPCHK(hipMemsetAsync(du, 0, size_t(n) * 8, stream));
            sync();
            info.converged = 1;
            info.device_ms = ms(t0);
            return;
        }
        double rel = std::sqrt(stHost->rr) / bnorm;
        const int prec = o.precision;
        const int batch = o.check_every > 0 ? o.check_every : 8;
        // the reference's algorithm keeps its operator; the FP64 and mixed solves use the exact
        // child-scale form, and the mixed smoother the matrix cores when asked
        const int form = prec == PSX_PREC_REF ? 0 : (OF_CHILD | (prec == PSX_PREC_MIXED && o.mfma_smoother ? OF_MFMA : 0));
        const int degree = o.cheb_degree > 0 ? o.cheb_degree : 2;
        if (degree > CHEB_MAX) throw std::invalid_argument("Chebyshev degree above 6");
        double rmax = 1;
        if (prec == PSX_PREC_MIXED) {
            // FP32 units for the whole solve: the first residual scaled to order one
            launchMaxAbs64(r, n, &st->rmax, red, stream);
            readState();
            rmax = stHost->rmax > 0 ? stHost->rmax : 1;
        }
        CgState s0{};
        s0.rz = 1;
        s0.bnorm = bnorm;
        s0.tol = o.tol;
        s0.maxIter = o.max_iter;
        s0.scale = prec == PSX_PREC_MIXED ? rmax : 1;
        s0.rmax = rmax;
        s0.rel = rel;
        s0.trueRel = rel;
        s0.peak = rel;
        s0.noFold = 1;
        s0.foldEvery = o.fold_every > 0 ? o.fold_every : 20;
        s0.stop = (rel <= o.tol || o.max_iter <= 0) ? 1 : 0;
        HIPCHK(hipMemcpyAsync(st, &s0, sizeof s0, hipMemcpyHostToDevice, stream));
        int* stop = &st->stop;
        float* r32 = lv[0].r32;
        if (prec == PSX_PREC_MIXED) {
            HIPCHK(hipMemsetAsync(p32, 0, size_t(n) * 4, stream));
            HIPCHK(hipMemsetAsync(q32, 0, size_t(n) * 4, stream));
            launchToF32(r, r32, n, &st->rmax, stream);
        } else {
            HIPCHK(hipMemsetAsync(p, 0, size_t(n) * 8, stream));
            HIPCHK(hipMemsetAsync(q, 0, size_t(n) * 8, stream));
        }
        while (!s0.stop) {
            for (int k = 0; k < batch; k++) {
                if (prec == PSX_PREC_MIXED) {
                    const float* z = vcycle<float>(0, r32, stop, form, degree);
                    launchDotsRZ<float>(r32, z, q32, n, st, red, stream);
                    launchUpdateP<float>(z, p32, n, st, stream);
                    op<float>(0, AM_APPLY, p32, q32, nullptr, nullptr, 0, 0, stop);
                    launchDotPQ<float>(p32, q32, n, st, red, stream);
                    launchUpdateXRMixed(du, p32, r, r32, q32, n, st, red, stream);
                    // reliable update: the FP64 true residual replaces the recurrence's
                    op<double>(0, AM_APPLY, du, q, nullptr, nullptr, 0, 0, &st->noFold);
                    launchFoldR(df, q, D0.invD64, r, r32, n, st, red, stream);
                    continue;
                }
                const double* z;
                if (prec == PSX_PREC_FP64) z = vcycle<double>(0, r, stop, form, degree);
                else {
                    precondition32(r, lv[0].b64, stop, degree);
                    z = lv[0].b64;
                }
                launchDotsRZ<double>(r, z, q, n, st, red, stream);
                launchUpdateP<double>(z, p, n, st, stream);
                op<double>(0, AM_APPLY, p, q, nullptr, nullptr, 0, 0, stop);
                launchDotPQ<double>(p, q, n, st, red, stream);
                launchUpdateXR<double>(du, p, r, q, n, st, red, stream);
            }
            readState();
            if (stHost->stop) break;
            if (o.on_progress && o.on_progress(stHost->iters, stHost->rel, o.user)) {
                info.cancelled = 1;
                break;
            }
        }
        readState();
        info.iterations = stHost->iters;
        info.broken = stHost->broken;
        info.folds = stHost->folds;
        // the final true residual against the equilibrium equations (voxel_fea.cpp:920-925)
        op<double>(0, AM_APPLY, du, q, nullptr, nullptr, 0, 0, nullptr);
        launchResidual64(df, q, D0.invD64, r, n, &st->rr, red, stream);
        readState();
        info.residual = std::sqrt(stHost->rr) / bnorm;
        info.converged = !info.cancelled && std::isfinite(info.residual) && info.residual <= o.tol * 10;
        info.device_ms = ms(t0);
    }

    static double ms(std::chrono::steady_clock::time_point t0) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

private:
    std::vector<void*> allocs;
};

}  // namespace psx

// =================================================================================================
// C ABI

struct psx_model_s {
    std::unique_ptr<psx::Model> m;
};

namespace {
thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

template <class F> int guarded(F&& fn) {
    try {
        return fn();
    } catch (const psx::NoDevice& e) {
        return fail(PSX_ERR_NO_DEVICE, e.what());
    } catch (const psx::DeviceError& e) {
        return fail(PSX_ERR_DEVICE, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(PSX_ERR_ARGUMENT, e.what());
    } catch (const std::runtime_error& e) {
        const std::string w = e.what();
        return fail(w.find("direct solver") != std::string::npos ? PSX_ERR_NO_COARSE : PSX_ERR_INTERNAL, w);
    } catch (const std::exception& e) {
        return fail(PSX_ERR_INTERNAL, e.what());
    } catch (...) {
        return fail(PSX_ERR_INTERNAL, "unknown error");
    }
}

psx_solve_opts defaults(const psx_solve_opts* o) {
    psx_solve_opts d{};
    d.tol = 1e-6;
    d.max_iter = 500;
    d.precision = PSX_PREC_FP64;
    if (o) {
        d = *o;
        if (!(d.tol > 0)) d.tol = 1e-6;
        if (d.precision < PSX_PREC_FP64 || d.precision > PSX_PREC_MIXED) throw std::invalid_argument("unknown precision");
    }
    return d;
}

void checkFinite(const double* v, int64_t n, const char* what) {
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(v[i])) throw std::invalid_argument(std::string(what) + " must be finite numbers.");
}
}  // namespace

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
        out->n_elems = int64_t(L.elemVox.size());
        out->n_own = L.nOwn;
        out->lmax = L.lmax;
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

int psx_apply_bf16(psx_model m, int level, const float* x, float* y) {
    return guarded([&] {
        if (!m || !x || !y || level < 0 || level >= int(m->m->H.levels.size())) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        auto& V = M.lv[level];
        if (!V.aop) throw std::invalid_argument("the level has own element matrices: no matrix-core operator");
        const int64_t n = V.dl.nDof;
        HIPCHK(hipMemcpyAsync(V.a32, x, size_t(n) * 4, hipMemcpyHostToDevice, M.stream));
        psx::launchFineMfma(V.dl, V.aop, psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr, M.stream);
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

int psx_reactions(psx_model m, const double* u, const double* f, double out[3]) {
    return guarded([&] {
        if (!m || !u || !out) throw std::invalid_argument("bad arguments");
        psx::Model& M = *m->m;
        const psx::Level& L = M.H.levels[0];
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
        char line[320];
        for (size_t l = 0; l < M.lv.size(); l++) {
            auto& V = M.lv[l];
            const psx::Level& L = M.H.levels[l];
            const double t64 = timeIt([&] { M.op<double>(l, psx::AM_APPLY, V.a64, V.b64, nullptr, nullptr, 0, 0, nullptr); });
            const double t32 = timeIt([&] { M.op<float>(l, psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr); });
            double tbf = 0, c64 = 0, c32 = 0;
            if (V.aop) tbf = timeIt([&] { psx::launchFineMfma(V.dl, V.aop, psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr, M.stream); });
            if (V.dl.childForm) {
                c64 = timeIt([&] { M.op<double>(l, psx::AM_APPLY, V.a64, V.b64, nullptr, nullptr, 0, 0, nullptr, psx::OF_CHILD); });
                c32 = timeIt([&] { M.op<float>(l, psx::AM_APPLY, V.a32, V.b32, nullptr, nullptr, 0, 0, nullptr, psx::OF_CHILD); });
            }
            std::snprintf(line, sizeof line,
                          "L%zu nodes=%lld elems=%zu own matrices=%lld: matvec64 %.4f ms  matvec32 %.4f ms  matvec-bf16 %.4f ms  child-form64 %.4f ms  child-form32 %.4f ms\n",
                          l, (long long)L.nNodes, L.elemVox.size(), (long long)L.nOwn, t64, t32, tbf, c64, c32);
            out += line;
        }
        const double tv64 = timeIt([&] { M.vcycle<double>(0, M.lv[0].r64, nullptr, psx::OF_CHILD); });
        const double tv32 = timeIt([&] { M.vcycle<float>(0, M.lv[0].r32, nullptr, psx::OF_CHILD); });
        const double tv3 = timeIt([&] { M.vcycle<float>(0, M.lv[0].r32, nullptr, psx::OF_CHILD, 3); });
        const double tvbf = timeIt([&] { M.vcycle<float>(0, M.lv[0].r32, nullptr, psx::OF_CHILD | psx::OF_MFMA); });
        const double tvref = timeIt([&] { M.vcycle<float>(0, M.lv[0].r32, nullptr, 0); });
        std::snprintf(line, sizeof line,
                      "V-cycle FP64 %.4f ms  FP32 %.4f ms  FP32 degree-3 %.4f ms  FP32+BF16 matrix cores %.4f ms  FP32 stored own matrices %.4f ms\n",
                      tv64, tv32, tv3, tvbf, tvref);
        out += line;
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        std::snprintf(buf, len, "%s", out.c_str());
        return PSX_OK;
    });
}

int64_t psx_device_bytes(psx_model m) { return m ? m->m->bytes : 0; }

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
        out->n_elems = int64_t(L.elemVox.size());
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
