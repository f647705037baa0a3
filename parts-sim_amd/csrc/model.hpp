// The device-resident model: one part's hierarchy uploaded to HBM, its work vectors, and the
// host orchestration of operator applications, V-cycles and conjugate-gradient runs on the model's
// own HIP stream. A conjugate-gradient run queues iterations without waiting: every kernel of an
// iteration returns at once when the device-side state says stop (converged, broken, out of
// iterations), so the host reads the state back only every `check_every` iterations.
#pragma once

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
inline Cheb chebyshev(double lmax, int degree) {
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

class Model;
// builds the hierarchy of d with the device kernels of build.hip on m's stream (sets m.H's level
// sizes, uniform elements and coarsest map, and every device buffer of m.lv)
void buildOnDevice(Model& m, const psx_model_desc& d);
// the Lanczos start vector of a level on the device: the LCG sequence of lanczosStart with held
// DOFs 0, normalised (tmp: one device double)
void launchLanczosStart(const double* invD, int64_t n, double* v, double* tmp, Reduce red, hipStream_t s);

class Model {
public:
    Hierarchy H;
    hipStream_t stream = nullptr;
    int device = 0;
    int64_t bytes = 0;

    struct Lvl {
        DevLevel dl{};
        bool childCheap = false;  // child-scale form preferred (OF_AUTO): most own elements two-valued
        double *r64 = nullptr, *a64 = nullptr, *b64 = nullptr, *d64 = nullptr;
        float *r32 = nullptr, *a32 = nullptr, *b32 = nullptr, *d32 = nullptr;
    };
    std::vector<Lvl> lv;
    int32_t* freeIdx = nullptr;
    int32_t* denseIdx = nullptr;  // coarsest DOF -> dense index, -1 when held
    double* Ainv = nullptr;
    // level-0 conjugate-gradient vectors
    double *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *f = nullptr, *u = nullptr;
    float *p32 = nullptr, *q32 = nullptr, *x32 = nullptr;
    // the second buffer of the search direction: an iteration whose direction update is fused into
    // the outer operator writes the new direction beside the old one
    double* pAlt = nullptr;
    float* p32Alt = nullptr;
    double* stressS = nullptr;  // corner stress matrices scaled by E / h (8 x 144)
    CgState* st = nullptr;      // one per right-hand side
    CgState* stHost = nullptr;  // pinned mirror
    Reduce red{};
    bool smoothersReady = false;
    // right-hand sides the work vectors hold (instance-major, each level's nDof per instance) and
    // the number the current launches run
    int cap = 0;
    int ni = 1;
    int preDegree = 0;  // Chebyshev degree of the V-cycle's pre-smoothing (0 = the post-smoothing's)
    // the last captured batch of conjugate-gradient iterations and what it was captured for
    struct GraphKey {
        int prec = -1, degree = 0, pre = 0, form = 0, k = 0, batch = 0;
        const double* df = nullptr;
        double* du = nullptr;
        bool operator==(const GraphKey& o) const {
            return prec == o.prec && degree == o.degree && pre == o.pre && form == o.form && k == o.k && batch == o.batch && df == o.df && du == o.du;
        }
    };
    hipGraphExec_t graph = nullptr;
    GraphKey graphKey{};
    void dropGraph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graphKey = GraphKey{};
    }

    // hierarchy built by the device kernels (build.hip): H holds the level sizes and uniform
    // elements only; the level-0 held DOFs stay on the device (bc0) until a host-side caller needs
    // them (hostLevel0)
    bool deviceBuilt = false;
    uint8_t* bc0 = nullptr;

    explicit Model(const psx_model_desc& d) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw NoDevice("no usable GPU was found");
        HIPCHK(hipGetDevice(&device));
        if (d.build_on_device) {
            HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            try {
                buildOnDevice(*this, d);
                halfOwn();
            } catch (...) {
                release();
                throw;
            }
            return;
        }
        buildHierarchy(H, d.nx, d.ny, d.nz, d.density, d.nu, d.bc, d.coarsest_max_dof > 0 ? d.coarsest_max_dof : 1100, d.diag_add, d.max_levels);
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        upload();
        halfOwn();
    }
    // FP16 copies of the stored own matrices (the mixed V-cycle's smoothing reads them, OF_AUTO)
    void halfOwn() {
        for (size_t l = 0; l < lv.size(); l++) {
            DevLevel& D = lv[l].dl;
            const int64_t nOwn = H.levels[l].nOwn;
            if (nOwn <= 0 || !D.Kown) continue;
            uint16_t* kh = alloc<uint16_t>(nOwn * 576);
            float* sc = alloc<float>(nOwn);
            launchOwnToHalf(D.Kown, nOwn, kh, sc, stream);
            HIPCHK(hipGetLastError());
            D.KownH = kh;
            D.ownScale = sc;
        }
    }
    // Device buffers come from the stream-ordered allocator: plain hipFree waits for the whole
    // device, which would stall every other model's stream (concurrent serving requests).
    ~Model() { release(); }
    void release() {
        dropGraph();
        if (stream) {
            for (void* ptr : allocs) (void)hipFreeAsync(ptr, stream);
            (void)hipStreamSynchronize(stream);
        }
        if (stream) {
            for (void* ptr : work) (void)hipFreeAsync(ptr, stream);
            (void)hipStreamSynchronize(stream);
        }
        if (stHost) (void)hipHostFree(stHost);
        if (stream) (void)hipStreamDestroy(stream);
        allocs.clear();
        work.clear();
        stHost = nullptr;
        stream = nullptr;
    }

    template <class T> T* alloc(int64_t count, bool perRhs = false) {
        void* ptr = nullptr;
        const size_t b = size_t(std::max<int64_t>(count, 1)) * sizeof(T);
        HIPCHK(hipMallocAsync(&ptr, b, stream));
        HIPCHK(hipMemsetAsync(ptr, 0, b, stream));
        (perRhs ? work : allocs).push_back(ptr);
        (perRhs ? workBytes : bytes) += int64_t(b);
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
            uploadConstants(l);
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
            {
                // per-voxel form where most nodes have differing voxels (PSX_DENSE=0 / 1 forces it
                // off / on for the levels without own elements)
                int64_t irregular = 0;
                for (int64_t nn = 0; nn < L.nNodes; nn++) irregular += L.diffMask[nn] != 0 ? 1 : 0;
                const char* e = std::getenv("PSX_DENSE");
                D.dense = L.nOwn == 0 && (e && e[0] ? e[0] == '1' : 2 * irregular > L.nNodes);
            }
            if (D.childForm) {
                int64_t two = 0;
                for (int64_t e = 0; e < L.nOwn; e++) two += L.childTwo[size_t(e) * 4 + 2] >= 0 ? 1 : 0;
                lv[l].childCheap = 2 * two >= L.nOwn;
                D.childS = upload(L.childS.data(), int64_t(L.childS.size()));
                D.childTwo = upload(L.childTwo.data(), int64_t(L.childTwo.size()));
                uploadChildConstants(l);
            }
            std::vector<float> inv32(L.invDiag.begin(), L.invDiag.end());
            D.invD64 = upload(L.invDiag.data(), L.nDof);
            D.invD32 = upload(inv32.data(), L.nDof);
            D.diagAdd = L.diagAdd.empty() ? nullptr : upload(L.diagAdd.data(), L.nDof);
        }
        freeIdx = upload(H.freeIdx.data(), H.m);
        {
            std::vector<int32_t> dense(H.levels.back().nDof, -1);
            for (int64_t q = 0; q < H.m; q++) dense[H.freeIdx[q]] = int32_t(q);
            denseIdx = upload(dense.data(), int64_t(dense.size()));
        }
        Ainv = upload(H.Ainv.data(), H.m * H.m);
        finishSetup();
    }
    void finishSetup() {
        stressS = alloc<double>(8 * 144);
        red.counter = alloc<unsigned>(1);
        // the child-scale level's own-element voxels (the matrix-core operator's corner-row pass)
        for (size_t l = 0; l < lv.size(); l++) {
            DevLevel& D = lv[l].dl;
            if (!D.childForm || H.levels[l].nOwn == 0) continue;
            int32_t* ov = alloc<int32_t>(H.levels[l].nOwn);
            launchOwnVox(D.words, int64_t(D.nx) * D.ny * D.nz, ov, stream);
            D.ownVox = ov;
            D.ownTStride = 24 * H.levels[l].nOwn;
        }
        reserve(1);
    }
    // the level's uniform element, its stencil and symmetry classes (from H.levels[l].Kb)
    void uploadConstants(size_t l) {
        const Level& L = H.levels[l];
        DevLevel& D = lv[l].dl;
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
        const double dev = symmetryClasses(L.Kb, D.sym64);
        for (int q = 0; q < 22; q++) D.sym32[q] = float(D.sym64[q]);
        D.symOk = dev < 1e-12;
        const auto A = elementOperands(L.Kb);
        D.elA = reinterpret_cast<const uint4*>(upload(A.data(), int64_t(A.size())));
    }
    // Kb as the A operands of the element-by-element matrix-core product (fine_el.hip): per half h
    // (output corners at z offset h), piece p and lane l, 8 BF16: row l % 16 = 4 component + corner
    // (x + 2 y of the output corner; component 3 zero), K slice 8 (l / 16) .. + 7 = input corners
    // (0, g & 1, g >> 1) and (1, g & 1, g >> 1), g = l / 16, with components 0..3 (the fourth zero).
    // The pieces split each entry v in turn, hi = bf16(v), mid = bf16(v - hi), lo = bf16(v - hi - mid),
    // rounded to nearest.
    static std::vector<uint32_t> elementOperands(const double* Kb) {
        auto bf = [](double v) -> uint32_t {
            const float f = float(v);
            uint32_t u;
            std::memcpy(&u, &f, 4);
            return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
        };
        auto val = [](uint32_t b) {
            const uint32_t u = b << 16;
            float f;
            std::memcpy(&f, &u, 4);
            return double(f);
        };
        auto corner = [](int x, int y, int z) { return (y ? (x ? 2 : 3) : (x ? 1 : 0)) + 4 * z; };
        std::vector<uint32_t> out(size_t(2) * 3 * 64 * 4);
        for (int h = 0; h < 2; h++)
            for (int l = 0; l < 64; l++) {
                const int m = l & 15, kg = l >> 4, d = m >> 2, r = m & 3, ao = corner(r & 1, r >> 1, h);
                uint32_t pc[3][8];
                for (int t = 0; t < 8; t++) {
                    const int c = t & 3, ai = corner(t >> 2, kg & 1, kg >> 1);
                    double v = (d < 3 && c < 3) ? Kb[(3 * ao + d) * 24 + 3 * ai + c] : 0.0;
                    for (int p = 0; p < 3; p++) {
                        pc[p][t] = bf(v);
                        v -= val(pc[p][t]);
                    }
                }
                for (int p = 0; p < 3; p++)
                    for (int w = 0; w < 4; w++) out[((size_t(h) * 3 + p) * 64 + l) * 4 + w] = pc[p][2 * w] | (pc[p][2 * w + 1] << 16);
            }
        return out;
    }
    // the 8 child products M_c as A operands of the own-element matrix-core pass (fine_el.hip), per
    // child c, half h (output corners 4 h .. 4 h + 3, numbered x + 2 y + 4 z), piece p and lane l:
    // row l % 16 = 4 (output corner - 4 h) + component (component 3 zero), K slice 8 (l / 16) .. + 7 =
    // input corners 2 g, 2 g + 1 (g = l / 16) with components 0..3 (the fourth zero); two BF16 pieces
    // hi = bf16(v), lo = bf16(v - hi), rounded to nearest
    static std::vector<uint32_t> childOperands(const double* M) {
        auto bf = [](double v) -> uint32_t {
            const float f = float(v);
            uint32_t u;
            std::memcpy(&u, &f, 4);
            return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
        };
        auto val = [](uint32_t b) {
            const uint32_t u = b << 16;
            float f;
            std::memcpy(&f, &u, 4);
            return double(f);
        };
        auto corner = [](int q) { const int x = q & 1, y = (q >> 1) & 1, z = q >> 2; return (y ? (x ? 2 : 3) : (x ? 1 : 0)) + 4 * z; };
        std::vector<uint32_t> out(size_t(8) * 2 * 2 * 64 * 4);
        for (int c = 0; c < 8; c++)
            for (int h = 0; h < 2; h++)
                for (int l = 0; l < 64; l++) {
                    const int m = l & 15, kg = l >> 4, d = m & 3, ao = corner(4 * h + (m >> 2));
                    uint32_t pc[2][8];
                    for (int t = 0; t < 8; t++) {
                        const int ci = t & 3, ai = corner(2 * kg + (t >> 2));
                        double v = (d < 3 && ci < 3) ? M[size_t(c) * 576 + size_t(3 * ao + d) * 24 + 3 * ai + ci] : 0.0;
                        for (int p = 0; p < 2; p++) {
                            pc[p][t] = bf(v);
                            v -= val(pc[p][t]);
                        }
                    }
                    for (int p = 0; p < 2; p++)
                        for (int w = 0; w < 4; w++) out[(((size_t(c) * 2 + h) * 2 + p) * 64 + l) * 4 + w] = pc[p][2 * w] | (pc[p][2 * w + 1] << 16);
                }
        return out;
    }
    // child-scale form: corner-0 rows of the 8 Galerkin products and the subset table (from
    // H.levels[l].childM / childTab)
    void uploadChildConstants(size_t l) {
        const Level& L = H.levels[l];
        DevLevel& D = lv[l].dl;
        double rows[8 * 72];
        float rows32[8 * 72];
        for (int c = 0; c < 8; c++)
            for (int u = 0; u < 72; u++) rows[c * 72 + u] = L.childM[size_t(c) * 576 + u];  // rows 0..2 of M_c
        for (int u = 0; u < 8 * 72; u++) rows32[u] = float(rows[u]);
        const auto CA = childOperands(L.childM.data());
        D.childA = reinterpret_cast<const uint4*>(upload(CA.data(), int64_t(CA.size())));
        D.crow64 = upload(rows, 8 * 72);
        D.crow32 = upload(rows32, 8 * 72);
        std::vector<float> tab32(L.childTab.begin(), L.childTab.end());
        D.ctab64 = upload(L.childTab.data(), int64_t(L.childTab.size()));
        D.ctab32 = upload(tab32.data(), int64_t(tab32.size()));
        // FP16 copy, each row of 64 with its own power-of-two scale (the rule of the FP16 own matrices)
        const size_t nRows = L.childTab.size() / 64;
        std::vector<uint16_t> tab16(L.childTab.size());
        std::vector<float> rowScale(nRows);
        for (size_t b = 0; b < nRows; b++) {
            double mx = 0;
            for (size_t i = 0; i < 64; i++) mx = std::max(mx, std::abs(L.childTab[64 * b + i]));
            const double sc = mx > 0 ? std::exp2(std::ceil(std::log2(mx / 16384.0))) : 1.0;
            rowScale[b] = float(sc);
            for (size_t i = 0; i < 64; i++) {
                const _Float16 h = static_cast<_Float16>(L.childTab[64 * b + i] / sc);
                std::memcpy(&tab16[64 * b + i], &h, sizeof h);
            }
        }
        D.ctab16 = upload(tab16.data(), int64_t(tab16.size()));
        D.ctabRowScale = upload(rowScale.data(), int64_t(nRows));
    }
    // level 0's held DOFs, active nodes and element list (element -> voxel, in voxel order) on the
    // host (reactions, the break and topology loops); a device-built model downloads its voxel words
    // and held DOFs on first use and derives them
    const Level& hostLevel0() {
        Level& L = H.levels[0];
        if (!deviceBuilt || !L.active.empty()) return L;
        std::vector<int32_t> words(size_t(L.nVox));
        L.bc.resize(size_t(L.nDof));
        HIPCHK(hipMemcpyAsync(words.data(), lv[0].dl.words, size_t(L.nVox) * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(L.bc.data(), bc0, size_t(L.nDof), hipMemcpyDeviceToHost, stream));
        sync();
        L.elemVox.clear();
        L.elemVox.reserve(size_t(L.nElem));
        for (int64_t v = 0; v < L.nVox; v++)
            if (words[size_t(v)] != 0) L.elemVox.push_back(int32_t(v));
        L.active.assign(size_t(L.nNodes), 0);
        for (int k = 0; k < L.NZ; k++)
            for (int j = 0; j < L.NY; j++)
                for (int i = 0; i < L.NX; i++) {
                    bool act = false;
                    for (int v = 0; v < 8 && !act; v++) {
                        const int vi = i - 1 + (v & 1), vj = j - 1 + ((v >> 1) & 1), vk = k - 1 + (v >> 2);
                        if (vi < 0 || vj < 0 || vk < 0 || vi >= L.nx || vj >= L.ny || vk >= L.nz) continue;
                        act = words[size_t(vi + int64_t(L.nx) * (vj + int64_t(L.ny) * vk))] != 0;
                    }
                    L.active[size_t(L.node(i, j, k))] = act ? 1 : 0;
                }
        return L;
    }

    // work vectors for k right-hand sides (grown, never shrunk; the contents do not survive)
    void reserve(int k) {
        if (k <= cap) return;
        sync();
        dropGraph();  // its kernels point into the buffers about to be replaced
        for (void* ptr : work) HIPCHK(hipFreeAsync(ptr, stream));
        work.clear();
        workBytes = 0;
        for (size_t l = 0; l < H.levels.size(); l++) {
            const int64_t nl = H.levels[l].nDof * int64_t(k);
            lv[l].r64 = alloc<double>(nl, true);
            lv[l].a64 = alloc<double>(nl, true);
            lv[l].b64 = alloc<double>(nl, true);
            lv[l].d64 = alloc<double>(nl, true);
            lv[l].r32 = alloc<float>(nl, true);
            lv[l].a32 = alloc<float>(nl, true);
            lv[l].b32 = alloc<float>(nl, true);
            lv[l].d32 = alloc<float>(nl, true);
            if (lv[l].dl.ownVox) lv[l].dl.ownT = alloc<float>(lv[l].dl.ownTStride * k, true);
        }
        const int64_t n = nDof(), nk = n * k;
        x = alloc<double>(nk, true);
        r = alloc<double>(nk, true);
        p = alloc<double>(nk, true);
        q = alloc<double>(nk, true);
        f = alloc<double>(nk, true);
        u = alloc<double>(nk, true);
        p32 = alloc<float>(nk, true);
        pAlt = alloc<double>(nk, true);
        p32Alt = alloc<float>(nk, true);
        x32 = alloc<float>(nk, true);
        q32 = alloc<float>(nk, true);
        st = alloc<CgState>(k, true);
        {
            // per right-hand side: 4 partials per reduction block and the operator's fused dots (2 per
            // tile of 32 x 8 nodes x the shortest z-chunk, 2 planes)
            const Level& L0 = H.levels[0];
            const int64_t tiles = int64_t((L0.NX + 31) / 32) * ((L0.NY + 7) / 8) * ((L0.NZ + 1) / 2);
            red.stride = (std::max<int64_t>(4 * (n / RED_TILE + 2), 2 * tiles + 16) + 31) / 32 * 32;
            red.partials = alloc<double>(red.stride * k, true);
        }
        if (stHost) HIPCHK(hipHostFree(stHost));
        stHost = nullptr;
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&stHost), sizeof(CgState) * size_t(k), hipHostMallocDefault));
        std::memset(stHost, 0, sizeof(CgState) * size_t(k));
        cap = k;
        sync();
    }
    int64_t totalBytes() const { return bytes + workBytes; }

    // ----- operator dispatch: the z-marching kernel on levels without own matrices -----
    bool useFine = [] {
        const char* e = std::getenv("PSX_FINE");
        return !(e && e[0] == '0');
    }();
    // form: 0 = the reference's operator (own matrices as stored FP32 entries); OF_CHILD = own
    // elements in child-scale form where the level has it (exact Galerkin products); OF_AUTO = the
    // child-scale form only where it is the cheaper one (the FP64 and mixed solves: a V-cycle needs
    // no exact coarse operators); OF_MFMA = the BF16 matrix-core operator (FP32 vectors) where the
    // level allows it
    template <class T>
    // fd: a conjugate-gradient dot product over the level's DOFs (FD_PQ: x.y, FD_RZ: r.y and y.q),
    // fused into the operator kernel where it can be, else run after it; either way finished into
    // the CG state
    void op(size_t l, int mode, const T* xx, T* yy, const T* rr, const T* dd, double c1, double c2, const int* stop, int form = 0,
            const DotFusion* fd = nullptr) {
        const DevLevel& D = lv[l].dl;
        // the element-by-element matrix-core form: FP32 levels without own elements
        // (the child-scale level too: its own elements' corner rows come from a pass of their own)
        const int el = sizeof(T) == 4 && (H.levels[l].nOwn == 0 || (D.childForm && D.ownVox)) ? ((form & OF_EL3) ? 3 : (form & OF_MFMA) ? 2 : 0) : 0;
        if (!useFine) launchApply<T>(D, mode, xx, yy, rr, dd, c1, c2, stop, stream, ni);
        else if (D.symOk && (H.levels[l].nOwn == 0 || (D.childForm && ((form & OF_CHILD) || ((form & OF_AUTO) && lv[l].childCheap))))) {
            // the mixed V-cycle's smoothing (OF_AUTO, FP32) reads the child-scale subset table in FP16
            const bool half = (form & OF_AUTO) != 0;
            if (fd) {
                const unsigned nb = launchFineDot<T>(D, mode, xx, yy, rr, dd, c1, c2, stop, *fd, stream, ni, el, half);
                if (nb) {
                    if (fd->kind == FD_PQ) launchFinishPQ(fd->red, nb, st, stream, ni);
                    else launchFinishRZ(fd->red, nb, st, stream, ni);
                    return;
                }
            } else {
                launchFine<T>(D, mode, xx, yy, rr, dd, c1, c2, stop, stream, ni, el, half);
            }
        } else if constexpr (sizeof(T) == 4) {
            if (form & OF_MFMA_OWN) launchApplyOwnMf(D, mode, xx, yy, rr, dd, c1, c2, stop, stream, ni);
            else launchApplyOwn<T>(D, mode, xx, yy, rr, dd, c1, c2, stop, stream, ni, (form & OF_AUTO) != 0);
        } else {
            launchApplyOwn<T>(D, mode, xx, yy, rr, dd, c1, c2, stop, stream, ni);
        }
        if (fd) dotAfter<T>(*fd, D.nDof, xx, yy, rr);
    }
    // whether y = K p on level 0 (with this form) is the VALU z-march, which can fuse the direction update
    bool fusesDirection(int form) const {
        const DevLevel& D = lv[0].dl;
        return useFine && D.symOk && H.levels[0].nOwn == 0 && !(form & OF_EL3) && !directionUnfused();
    }
    static bool directionUnfused() {
        static const bool off = [] {
            const char* e = std::getenv("PSX_FUSE_P");
            return e && e[0] == '0';
        }();
        return off;
    }
    template <class T> void dotAfter(const DotFusion& fd, int64_t n, const T* xx, const T* yy, const T* rr) {
        if (fd.kind == FD_PQ) launchDotPQ<T>(xx, yy, n, st, red, stream, ni);
        else launchDotsRZ<T>(rr, yy, static_cast<const T*>(fd.q), n, st, red, stream, ni);
    }

    // ----- multigrid -----

    // largest eigenvalue of D^-1 K per level: 10 Lanczos steps on the device from the reference's
    // LCG start vector (voxel_fea.cpp:579-610)
    void estimateSmoothers() {
        const int saved = ni;
        ni = 1;
        for (size_t l = 0; l + 1 < lv.size(); l++) {
            Level& L = H.levels[l];
            if (L.lmax > 0) continue;
            const DevLevel& D = lv[l].dl;
            const int64_t n = L.nDof;
            double *v = lv[l].a64, *vPrev = lv[l].d64, *w = lv[l].b64, *xx = lv[l].r64;
            if (deviceBuilt) {
                launchLanczosStart(D.invD64, n, v, &st->spare[0], red, stream);
            } else {
                const auto v0 = lanczosStart(L);
                HIPCHK(hipMemcpyAsync(v, v0.data(), size_t(n) * 8, hipMemcpyHostToDevice, stream));
            }
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
        ni = saved;
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
    // rz: the conjugate gradients' r.z and z.q (q = the previous K p), fused into the last step
    // firstReady: buffer a of level l already holds the first Chebyshev step c0 D^-1 rIn (the
    // conjugate gradients' residual update writes it), so the cycle starts from it
    template <class T>
    T* vcycle(size_t l, const T* rIn, const int* stop, int form = 0, int degree = 2, const DotFusion* rz = nullptr, bool firstReady = false) {
        Vecs<T> V = vecs<T>(l);
        const DevLevel& D = lv[l].dl;
        if (l + 1 == lv.size()) {
            launchCoarse<T>(freeIdx, denseIdx, H.m, D.nDof, Ainv, rIn, V.b, stop, stream, ni);
            if (rz) dotAfter<T>(*rz, D.nDof, nullptr, V.b, rIn);
            return V.b;
        }
        const Cheb c = chebyshev(H.levels[l].lmax, degree);
        const DevLevel& C = lv[l + 1].dl;
        // three buffers rotate through the iterates: z (current), zp (previous, nullptr for the
        // zero start) and the one being written
        T* buf[3] = {V.a, V.b, V.d};
        T *z = buf[0], *zp = nullptr;
        auto freeBuf = [&] {
            for (T* b : buf)
                if (b != z && b != zp) return b;
            return buf[2];
        };
        auto step = [&](int mode, double c1, double c2, const DotFusion* fd = nullptr) {
            T* y = freeBuf();
            op<T>(l, mode, z, y, rIn, zp, c1, c2, stop, form, fd);
            zp = z;
            z = y;
        };
        if (!firstReady) launchChebFirst<T>(D, rIn, z, c.c0, stop, stream, ni);
        const int pre = preDegree > 0 ? std::min(preDegree, degree) : degree;
        for (int j = 1; j < pre; j++) step(AM_CHEB2, c.c1[j], c.c2[j]);
        zp = nullptr;
        T* t = freeBuf();
        op<T>(l, AM_RESID, z, t, rIn, nullptr, 0, 0, stop, form);
        Vecs<T> W = vecs<T>(l + 1);
        // the restriction also writes the next level's first Chebyshev step (its buffer a), unless that
        // level is the coarsest
        const bool smoothNext = l + 2 < lv.size();
        launchRestrict<T>(D, C, t, W.r, stop, stream, ni, smoothNext ? W.a : nullptr, smoothNext ? chebyshev(H.levels[l + 1].lmax, degree).c0 : 0.0);
        const T* zc = vcycle<T>(l + 1, W.r, stop, form, degree, nullptr, smoothNext);
        launchProlong<T>(D, C, zc, z, stop, stream, ni);
        step(AM_CHEB1, c.c0, 0, degree == 1 ? rz : nullptr);
        for (int j = 1; j < degree; j++) step(AM_CHEB2, c.c1[j], c.c2[j], j + 1 == degree ? rz : nullptr);
        return z;
    }

    // z = M r on level 0 with the FP32 cycle and r scaled to order one (voxel_fea.cpp:835-861)
    void precondition32(const double* dr, double* dz, const int* stop, int degree = 2) {
        const int64_t n = nDof();
        launchMaxAbs64(dr, n, &st->rmax, red, stream, ni);
        launchToF32(dr, lv[0].r32, n, &st->rmax, stream, ni);
        const float* z32 = vcycle<float>(0, lv[0].r32, stop, 0, degree);
        launchFromF32(z32, dz, n, &st->rmax, stream, ni);
    }

    // the corner stress matrices scaled by E / h, uploaded into stressS
    void stressMatrices(double E, double h) {
        if (!(E > 0) || !(h > 0)) throw std::invalid_argument("E and h must be positive");
        std::vector<double> S(8 * 144);
        for (int a = 0; a < 8; a++)
            for (int q = 0; q < 144; q++) S[size_t(a) * 144 + size_t(q)] = H.el.corner[a][q] * E / h;
        HIPCHK(hipMemcpyAsync(stressS, S.data(), S.size() * 8, hipMemcpyHostToDevice, stream));
        sync();
    }

    // one V-cycle for each of k vectors (instance-major) in one set of launches; each result equals
    // the vector's lone precondition32
    void preconditionMulti(int k, const double* dr, double* dz) {
        reserve(k);
        ni = k;
        struct Restore {
            int& n;
            ~Restore() { n = 1; }
        } restore{ni};
        precondition32(dr, dz, nullptr);
    }

    // ----- conjugate gradients -----

    void readState() {
        HIPCHK(hipMemcpyAsync(stHost, st, sizeof(CgState) * size_t(ni), hipMemcpyDeviceToHost, stream));
        sync();
    }

    void solveDevice(const double* df, const double* dx0, const psx_solve_opts& o, double* du, psx_solve_info& info) { solveMulti(1, df, dx0, o, du, &info); }

    // One conjugate-gradient run over k right-hand sides, in phases: begin (start vector, initial
    // residual and device state), batches of iterations (enqueueBatch, then poll), end (the last fold
    // and the true residual). solveMulti runs them on the model's stream; a packed solve
    // (packed.hpp) runs the batches of several models' runs as shared launches.
    struct Run {
        int k = 1;
        const double* df = nullptr;
        const double* dx0 = nullptr;
        double* du = nullptr;
        psx_solve_opts o{};
        psx_solve_info* info = nullptr;
        std::chrono::steady_clock::time_point t0;
        std::vector<double> bnorm;
        int prec = 0, batch = 8, form = 0, degree = 2;
        bool fuseFirst = false;
        double c0first = 0;
        bool allStopped = true, cancelled = false;
    };

    // k independent right-hand sides in one set of launches: df, du (and dx0 when not null) hold k
    // vectors of nDof, instance-major; every kernel runs all k (blockIdx.y) and each instance stops
    // on its own state, so the result of each right-hand side is the one a lone solve gives
    void solveMulti(int k, const double* df, const double* dx0, const psx_solve_opts& o, double* du, psx_solve_info* info) {
        Run R;
        struct Reset {
            Model& m;
            ~Reset() { m.resetRun(); }
        } reset{*this};
        if (!begin(R, k, df, dx0, o, du, info)) return;
        // the batch as a HIP graph, captured once and replayed (kernels only: no copies or memsets);
        // it saves the host one launch per kernel, which bounds small parts
        const GraphKey key{R.prec, R.degree, preDegree, R.form, k, R.batch, df, du};
        const bool useGraph = !o.eager && !graphsOff();
        if (useGraph && !(graph && graphKey == key)) {
            dropGraph();
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            try {
                enqueueBatch(R);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamEndCapture(stream, &g));
            const hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                graph = nullptr;
                throw DeviceError(std::string("hipGraphInstantiate failed: ") + hipGetErrorString(e));
            }
            graphKey = key;
        }
        for (;;) {
            if (useGraph) HIPCHK(hipGraphLaunch(graph, stream));
            else enqueueBatch(R);
            readState();
            if (poll(R, true)) break;
        }
        end(R);
    }

    // the run's launch counts back to one right-hand side and the pre-smoothing to the post-smoothing's
    void resetRun() {
        ni = 1;
        preDegree = 0;
    }

    // x = x0, the initial residual and the per-instance state; false when no instance iterates (info
    // is then complete)
    bool begin(Run& R, int k, const double* df, const double* dx0, const psx_solve_opts& o, double* du, psx_solve_info* info) {
        R = Run{};
        R.t0 = std::chrono::steady_clock::now();
        R.k = k;
        R.df = df;
        R.dx0 = dx0;
        R.du = du;
        R.o = o;
        R.info = info;
        if (k < 1) throw std::invalid_argument("no right-hand sides");
        if (!smoothersReady) estimateSmoothers();
        reserve(k);
        ni = k;
        const int64_t n = nDof();
        const size_t nk = size_t(n) * size_t(k);
        const DevLevel& D0 = lv[0].dl;
        for (int i = 0; i < k; i++) info[i] = psx_solve_info{};
        // x = x0 on free DOFs; bnorm = |f_free|; r = f - K x (FP64)
        launchMaskCopy64(dx0, D0.invD64, du, n, stream, k);
        launchNormFree64(df, D0.invD64, n, &st->bnorm, red, stream, k);
        op<double>(0, AM_APPLY, du, q, nullptr, nullptr, 0, 0, nullptr);
        launchResidual64(df, q, D0.invD64, r, n, &st->rr, red, stream, k);
        readState();
        std::vector<double> rel(k), rmax(k, 1.0);
        R.bnorm.assign(k, 0.0);
        bool anyLoad = false;
        for (int i = 0; i < k; i++) {
            R.bnorm[i] = std::sqrt(stHost[i].bnorm);
            rel[i] = R.bnorm[i] > 0 ? std::sqrt(stHost[i].rr) / R.bnorm[i] : 0;
            anyLoad = anyLoad || R.bnorm[i] > 0;
            // no load on the free DOFs: u = 0 (voxel_fea.cpp:872-876)
            if (R.bnorm[i] == 0) HIPCHK(hipMemsetAsync(du + size_t(i) * n, 0, size_t(n) * 8, stream));
        }
        if (!anyLoad) {
            sync();
            for (int i = 0; i < k; i++) {
                info[i].converged = 1;
                info[i].device_ms = ms(R.t0);
            }
            return false;
        }
        R.prec = o.precision;
        R.batch = o.check_every > 0 ? o.check_every : 8;
        // the reference's algorithm keeps its operator; the FP64 and mixed solves use the exact
        // child-scale form, and the mixed smoother the matrix cores when asked (one right-hand side)
        // mfma_smoother bit 0: the V-cycle's levels without own elements as the element-by-element
        // matrix-core product (two BF16 pieces); bit 1: own-element levels; bit 2: the outer FP32
        // operator as the three-piece element product
        const bool mx = R.prec == PSX_PREC_MIXED;
        R.form = R.prec == PSX_PREC_REF ? 0
                                        : (OF_AUTO | (mx && (o.mfma_smoother & 1) ? OF_MFMA : 0) | (mx && (o.mfma_smoother & 2) ? OF_MFMA_OWN : 0) |
                                           (mx && (o.mfma_smoother & 4) ? OF_EL3 : 0));
        R.degree = o.cheb_degree > 0 ? o.cheb_degree : 2;
        if (R.degree > CHEB_MAX) throw std::invalid_argument("Chebyshev degree above 6");
        preDegree = o.pre_degree > 0 ? std::min(o.pre_degree, R.degree) : 0;
        // mixed: the residual updates also write the next V-cycle's first Chebyshev step on level 0
        R.fuseFirst = R.prec == PSX_PREC_MIXED && lv.size() > 1;
        R.c0first = R.fuseFirst ? chebyshev(H.levels[0].lmax, R.degree).c0 : 0.0;
        if (R.prec == PSX_PREC_MIXED) {
            // FP32 units for the whole solve: each first residual scaled to order one
            launchMaxAbs64(r, n, &st->rmax, red, stream, k);
            readState();
            for (int i = 0; i < k; i++) rmax[i] = stHost[i].rmax > 0 ? stHost[i].rmax : 1;
        }
        std::vector<CgState> s0(k);
        R.allStopped = true;
        for (int i = 0; i < k; i++) {
            CgState& c = s0[i];
            c = CgState{};
            c.rz = 1;
            c.bnorm = R.bnorm[i];
            c.tol = o.tol;
            c.maxIter = o.max_iter;
            c.scale = R.prec == PSX_PREC_MIXED ? rmax[i] : 1;
            c.rmax = rmax[i];
            c.rel = rel[i];
            c.trueRel = rel[i];
            c.peak = rel[i];
            c.noFold = 1;
            c.foldEvery = o.fold_every > 0 ? o.fold_every : 20;
            c.stop = (R.bnorm[i] == 0 || rel[i] <= o.tol || o.max_iter <= 0) ? 1 : 0;
            R.allStopped = R.allStopped && c.stop;
        }
        HIPCHK(hipMemcpyAsync(st, s0.data(), sizeof(CgState) * size_t(k), hipMemcpyHostToDevice, stream));
        if (R.prec == PSX_PREC_MIXED) {
            HIPCHK(hipMemsetAsync(p32, 0, nk * 4, stream));
            HIPCHK(hipMemsetAsync(q32, 0, nk * 4, stream));
            HIPCHK(hipMemsetAsync(x32, 0, nk * 4, stream));
            launchToF32(r, lv[0].r32, n, &st->rmax, stream, k);
            // the first V-cycle's first Chebyshev step; later ones come from the residual updates
            if (R.fuseFirst) launchChebFirst<float>(D0, lv[0].r32, lv[0].a32, R.c0first, &st->stop, stream, k);
        } else {
            HIPCHK(hipMemsetAsync(p, 0, nk * 8, stream));
            HIPCHK(hipMemsetAsync(q, 0, nk * 8, stream));
        }
        if (R.allStopped) {
            end(R);
            return false;
        }
        // the state and the cleared vectors are in place before any other stream runs the batches
        sync();
        return true;
    }

    // one batch of iterations: every kernel returns at once for an instance whose state says stop,
    // so the same launches serve every batch (kernels only: it is captured and recorded)
    void enqueueBatch(const Run& R) {
        const int64_t n = nDof();
        const int k = R.k;
        const DevLevel& D0 = lv[0].dl;
        int* stop = &st->stop;
        float* r32 = lv[0].r32;
        const int prec = R.prec, form = R.form, degree = R.degree;
        // The direction update p = z + beta p joins the outer operator (one pass over z and p instead
        // of two) where that operator is the VALU z-march: the new direction goes to the other buffer,
        // so the buffers alternate. An odd batch updates in place in its last iteration, which leaves
        // the direction in its home buffer at every batch boundary (the batch replays as one graph).
        const bool fuseP = fusesDirection(prec == PSX_PREC_MIXED ? form & ~(OF_MFMA | OF_MFMA_OWN) : form);
        for (int b = 0; b < R.batch; b++) {
            const bool fused = fuseP && (b + 1 < R.batch || R.batch % 2 == 0);
            if (prec == PSX_PREC_MIXED) {
                float* pOld = b % 2 && fuseP ? p32Alt : p32;
                float* pNew = fused ? (pOld == p32 ? p32Alt : p32) : pOld;
                const DotFusion fz{FD_RZ, q32, red};
                const float* z = vcycle<float>(0, r32, stop, form & ~OF_EL3, degree, &fz, R.fuseFirst);
                DotFusion fp{FD_PQ, nullptr, red};
                if (fused) {
                    fp.pz = z;
                    fp.pw = pNew;
                } else {
                    launchUpdateP<float>(z, pOld, n, st, stream, k);
                }
                // the outer operator: the FP32 VALU form, or the three-piece element product (OF_EL3)
                op<float>(0, AM_APPLY, pOld, q32, nullptr, nullptr, 0, 0, stop, form & ~(OF_MFMA | OF_MFMA_OWN), &fp);
                launchUpdateXR32Fold(x32, pNew, r32, q32, n, st, red, stream, k, R.fuseFirst ? lv[0].a32 : nullptr, D0.invD32, R.c0first);
                // reliable update: the correction joins the FP64 solution and the FP64 true
                // residual replaces the recurrence's
                launchFoldX(R.du, x32, n, st, stream, k);
                op<double>(0, AM_APPLY, R.du, q, nullptr, nullptr, 0, 0, &st->noFold);
                // (the FP64 residual itself is not kept: nothing reads it before the run's final one; the
                // held DOFs come from the FP32 inverse diagonal)
                launchFoldR(R.df, q, nullptr, nullptr, r32, n, st, red, stream, k, R.fuseFirst ? lv[0].a32 : nullptr, D0.invD32, R.c0first);
                continue;
            }
            const double* z;
            const DotFusion fz{FD_RZ, q, red};
            if (prec == PSX_PREC_FP64) z = vcycle<double>(0, r, stop, form, degree, &fz);
            else {
                precondition32(r, lv[0].b64, stop, degree);
                z = lv[0].b64;
                launchDotsRZ<double>(r, z, q, n, st, red, stream, k);
            }
            double* pOld = b % 2 && fuseP ? pAlt : p;
            double* pNew = fused ? (pOld == p ? pAlt : p) : pOld;
            DotFusion fp{FD_PQ, nullptr, red};
            if (fused) {
                fp.pz = z;
                fp.pw = pNew;
            } else {
                launchUpdateP<double>(z, pOld, n, st, stream, k);
            }
            op<double>(0, AM_APPLY, pOld, q, nullptr, nullptr, 0, 0, stop, form, &fp);
            launchUpdateXR<double>(R.du, pNew, r, q, n, st, red, stream, k);
        }
    }

    // after a batch (the state already read back): true when every instance has stopped or the
    // progress callback (when asked) cancels
    bool poll(Run& R, bool progress) {
        R.allStopped = true;
        int itMax = 0;
        double relMax = 0;
        for (int i = 0; i < R.k; i++) {
            R.allStopped = R.allStopped && stHost[i].stop;
            itMax = std::max(itMax, stHost[i].iters);
            relMax = std::max(relMax, stHost[i].rel);
        }
        if (R.allStopped) return true;
        if (progress && R.o.on_progress && R.o.on_progress(itMax, relMax, R.o.user)) {
            R.cancelled = true;
            return true;
        }
        return false;
    }

    void end(Run& R) {
        const int k = R.k;
        const int64_t n = nDof();
        const DevLevel& D0 = lv[0].dl;
        psx_solve_info* info = R.info;
        // host source of the flag copies below: alive until readState has waited for them
        std::vector<CgState> s0(k);
        if (R.prec == PSX_PREC_MIXED) {
            // a cancelled run still folds in its last correction
            for (int i = 0; i < k; i++) s0[i].noFold = 0;
            for (int i = 0; i < k; i++)
                HIPCHK(hipMemcpyAsync(&st[i].noFold, &s0[i].noFold, sizeof(int), hipMemcpyHostToDevice, stream));
            launchFoldX(R.du, x32, n, st, stream, k);
        }
        readState();
        for (int i = 0; i < k; i++) {
            info[i].iterations = stHost[i].iters;
            info[i].broken = stHost[i].broken;
            info[i].folds = stHost[i].folds;
        }
        // the final true residual against the equilibrium equations (voxel_fea.cpp:920-925)
        op<double>(0, AM_APPLY, R.du, q, nullptr, nullptr, 0, 0, nullptr);
        launchResidual64(R.df, q, D0.invD64, r, n, &st->rr, red, stream, k);
        readState();
        for (int i = 0; i < k; i++) {
            if (R.bnorm[i] == 0) {
                info[i].converged = 1;
                info[i].device_ms = ms(R.t0);
                continue;
            }
            info[i].cancelled = R.cancelled;
            info[i].residual = std::sqrt(stHost[i].rr) / R.bnorm[i];
            info[i].converged = !R.cancelled && std::isfinite(info[i].residual) && info[i].residual <= R.o.tol * 10;
            info[i].device_ms = ms(R.t0);
        }
    }

    // PSX_GRAPHS=0 launches every kernel from the host
    static bool graphsOff() {
        static const bool off = [] {
            const char* e = std::getenv("PSX_GRAPHS");
            return e && e[0] == '0';
        }();
        return off;
    }

    static double ms(std::chrono::steady_clock::time_point t0) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

private:
    std::vector<void*> allocs, work;
    int64_t workBytes = 0;
};

}  // namespace psx
