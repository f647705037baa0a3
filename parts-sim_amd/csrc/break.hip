// The break test on the device (breakTest, static_study.cpp:197-279): per step one warm-started solve,
// the nodal stresses, each voxel's failure measure (its worst corner), the largest measure and
// displacement, the crack and the prune of the voxels it cuts off, with the model's arrays resident
// in HBM; the host builds each step's hierarchy and orders the crack.
//
// The prune (pruneFloating, voxel_fea.cpp:1086-1119) is a level-synchronous breadth-first search:
// the held voxels seed a queue, and each level's kernel visits its slice of the queue, claims
// unvisited solid face neighbours with an atomic OR on a bit per voxel and appends them. The slice
// bounds live in device memory (a one-thread kernel records the queue's end after each level), so
// levels run back to back and the host checks for an empty level once per batch. The set of voxels
// reached does not depend on the order of the visits, so the result equals the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <stdexcept>
#include <vector>

#include "../../include/psx.h"
#include "abi.hpp"

#pragma clang fp contract(off)

namespace psx {

namespace {

constexpr int BLOCK = 256;
constexpr int BFS_BLOCKS = 1024;
constexpr int BFS_BATCH = 16;

inline unsigned grid(int64_t n) { return unsigned(std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK)); }

struct Grid3 {
    int nx, ny, nz;
};

__global__ __launch_bounds__(BLOCK) void kPruneSeed(Grid3 g, const float* __restrict__ dens, const uint8_t* __restrict__ held, uint32_t* __restrict__ bits,
                                                    int32_t* __restrict__ queue, int* __restrict__ cursor) {
    const int64_t e = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    const int64_t nV = int64_t(g.nx) * g.ny * g.nz;
    if (e >= nV || !(dens[e] > 0)) return;
    const int i = int(e % g.nx), j = int((e / g.nx) % g.ny), k = int(e / (int64_t(g.nx) * g.ny));
    const int64_t NX = g.nx + 1, NXY = NX * (g.ny + 1);
    const int64_t n0 = i + NX * j + NXY * k;
    const bool h = held[n0] | held[n0 + 1] | held[n0 + NX] | held[n0 + NX + 1] | held[n0 + NXY] | held[n0 + NXY + 1] | held[n0 + NXY + NX] |
                   held[n0 + NXY + NX + 1];
    if (!h) return;
    atomicOr(&bits[e >> 5], 1u << (e & 31));
    queue[atomicAdd(&cursor[0], 1)] = int32_t(e);
}

// the level-end marks form a ring: level l reads marks[l], marks[l + 1] and its mark kernel writes
// marks[l + 2], so a ring of 64 covers any batch of levels queued between two host checks
constexpr int MARK_RING = 64;
static_assert(MARK_RING >= 2 * (BFS_BATCH + 2), "the ring must hold a batch of levels");

// marks[level + 1] = the queue's end
__global__ void kPruneMark(int* __restrict__ marks, int level, const int* __restrict__ cursor) {
    if (threadIdx.x == 0 && blockIdx.x == 0) marks[(level + 1) % MARK_RING] = cursor[0];
}

__global__ __launch_bounds__(BLOCK) void kPruneExpand(Grid3 g, const float* __restrict__ dens, uint32_t* __restrict__ bits, int32_t* __restrict__ queue,
                                                      int* __restrict__ cursor, const int* __restrict__ marks, int level) {
    const int lo = marks[level % MARK_RING], hi = marks[(level + 1) % MARK_RING];
    const int64_t sxy = int64_t(g.nx) * g.ny;
    for (int64_t t = lo + int64_t(blockIdx.x) * BLOCK + threadIdx.x; t < hi; t += int64_t(gridDim.x) * BLOCK) {
        const int64_t e = queue[t];
        const int i = int(e % g.nx), j = int((e / g.nx) % g.ny), k = int(e / sxy);
        const int64_t nb[6] = {i > 0 ? e - 1 : -1, i < g.nx - 1 ? e + 1 : -1, j > 0 ? e - g.nx : -1, j < g.ny - 1 ? e + g.nx : -1,
                               k > 0 ? e - sxy : -1, k < g.nz - 1 ? e + sxy : -1};
        for (int a = 0; a < 6; a++) {
            const int64_t m = nb[a];
            if (m < 0 || !(dens[m] > 0)) continue;
            const uint32_t bit = 1u << (m & 31);
            if (bits[m >> 5] & bit) continue;  // already claimed (a stale read only costs the atomic)
            if (atomicOr(&bits[m >> 5], bit) & bit) continue;
            queue[atomicAdd(&cursor[0], 1)] = int32_t(m);
        }
    }
}

// out = the reached solid voxels' densities; counts removed (cursor[1]) and kept (cursor[2]) voxels;
// the solid voxels not reached are appended to `lost` (cursor[3]) when given
__global__ __launch_bounds__(BLOCK) void kPruneFinish(int64_t nV, const float* __restrict__ dens, const uint32_t* __restrict__ bits, float* __restrict__ out,
                                                      int* __restrict__ cursor, int32_t* __restrict__ lost) {
    const int64_t e = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= nV) return;
    const float d = dens[e];
    const bool solid = d > 0, reached = (bits[e >> 5] >> (e & 31)) & 1u;
    out[e] = solid && reached ? d : 0.f;
    if (solid && !reached) {
        atomicAdd(&cursor[1], 1);
        if (lost) lost[atomicAdd(&cursor[3], 1)] = int32_t(e);
    }
    if (solid && reached) atomicAdd(&cursor[2], 1);
}

__global__ __launch_bounds__(BLOCK) void kZeroWords(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t t = int64_t(blockIdx.x) * BLOCK + threadIdx.x; t < n; t += int64_t(gridDim.x) * BLOCK) p[t] = 0u;
}

// where the prune's scratch lives: a bit per voxel, the queue (one entry per voxel), the ring of
// level marks and the counters
struct PruneScratch {
    uint32_t* bits = nullptr;
    int32_t* queue = nullptr;
    int* marks = nullptr;   // MARK_RING entries
    int* cursor = nullptr;  // [0] queue end, [1] removed, [2] kept, [3] lost entries
};

// the breadth-first search on scratch S; host receives the cursor and two marks (4 ints, pinned or
// not). out = the pruned densities; returns the count removed; kept receives the solid voxels left;
// lostHost (with lost, nV entries of device scratch) the removed voxels in increasing order
int runPrune(hipStream_t s, const PruneScratch& S, int64_t nV, int* host, Grid3 g, const float* dens, const uint8_t* held, float* out, int* kept,
             int32_t* lost = nullptr, std::vector<int32_t>* lostHost = nullptr) {
    // cleared by kernels (no memset nodes, for callers that record the stream)
    kZeroWords<<<unsigned(std::min<int64_t>(2048, (nV / 32 + BLOCK) / BLOCK + 1)), BLOCK, 0, s>>>(S.bits, (nV + 31) / 32);
    kZeroWords<<<1, BLOCK, 0, s>>>(reinterpret_cast<uint32_t*>(S.cursor), 4);
    kZeroWords<<<1, BLOCK, 0, s>>>(reinterpret_cast<uint32_t*>(S.marks), MARK_RING);
    kPruneSeed<<<grid(nV), BLOCK, 0, s>>>(g, dens, held, S.bits, S.queue, S.cursor);
    kPruneMark<<<1, 64, 0, s>>>(S.marks, 0, S.cursor);
    // level l visits queue[marks[l], marks[l + 1]); every voxel enters the queue once, so at most
    // nV levels are non-empty
    int64_t level = 0;
    for (;;) {
        for (int b = 0; b < BFS_BATCH && level < nV; b++, level++) {
            kPruneExpand<<<BFS_BLOCKS, BLOCK, 0, s>>>(g, dens, S.bits, S.queue, S.cursor, S.marks, int(level));
            kPruneMark<<<1, 64, 0, s>>>(S.marks, int(level + 1), S.cursor);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(host, S.marks + level % MARK_RING, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(host + 1, S.marks + (level + 1) % MARK_RING, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (host[0] == host[1] || level >= nV) break;  // the last level added nothing
    }
    kPruneFinish<<<grid(nV), BLOCK, 0, s>>>(nV, dens, S.bits, out, S.cursor, lost);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, S.cursor, 16, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (kept) *kept = host[2];
    if (lost && lostHost) {
        lostHost->resize(size_t(host[3]));
        if (host[3] > 0) HIPCHK(hipMemcpyAsync(lostHost->data(), lost, size_t(host[3]) * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::sort(lostHost->begin(), lostHost->end());
    }
    return host[1];
}

// the prune's scratch in stream-ordered allocations of the caller's stream
struct Pruner {
    hipStream_t s;
    int64_t nV;
    PruneScratch S;
    int* host = nullptr;  // pinned copies of the cursor and two marks
    Pruner(hipStream_t st, int64_t n) : s(st), nV(n) {
        HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&S.bits), size_t((n + 31) / 32) * 4, s));
        HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&S.queue), size_t(n) * 4, s));
        HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&S.marks), size_t(MARK_RING) * 4, s));
        HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&S.cursor), 16, s));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&host), 32, hipHostMallocDefault));
    }
    ~Pruner() {
        if (S.bits) (void)hipFreeAsync(S.bits, s);
        if (S.queue) (void)hipFreeAsync(S.queue, s);
        if (S.marks) (void)hipFreeAsync(S.marks, s);
        if (S.cursor) (void)hipFreeAsync(S.cursor, s);
        (void)hipStreamSynchronize(s);
        if (host) (void)hipHostFree(host);
    }
    int run(Grid3 g, const float* dens, const uint8_t* held, float* out, int* kept, int32_t* lost = nullptr, std::vector<int32_t>* lostHost = nullptr) {
        return runPrune(s, S, nV, host, g, dens, held, out, kept, lost, lostHost);
    }
};

// failure measure per element (its worst corner, at least 0) and the largest one (float bits of a
// non-negative float order as unsigned integers)
__global__ __launch_bounds__(BLOCK) void kMeasure(const int32_t* __restrict__ elemVox, int64_t nE, int nx, int ny, const float* __restrict__ nodal,
                                                  float* __restrict__ measure, unsigned* __restrict__ smaxBits) {
    const int64_t q = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    float m = 0.f;
    if (q < nE) {
        const int v = elemVox[q];
        const int i = v % nx, j = (v / nx) % ny, k = v / (nx * ny);
        const int64_t NX = nx + 1, NXY = NX * (ny + 1);
        const int64_t n0 = i + NX * j + NXY * k;
        const int64_t off[8] = {0, 1, 1 + NX, NX, NXY, 1 + NXY, 1 + NX + NXY, NX + NXY};
        for (int a = 0; a < 8; a++) m = fmaxf(m, nodal[n0 + off[a]]);
        m = m > 0.f ? m : 0.f;  // +0 for a -0 corner, so the bits order as the values
        measure[q] = m;
    }
    __shared__ float sh[BLOCK];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(smaxBits, __float_as_uint(sh[0]));
}

// the elements within 95 % of the largest measure and their measures (unordered)
__global__ __launch_bounds__(BLOCK) void kCandidates(const float* __restrict__ measure, int64_t nE, double smax, int32_t* __restrict__ cand,
                                                     float* __restrict__ candM, int* __restrict__ count) {
    const int64_t q = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (q >= nE) return;
    const float m = measure[q];
    if (m < 0.95 * smax) return;
    const int t = atomicAdd(count, 1);
    cand[t] = int32_t(q);
    candM[t] = m;
}

// u = s u0 (the physical displacements) and the largest |u|^2 over the nodes (double bits of a
// non-negative double order as unsigned integers)
__global__ __launch_bounds__(BLOCK) void kScaleDisp(const double* __restrict__ u0, double scale, int64_t nNodes, double* __restrict__ u,
                                                    unsigned long long* __restrict__ best) {
    const int64_t n = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    double b = 0;
    if (n < nNodes) {
        const double a = u0[3 * n] * scale, c = u0[3 * n + 1] * scale, d = u0[3 * n + 2] * scale;
        u[3 * n] = a;
        u[3 * n + 1] = c;
        u[3 * n + 2] = d;
        b = a * a + c * c + d * d;
    }
    __shared__ double sh[BLOCK];
    sh[threadIdx.x] = b;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(best, static_cast<unsigned long long>(__double_as_longlong(sh[0])));
}

__global__ __launch_bounds__(BLOCK) void kZeroVoxels(const int32_t* __restrict__ ids, int n, float* __restrict__ dens) {
    const int t = int(blockIdx.x) * BLOCK + threadIdx.x;
    if (t < n) dens[ids[t]] = 0.f;
}

// the step's fields at lambda: float(u lambda), float(vm lambda)
__global__ __launch_bounds__(BLOCK) void kFields(const double* __restrict__ u, const float* __restrict__ vm, int64_t nNodes, double lambda,
                                                 float* __restrict__ uOut, float* __restrict__ vmOut) {
    const int64_t n = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (n >= nNodes) return;
    for (int d = 0; d < 3; d++) uOut[3 * n + d] = float(u[3 * n + d] * lambda);
    vmOut[n] = float(double(vm[n]) * lambda);
}

struct DevBufs {
    hipStream_t s = nullptr;
    std::vector<void*> p;
    template <class T> T* get(size_t count) {
        void* q = nullptr;
        HIPCHK(hipMallocAsync(&q, std::max<size_t>(count, 1) * sizeof(T), s));
        HIPCHK(hipMemsetAsync(q, 0, std::max<size_t>(count, 1) * sizeof(T), s));
        p.push_back(q);
        return static_cast<T*>(q);
    }
    ~DevBufs() {
        for (void* q : p) (void)hipFreeAsync(q, s);
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

// static_study.cpp:23-32, host side over the level's active nodes (the reference's sum order)
double lostLoadFraction(const Level& L, const double* f) {
    double lost = 0, total = 0;
    for (size_t n = 0; n < L.active.size(); n++) {
        const double m = std::sqrt(f[3 * n] * f[3 * n] + f[3 * n + 1] * f[3 * n + 1] + f[3 * n + 2] * f[3 * n + 2]);
        total += m;
        if (!L.active[n]) lost += m;
    }
    return total > 0 ? lost / total : 0;
}

}  // namespace

// Scratch of the allocation-free prune (psx_prune_floating_device_noalloc), one copy per GPU that runs
// it, held until the process ends: 64 words of counters, the ring of level marks (64 words), then the
// visited bits and the queue of a part of at most PRUNE_NOALLOC_MAX_VOXELS voxels. Every call on a
// GPU uses that GPU's copy, so calls that overlap on one GPU race.
constexpr int64_t PRUNE_NOALLOC_MAX_VOXELS = int64_t(1) << 23;
constexpr int64_t NOALLOC_MARKS = 64, NOALLOC_BITS = 128, NOALLOC_QUEUE = NOALLOC_BITS + PRUNE_NOALLOC_MAX_VOXELS / 32;
constexpr int64_t NOALLOC_WORDS = NOALLOC_QUEUE + PRUNE_NOALLOC_MAX_VOXELS;
static_assert(NOALLOC_MARKS >= MARK_RING && NOALLOC_BITS - NOALLOC_MARKS >= MARK_RING, "ring and counters fit their slots");
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

namespace {

constexpr int NOALLOC_DEVICES = 64;
std::mutex g_scratchMu;
uint32_t* g_scratchPtr[NOALLOC_DEVICES] = {};

// the device address of the current GPU's scratch copy, looked up on the first call on that GPU (the
// array's name in host code is a host-side stand-in, never a device address)
uint32_t* scratchOnCurrentDevice() {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_DEVICES) throw std::invalid_argument("device index beyond the allocation-free prune's table");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) throw DeviceError(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratchPtr[dev] = static_cast<uint32_t*>(ptr);
    }
    return g_scratchPtr[dev];
}

}  // namespace

}  // namespace psx

extern "C" {

int psx_prune_floating_device_noalloc(int nx, int ny, int nz, const float* d_density, const uint8_t* d_held, float* d_out, int* removed) {
    using namespace psx;
    return abi::guarded([&] {
        if (!d_density || !d_held || !d_out || nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("bad arguments");
        const int64_t nV = int64_t(nx) * ny * nz;
        if (nV > PRUNE_NOALLOC_MAX_VOXELS)
            throw std::invalid_argument("grid of " + std::to_string(nV) + " voxels exceeds the allocation-free prune's limit of " +
                                        std::to_string(PRUNE_NOALLOC_MAX_VOXELS) + " voxels");
        uint32_t* base = scratchOnCurrentDevice();
        PruneScratch S;
        S.cursor = reinterpret_cast<int*>(base);
        S.marks = reinterpret_cast<int*>(base + NOALLOC_MARKS);
        S.bits = base + NOALLOC_BITS;
        S.queue = reinterpret_cast<int32_t*>(base + NOALLOC_QUEUE);
        int host[4] = {0, 0, 0, 0};  // read back through pageable copies: no pinned allocation
        const int r = runPrune(nullptr, S, nV, host, Grid3{nx, ny, nz}, d_density, d_held, d_out, nullptr);
        if (removed) *removed = r;
        return PSX_OK;
    });
}

int psx_prune_floating_device(int nx, int ny, int nz, const float* d_density, const uint8_t* d_held, float* d_out, int* removed) {
    using namespace psx;
    return abi::guarded([&] {
        if (!d_density || !d_held || !d_out || nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("bad arguments");
        const int64_t nV = int64_t(nx) * ny * nz;
        if (nV > int64_t(INT32_MAX) - 2) throw std::invalid_argument("grid too large for 32-bit voxel indices");
        hipStream_t s = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        int r = 0;
        {
            Pruner P(s, nV);
            r = P.run(Grid3{nx, ny, nz}, d_density, d_held, d_out, nullptr);
        }
        HIPCHK(hipStreamDestroy(s));
        if (removed) *removed = r;
        return PSX_OK;
    });
}

int psx_break_test(const psx_model_desc* part, const double* f, const psx_break_opts* opts, psx_break_result* out) {
    using namespace psx;
    return abi::guarded([&] {
        if (!part || !f || !opts || !out || !part->density || !part->bc) throw std::invalid_argument("bad arguments");
        if (part->diag_add) throw std::invalid_argument("the break test takes no diagonal shift");
        psx_break_opts o = *opts;
        if (!(o.strength > 0)) throw std::invalid_argument("Failure strength must be positive.");
        if (!(o.E > 0) || !(o.h > 0)) throw std::invalid_argument("E and h must be positive");
        if (o.tol <= 0) o.tol = 1e-6;
        if (o.max_iter <= 0) o.max_iter = 3000;
        if (o.precision < PSX_PREC_FP64 || o.precision > PSX_PREC_MIXED) throw std::invalid_argument("unknown precision");
        const int nx = part->nx, ny = part->ny, nz = part->nz;
        if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("The structural grid must have three positive dimensions.");
        const int64_t nV = int64_t(nx) * ny * nz, NX = nx + 1, NY = ny + 1, nNodes = NX * NY * (nz + 1), nDof = 3 * nNodes;
        if (nV > int64_t(INT32_MAX) - 2) throw std::invalid_argument("grid too large for 32-bit voxel indices");
        abi::checkFinite(f, nDof, "Forces");
        *out = psx_break_result{};
        // heldNodes (static_study.cpp:17-21)
        std::vector<uint8_t> held(static_cast<size_t>(nNodes));
        for (int64_t n = 0; n < nNodes; n++) held[size_t(n)] = part->bc[3 * n] | part->bc[3 * n + 1] | part->bc[3 * n + 2];
        DevBufs D;
        HIPCHK(hipStreamCreateWithFlags(&D.s, hipStreamNonBlocking));
        const Grid3 g{nx, ny, nz};
        uint8_t* dHeld = D.get<uint8_t>(size_t(nNodes));
        float* dCur = D.get<float>(size_t(nV));
        float* dNext = D.get<float>(size_t(nV));
        int32_t* dLost = D.get<int32_t>(size_t(nV));
        int32_t* dIds = D.get<int32_t>(size_t(nV));
        float* dMeasure = D.get<float>(size_t(nV));
        float* dCandM = D.get<float>(size_t(nV));
        double* dF = D.get<double>(size_t(nDof));
        double* dX = D.get<double>(size_t(nDof));
        double* dU = D.get<double>(size_t(nDof));
        float* dVm = D.get<float>(size_t(nNodes));
        float* dP1 = D.get<float>(size_t(nNodes));
        float* dP3 = D.get<float>(size_t(nNodes));
        unsigned* dSmax = D.get<unsigned>(2);
        unsigned long long* dBest = D.get<unsigned long long>(1);
        int* dCount = D.get<int>(1);
        float* dUf = o.fields ? D.get<float>(size_t(nDof)) : nullptr;
        float* dVmf = o.fields ? D.get<float>(size_t(nNodes)) : nullptr;
        HIPCHK(hipMemcpyAsync(dHeld, held.data(), size_t(nNodes), hipMemcpyHostToDevice, D.s));
        HIPCHK(hipMemcpyAsync(dNext, part->density, size_t(nV) * 4, hipMemcpyHostToDevice, D.s));
        HIPCHK(hipMemcpyAsync(dF, f, size_t(nDof) * 8, hipMemcpyHostToDevice, D.s));
        Pruner P(D.s, nV);
        int kept = 0;
        out->removed = P.run(g, dNext, dHeld, dCur, &kept);
        std::vector<float> density(static_cast<size_t>(nV));
        HIPCHK(hipMemcpyAsync(density.data(), dCur, size_t(nV) * 4, hipMemcpyDeviceToHost, D.s));
        HIPCHK(hipStreamSynchronize(D.s));
        out->reason = PSX_BREAK_STEP_LIMIT;
        bool warm = false;
        std::vector<float> hMeasure, hUf, hVmf;
        std::vector<int32_t> cand, cracked, detached;
        for (int step = 0; step < o.max_steps; step++) {
            if (kept == 0) {
                out->reason = step > 0 ? PSX_BREAK_SEPARATED : PSX_BREAK_UNSUPPORTED;
                break;
            }
            psx_model_desc desc = *part;
            desc.density = density.data();
            desc.build_on_device = 1;
            Model M(desc);
            const Level& L = M.hostLevel0();
            psx_solve_opts so{};
            so.tol = o.tol;
            so.max_iter = o.max_iter;
            so.precision = o.precision;
            psx_solve_info info{};
            M.solveDevice(dF, warm ? dX : nullptr, so, dX, info);
            warm = true;
            HIPCHK(hipStreamSynchronize(M.stream));
            const int64_t nE = int64_t(L.elemVox.size());
            if (nE == 0 || lostLoadFraction(L, f) > 0.5) {
                out->reason = step > 0 ? PSX_BREAK_SEPARATED : PSX_BREAK_UNSUPPORTED;
                break;
            }
            if (!info.converged) {
                out->reason = PSX_BREAK_NOT_CONVERGED;
                break;
            }
            // physical displacements and their largest magnitude, the nodal stresses
            HIPCHK(hipMemsetAsync(dBest, 0, 8, D.s));
            HIPCHK(hipMemsetAsync(dSmax, 0, 8, D.s));
            HIPCHK(hipMemsetAsync(dCount, 0, 4, D.s));
            kScaleDisp<<<grid(nNodes), BLOCK, 0, D.s>>>(dX, 1 / (o.E * o.h), nNodes, dU, dBest);
            HIPCHK(hipMemcpyAsync(dIds, L.elemVox.data(), size_t(nE) * 4, hipMemcpyHostToDevice, D.s));
            HIPCHK(hipStreamSynchronize(D.s));
            M.stressMatrices(o.E, o.h);
            launchStresses(M.lv[0].dl, M.stressS, dU, o.principal ? 1 : 0, dVm, dP1, dP3, M.stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(M.stream));
            kMeasure<<<grid(nE), BLOCK, 0, D.s>>>(dIds, nE, nx, ny, o.principal ? dP1 : dVm, dMeasure, dSmax);
            HIPCHK(hipGetLastError());
            unsigned smaxBits = 0;
            unsigned long long bestBits = 0;
            HIPCHK(hipMemcpyAsync(&smaxBits, dSmax, 4, hipMemcpyDeviceToHost, D.s));
            HIPCHK(hipMemcpyAsync(&bestBits, dBest, 8, hipMemcpyDeviceToHost, D.s));
            HIPCHK(hipStreamSynchronize(D.s));
            float smaxF;
            std::memcpy(&smaxF, &smaxBits, 4);
            const double smax = smaxF;
            if (!(smax > 0)) {
                out->reason = PSX_BREAK_NO_STRESS;
                break;
            }
            double best;
            std::memcpy(&best, &bestBits, 8);
            const double lambda = o.strength / smax;
            out->peak = std::max(out->peak, lambda);
            const size_t cap = std::max<size_t>(1, size_t(std::lround(0.02 * std::pow(double(nE), 2.0 / 3))));
            // the crack: element indices, worst first
            std::vector<int32_t> sel;
            if (o.tie_order == 0) {
                hMeasure.resize(size_t(nE));
                HIPCHK(hipMemcpyAsync(hMeasure.data(), dMeasure, size_t(nE) * 4, hipMemcpyDeviceToHost, D.s));
                HIPCHK(hipStreamSynchronize(D.s));
                sel = crackSelect(hMeasure.data(), size_t(nE), smax, cap);
            } else {
                kCandidates<<<grid(nE), BLOCK, 0, D.s>>>(dMeasure, nE, smax, dLost, dCandM, dCount);
                HIPCHK(hipGetLastError());
                int nc = 0;
                HIPCHK(hipMemcpyAsync(&nc, dCount, 4, hipMemcpyDeviceToHost, D.s));
                HIPCHK(hipStreamSynchronize(D.s));
                cand.resize(size_t(nc));
                hMeasure.resize(size_t(nc));
                if (nc > 0) {
                    HIPCHK(hipMemcpyAsync(cand.data(), dLost, size_t(nc) * 4, hipMemcpyDeviceToHost, D.s));
                    HIPCHK(hipMemcpyAsync(hMeasure.data(), dCandM, size_t(nc) * 4, hipMemcpyDeviceToHost, D.s));
                }
                HIPCHK(hipStreamSynchronize(D.s));
                // worst first, equal measures by increasing element (= voxel) index
                std::vector<size_t> idx(cand.size());
                std::iota(idx.begin(), idx.end(), size_t(0));
                std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
                    return hMeasure[a] != hMeasure[b] ? hMeasure[a] > hMeasure[b] : cand[a] < cand[b];
                });
                for (size_t t = 0; t < idx.size() && sel.size() < cap; t++) sel.push_back(cand[idx[t]]);
            }
            cracked.resize(sel.size());
            for (size_t t = 0; t < sel.size(); t++) cracked[t] = L.elemVox[size_t(sel[t])];
            // next = density without the crack, then pruned; detached = cut off by the crack
            for (int32_t e : cracked) density[size_t(e)] = 0.f;
            HIPCHK(hipMemcpyAsync(dIds, cracked.data(), cracked.size() * 4, hipMemcpyHostToDevice, D.s));
            kZeroVoxels<<<grid(int64_t(cracked.size())), BLOCK, 0, D.s>>>(dIds, int(cracked.size()), dCur);
            P.run(g, dCur, dHeld, dNext, &kept, dLost, &detached);
            std::swap(dCur, dNext);
            for (int32_t e : detached) density[size_t(e)] = 0.f;
            psx_break_step bs{};
            bs.step = step;
            bs.lambda = lambda;
            bs.max_disp = std::sqrt(best) * lambda;
            bs.voxels = int(nE);
            bs.cracked = cracked.data();
            bs.n_cracked = int(cracked.size());
            bs.detached = detached.data();
            bs.n_detached = int(detached.size());
            if (o.fields) {
                kFields<<<grid(nNodes), BLOCK, 0, D.s>>>(dU, dVm, nNodes, lambda, dUf, dVmf);
                HIPCHK(hipGetLastError());
                hUf.resize(size_t(nDof));
                hVmf.resize(size_t(nNodes));
                HIPCHK(hipMemcpyAsync(hUf.data(), dUf, size_t(nDof) * 4, hipMemcpyDeviceToHost, D.s));
                HIPCHK(hipMemcpyAsync(hVmf.data(), dVmf, size_t(nNodes) * 4, hipMemcpyDeviceToHost, D.s));
                HIPCHK(hipStreamSynchronize(D.s));
                bs.u = hUf.data();
                bs.node_vm = hVmf.data();
                bs.active_node = L.active.data();
            }
            out->steps = step + 1;
            if (o.on_step && o.on_step(&bs, o.user)) {
                out->reason = PSX_BREAK_STOPPED;
                break;
            }
        }
        return PSX_OK;
    });
}

}  // extern "C"
