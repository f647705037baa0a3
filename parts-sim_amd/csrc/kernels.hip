// HIP kernels of the voxel FEA solve for gfx950 (MI355X): the matrix-free hex8 operator with its
// fused Chebyshev / residual epilogues, grid transfers, the coarsest dense solve, and the
// conjugate-gradient vector steps with deterministic FP64 reductions.
//
// Operator (native/src/fea/voxel_fea.cpp:390-504 is the reference form): one thread per node gathers
// its 27 neighbours. A node whose 8 voxels are all s * Kb uses the assembled 27-point stencil
// (243 multiply-adds); any other node sums the rows of its up to 8 voxels (72 multiply-adds each),
// from s * Kb or from the voxel's own matrix. Nothing is scattered, so there are no write conflicts.
//
// Reductions: every block reduces with wavefront shuffles (64 lanes) and LDS and writes one partial;
// a one-block finish kernel sums the partials in block order. The grid size depends only on the
// vector length, so every sum is bitwise reproducible and independent of what else runs on the GPU.
//
// Every solve kernel also serves multi-instance launches (gridDim.y right-hand sides of one level):
// the block's y index selects the instance's slice of each vector and its CgState. The solve path's
// kernels are bodies (run(VBlk, ...)) behind named kernels, so packed launches (pack_dev.hpp) can
// run them for several models at once.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <type_traits>

#include "device.hpp"
#include "pack_dev.hpp"

namespace psx {

namespace {

constexpr int BLOCK = 256;

__device__ constexpr int d_corner(int x, int y, int z) { return (y ? (x ? 2 : 3) : (x ? 1 : 0)) + 4 * z; }
// corner offsets as compile-time tables for the unrolled loops
__device__ constexpr int kCX(int b) { return (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0; }
__device__ constexpr int kCY(int b) { return (b == 2 || b == 3 || b == 6 || b == 7) ? 1 : 0; }
__device__ constexpr int kCZ(int b) { return b >= 4 ? 1 : 0; }

template <class T> struct Prec;
template <> struct Prec<double> {
    __device__ static const double* Kb(const DevLevel& L) { return L.Kb64; }
    __device__ static const double* St(const DevLevel& L) { return L.St64; }
    __device__ static const double* invD(const DevLevel& L) { return L.invD64; }
};
template <> struct Prec<float> {
    __device__ static const float* Kb(const DevLevel& L) { return L.Kb32; }
    __device__ static const float* St(const DevLevel& L) { return L.St32; }
    __device__ static const float* invD(const DevLevel& L) { return L.invD32; }
};

__device__ __forceinline__ bool stopped(const int* stop) { return stop && __builtin_nontemporal_load(stop) != 0; }
// multi-instance launches: the instance's stop flag and the start of its slice of a vector
__device__ __forceinline__ const int* instFlag(VBlk vb, const int* f) { return f ? f + int64_t(vb.y) * CG_INTS : nullptr; }
__device__ __forceinline__ int64_t instOffset(VBlk vb, int64_t n) { return int64_t(vb.y) * n; }

// ------------------------------------------------------------------------------------------------
// reductions

__device__ __forceinline__ double waveSum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sums v[0..NV) over the block of NTH threads; the result is valid in thread 0
template <int NV, int NTH = BLOCK> __device__ __forceinline__ void blockSum(double (&v)[NV]) {
    __shared__ double sh[NV][NTH / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = waveSum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k][wid] = v[k];
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = waveSum(lane < NTH / 64 ? sh[k][lane] : 0.0);
    }
    __syncthreads();
}

// Grid reductions in two kernels: every block writes its partial sums (plain stores, no fence or
// atomic per block), then a one-block kernel adds the partials in block order and applies fin().
// The kernel boundary makes the partials visible; the order is fixed, so results are bitwise
// reproducible.
template <int NV> __device__ __forceinline__ void blockPartials(VBlk vb, double (&v)[NV], Reduce red) {
    blockSum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) red.partials[int64_t(vb.y) * red.stride + k * vb.gx + vb.x] = v[k];
}

// guard: skip when *guard != 0 (the producer kernel returned early for the same reason)
// one block of FIN_BLOCK threads per instance (gridDim.x = instances): the partials are read by one
// compute unit, so the wider block keeps more loads in flight
constexpr int FIN_BLOCK = 1024;
template <int NV, class Fin> struct Finish {
    static constexpr int NTH = FIN_BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, Reduce red, unsigned nb, const int* guard, Fin fin);
};
template <int NV, class Fin> __global__ __launch_bounds__(FIN_BLOCK) void kFinish(Reduce red, unsigned nb, const int* guard, Fin fin) {
    Finish<NV, Fin>::run(ownBlock(), red, nb, guard, fin);
}
template <int NV, class Fin> __device__ void Finish<NV, Fin>::run(VBlk vb, Reduce red, unsigned nb, const int* guard, Fin fin) {
    const int inst = vb.x;
    if (guard && guard[int64_t(inst) * CG_INTS]) return;
    red.partials += int64_t(inst) * red.stride;
    // 8 independent partial sums per thread keep 8 loads in flight; they are combined in a fixed
    // order, so the result does not depend on timing
    constexpr int U = 8;
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double a[U];
#pragma unroll
        for (int u = 0; u < U; u++) a[u] = 0;
        for (unsigned b0 = 0; b0 < nb; b0 += FIN_BLOCK * U) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const unsigned b = b0 + u * FIN_BLOCK + threadIdx.x;
                if (b < nb) a[u] += red.partials[k * nb + b];
            }
        }
        s[k] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    }
    blockSum<NV, FIN_BLOCK>(s);
    if (threadIdx.x == 0) fin(s, inst);
}

// reductions: each block owns a contiguous tile of RED_TILE elements, 8 per thread, loads issued
// together (unrolled) so every thread keeps 8 independent requests in flight
inline unsigned reduceGrid(int64_t n) {
    const int64_t g = (n + RED_TILE - 1) / RED_TILE;
    return unsigned(g < 1 ? 1 : g);
}
#define TILE_LOOP(body)                                                                  \
    _Pragma("unroll") for (int k_ = 0; k_ < RED_TILE / BLOCK; k_++) {                    \
        const int64_t i = int64_t(vb.x) * RED_TILE + threadIdx.x + k_ * BLOCK;          \
        if (i < n) { body }                                                              \
    }
inline unsigned flatGrid(int64_t n) { return unsigned((n + BLOCK - 1) / BLOCK); }

// ------------------------------------------------------------------------------------------------
// operator

// acc = (K x) at node n = (i, j, k)
template <class T>
__device__ __forceinline__ void applyNode(const DevLevel& L, const T* __restrict__ x, int i, int j, int k, int64_t n, T acc[3]) {
    acc[0] = acc[1] = acc[2] = T(0);
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
    const float s0 = L.nodeScale[n];
    if (s0 > 0.f) {
        const T* __restrict__ S = Prec<T>::St(L);
#pragma unroll
        for (int m = 0; m < 27; m++) {
            const int ox = m % 3 - 1, oy = (m / 3) % 3 - 1, oz = m / 9 - 1;
            const T* xm = x + 3 * (n + ox + oy * NX + oz * NXY);
            const T x0 = xm[0], x1 = xm[1], x2 = xm[2];
#pragma unroll
            for (int d = 0; d < 3; d++) acc[d] += S[m * 9 + d * 3] * x0 + S[m * 9 + d * 3 + 1] * x1 + S[m * 9 + d * 3 + 2] * x2;
        }
        const T s = T(s0);
        acc[0] *= s;
        acc[1] *= s;
        acc[2] *= s;
    } else {
#pragma unroll
        for (int v = 0; v < 8; v++) {
            const int di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
            const int ei = i - 1 + di, ej = j - 1 + dj, ek = k - 1 + dk;
            if (ei < 0 || ej < 0 || ek < 0 || ei >= L.nx || ej >= L.ny || ek >= L.nz) continue;
            const int32_t w = L.words[ei + int64_t(L.nx) * (ej + int64_t(L.ny) * ek)];
            if (w == 0) continue;
            const int a = d_corner(1 - di, 1 - dj, 1 - dk);
            const int64_t nb = n - (1 - di) - (1 - dj) * NX - (1 - dk) * NXY;  // the voxel's corner 0
            T xe[24];
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const T* xm = x + 3 * (nb + kCX(b) + kCY(b) * NX + kCZ(b) * NXY);
                xe[3 * b] = xm[0];
                xe[3 * b + 1] = xm[1];
                xe[3 * b + 2] = xm[2];
            }
            T b0 = 0, b1 = 0, b2 = 0;
            if (w > 0) {
                const T* __restrict__ K = Prec<T>::Kb(L) + 72 * a;
#pragma unroll
                for (int c = 0; c < 24; c++) {
                    b0 += K[c] * xe[c];
                    b1 += K[24 + c] * xe[c];
                    b2 += K[48 + c] * xe[c];
                }
                const T s = T(__int_as_float(w));
                acc[0] += s * b0;
                acc[1] += s * b1;
                acc[2] += s * b2;
            } else {
                const float4* __restrict__ K4 = reinterpret_cast<const float4*>(L.Kown + int64_t(w & 0x7fffffff) * 576 + 72 * a);
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    const float4 r0 = K4[q], r1 = K4[6 + q], r2 = K4[12 + q];
                    b0 += T(r0.x) * xe[4 * q] + T(r0.y) * xe[4 * q + 1] + T(r0.z) * xe[4 * q + 2] + T(r0.w) * xe[4 * q + 3];
                    b1 += T(r1.x) * xe[4 * q] + T(r1.y) * xe[4 * q + 1] + T(r1.z) * xe[4 * q + 2] + T(r1.w) * xe[4 * q + 3];
                    b2 += T(r2.x) * xe[4 * q] + T(r2.y) * xe[4 * q + 1] + T(r2.z) * xe[4 * q + 2] + T(r2.w) * xe[4 * q + 3];
                }
                acc[0] += b0;
                acc[1] += b1;
                acc[2] += b2;
            }
        }
    }
    if (L.diagAdd) {
#pragma unroll
        for (int d = 0; d < 3; d++) acc[d] += T(L.diagAdd[3 * n + d]) * x[3 * n + d];
    }
}

template <class T, int MODE> struct Apply {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, DevLevel L, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ r, const T* __restrict__ zp, T c1, T c2,
                               const int* stop);
};
template <class T, int MODE>
__global__ __launch_bounds__(BLOCK) void kApply(DevLevel L, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ r, const T* __restrict__ zp,
                                                 T c1, T c2, const int* stop) {
    Apply<T, MODE>::run(ownBlock(), L, x, y, r, zp, c1, c2, stop);
}
template <class T, int MODE>
__device__ void Apply<T, MODE>::run(VBlk vb, DevLevel L, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ r, const T* __restrict__ zp,
                                    T c1, T c2, const int* stop) {
    stop = instFlag(vb, stop);
    if (stopped(stop)) return;
    {
        const int64_t io = instOffset(vb, L.nDof);
        x += io;
        y += io;
        if (r) r += io;
        if (zp) zp += io;
    }
    const int64_t n = int64_t(vb.x) * BLOCK + threadIdx.x;
    if (n >= L.nNodes) return;
    const unsigned nn = unsigned(n), NX = unsigned(L.NX), NY = unsigned(L.NY);
    const int i = int(nn % NX), j = int((nn / NX) % NY), k = int(nn / (NX * NY));
    T acc[3];
    applyNode<T>(L, x, i, j, k, n, acc);
    const T* __restrict__ invD = Prec<T>::invD(L);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int64_t q = 3 * n + d;
        const T id = invD[q];
        if constexpr (MODE == AM_APPLY) y[q] = id != T(0) ? acc[d] : T(0);
        else if constexpr (MODE == AM_RAW) y[q] = acc[d];
        else if constexpr (MODE == AM_RESID) y[q] = id != T(0) ? r[q] - acc[d] : T(0);
        else if constexpr (MODE == AM_CHEB1) y[q] = x[q] + c1 * id * (r[q] - acc[d]);
        else y[q] = x[q] + c1 * (zp ? x[q] - zp[q] : x[q]) + c2 * id * (r[q] - acc[d]);
    }
}

// The same product for levels with own element matrices, 8 lanes per node: lane v takes voxel v
// (3 rows x 24 columns from s * Kb in LDS or from the voxel's own matrix), the 8 partial rows are
// summed with shuffles inside the aligned 8-lane group, lane 0 writes. Coarse levels are small:
// this spreads one node's dependent loads over 8 lanes instead of one thread walking 8 voxels.
template <class T, int MODE, bool H16> struct ApplyOwn {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, DevLevel L, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ r, const T* __restrict__ zp, T c1, T c2,
                               const int* stop, int ninst);
};
template <class T, int MODE, bool H16 = false>
__global__ __launch_bounds__(BLOCK) void kApplyOwn(DevLevel L, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ r, const T* __restrict__ zp,
                                                   T c1, T c2, const int* stop, int ninst) {
    ApplyOwn<T, MODE, H16>::run(ownBlock(), L, x, y, r, zp, c1, c2, stop, ninst);
}
template <class T, int MODE, bool H16>
__device__ void ApplyOwn<T, MODE, H16>::run(VBlk vb, DevLevel L, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ r,
                                            const T* __restrict__ zp, T c1, T c2, const int* stop, int ninst) {
    // Several right-hand sides: this lane's element rows (72 coefficients) are loaded once and
    // applied to every instance in turn (instance-major vectors of nDof each, a stop flag per
    // instance), so the own matrices stream from HBM once per launch instead of once per instance.
    __shared__ T Kb[576];
    for (int q = threadIdx.x; q < 576; q += BLOCK) Kb[q] = Prec<T>::Kb(L)[q];
    __syncthreads();
    const int64_t n = int64_t(vb.x) * (BLOCK / 8) + threadIdx.x / 8;
    const int v = threadIdx.x & 7;
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
    bool has = false, inVox = false;
    int32_t w = 0;
    int64_t nb = 0;
    int a = 0;
    if (n < L.nNodes) {
        const unsigned nn = unsigned(n), NXu = unsigned(L.NX), NYu = unsigned(L.NY);
        const int i = int(nn % NXu), j = int((nn / NXu) % NYu), k = int(nn / (NXu * NYu));
        const int di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
        const int ei = i - 1 + di, ej = j - 1 + dj, ek = k - 1 + dk;
        if (ei >= 0 && ej >= 0 && ek >= 0 && ei < L.nx && ej < L.ny && ek < L.nz) {
            w = L.words[ei + int64_t(L.nx) * (ej + int64_t(L.ny) * ek)];
            inVox = true;
            a = d_corner(1 - di, 1 - dj, 1 - dk);
            nb = ei + NX * ej + NXY * ek;
        }
    }
    // the first instance's operands do not depend on the voxel word: the element's corner values and
    // (one lane per node) the node's own vectors are loaded beside it, so a wavefront waits for the
    // word, then the element rows, instead of four memory round trips in turn
    struct alignas(sizeof(T)) V3 {
        T c[3];
    };
    const T* __restrict__ invD = Prec<T>::invD(L);
    T xe[24];
    V3 idv{}, xv{}, rv{}, zv{};
    auto loadCorners = [&](int64_t io) {
        const T* __restrict__ xi = x + io;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const T* xm = xi + 3 * (nb + kCX(b) + kCY(b) * NX + kCZ(b) * NXY);
            xe[3 * b] = xm[0];
            xe[3 * b + 1] = xm[1];
            xe[3 * b + 2] = xm[2];
        }
    };
    auto loadNode = [&](int64_t io) {
        idv = *reinterpret_cast<const V3*>(invD + 3 * n);
        if constexpr (MODE == AM_CHEB1 || MODE == AM_CHEB2) xv = *reinterpret_cast<const V3*>(x + io + 3 * n);
        if constexpr (MODE == AM_RESID || MODE == AM_CHEB1 || MODE == AM_CHEB2) rv = *reinterpret_cast<const V3*>(r + io + 3 * n);
        if constexpr (MODE == AM_CHEB2)
            if (zp) zv = *reinterpret_cast<const V3*>(zp + io + 3 * n);
    };
    // (measured on the 512x128x128 block: level 2 FP32 28.7 -> 27.8 us, the mixed solve 105.0 ->
    // 104.2 ms. FP64 loads them after the rows: its prefetch registers cost the later Chebyshev step a
    // wavefront per SIMD, and the FP64 solve did not gain)
    constexpr bool PREFETCH = sizeof(T) == 4;
    if constexpr (PREFETCH) {
        if (inVox) loadCorners(0);
        if (v == 0 && n < L.nNodes) loadNode(0);
    }
    has = w != 0;
    // the node's 3 rows of the voxel's element: s * Kb (scaled after the sum) or the own matrix
    T K[72];
    // s scales the summed rows: s for s * Kb, 1 for a stored own matrix, its power-of-two scale for an
    // FP16 one
    T s = w > 0 ? T(__int_as_float(w)) : T(1);
    if (has) {
        if (w > 0) {
#pragma unroll
            for (int c = 0; c < 72; c++) K[c] = Kb[72 * a + c];
        } else if constexpr (H16) {
            const int64_t e = int64_t(w & 0x7fffffff);
            const uint4* __restrict__ H4 = reinterpret_cast<const uint4*>(L.KownH + e * 576 + 72 * a);  // 144-byte rows: 16-byte aligned
            s = T(L.ownScale[e]);
#pragma unroll
            for (int q = 0; q < 9; q++) {
                const uint4 u = H4[q];
                const uint32_t wd[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                for (int h = 0; h < 4; h++) {
                    K[8 * q + 2 * h] = T(__half2float(__ushort_as_half(uint16_t(wd[h] & 0xffffu))));
                    K[8 * q + 2 * h + 1] = T(__half2float(__ushort_as_half(uint16_t(wd[h] >> 16))));
                }
            }
        } else {
            const float4* __restrict__ K4 = reinterpret_cast<const float4*>(L.Kown + int64_t(w & 0x7fffffff) * 576 + 72 * a);
#pragma unroll
            for (int q = 0; q < 18; q++) {
                const float4 r4 = K4[q];
                K[4 * q] = T(r4.x);
                K[4 * q + 1] = T(r4.y);
                K[4 * q + 2] = T(r4.z);
                K[4 * q + 3] = T(r4.w);
            }
        }
    }
    for (int inst = 0; inst < ninst; inst++) {
        if (stopped(stop ? stop + int64_t(inst) * CG_INTS : nullptr)) continue;  // uniform
        const int64_t io = int64_t(inst) * L.nDof;
        const T* __restrict__ xi = x + io;
        const bool load = inst > 0 || !PREFETCH;
        T b0 = 0, b1 = 0, b2 = 0;
        if (has) {
            if (load) loadCorners(io);
#pragma unroll
            for (int c = 0; c < 24; c++) {
                b0 += K[c] * xe[c];
                b1 += K[24 + c] * xe[c];
                b2 += K[48 + c] * xe[c];
            }
            b0 *= s;
            b1 *= s;
            b2 *= s;
        }
#pragma unroll
        for (int o = 4; o >= 1; o >>= 1) {
            b0 += __shfl_xor(b0, o, 64);
            b1 += __shfl_xor(b1, o, 64);
            b2 += __shfl_xor(b2, o, 64);
        }
        if (v == 0 && n < L.nNodes) {
            T acc[3] = {b0, b1, b2};
            if (L.diagAdd) {
#pragma unroll
                for (int d = 0; d < 3; d++) acc[d] += T(L.diagAdd[3 * n + d]) * xi[3 * n + d];
            }
            if (load) loadNode(io);
            // the node's result as one three-component store
            V3 out;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const T id = idv.c[d];
                if constexpr (MODE == AM_APPLY) out.c[d] = id != T(0) ? acc[d] : T(0);
                else if constexpr (MODE == AM_RAW) out.c[d] = acc[d];
                else if constexpr (MODE == AM_RESID) out.c[d] = id != T(0) ? rv.c[d] - acc[d] : T(0);
                else if constexpr (MODE == AM_CHEB1) out.c[d] = xv.c[d] + c1 * id * (rv.c[d] - acc[d]);
                else out.c[d] = xv.c[d] + c1 * (zp ? xv.c[d] - zv.c[d] : xv.c[d]) + c2 * id * (rv.c[d] - acc[d]);
            }
            *reinterpret_cast<V3*>(y + io + 3 * n) = out;
        }
    }
}

// The own-element product on the matrix cores (FP32 vectors: the mixed V-cycle's smoothing steps on
// levels with own element matrices), one wavefront per node and up to 16 instances per pass. For each
// of the node's 8 voxels one v_mfma_f32_16x16x32_bf16 accumulates D += A B: A = the node's 3 rows of the
// voxel's element matrix (s Kb or the own matrix; rows 3..15 and columns 24..31 zero) and B = the
// element's 24 corner values minus the node's own, one instance per column. The element rows sum to
// zero over a rigid translation, so the differences carry the product (the BF16 rounding acts on them,
// not on the values). BF16 operands, FP32 accumulation. Lane l holds A row l % 16 and B column l % 16
// over the K slice 8 (l / 16) .. + 7; after the 8 voxels lanes 0..15 hold D rows 0..3 of column l, the
// node's 3 rows for instance l. An instance's column is computed the same way for any instance count,
// so a batch equals its lone solves bit for bit.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

template <int MODE> struct ApplyOwnMf {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, DevLevel L, const float* __restrict__ x, float* __restrict__ y, const float* __restrict__ r,
                               const float* __restrict__ zp, float c1, float c2, const int* stop, int ninst);
};
template <int MODE>
__global__ __launch_bounds__(BLOCK) void kApplyOwnMf(DevLevel L, const float* __restrict__ x, float* __restrict__ y, const float* __restrict__ r,
                                                    const float* __restrict__ zp, float c1, float c2, const int* stop, int ninst) {
    ApplyOwnMf<MODE>::run(ownBlock(), L, x, y, r, zp, c1, c2, stop, ninst);
}
template <int MODE>
__device__ void ApplyOwnMf<MODE>::run(VBlk vb, DevLevel L, const float* __restrict__ x, float* __restrict__ y, const float* __restrict__ r,
                                      const float* __restrict__ zp, float c1, float c2, const int* stop, int ninst) {
    __shared__ float Kb[576];
    for (int q = threadIdx.x; q < 576; q += BLOCK) Kb[q] = L.Kb32[q];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t n = int64_t(vb.x) * (BLOCK / 64) + (threadIdx.x >> 6);
    if (n >= L.nNodes) return;  // the whole wavefront
    const unsigned nn = unsigned(n), NXu = unsigned(L.NX), NYu = unsigned(L.NY);
    const int i = int(nn % NXu), j = int((nn / NXu) % NYu), k = int(nn / (NXu * NYu));
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
    const int col = lane & 15, kb = lane >> 4;
    for (int g0 = 0; g0 < ninst; g0 += 16) {
        const int inst = g0 + col;
        const bool act = inst < ninst && !stopped(stop ? stop + int64_t(inst) * CG_INTS : nullptr);
        const int64_t io = int64_t(act ? inst : 0) * L.nDof;
        const float* __restrict__ xi = x + io;
        float xn[3] = {0.f, 0.f, 0.f};
        if (act) {
            xn[0] = xi[3 * n];
            xn[1] = xi[3 * n + 1];
            xn[2] = xi[3 * n + 2];
        }
        floatx4_t acc = {0.f, 0.f, 0.f, 0.f};
        for (int v = 0; v < 8; v++) {  // wave-uniform: one node per wavefront
            const int di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
            const int ei = i - 1 + di, ej = j - 1 + dj, ek = k - 1 + dk;
            if (ei < 0 || ej < 0 || ek < 0 || ei >= L.nx || ej >= L.ny || ek >= L.nz) continue;
            const int32_t w = L.words[ei + int64_t(L.nx) * (ej + int64_t(L.ny) * ek)];
            if (w == 0) continue;
            const int a = d_corner(1 - di, 1 - dj, 1 - dk);
            float ta[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (col < 3 && kb < 3) {
                const int at = (3 * a + col) * 24 + 8 * kb;  // 16-byte aligned (rows of 24 floats)
                if (w > 0) {
                    const float s = __int_as_float(w);
#pragma unroll
                    for (int t = 0; t < 8; t++) ta[t] = s * Kb[at + t];
                } else {
                    const float4* __restrict__ K4 = reinterpret_cast<const float4*>(L.Kown + int64_t(w & 0x7fffffff) * 576 + at);
                    const float4 p = K4[0], q = K4[1];
                    ta[0] = p.x, ta[1] = p.y, ta[2] = p.z, ta[3] = p.w, ta[4] = q.x, ta[5] = q.y, ta[6] = q.z, ta[7] = q.w;
                }
            }
            float tb[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (act && kb < 3) {
                const int64_t nb = ei + NX * ej + NXY * ek;  // the element's corner-0 node
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    // K index 8 kb + t = corner q, component cc
                    const int kk = 8 * kb + t, q = kk / 3, cc = kk - 3 * q;
                    const int64_t node = nb + kCX(q) + kCY(q) * NX + kCZ(q) * NXY;
                    tb[t] = xi[3 * node + cc] - xn[cc];
                }
            }
            bf16x8_t A, B;
#pragma unroll
            for (int t = 0; t < 8; t++) {
                A[t] = (__bf16)ta[t];
                B[t] = (__bf16)tb[t];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
        }
        if (kb == 0 && act) {
            float out[3] = {acc[0], acc[1], acc[2]};
            if (L.diagAdd) {
#pragma unroll
                for (int d = 0; d < 3; d++) out[d] += float(L.diagAdd[3 * n + d]) * xn[d];
            }
            const float* __restrict__ idv = L.invD32 + 3 * n;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const float id = idv[d];
                float o;
                if constexpr (MODE == AM_APPLY) o = id != 0.f ? out[d] : 0.f;
                else if constexpr (MODE == AM_RAW) o = out[d];
                else if constexpr (MODE == AM_RESID) o = id != 0.f ? r[io + 3 * n + d] - out[d] : 0.f;
                else if constexpr (MODE == AM_CHEB1) o = xn[d] + c1 * id * (r[io + 3 * n + d] - out[d]);
                else o = xn[d] + c1 * (zp ? xn[d] - zp[io + 3 * n + d] : xn[d]) + c2 * id * (r[io + 3 * n + d] - out[d]);
                y[io + 3 * n + d] = o;
            }
        }
    }
}

// FP16 copy of the own matrices, one wavefront per element: the element's largest magnitude sets a
// power-of-two scale 2^e with max / 2^e in (2^13, 2^14] (the rule of the reference's GPU solver,
// gpu_fea.cpp:291-299, per element instead of per level so soft elements keep their precision)
__global__ __launch_bounds__(64) void kOwnToHalf(const float* __restrict__ Kown, int64_t n, uint16_t* __restrict__ half, float* __restrict__ scale) {
    const int64_t e = blockIdx.x;
    if (e >= n) return;
    const float* __restrict__ K = Kown + e * 576;
    float m = 0.f;
    for (int q = threadIdx.x; q < 576; q += 64) m = fmaxf(m, fabsf(K[q]));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    const float sc = m > 0.f ? exp2f(ceilf(log2f(m / 16384.f))) : 1.f;
    const float inv = 1.f / sc;  // exact: a power of two
    for (int q = threadIdx.x; q < 576; q += 64) half[e * 576 + q] = __half_as_ushort(__float2half_rn(K[q] * inv));
    if (threadIdx.x == 0) scale[e] = sc;
}

template <class T> struct ChebFirst {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, const T* __restrict__ invD, const T* __restrict__ r, T* __restrict__ z, T c0, int64_t n, const int* stop) {
        stop = instFlag(vb, stop);
        if (stopped(stop)) return;
        const int64_t q = int64_t(vb.x) * BLOCK + threadIdx.x;
        if (q >= n) return;
        const int64_t io = instOffset(vb, n);
        z[io + q] = c0 * invD[q] * r[io + q];
    }
};
template <class T> __global__ __launch_bounds__(BLOCK) void kChebFirst(const T* __restrict__ invD, const T* __restrict__ r, T* __restrict__ z, T c0, int64_t n,
                                                                       const int* stop) {
    ChebFirst<T>::run(ownBlock(), invD, r, z, c0, n, stop);
}

// ------------------------------------------------------------------------------------------------
// grid transfers (voxel_fea.cpp:738-815): full weighting P^T and trilinear P, coarse node I <-> fine 2I

// z0 (when not null): also the coarse level's first Chebyshev step z0 = c0 D^-1 rc, as kChebFirst
// computes it (the coarse cycle then starts from it)
template <class T> struct Restrict {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, DevLevel F, DevLevel C, const T* __restrict__ rf, T* __restrict__ rc, const int* stop, T* __restrict__ z0,
                               T c0);
};
template <class T>
__global__ __launch_bounds__(BLOCK) void kRestrict(DevLevel F, DevLevel C, const T* __restrict__ rf, T* __restrict__ rc, const int* stop,
                                                    T* __restrict__ z0, T c0) {
    Restrict<T>::run(ownBlock(), F, C, rf, rc, stop, z0, c0);
}
template <class T>
__device__ void Restrict<T>::run(VBlk vb, DevLevel F, DevLevel C, const T* __restrict__ rf, T* __restrict__ rc, const int* stop, T* __restrict__ z0,
                                 T c0) {
    stop = instFlag(vb, stop);
    if (stopped(stop)) return;
    rf += instOffset(vb, F.nDof);
    rc += instOffset(vb, C.nDof);
    if (z0) z0 += instOffset(vb, C.nDof);
    const int64_t cn = int64_t(vb.x) * BLOCK + threadIdx.x;
    if (cn >= C.nNodes) return;
    const unsigned u = unsigned(cn), CNX = unsigned(C.NX), CNY = unsigned(C.NY);
    const int I = int(u % CNX), J = int((u / CNX) % CNY), K = int(u / (CNX * CNY));
    const T* __restrict__ cinv = Prec<T>::invD(C);
    const T i0 = cinv[3 * cn], i1 = cinv[3 * cn + 1], i2 = cinv[3 * cn + 2];
    T s0 = 0, s1 = 0, s2 = 0;
    {
        // the 27 fine nodes around 2I, all loads issued together and beside the coarse inverse
        // diagonal (a held coarse DOF discards its sum below): a neighbour outside the fine grid
        // (2I itself is outside for the last coarse node of an odd fine count) reads the nearest
        // fine node, clamped per axis, with weight 0 — an exact zero added to a sum that starts at
        // +0 leaves its bits unchanged, so the result is the sum over the in-range nodes in order
        struct alignas(sizeof(T)) V3 {
            T a, b, c;
        };
        // a 64-bit base per fine plane, 32-bit offsets in nodes within it (a plane holds < 2^31 / 3
        // nodes)
        const int64_t FNXY = int64_t(F.NX) * F.NY;
        const V3* __restrict__ rv = reinterpret_cast<const V3*>(rf);
#pragma unroll
        for (int dk = -1; dk <= 1; dk++) {
            const int kf = 2 * K + dk;
            const bool kin = kf >= 0 && kf < F.NZ;
            const int kc = min(max(kf, 0), F.NZ - 1);
            const T wk = dk == 0 ? T(1) : T(0.5);
            const V3* __restrict__ plane = rv + FNXY * kc;
#pragma unroll
            for (int dj = -1; dj <= 1; dj++) {
                const int jf = 2 * J + dj;
                const bool jin = kin && jf >= 0 && jf < F.NY;
                const int jc = min(max(jf, 0), F.NY - 1);
                const T wj = dj == 0 ? wk : T(0.5) * wk;
                const int32_t row = int32_t(F.NX) * jc;
#pragma unroll
                for (int di = -1; di <= 1; di++) {
                    const int fi = 2 * I + di;
                    const bool in = jin && fi >= 0 && fi < F.NX;
                    const int ic = min(max(fi, 0), F.NX - 1);
                    const T w = in ? (di == 0 ? wj : T(0.5) * wj) : T(0);
                    const V3 v = plane[row + ic];
                    s0 += w * v.a;
                    s1 += w * v.b;
                    s2 += w * v.c;
                }
            }
        }
    }
    const T v0 = i0 != T(0) ? s0 : T(0), v1 = i1 != T(0) ? s1 : T(0), v2 = i2 != T(0) ? s2 : T(0);
    rc[3 * cn] = v0;
    rc[3 * cn + 1] = v1;
    rc[3 * cn + 2] = v2;
    if (z0) {
        z0[3 * cn] = c0 * i0 * v0;
        z0[3 * cn + 1] = c0 * i1 * v1;
        z0[3 * cn + 2] = c0 * i2 * v2;
    }
}

// zf += P zc on the fine level's free DOFs (held ones keep their value), one thread per fine node
// gathering its 1, 2, 4 or 8 coarse nodes
template <class T> struct Prolong {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, DevLevel F, DevLevel C, const T* __restrict__ zc, T* __restrict__ zf, const int* stop);
};
template <class T>
__global__ __launch_bounds__(BLOCK) void kProlong(DevLevel F, DevLevel C, const T* __restrict__ zc, T* __restrict__ zf, const int* stop) {
    Prolong<T>::run(ownBlock(), F, C, zc, zf, stop);
}
template <class T> __device__ void Prolong<T>::run(VBlk vb, DevLevel F, DevLevel C, const T* __restrict__ zc, T* __restrict__ zf, const int* stop) {
    stop = instFlag(vb, stop);
    if (stopped(stop)) return;
    zc += instOffset(vb, C.nDof);
    zf += instOffset(vb, F.nDof);
    const int64_t fn = int64_t(vb.x) * BLOCK + threadIdx.x;
    if (fn >= F.nNodes) return;
    // held / inactive DOFs from the node info word (bits 8 + d: inverse diagonal 0)
    const uint32_t info = F.nodeInfo[fn];
    const bool f0 = !(info & 0x100u), f1 = !(info & 0x200u), f2 = !(info & 0x400u);
    if (!(f0 || f1 || f2)) return;
    // a node's three components in one load (dwordx3 / 3 x dwordx2)
    struct alignas(sizeof(T)) V3 {
        T a, b, c;
    };
    // the fine node itself, loaded ahead of the gathers (held components are written back as read)
    V3* __restrict__ dst = reinterpret_cast<V3*>(zf + 3 * fn);
    V3 v = *dst;
    const unsigned u = unsigned(fn), FNX = unsigned(F.NX), FNY = unsigned(F.NY);
    const int i = int(u % FNX), j = int((u / FNX) % FNY), k = int(u / (FNX * FNY));
    const int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
    const T w = T(1.0 / (ni * nj * nk));
    // the lowest coarse corner once in 64 bits, the other corners at 32-bit row / plane strides in
    // nodes (a plane holds < 2^31 / 3 nodes)
    const V3* __restrict__ base =
        reinterpret_cast<const V3*>(zc) + ((i >> 1) + int64_t(C.NX) * ((j >> 1) + int64_t(C.NY) * (k >> 1)));
    const int32_t oy = int32_t(C.NX), oz = int32_t(C.NX) * int32_t(C.NY);
    T s0 = 0, s1 = 0, s2 = 0;
    for (int c = 0; c < nk; c++) {
        for (int b = 0; b < nj; b++) {
            for (int a = 0; a < ni; a++) {
                const V3 q = base[a + b * oy + c * oz];
                s0 += w * q.a;
                s1 += w * q.b;
                s2 += w * q.c;
            }
        }
    }
    v.a = f0 ? v.a + s0 : v.a;
    v.b = f1 ? v.b + s1 : v.b;
    v.c = f2 ? v.c + s2 : v.c;
    *dst = v;
}

// The same, one thread per coarse node writing its 2 x 2 x 2 block of fine nodes (the 8 coarse
// nodes around the block loaded once for all 8): large levels, where it saves the per-fine-node
// gathers (level 0 of the 512x128x128 block: 73.4 -> 58.8 us, 45.3 us with the fine nodes' loads
// issued beside the coarse ones); small ones keep one thread per fine
// node (their few coarse nodes leave too few threads: 5.5 -> 8.8 us on a 18 785-node level)
template <class T> struct ProlongCells {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, DevLevel F, DevLevel C, const T* __restrict__ zc, T* __restrict__ zf, const int* stop);
};
template <class T>
__global__ __launch_bounds__(BLOCK) void kProlongCells(DevLevel F, DevLevel C, const T* __restrict__ zc, T* __restrict__ zf, const int* stop) {
    ProlongCells<T>::run(ownBlock(), F, C, zc, zf, stop);
}
template <class T> __device__ void ProlongCells<T>::run(VBlk vb, DevLevel F, DevLevel C, const T* __restrict__ zc, T* __restrict__ zf, const int* stop) {
    stop = instFlag(vb, stop);
    if (stopped(stop)) return;
    zc += instOffset(vb, C.nDof);
    zf += instOffset(vb, F.nDof);
    // one thread per coarse node (I, J, K): the fine nodes (2I + a, 2J + b, 2K + c), a, b, c in
    // {0, 1}, inside the fine grid. Fine node i takes coarse node i / 2 when i is even and the mean
    // of (i - 1) / 2 and (i + 1) / 2 when it is odd, per axis (voxel_fea.cpp:739-815); the 8 coarse
    // nodes around the block are loaded once for its 8 fine nodes
    const int64_t cn = int64_t(vb.x) * BLOCK + threadIdx.x;
    if (cn >= C.nNodes) return;
    const unsigned u = unsigned(cn), CNX = unsigned(C.NX), CNY = unsigned(C.NY);
    const int I = int(u % CNX), J = int((u / CNX) % CNY), K = int(u / (CNX * CNY));
    // a node's three components in one access (dwordx3 / 3 x dwordx2)
    struct alignas(sizeof(T)) V3 {
        T a, b, c;
    };
    const int64_t CXY = int64_t(C.NX) * C.NY, FNX = F.NX, FXY = int64_t(F.NX) * F.NY;
    // the coarse corners (x + 2 y + 4 z offsets) that odd fine neighbours need: present when that
    // fine node lies inside the grid (then the coarse node one further lies inside the coarse grid)
    const bool ox = 2 * I + 1 < F.NX, oy = 2 * J + 1 < F.NY, oz = 2 * K + 1 < F.NZ;
    // (the last coarse node of an odd fine cell count has no even fine node either)
    const bool ex = 2 * I < F.NX, ey = 2 * J < F.NY, ez = 2 * K < F.NZ;
    const V3* __restrict__ cb = reinterpret_cast<const V3*>(zc) + cn;
    V3 q[8];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int mx = m & 1, my = (m >> 1) & 1, mz = m >> 2;
        const bool in = (!mx || ox) && (!my || oy) && (!mz || oz);
        q[m] = in ? cb[mx + my * int64_t(C.NX) + mz * CXY] : V3{T(0), T(0), T(0)};
    }
    // the fine nodes' info words and values, loaded beside the coarse ones (not after each node's
    // held test)
    V3 fv[8];
    uint32_t finfo[8];
#pragma unroll
    for (int f = 0; f < 8; f++) {
        const int a = f & 1, b = (f >> 1) & 1, c = f >> 2;
        finfo[f] = 0x700u;
        if (!(a ? ox : ex) || !(b ? oy : ey) || !(c ? oz : ez)) continue;
        const int64_t fn = (2 * I + a) + FNX * (2 * J + b) + FXY * (2 * K + c);
        finfo[f] = F.nodeInfo[fn];
        fv[f] = *reinterpret_cast<const V3*>(zf + 3 * fn);
    }
#pragma unroll
    for (int f = 0; f < 8; f++) {
        const int a = f & 1, b = (f >> 1) & 1, c = f >> 2;
        if (!(a ? ox : ex) || !(b ? oy : ey) || !(c ? oz : ez)) continue;
        const int64_t fn = (2 * I + a) + FNX * (2 * J + b) + FXY * (2 * K + c);
        // held / inactive DOFs from the node info word (bits 8 + d: inverse diagonal 0)
        const uint32_t info = finfo[f];
        const bool f0 = !(info & 0x100u), f1 = !(info & 0x200u), f2 = !(info & 0x400u);
        if (!(f0 || f1 || f2)) continue;
        const T w = T(1.0 / ((a + 1) * (b + 1) * (c + 1)));
        T s0 = 0, s1 = 0, s2 = 0;
        // the coarse nodes in z, y, x order (the per-fine-node sum of the transfer's definition)
#pragma unroll
        for (int cz = 0; cz <= c; cz++)
#pragma unroll
            for (int by = 0; by <= b; by++)
#pragma unroll
                for (int ax = 0; ax <= a; ax++) {
                    const V3& t = q[ax + 2 * by + 4 * cz];
                    s0 += w * t.a;
                    s1 += w * t.b;
                    s2 += w * t.c;
                }
        V3 v = fv[f];
        v.a = f0 ? v.a + s0 : v.a;
        v.b = f1 ? v.b + s1 : v.b;
        v.c = f2 ? v.c + s2 : v.c;
        *reinterpret_cast<V3*>(zf + 3 * fn) = v;
    }
}

// z = Ainv r on the coarsest level's free DOFs and 0 on every held one (the whole vector is
// written, so nothing left in the buffer by another use leaks into the cycle): r gathered into LDS,
// one wavefront per DOF
constexpr int COARSE_MAX = 3072;
template <class T> struct Coarse {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, const int32_t* __restrict__ freeIdx, const int32_t* __restrict__ dense, int m, int64_t nDof,
                               const double* __restrict__ Ainv, const T* __restrict__ r, T* __restrict__ z, const int* stop);
};
template <class T>
__global__ __launch_bounds__(BLOCK) void kCoarse(const int32_t* __restrict__ freeIdx, const int32_t* __restrict__ dense, int m, int64_t nDof,
                                                 const double* __restrict__ Ainv, const T* __restrict__ r, T* __restrict__ z, const int* stop) {
    Coarse<T>::run(ownBlock(), freeIdx, dense, m, nDof, Ainv, r, z, stop);
}
template <class T>
__device__ void Coarse<T>::run(VBlk vb, const int32_t* __restrict__ freeIdx, const int32_t* __restrict__ dense, int m, int64_t nDof,
                               const double* __restrict__ Ainv, const T* __restrict__ r, T* __restrict__ z, const int* stop) {
    stop = instFlag(vb, stop);
    if (stopped(stop)) return;
    r += instOffset(vb, nDof);
    z += instOffset(vb, nDof);
    __shared__ double rl[COARSE_MAX];
    for (int q = threadIdx.x; q < m; q += BLOCK) rl[q] = double(r[freeIdx[q]]);
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t q = int64_t(vb.x) * (BLOCK / 64) + wid;  // one wavefront per DOF
    if (q >= nDof) return;
    const int row = dense[q];
    double s = 0;
    if (row >= 0) {
        const double* a = Ainv + int64_t(row) * m;
        double s1 = 0, s2 = 0, s3 = 0;
        int c = lane;
        for (; c + 192 < m; c += 256) {
            s += a[c] * rl[c];
            s1 += a[c + 64] * rl[c + 64];
            s2 += a[c + 128] * rl[c + 128];
            s3 += a[c + 192] * rl[c + 192];
        }
        for (; c < m; c += 64) s += a[c] * rl[c];
        s = waveSum((s + s1) + (s2 + s3));
    }
    if (lane == 0) z[q] = T(s);
}

// ------------------------------------------------------------------------------------------------
// conjugate-gradient vector steps

template <class T> struct DotPQ {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, const T* __restrict__ p, const T* __restrict__ q, int64_t n, CgState* st, Reduce red) {
        st += vb.y;
        if (st->stop) return;
        p += instOffset(vb, n);
        q += instOffset(vb, n);
        double v[1] = {0};
        TILE_LOOP(v[0] += double(p[i]) * double(q[i]);)
        blockPartials<1>(vb, v, red);
    }
};
template <class T>
__global__ __launch_bounds__(BLOCK) void kDotPQ(const T* __restrict__ p, const T* __restrict__ q, int64_t n, CgState* st, Reduce red) {
    DotPQ<T>::run(ownBlock(), p, q, n, st, red);
}

// z0 (when not null): the next V-cycle's first Chebyshev step z0 = c0 D^-1 r from the new r, as
// kChebFirst computes it (the V-cycle then starts from it instead of reading r again)
template <class T> struct UpdateXR {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, T* __restrict__ x, const T* __restrict__ p, T* __restrict__ r, const T* __restrict__ q, int64_t n, CgState* st,
                               Reduce red, T* __restrict__ z0, const T* __restrict__ invD, T c0);
};
template <class T>
__global__ __launch_bounds__(BLOCK) void kUpdateXR(T* __restrict__ x, const T* __restrict__ p, T* __restrict__ r, const T* __restrict__ q, int64_t n,
                                                   CgState* st, Reduce red, T* __restrict__ z0, const T* __restrict__ invD, T c0) {
    UpdateXR<T>::run(ownBlock(), x, p, r, q, n, st, red, z0, invD, c0);
}
template <class T>
__device__ void UpdateXR<T>::run(VBlk vb, T* __restrict__ x, const T* __restrict__ p, T* __restrict__ r, const T* __restrict__ q, int64_t n, CgState* st,
                                 Reduce red, T* __restrict__ z0, const T* __restrict__ invD, T c0) {
    st += vb.y;
    if (st->stop) return;
    {
        const int64_t io = instOffset(vb, n);
        x += io;
        p += io;
        r += io;
        q += io;
        if (z0) z0 += io;
    }
    const T a = T(st->alpha);
    double v[1] = {0};
    TILE_LOOP(
        x[i] += a * p[i];
        const T ri = r[i] - a * q[i];
        r[i] = ri;
        if (z0) z0[i] = c0 * invD[i] * ri;
        v[0] += double(ri) * double(ri);)
    blockPartials<1>(vb, v, red);
}

template <class T> struct DotsRZ {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, const T* __restrict__ r, const T* __restrict__ z, const T* __restrict__ q, int64_t n, CgState* st, Reduce red) {
        st += vb.y;
        if (st->stop) return;
        {
            const int64_t io = instOffset(vb, n);
            r += io;
            z += io;
            q += io;
        }
        double v[2] = {0, 0};
        TILE_LOOP(
            const double zi = double(z[i]);
            v[0] += double(r[i]) * zi;
            v[1] += zi * double(q[i]);)
        blockPartials<2>(vb, v, red);
    }
};
template <class T>
__global__ __launch_bounds__(BLOCK) void kDotsRZ(const T* __restrict__ r, const T* __restrict__ z, const T* __restrict__ q, int64_t n, CgState* st,
                                                 Reduce red) {
    DotsRZ<T>::run(ownBlock(), r, z, q, n, st, red);
}

template <class T> struct UpdateP {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, const T* __restrict__ z, T* __restrict__ p, int64_t n, const CgState* st) {
        st += vb.y;
        if (st->stop) return;
        const int64_t i = int64_t(vb.x) * BLOCK + threadIdx.x;
        if (i >= n) return;
        const int64_t io = instOffset(vb, n);
        p[io + i] = z[io + i] + T(st->beta) * p[io + i];
    }
};
template <class T>
__global__ __launch_bounds__(BLOCK) void kUpdateP(const T* __restrict__ z, T* __restrict__ p, int64_t n, const CgState* st) {
    UpdateP<T>::run(ownBlock(), z, p, n, st);
}

// mixed precision with FP32 recurrences: the correction x32 and the residual r32 in units of
// 1 / rmax; the correction joins the FP64 solution at every fold
// The fold kernels run in every iteration and return at once unless a fold is due, so they use a
// small grid (grid-stride loops): an early return then costs a few hundred workgroups, not one per
// 256 DOFs.
constexpr unsigned FOLD_GRID = 1024;
struct FoldX {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, double* __restrict__ u, float* __restrict__ x, int64_t n, const CgState* st) {
        st += vb.y;
        if (st->noFold) return;
        u += instOffset(vb, n);
        x += instOffset(vb, n);
        const double rmax = st->rmax;
        for (int64_t i = int64_t(vb.x) * BLOCK + threadIdx.x; i < n; i += int64_t(vb.gx) * BLOCK) {
            u[i] += double(x[i]) * rmax;
            x[i] = 0.f;
        }
    }
};
__global__ __launch_bounds__(BLOCK) void kFoldX(double* __restrict__ u, float* __restrict__ x, int64_t n, const CgState* st) {
    FoldX::run(ownBlock(), u, x, n, st);
}

// z0 (when not null): the first Chebyshev step from the new r32, as kUpdateXR writes it
struct FoldR {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, const double* __restrict__ f, const double* __restrict__ Kx, const double* __restrict__ invD, double* __restrict__ r64,
                               float* __restrict__ r32, int64_t n, CgState* st, Reduce red, float* __restrict__ z0, const float* __restrict__ invD32,
                               float c0);
};
__global__ __launch_bounds__(BLOCK) void kFoldR(const double* __restrict__ f, const double* __restrict__ Kx, const double* __restrict__ invD,
                                                double* __restrict__ r64, float* __restrict__ r32, int64_t n, CgState* st, Reduce red,
                                                float* __restrict__ z0, const float* __restrict__ invD32, float c0) {
    FoldR::run(ownBlock(), f, Kx, invD, r64, r32, n, st, red, z0, invD32, c0);
}
__device__ void FoldR::run(VBlk vb, const double* __restrict__ f, const double* __restrict__ Kx, const double* __restrict__ invD, double* __restrict__ r64,
                           float* __restrict__ r32, int64_t n, CgState* st, Reduce red, float* __restrict__ z0, const float* __restrict__ invD32, float c0) {
    st += vb.y;
    if (st->noFold) return;
    f += instOffset(vb, n);
    Kx += instOffset(vb, n);
    if (r64) r64 += instOffset(vb, n);
    r32 += instOffset(vb, n);
    if (z0) z0 += instOffset(vb, n);
    const double inv = 1.0 / st->rmax;
    double v[1] = {0};
    // tiles of RED_TILE elements, block b taking tiles b, b + (blocks of the launch), ...
    for (int64_t t0 = int64_t(vb.x) * RED_TILE; t0 < n; t0 += int64_t(vb.gx) * RED_TILE) {
#pragma unroll
        for (int k = 0; k < RED_TILE / BLOCK; k++) {
            const int64_t i = t0 + threadIdx.x + k * BLOCK;
            if (i < n) {
                // held DOFs: a zero inverse diagonal (the FP32 copy where the FP64 one is not given);
                // f and Kx are read whatever it holds, so all three loads are in flight together
                const double fi = f[i], ki = Kx[i];
                const bool free = invD ? invD[i] != 0 : invD32[i] != 0.f;
                const double ri = free ? fi - ki : 0.0;
                if (r64) r64[i] = ri;
                const float rf = float(ri * inv);
                r32[i] = rf;
                if (z0) z0[i] = c0 * invD32[i] * rf;
                v[0] += ri * ri;
            }
        }
    }
    blockPartials<1>(vb, v, red);
}

__global__ __launch_bounds__(BLOCK) void kResidual64(const double* __restrict__ f, const double* __restrict__ Kx, const double* __restrict__ invD,
                                                     double* __restrict__ r, int64_t n, Reduce red) {
    const VBlk vb = ownBlock();
    f += instOffset(vb, n);
    Kx += instOffset(vb, n);
    r += instOffset(vb, n);
    double v[1] = {0};
    TILE_LOOP(
        const double ri = invD[i] != 0 ? f[i] - Kx[i] : 0.0;
        r[i] = ri;
        v[0] += ri * ri;)
    blockPartials<1>(vb, v, red);
}

__global__ __launch_bounds__(BLOCK) void kNormFree64(const double* __restrict__ f, const double* __restrict__ invD, int64_t n, Reduce red) {
    const VBlk vb = ownBlock();
    f += instOffset(vb, n);
    double v[1] = {0};
    TILE_LOOP(if (invD[i] != 0) v[0] += f[i] * f[i];)
    blockPartials<1>(vb, v, red);
}

__global__ __launch_bounds__(BLOCK) void kMaskCopy64(const double* __restrict__ x0, const double* __restrict__ invD, double* __restrict__ x, int64_t n) {
    const VBlk vb = ownBlock();
    const int64_t i = int64_t(vb.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int64_t io = instOffset(vb, n);
    x[io + i] = (x0 && invD[i] != 0) ? x0[io + i] : 0.0;
}

struct MaxAbs64 {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, const double* __restrict__ v, int64_t n, Reduce red);
};
__global__ __launch_bounds__(BLOCK) void kMaxAbs64(const double* __restrict__ v, int64_t n, Reduce red) { MaxAbs64::run(ownBlock(), v, n, red); }
__device__ void MaxAbs64::run(VBlk vb, const double* __restrict__ v, int64_t n, Reduce red) {
    v += instOffset(vb, n);
    double m = 0;
    TILE_LOOP(m = fmax(m, fabs(v[i]));)
    __shared__ double sh[BLOCK / 64];
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = sh[0];
        for (int w = 1; w < BLOCK / 64; w++) b = fmax(b, sh[w]);
        red.partials[int64_t(vb.y) * red.stride + vb.x] = b;
    }
}

// one block per instance; out steps by CgState
struct MaxFinish {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, Reduce red, unsigned nb, double* out);
};
__global__ __launch_bounds__(BLOCK) void kMaxFinish(Reduce red, unsigned nb, double* out) { MaxFinish::run(ownBlock(), red, nb, out); }
__device__ void MaxFinish::run(VBlk vb, Reduce red, unsigned nb, double* out) {
    red.partials += int64_t(vb.x) * red.stride;
    out += int64_t(vb.x) * CG_DBLS;
    double m = 0;
    for (unsigned b = threadIdx.x; b < nb; b += BLOCK) m = fmax(m, red.partials[b]);
    __shared__ double sh[BLOCK / 64];
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = sh[0];
        for (int w = 1; w < BLOCK / 64; w++) b = fmax(b, sh[w]);
        *out = b;
    }
}

// scaleFrom points into CgState (steps by instance)
struct ToF32 {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, const double* __restrict__ x, float* __restrict__ y, int64_t n, const double* scaleFrom) {
        const int64_t i = int64_t(vb.x) * BLOCK + threadIdx.x;
        if (i >= n) return;
        const int64_t io = instOffset(vb, n);
        const double s = scaleFrom[int64_t(vb.y) * CG_DBLS];
        y[io + i] = s > 0 ? float(x[io + i] * (1.0 / s)) : 0.f;
    }
};
__global__ __launch_bounds__(BLOCK) void kToF32(const double* __restrict__ x, float* __restrict__ y, int64_t n, const double* scaleFrom) {
    ToF32::run(ownBlock(), x, y, n, scaleFrom);
}

struct FromF32 {
    static constexpr int NTH = BLOCK, WPE = 0;
    __device__ static void run(VBlk vb, const float* __restrict__ z, double* __restrict__ y, int64_t n, const double* scaleFrom) {
        const int64_t i = int64_t(vb.x) * BLOCK + threadIdx.x;
        if (i >= n) return;
        const int64_t io = instOffset(vb, n);
        y[io + i] = double(z[io + i]) * scaleFrom[int64_t(vb.y) * CG_DBLS];
    }
};
__global__ __launch_bounds__(BLOCK) void kFromF32(const float* __restrict__ z, double* __restrict__ y, int64_t n, const double* scaleFrom) {
    FromF32::run(ownBlock(), z, y, n, scaleFrom);
}

// Lanczos on D^-1/2 K D^-1/2
__global__ __launch_bounds__(BLOCK) void kLzX(const double* __restrict__ v, const double* __restrict__ invD, double* __restrict__ x, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    x[i] = sqrt(invD[i]) * v[i];
}

__global__ __launch_bounds__(BLOCK) void kLzW(double* __restrict__ w, const double* __restrict__ invD, const double* __restrict__ vPrev,
                                              const double* __restrict__ v, int64_t n, CgState* st, int j, Reduce red) {
    const VBlk vb = ownBlock();
    const double b = j > 0 ? st->lz[16 + j - 1] : 0.0;
    double s[1] = {0};
    TILE_LOOP(
        const double wi = sqrt(invD[i]) * w[i] - b * vPrev[i];
        w[i] = wi;
        s[0] += wi * v[i];)
    blockPartials<1>(vb, s, red);
}

__global__ __launch_bounds__(BLOCK) void kLzAxpy(double* __restrict__ w, const double* __restrict__ v, int64_t n, CgState* st, int j, Reduce red) {
    const VBlk vb = ownBlock();
    const double a = st->lz[j];
    double s[1] = {0};
    TILE_LOOP(
        const double wi = w[i] - a * v[i];
        w[i] = wi;
        s[0] += wi * wi;)
    blockPartials<1>(vb, s, red);
}

__global__ __launch_bounds__(BLOCK) void kLzNext(const double* __restrict__ w, double* __restrict__ v, double* __restrict__ vPrev, int64_t n,
                                                 const CgState* st, int j) {
    const int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double b = st->lz[16 + j];
    const double vi = v[i];
    vPrev[i] = vi;
    v[i] = b > 0 ? w[i] / b : 0.0;
}

// finalizers of the reductions (thread 0 of the one-block kernel)
struct FinPQ {
    CgState* st0;
    __device__ void operator()(const double* s, int inst) const {
        CgState* st = st0 + inst;
        st->pq = s[0];
        if (!(s[0] > 0)) {  // loss of positive-definiteness (a mechanism)
            st->broken = 1;
            st->stop = 1;
        } else st->alpha = st->rz / s[0];
    }
};
struct FinXR {
    CgState* st0;
    __device__ void operator()(const double* s, int inst) const {
        CgState* st = st0 + inst;
        st->rr = s[0];
        st->iters += 1;
        st->sinceFold += 1;
        st->rel = sqrt(s[0]) * st->scale / st->bnorm;
        if (!(st->rel > st->tol) || st->iters >= st->maxIter || !isfinite(st->rel)) st->stop = 1;
    }
};
struct FinRZ {
    CgState* st0;
    __device__ void operator()(const double* s, int inst) const {
        CgState* st = st0 + inst;
        st->rzNew = s[0];
        st->zq = s[1];
        const double b = st->rz != 0 ? -st->alpha * s[1] / st->rz : 0.0;
        st->beta = b > 0 ? b : 0.0;
        st->rz = s[0];
    }
};
struct FinXRFold {
    CgState* st0;
    __device__ void operator()(const double* s, int inst) const {
        CgState* st = st0 + inst;
        st->rr = s[0];
        st->iters += 1;
        st->sinceFold += 1;
        const double est = sqrt(s[0]) * st->scale / st->bnorm;
        st->rel = est;
        st->peak = fmax(st->peak, est);
        const bool fold = !(est > st->tol) || !(est > 0.1 * st->peak) || st->sinceFold >= st->foldEvery || st->iters >= st->maxIter || !isfinite(est);
        st->noFold = fold ? 0 : 1;
    }
};
struct FinFoldR {
    CgState* st0;
    __device__ void operator()(const double* s, int inst) const {
        CgState* st = st0 + inst;
        const double rel = sqrt(s[0]) / st->bnorm;
        st->trueRel = rel;
        st->rel = rel;
        st->peak = rel;
        st->sinceFold = 0;
        st->folds += 1;
        st->noFold = 1;
        if (!(rel > st->tol) || st->iters >= st->maxIter || !isfinite(rel)) st->stop = 1;
    }
};
struct FinStore {
    double* out;  // into CgState: steps by instance
    __device__ void operator()(const double* s, int inst) const { out[int64_t(inst) * CG_DBLS] = s[0]; }
};
struct FinLzA {
    CgState* st;
    int j;
    __device__ void operator()(const double* s, int) const { st->lz[j] = s[0]; }
};
struct FinLzB {
    CgState* st;
    int j;
    __device__ void operator()(const double* s, int) const { st->lz[16 + j] = sqrt(s[0]); }
};

// one finish block per instance
template <int NV, class Fin> void finish(int ninst, hipStream_t s, Reduce red, unsigned nb, const int* guard, Fin fin) {
    go<Finish<NV, Fin>>(kFinish<NV, Fin>, dim3(unsigned(ninst)), s, red, nb, guard, fin);
}
// launchers whose kernels a packed solve never records (set-up and final steps) refuse to run while
// this thread records, so nothing is issued out of order
void notRecording() {
    if (packRecorder()) throw std::logic_error("a set-up kernel was launched while a packed batch was being recorded");
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// launchers

template <class T>
void launchApply(const DevLevel& L, int mode, const T* x, T* y, const T* r, const T* d, double c1, double c2, const int* stop, hipStream_t s, int ninst) {
    const dim3 g(flatGrid(L.nNodes), ninst);
    auto one = [&](auto m) {
        constexpr int M = decltype(m)::value;
        go<Apply<T, M>>(kApply<T, M>, g, s, L, x, y, r, d, T(c1), T(c2), stop);
    };
    switch (mode) {
        case AM_APPLY: one(std::integral_constant<int, AM_APPLY>{}); break;
        case AM_RAW: one(std::integral_constant<int, AM_RAW>{}); break;
        case AM_RESID: one(std::integral_constant<int, AM_RESID>{}); break;
        case AM_CHEB1: one(std::integral_constant<int, AM_CHEB1>{}); break;
        default: one(std::integral_constant<int, AM_CHEB2>{}); break;
    }
}

template <class T>
void launchApplyOwn(const DevLevel& L, int mode, const T* x, T* y, const T* r, const T* d, double c1, double c2, const int* stop, hipStream_t s, int ninst,
                    bool half) {
    // every instance in one grid: the kernel loops over them (the element rows are read once)
    const dim3 g(unsigned((L.nNodes + BLOCK / 8 - 1) / (BLOCK / 8)));
    auto one = [&](auto m, auto h) {
        constexpr int M = decltype(m)::value;
        constexpr bool H = decltype(h)::value;
        go<ApplyOwn<T, M, H>>(kApplyOwn<T, M, H>, g, s, L, x, y, r, d, T(c1), T(c2), stop, ninst);
    };
    auto modes = [&](auto h) {
        switch (mode) {
            case AM_APPLY: one(std::integral_constant<int, AM_APPLY>{}, h); break;
            case AM_RAW: one(std::integral_constant<int, AM_RAW>{}, h); break;
            case AM_RESID: one(std::integral_constant<int, AM_RESID>{}, h); break;
            case AM_CHEB1: one(std::integral_constant<int, AM_CHEB1>{}, h); break;
            default: one(std::integral_constant<int, AM_CHEB2>{}, h); break;
        }
    };
    if constexpr (sizeof(T) == 4) {
        if (half && L.KownH) {
            modes(std::true_type{});
            return;
        }
    }
    modes(std::false_type{});
}

void launchOwnToHalf(const float* Kown, int64_t n, uint16_t* half, float* scale, hipStream_t s) {
    notRecording();
    if (n > 0) kOwnToHalf<<<unsigned(n), 64, 0, s>>>(Kown, n, half, scale);
}

void launchApplyOwnMf(const DevLevel& L, int mode, const float* x, float* y, const float* r, const float* d, double c1, double c2, const int* stop,
                      hipStream_t s, int ninst) {
    const dim3 g(unsigned((L.nNodes + BLOCK / 64 - 1) / (BLOCK / 64)));
    const float f1 = float(c1), f2 = float(c2);
    auto one = [&](auto m) {
        constexpr int M = decltype(m)::value;
        go<ApplyOwnMf<M>>(kApplyOwnMf<M>, g, s, L, x, y, r, d, f1, f2, stop, ninst);
    };
    switch (mode) {
        case AM_APPLY: one(std::integral_constant<int, AM_APPLY>{}); break;
        case AM_RAW: one(std::integral_constant<int, AM_RAW>{}); break;
        case AM_RESID: one(std::integral_constant<int, AM_RESID>{}); break;
        case AM_CHEB1: one(std::integral_constant<int, AM_CHEB1>{}); break;
        default: one(std::integral_constant<int, AM_CHEB2>{}); break;
    }
}

template <class T> void launchChebFirst(const DevLevel& L, const T* r, T* z, double c0, const int* stop, hipStream_t s, int ninst) {
    const T* invD = sizeof(T) == 8 ? reinterpret_cast<const T*>(L.invD64) : reinterpret_cast<const T*>(L.invD32);
    go<ChebFirst<T>>(kChebFirst<T>, dim3(flatGrid(L.nDof), ninst), s, invD, r, z, T(c0), L.nDof, stop);
}

template <class T>
void launchRestrict(const DevLevel& F, const DevLevel& C, const T* rf, T* rc, const int* stop, hipStream_t s, int ninst, T* z0, double c0) {
    go<Restrict<T>>(kRestrict<T>, dim3(flatGrid(C.nNodes), ninst), s, F, C, rf, rc, stop, z0, T(c0));
}

template <class T> void launchProlong(const DevLevel& F, const DevLevel& C, const T* zc, T* zf, const int* stop, hipStream_t s, int ninst) {
    if (C.nNodes >= (int64_t(1) << 16)) go<ProlongCells<T>>(kProlongCells<T>, dim3(flatGrid(C.nNodes), ninst), s, F, C, zc, zf, stop);
    else go<Prolong<T>>(kProlong<T>, dim3(flatGrid(F.nNodes), ninst), s, F, C, zc, zf, stop);
}

template <class T>
void launchCoarse(const int32_t* freeIdx, const int32_t* dense, int64_t m, int64_t nDof, const double* Ainv, const T* r, T* z, const int* stop, hipStream_t s,
                  int ninst) {
    go<Coarse<T>>(kCoarse<T>, dim3(unsigned((nDof + BLOCK / 64 - 1) / (BLOCK / 64)), ninst), s, freeIdx, dense, int(m), nDof, Ainv, r, z, stop);
}

template <class T> void launchDotPQ(const T* p, const T* q, int64_t n, CgState* st, Reduce red, hipStream_t s, int ninst) {
    const unsigned g = reduceGrid(n);
    go<DotPQ<T>>(kDotPQ<T>, dim3(g, ninst), s, p, q, n, st, red);
    finish<1, FinPQ>(ninst, s, red, g, &st->stop, FinPQ{st});
}
template <class T> void launchUpdateXR(T* x, const T* p, T* r, const T* q, int64_t n, CgState* st, Reduce red, hipStream_t s, int ninst) {
    const unsigned g = reduceGrid(n);
    go<UpdateXR<T>>(kUpdateXR<T>, dim3(g, ninst), s, x, p, r, q, n, st, red, nullptr, nullptr, T(0));
    finish<1, FinXR>(ninst, s, red, g, &st->stop, FinXR{st});
}
template <class T> void launchDotsRZ(const T* r, const T* z, const T* q, int64_t n, CgState* st, Reduce red, hipStream_t s, int ninst) {
    const unsigned g = reduceGrid(n);
    go<DotsRZ<T>>(kDotsRZ<T>, dim3(g, ninst), s, r, z, q, n, st, red);
    finish<2, FinRZ>(ninst, s, red, g, &st->stop, FinRZ{st});
}
template <class T> void launchUpdateP(const T* z, T* p, int64_t n, CgState* st, hipStream_t s, int ninst) {
    go<UpdateP<T>>(kUpdateP<T>, dim3(flatGrid(n), ninst), s, z, p, n, st);
}

// finishers for dot products whose block partials another kernel wrote ([k * nb + block])
void launchFinishPQ(Reduce red, unsigned nb, CgState* st, hipStream_t s, int ninst) { finish<1, FinPQ>(ninst, s, red, nb, &st->stop, FinPQ{st}); }
void launchFinishRZ(Reduce red, unsigned nb, CgState* st, hipStream_t s, int ninst) { finish<2, FinRZ>(ninst, s, red, nb, &st->stop, FinRZ{st}); }

void launchUpdateXR32Fold(float* x, const float* p, float* r, const float* q, int64_t n, CgState* st, Reduce red, hipStream_t s, int ninst, float* z0,
                          const float* invD32, double c0) {
    const unsigned g = reduceGrid(n);
    go<UpdateXR<float>>(kUpdateXR<float>, dim3(g, ninst), s, x, p, r, q, n, st, red, z0, invD32, float(c0));
    finish<1, FinXRFold>(ninst, s, red, g, &st->stop, FinXRFold{st});
}
void launchFoldX(double* u, float* x, int64_t n, CgState* st, hipStream_t s, int ninst) {
    go<FoldX>(kFoldX, dim3(std::min(flatGrid(n), FOLD_GRID), ninst), s, u, x, n, st);
}

void launchFoldR(const double* f, const double* Kx, const double* invD, double* r64, float* r32, int64_t n, CgState* st, Reduce red, hipStream_t s,
                 int ninst, float* z0, const float* invD32, double c0) {
    if (!invD && !invD32) throw std::invalid_argument("launchFoldR: an inverse diagonal marks the held DOFs");
    // four times the fold grid: a fold streams 28 bytes a DOF (208.9 -> 147.3 us on 25.6 M DOFs against
    // 1024 blocks) while a launch without a fold still returns in the same 4.5 us
    const unsigned g = std::min(reduceGrid(n), 4 * FOLD_GRID);
    go<FoldR>(kFoldR, dim3(g, ninst), s, f, Kx, invD, r64, r32, n, st, red, z0, invD32, float(c0));
    finish<1, FinFoldR>(ninst, s, red, g, &st->noFold, FinFoldR{st});
}

void launchResidual64(const double* f, const double* Kx, const double* invD, double* r, int64_t n, double* out, Reduce red, hipStream_t s, int ninst) {
    notRecording();
    const unsigned g = reduceGrid(n);
    kResidual64<<<dim3(g, ninst), BLOCK, 0, s>>>(f, Kx, invD, r, n, red);
    kFinish<1, FinStore><<<ninst, FIN_BLOCK, 0, s>>>(red, g, nullptr, FinStore{out});
}
void launchNormFree64(const double* f, const double* invD, int64_t n, double* out, Reduce red, hipStream_t s, int ninst) {
    notRecording();
    const unsigned g = reduceGrid(n);
    kNormFree64<<<dim3(g, ninst), BLOCK, 0, s>>>(f, invD, n, red);
    kFinish<1, FinStore><<<ninst, FIN_BLOCK, 0, s>>>(red, g, nullptr, FinStore{out});
}
void launchMaskCopy64(const double* x0, const double* invD, double* x, int64_t n, hipStream_t s, int ninst) {
    notRecording();
    kMaskCopy64<<<dim3(flatGrid(n), ninst), BLOCK, 0, s>>>(x0, invD, x, n);
}
void launchMaxAbs64(const double* v, int64_t n, double* out, Reduce red, hipStream_t s, int ninst) {
    const unsigned g = reduceGrid(n);
    go<MaxAbs64>(kMaxAbs64, dim3(g, ninst), s, v, n, red);
    go<MaxFinish>(kMaxFinish, dim3(unsigned(ninst)), s, red, g, out);
}
void launchToF32(const double* x, float* y, int64_t n, const double* scaleFrom, hipStream_t s, int ninst) {
    go<ToF32>(kToF32, dim3(flatGrid(n), ninst), s, x, y, n, scaleFrom);
}
void launchFromF32(const float* z, double* y, int64_t n, const double* scaleFrom, hipStream_t s, int ninst) {
    go<FromF32>(kFromF32, dim3(flatGrid(n), ninst), s, z, y, n, scaleFrom);
}
void launchLzX(const double* v, const double* invD, double* x, int64_t n, hipStream_t s) {
    notRecording();
    kLzX<<<flatGrid(n), BLOCK, 0, s>>>(v, invD, x, n);
}
void launchLzW(double* w, const double* invD, const double* vPrev, const double* v, int64_t n, CgState* st, int j, Reduce red, hipStream_t s) {
    notRecording();
    const unsigned g = reduceGrid(n);
    kLzW<<<g, BLOCK, 0, s>>>(w, invD, vPrev, v, n, st, j, red);
    kFinish<1, FinLzA><<<1, FIN_BLOCK, 0, s>>>(red, g, nullptr, FinLzA{st, j});
}
void launchLzAxpy(double* w, const double* v, int64_t n, CgState* st, int j, Reduce red, hipStream_t s) {
    notRecording();
    const unsigned g = reduceGrid(n);
    kLzAxpy<<<g, BLOCK, 0, s>>>(w, v, n, st, j, red);
    kFinish<1, FinLzB><<<1, FIN_BLOCK, 0, s>>>(red, g, nullptr, FinLzB{st, j});
}
void launchLzNext(const double* w, double* v, double* vPrev, int64_t n, CgState* st, int j, hipStream_t s) {
    notRecording();
    kLzNext<<<flatGrid(n), BLOCK, 0, s>>>(w, v, vPrev, n, st, j);
}

#define PSX_INSTANTIATE(T)                                                                                                                   \
    template void launchApply<T>(const DevLevel&, int, const T*, T*, const T*, const T*, double, double, const int*, hipStream_t, int);     \
    template void launchChebFirst<T>(const DevLevel&, const T*, T*, double, const int*, hipStream_t, int);                                  \
    template void launchApplyOwn<T>(const DevLevel&, int, const T*, T*, const T*, const T*, double, double, const int*, hipStream_t, int, bool); \
    template void launchRestrict<T>(const DevLevel&, const DevLevel&, const T*, T*, const int*, hipStream_t, int, T*, double);           \
    template void launchProlong<T>(const DevLevel&, const DevLevel&, const T*, T*, const int*, hipStream_t, int);                          \
    template void launchCoarse<T>(const int32_t*, const int32_t*, int64_t, int64_t, const double*, const T*, T*, const int*, hipStream_t, int); \
    template void launchDotPQ<T>(const T*, const T*, int64_t, CgState*, Reduce, hipStream_t, int);                                        \
    template void launchUpdateXR<T>(T*, const T*, T*, const T*, int64_t, CgState*, Reduce, hipStream_t, int);                             \
    template void launchDotsRZ<T>(const T*, const T*, const T*, int64_t, CgState*, Reduce, hipStream_t, int);                             \
    template void launchUpdateP<T>(const T*, T*, int64_t, CgState*, hipStream_t, int);

PSX_INSTANTIATE(double)
PSX_INSTANTIATE(float)

}  // namespace psx
