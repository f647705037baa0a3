// Element-by-element operator on the matrix cores, for FP32 levels whose elements are all s * Kb
// (every finest level, uniform or per-voxel scales) and for the child-scale level below one: y = K x
// with the fused epilogues of kFine.
//
// K x = sum over voxels e of s_e Kb x_e, and Kb annihilates rigid translations, so the product is
// taken of x' = x - c, c one node's value per block (the tile's centre): only the spread of x over
// the tile enters the arithmetic, not its magnitude. x' is split into BF16 pieces (hi = bf16(x'),
// lo = bf16(x' - hi), ...), Kb likewise, and each element's 24 x 24 product is a sum of
// v_mfma_f32_16x16x32_bf16 products of the pieces with FP32 accumulation; the piece pairs whose
// weight is below 2^-8 per piece order are left out (two pieces each: hi.hi + hi.lo + lo.hi, an
// error of about 2^-17 |Kb| |x'|; three: six pairs, about 2^-25).
//
// Layout of the work: a block of 4 wavefronts owns 60 x 8 nodes in x-y and marches along z one
// element layer (the voxels between node planes k and k + 1) per step. A wavefront owns a strip of
// 15 x 8 nodes and computes the 16 x 9 elements around it, one element row (16 elements along x) per
// matrix-instruction group:
//   columns = the 16 elements; K = their 8 corners x 4 components (the fourth zero), read as two
//   adjacent nodes per lane from the LDS node planes (4 BF16 per node: one 16-byte read); rows =
//   4 components x 4 corners of one z side (two instructions per piece pair cover both sides), so
//   lane (element n, component c) holds in its 4 accumulators the element's 4 corners of that side.
// Each lane scales its element's corners by the voxel scale, and the node sums follow without a trip
// through LDS: the x neighbour's corner arrives by a row shift (DPP row_shr:1: lane n - 1 to lane n,
// the 16 lanes of a row being the 16 elements of one component), the y neighbour's from the previous
// element row's group and the z neighbour's from the previous layer, both in registers. Node planes
// are the only shared data: one barrier per layer.
// Nothing depends on which voxels differ in scale, so void surfaces and per-voxel densities cost the
// same as solid regions (no pair lists).
//
// The child-scale level: its uniform elements are s * Kb with that level's Kb and take the path above;
// an own element (voxel word < 0) enters the products with scale 0, and its product
// sum_c s_c M_c (x_e - x_e0) over the 8 child matrices (OwnRowsMf, one element per MFMA column) is
// written to a work buffer just before and added to its corners' nodes in the epilogue. The inverse
// diagonal there is the level's stored one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "device.hpp"
#include "pack_dev.hpp"
#include "stencil.hpp"

namespace psx {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256, NW = NT / 64;
constexpr int SX = 15;                                  // nodes per wavefront strip in x (16 elements)
constexpr int BX = SX * NW, R = 8;                       // block tile: 60 x 8 nodes
constexpr int PX = BX + 2, PY = R + 2, NP = PX * PY;     // node plane with halo (62 x 10)
constexpr int NQ = (NP + NT - 1) / NT;                   // nodes staged per thread
constexpr int VX = BX + 1, VY = R + 1, NV = VX * VY;     // voxel layer of the block (61 x 9)
constexpr int NVQ = (NV + NT - 1) / NT;                  // voxels staged per thread

// node-plane slots: a ring of 3 (plane k + 2 staged while planes k, k + 1 are read: one barrier per
// layer), or 2 for three pieces (staged after a second barrier), which keeps the block at 38 KB and 4
// blocks per compute unit
template <int XP> constexpr int slotsFor() { return XP >= 3 ? 2 : 3; }
template <int XP> struct ElShared {
    uint2 xq[slotsFor<XP>()][XP][NP];  // node planes: the pieces of x - c, 4 BF16 per node
    int32_t s[3][NV];         // voxel words of element layers k - 1 .. k + 1 (0: empty or outside)
    float4 ey[NW][R / 2 * SX];  // per wavefront: half its nodes' sums of plane k (3 components) and voxel-scale sum
};

__device__ __forceinline__ uint32_t bfBits(float v) { return uint32_t(__builtin_bit_cast(uint16_t, (__bf16)v)); }
__device__ __forceinline__ float bfValue(uint32_t b) { return __uint_as_float(b << 16); }
// lane n of each 16-lane row receives v of lane n - 1 (lane 0 of a row: 0)
__device__ __forceinline__ float fromLeft(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));
}

#define PSX_EL_PARAMS                                                                                                                          \
    DevLevel L, const float *__restrict__ x, float *__restrict__ y, const float *__restrict__ r, const float *__restrict__ zp, float c1, float c2, \
        const int *stop, int zChunk, int gx, int gy, int nTiles, double *__restrict__ dots, int64_t dotStride, const float *__restrict__ qdot

// OWNC: the child-scale level; its own elements enter with scale 0 in the products, and their corner
// rows (the element's own product in the difference form, written by OwnRows) are added to the sums
template <int MODE, bool DOT, int XP, int AP, bool OWNC> __device__ __forceinline__ void elBody(VBlk vb, PSX_EL_PARAMS) {
    if (stop) stop += int64_t(vb.y) * CG_INTS;
    if (stop && __builtin_nontemporal_load(stop) != 0) return;
    {
        const int64_t io = int64_t(vb.y) * L.nDof;
        x += io;
        y += io;
        if (r) r += io;
        if (zp) zp += io;
        if (qdot) qdot += io;
        if (DOT) dots += int64_t(vb.y) * dotStride;
    }
    const int t = int(vb.x);
    if (t >= nTiles) return;
    __shared__ ElShared<XP> sh;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, n = lane & 15, g = lane >> 4;
    const int X0 = (t % gx) * BX, Y0 = ((t / gx) % gy) * R;
    const int z0 = (t / (gx * gy)) * zChunk, z1 = min(L.NZ, z0 + zChunk);
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
    constexpr bool NEEDS_R = MODE == AM_RESID || MODE == AM_CHEB1 || MODE == AM_CHEB2;
    constexpr bool NEEDS_D = MODE == AM_CHEB2;
    constexpr bool NEEDS_INV = MODE == AM_CHEB1 || MODE == AM_CHEB2;
    // this wavefront's strip: nodes Xw .. Xw + 14; lane n's element is voxel Xw - 1 + n, its node
    // (after the row shift) Xw - 1 + n, so lanes 1..15 hold the strip's node sums; lane group g is the
    // component in the products (g = 3 holds none)
    const int Xw = X0 + SX * wid;
    const bool waveActive = Xw < L.NX;

    // the block's reference value c: the node at the tile's centre (uniform loads)
    float cref[3];
    {
        const int ic = min(X0 + BX / 2, L.NX - 1), jc = min(Y0 + R / 2, L.NY - 1), zc = min(z0 + zChunk / 2, L.NZ - 1);
        const int64_t n3 = 3 * (ic + NX * jc + NXY * zc);
        cref[0] = x[n3];
        cref[1] = x[n3 + 1];
        cref[2] = x[n3 + 2];
    }
    // matrix-core operands of Kb (see DevLevel::elA): A[h][piece] for the corners at z offset h
    bf16x8 A[2][AP];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int p = 0; p < AP; p++) {
            const uint4 v = L.elA[(h * 3 + p) * 64 + lane];
            typedef unsigned u4 __attribute__((ext_vector_type(4)));
            const u4 w = {v.x, v.y, v.z, v.w};
            A[h][p] = __builtin_bit_cast(bf16x8, w);
        }

    // staging: node e of the plane with halo (local (e % PX, e / PX) = node (X0 - 1, Y0 - 1) + that),
    // e = tid + q NT; offsets OUT_OF_RANGE outside the part (loads return 0)
    int32_t poff[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int e = tid + q * NT;
        const int gi = X0 - 1 + e % PX, gj = Y0 - 1 + e / PX;
        const bool in = e < NP && gi >= 0 && gj >= 0 && gi < L.NX && gj < L.NY;
        poff[q] = in ? int32_t(12 * (gi + NX * gj)) : OUT_OF_RANGE;
    }
    const uint32_t planeBytes = uint32_t(12 * NXY);
    auto fetchPlane = [&](int z, float (&px)[NQ][3]) {
        const bool zin = z >= 0 && z < L.NZ;
        const auto rs = planeRsrc(x + 3 * NXY * (zin ? z : 0), zin ? planeBytes : 0u);
#pragma unroll
        for (int q = 0; q < NQ; q++) bload3(rs, poff[q], px[q]);
    };
    constexpr int NS = slotsFor<XP>();
    auto storePlane = [&](int z, const float (&px)[NQ][3]) {
        const int sl = (z + 2 * NS) % NS;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int e = tid + q * NT;
            if (e >= NP) continue;
            float v[3];
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] = px[q][c] - cref[c];
#pragma unroll
            for (int p = 0; p < XP; p++) {
                uint32_t b[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    b[c] = bfBits(v[c]);
                    v[c] -= bfValue(b[c]);
                }
                sh.xq[sl][p][e] = uint2{b[0] | (b[1] << 16), b[2]};
            }
        }
    };
    // the voxel scales of a layer: voxel e of the block's layer (local (e % VX, e / VX) = voxel
    // (X0 - 1, Y0 - 1) + that), e = tid + q NT
    const int64_t nxy = int64_t(L.nx) * L.ny;
    int32_t voff[NVQ];
#pragma unroll
    for (int q = 0; q < NVQ; q++) {
        const int e = tid + q * NT;
        const int vi = X0 - 1 + e % VX, vj = Y0 - 1 + e / VX;
        const bool in = e < NV && vi >= 0 && vj >= 0 && vi < L.nx && vj < L.ny;
        voff[q] = in ? int32_t(4 * (vi + int64_t(L.nx) * vj)) : OUT_OF_RANGE;
    }
    auto fetchVox = [&](int k, int32_t (&pv)[NVQ]) {
        const bool kin = k >= 0 && k < L.nz;
        const auto rs = planeRsrc(L.words + nxy * (kin ? k : 0), kin ? uint32_t(4 * nxy) : 0u);
#pragma unroll
        for (int q = 0; q < NVQ; q++) pv[q] = int32_t(__builtin_amdgcn_raw_buffer_load_b32(rs, voff[q], 0, 0));
    };
    auto storeVox = [&](int k, const int32_t (&pv)[NVQ]) {
#pragma unroll
        for (int q = 0; q < NVQ; q++) {
            const int e = tid + q * NT;
            if (e < NV) sh.s[(k + 3) % 3][e] = pv[q];
        }
    };
    // the epilogue, one node per lane, in two passes of R / 2 rows: node q of a pass (local
    // (q % SX, q / SX) in the pass's rows), q = lane < R / 2 * SX; its offsets within a plane
    // (OUT_OF_RANGE past the part's edge)
    constexpr int HR = R / 2;
    static_assert(HR * SX <= 64, "one node per lane and pass");
    int32_t noff[2], ioff[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int xq = Xw + lane % SX, yq = Y0 + HR * u + lane / SX;
        const bool in = lane < HR * SX && xq < L.NX && yq < L.NY;
        noff[u] = in ? int32_t(12 * (xq + NX * yq)) : OUT_OF_RANGE;
        ioff[u] = in ? int32_t(2 * (xq + NX * yq)) : OUT_OF_RANGE;
    }
    // the lane's operand reads: node pair (local x SX wid + n, + 1) of local row i + (g & 1) in node
    // plane k + (g >> 1)
    const int bBase = (g & 1) * PX + SX * wid + n, zOff = g >> 1;

    // prologue: node planes z0 - 1, z0 and voxel layer z0 - 1 in LDS; plane z0 + 1 and layer z0 in
    // flight
    float pn[NQ][3];
    int32_t vn[NVQ];
    {
        float pa[NQ][3], pb[NQ][3];
        int32_t va[NVQ];
        fetchPlane(z0 - 1, pa);
        fetchPlane(z0, pb);
        fetchVox(z0 - 1, va);
        fetchPlane(z0 + 1, pn);
        fetchVox(z0, vn);
        storePlane(z0 - 1, pa);
        storePlane(z0, pb);
        storeVox(z0 - 1, va);
    }
    float carry[R], syPrev[R];
#pragma unroll
    for (int j = 0; j < R; j++) carry[j] = syPrev[j] = 0.f;
    double dot0 = 0, dot1 = 0;
    const float k00 = L.sym32[0];
    __syncthreads();

    for (int k = z0 - 1; k < z1; k++) {
        // node planes k (lower corners) and k + 1 (upper corners) of the strip's rows, component g
        float lo[R], up[R], sy[R];
#pragma unroll
        for (int j = 0; j < R; j++) lo[j] = up[j] = sy[j] = 0.f;
        if (waveActive) {
            const uint2* __restrict__ xs = &sh.xq[(k + zOff + 2 * NS) % NS][0][0];
            const int32_t* __restrict__ sv = &sh.s[(k + 3) % 3][SX * wid + n];
            // element row Y0 - 1 + i: its operands (two node pairs per lane and piece) and scale
            auto fetch = [&](int i, bf16x8 (&B)[XP], int32_t& wd) {
#pragma unroll
                for (int p = 0; p < XP; p++) {
                    const uint2 a = xs[p * NP + bBase + i * PX], b = xs[p * NP + bBase + i * PX + 1];
                    typedef unsigned u4 __attribute__((ext_vector_type(4)));
                    const u4 w = {a.x, a.y, b.x, b.y};
                    B[p] = __builtin_bit_cast(bf16x8, w);
                }
                wd = sv[i * VX];
            };
            // its products: C[h] = Kb (x_e - c) at the corners of z side h, the smaller piece
            // products first
            auto product = [&](const bf16x8 (&B)[XP], floatx4 (&C)[2]) {
#pragma unroll
                for (int h = 0; h < 2; h++) C[h] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ord = (AP > XP ? AP : XP) - 1; ord >= 0; ord--)
#pragma unroll
                    for (int a = 0; a < AP; a++) {
                        const int b = ord - a;
                        if (b >= 0 && b < XP) {
#pragma unroll
                            for (int h = 0; h < 2; h++) C[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[h][a], B[b], C[h], 0, 0, 0);
                        }
                    }
            };
            // its node sums: lane (n, g) holds the element's corners (x, y) = (r & 1, r >> 1) of
            // component g; the x-upper corners belong to the next lane's node
            auto gather = [&](int i, const floatx4 (&C)[2], int32_t wd) {
                // the voxel word: > 0 the scale of s * Kb, < 0 an own element (scale 0 here: its rows
                // join in the epilogue)
                const float s = wd > 0 ? __int_as_float(wd) : 0.f;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    float(&acc)[R] = h ? up : lo;
                    if (i >= 1) {  // y-lower corners: node row Y0 + i - 1
                        acc[i - 1] += s * C[h][0];
                        acc[i - 1] += fromLeft(s * C[h][1]);
                    }
                    if (i < R) {  // y-upper corners: node row Y0 + i
                        acc[i] += s * C[h][2];
                        acc[i] += fromLeft(s * C[h][3]);
                    }
                }
                if constexpr (NEEDS_INV && !OWNC) {
                    const float sx = s + fromLeft(s);  // the node's two voxels in x of this element row
                    if (i >= 1) sy[i - 1] += sx;
                    if (i < R) sy[i] += sx;
                }
            };
            // software pipeline over the element rows (rows past the part's last node row are computed
            // and not stored): the reads of row i + 1 and the sums of row i - 1 around the products
            // of row i
            bf16x8 B[2][XP];
            floatx4 C[2][2];
            int32_t sc[3];
            fetch(0, B[0], sc[0]);
#pragma unroll
            for (int i = 0; i <= R; i++) {
                if (i < R) fetch(i + 1, B[(i + 1) & 1], sc[(i + 1) % 3]);
                product(B[i & 1], C[i & 1]);
                if (i >= 1) gather(i - 1, C[(i - 1) & 1], sc[(i - 1) % 3]);
            }
            gather(R, C[R & 1], sc[R % 3]);
        }
        // node plane k is complete: the upper corners of layer k - 1 and the lower ones of layer k.
        // The (element, component) lanes hand the sums to one lane per node through the wavefront's
        // LDS rows (lane group 3, which holds no component, writes the voxel-scale sum)
        if (k >= z0 && waveActive) {
            float4* __restrict__ wy = sh.ey[wid];
            const int64_t pb = 3 * NXY * k;
            const auto rx = planeRsrc(x + pb, planeBytes), ry = planeRsrc(y + pb, planeBytes);
            const auto rr = planeRsrc(r ? r + pb : x, r ? planeBytes : 0u);
            const auto rd = planeRsrc(zp ? zp + pb : x, zp ? planeBytes : 0u);
            const auto rq = planeRsrc(qdot ? qdot + pb : x, qdot ? planeBytes : 0u);
            const auto ri = planeRsrc(L.nodeInfo + NXY * k, uint32_t(2 * NXY));
            const auto rinv = planeRsrc(OWNC ? L.invD32 + pb : x, OWNC ? planeBytes : 0u);
#pragma unroll
            for (int u = 0; u < 2; u++) {
                if (n >= 1) {
#pragma unroll
                    for (int jj = 0; jj < HR; jj++) {
                        const int j = HR * u + jj;
                        const float v = g < 3 ? carry[j] + lo[j] : (NEEDS_INV && !OWNC ? syPrev[j] + sy[j] : 0.f);
                        reinterpret_cast<float*>(wy)[4 * (jj * SX + n - 1) + g] = v;
                    }
                }
                __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wavefront's LDS writes done
                __builtin_amdgcn_wave_barrier();
                if (noff[u] != OUT_OF_RANGE) {
                    const float4 sum = wy[lane];
                    float xv[3] = {0.f, 0.f, 0.f}, rv[3] = {0.f, 0.f, 0.f}, dv[3] = {0.f, 0.f, 0.f}, qv[3] = {0.f, 0.f, 0.f};
                    // x itself only where the step uses it: the Chebyshev steps, the fused p.q, the diagonal shift
                    if (NEEDS_INV || (MODE == AM_APPLY && DOT) || L.diagAdd) bload3(rx, noff[u], xv);
                    if constexpr (NEEDS_R) bload3(rr, noff[u], rv);
                    if constexpr (NEEDS_D) bload3(rd, noff[u], dv);
                    if constexpr (NEEDS_INV && DOT) bload3(rq, noff[u], qv);
                    float iv[3] = {0.f, 0.f, 0.f};  // the stored inverse diagonal (levels with own elements)
                    if constexpr (NEEDS_INV && OWNC) bload3(rinv, noff[u], iv);
                    const uint32_t info = __builtin_amdgcn_raw_buffer_load_b16(ri, ioff[u], 0, 0);
                    float out0[3] = {sum.x, sum.y, sum.z};
                    if constexpr (OWNC) {
                        // the rows of the node's own elements (voxels of layers k - 1, k around it, in voxel
                        // order): the node is corner (1 - dx, 1 - dy, 1 - dz) of voxel (dx, dy, dz)
                        const int lxn = SX * wid + lane % SX, lyn = HR * u + lane / SX;
                        const float* __restrict__ ownT = L.ownT + int64_t(vb.y) * L.ownTStride;
#pragma unroll
                        for (int vq = 0; vq < 8; vq++) {
                            const int dx = vq & 1, dy = (vq >> 1) & 1, dz = vq >> 2;
                            const int32_t w = sh.s[(k - 1 + dz + 3) % 3][(lyn + dy) * VX + lxn + dx];
                            if (w < 0) {
                                const float* tr = ownT + 24 * int64_t(w & 0x7fffffff) + 3 * ((1 - dx) + 2 * (1 - dy) + 4 * (1 - dz));
                                out0[0] += tr[0];
                                out0[1] += tr[1];
                                out0[2] += tr[2];
                            }
                        }
                    }
                    float yv[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        float out = out0[d];
                        // the optional diagonal shift of this DOF (noff / 4: the node's first DOF in the plane)
                        const float dA = L.diagAdd ? float(L.diagAdd[pb + (noff[u] >> 2) + d]) : 0.f;
                        if (L.diagAdd) out += dA * xv[d];
                        const bool held = (info >> (8 + d)) & 1u;
                        float inv = 0.f;
                        if constexpr (NEEDS_INV && OWNC) inv = iv[d];
                        else if constexpr (NEEDS_INV) {
                            // inverse diagonal from the voxel scales: diag = Kb_00 sum_v s_v (+ the shift)
                            inv = held ? 0.f : 1.f / (L.diagAdd ? k00 * sum.w + dA : k00 * sum.w);
                        }
                        if constexpr (MODE == AM_APPLY) yv[d] = held ? 0.f : out;
                        else if constexpr (MODE == AM_RAW) yv[d] = out;
                        else if constexpr (MODE == AM_RESID) yv[d] = held ? 0.f : rv[d] - out;
                        else if constexpr (MODE == AM_CHEB1) yv[d] = xv[d] + c1 * inv * (rv[d] - out);
                        else yv[d] = xv[d] + c1 * (xv[d] - dv[d]) + c2 * inv * (rv[d] - out);
                        if constexpr (DOT) {
                            if constexpr (MODE == AM_APPLY) dot0 += double(xv[d]) * double(yv[d]);
                            if constexpr (NEEDS_INV) {
                                dot0 += double(rv[d]) * double(yv[d]);
                                dot1 += double(yv[d]) * double(qv[d]);
                            }
                        }
                    }
                    bstore3(ry, noff[u], yv);
                }
                // the next pass rewrites the rows: every lane's reads first
                __builtin_amdgcn_s_waitcnt(0xc07f);
                __builtin_amdgcn_wave_barrier();
            }
        }
#pragma unroll
        for (int j = 0; j < R; j++) {
            carry[j] = up[j];
            syPrev[j] = sy[j];
        }
        // staging: node plane k + 2 into the slot of plane k - 1 (ring of 3) or of plane k (2 slots:
        // after every wavefront's reads of this layer) and voxel layer k + 1 into that of layer k - 1
        // (read before the last barrier), then the next fetches
        if constexpr (NS == 2) __syncthreads();
        storePlane(k + 2, pn);
        storeVox(k + 1, vn);
        fetchPlane(k + 3, pn);
        fetchVox(k + 2, vn);
        __syncthreads();
    }
    if constexpr (DOT) {
        // block partials in a fixed order: wavefront butterflies, then the waves in order
        __shared__ double wsum[2][NW];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            dot0 += __shfl_xor(dot0, off, 64);
            dot1 += __shfl_xor(dot1, off, 64);
        }
        if (lane == 0) {
            wsum[0][wid] = dot0;
            wsum[1][wid] = dot1;
        }
        __syncthreads();
        if (tid == 0) {
            double a = 0, b = 0;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                a += wsum[0][w];
                b += wsum[1][w];
            }
            dots[t] = a;
            if constexpr (NEEDS_INV) dots[nTiles + t] = b;
        }
    }
}

#define PSX_EL_FWD L, x, y, r, zp, c1, c2, stop, zChunk, gx, gy, nTiles, dots, dotStride, qdot
template <int MODE, bool DOT, int XP, int AP, bool OWNC> struct FineEl {
    static constexpr int NTH = NT, WPE = 3;
    __device__ static void run(VBlk vb, PSX_EL_PARAMS) { elBody<MODE, DOT, XP, AP, OWNC>(vb, PSX_EL_FWD); }
};
// held to 168 registers: 3 wavefronts per SIMD
template <int MODE, bool DOT, int XP, int AP, bool OWNC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void kFineEl(PSX_EL_PARAMS) {
    elBody<MODE, DOT, XP, AP, OWNC>(ownBlock(), PSX_EL_FWD);
}

// The child-scale level's own elements for the matrix-core operator: each element's product
// K_e (x_e - x_e0) = sum_c s_c M_c (x_e - x_e0) (M_c the Galerkin product of the finer level's
// uniform element at child position c, s_c the child's scale, x_e0 the element's corner-0 node: the
// products annihilate translations), written to L.ownT as [own][corner x + 2 y + 4 z][component].
// Each wavefront takes OWN_TILES tiles of 16 elements, one element per MFMA column: per child c and
// half (output corners 0..3, 4..7) the BF16 products of A = M_c's pieces (L.childA) and B = the
// pieces of the element's corner differences, FP32 accumulation, then y += s_c (A B) per column.
// Blocks of 8 wavefronts share one LDS copy of the A operands. Measured on the 512x128x128 block's
// level 1 (133 834 own elements; the level's Chebyshev step with this pass): one tile per wavefront in
// 512-thread blocks 46.5 us, two tiles 48.8 us, one tile in 256-thread blocks 49.5 us, in 1024-thread
// blocks 47.4 us.
constexpr int OWN_TILES = 1;
struct OwnRowsMf {
    static constexpr int NTH = 512, WPE = 0;
    __device__ static void run(VBlk vb, DevLevel L, const float* __restrict__ x, const int* stop, int64_t nOwn) {
        if (stop) stop += int64_t(vb.y) * CG_INTS;
        if (stop && __builtin_nontemporal_load(stop) != 0) return;
        x += int64_t(vb.y) * L.nDof;
        float* __restrict__ T = L.ownT + int64_t(vb.y) * L.ownTStride;
        const int lane = threadIdx.x & 63, n = lane & 15, g = lane >> 4;
        // the A operands (32 KB) in LDS, read by every wavefront of the block for each child
        __shared__ uint4 sA[8 * 2 * 2 * 64];
        for (int i = threadIdx.x; i < 8 * 2 * 2 * 64; i += NTH) sA[i] = L.childA[i];
        __syncthreads();
        const int64_t tile0 = (int64_t(vb.x) * (NTH / 64) + (threadIdx.x >> 6)) * OWN_TILES;
        if (tile0 * 16 >= nOwn) return;  // the whole wavefront past the last element
        const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
        const uint32_t nx = uint32_t(L.nx), ny = uint32_t(L.ny);
        // B operands: lane (n, g) holds element n's corners 2 g, 2 g + 1 (x = 0, 1; y = g & 1;
        // z = g >> 1), components 0..2 and a zero, as differences from corner 0
        bf16x8 B[OWN_TILES][2];
        float sc[OWN_TILES][8];
        int64_t own[OWN_TILES];
#pragma unroll
        for (int t = 0; t < OWN_TILES; t++) {
            own[t] = (tile0 + t) * 16 + n;
            const bool live = own[t] < nOwn;
            // (voxel indices are 31-bit: the voxel's coordinates by 32-bit divisions)
            const uint32_t V = live ? uint32_t(L.ownVox[own[t]]) : 0u, Vy = V / nx;
            const int64_t m0 = (V - Vy * nx) + NX * (Vy % ny) + NXY * (Vy / ny);
            const int64_t mg = m0 + NX * (g & 1) + NXY * (g >> 1);
            float x0[3] = {0.f, 0.f, 0.f}, xa[3] = {0.f, 0.f, 0.f}, xb[3] = {0.f, 0.f, 0.f};
            if (live) {  // 12-byte node loads
                __builtin_memcpy(x0, x + 3 * m0, 12);
                __builtin_memcpy(xa, x + 3 * mg, 12);
                __builtin_memcpy(xb, x + 3 * (mg + 1), 12);
            }
            uint32_t hi[8], lo[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const float v = (q & 3) == 3 ? 0.f : ((q < 4 ? xa[q & 3] : xb[q & 3]) - x0[q & 3]);
                hi[q] = bfBits(v);
                lo[q] = bfBits(v - bfValue(hi[q]));
            }
            typedef unsigned u4 __attribute__((ext_vector_type(4)));
            const u4 wh = {hi[0] | (hi[1] << 16), hi[2] | (hi[3] << 16), hi[4] | (hi[5] << 16), hi[6] | (hi[7] << 16)};
            const u4 wl = {lo[0] | (lo[1] << 16), lo[2] | (lo[3] << 16), lo[4] | (lo[5] << 16), lo[6] | (lo[7] << 16)};
            B[t][0] = __builtin_bit_cast(bf16x8, wh);
            B[t][1] = __builtin_bit_cast(bf16x8, wl);
            const float4* cs = reinterpret_cast<const float4*>(L.childS + 8 * (live ? own[t] : 0));
            const float4 s0 = cs[0], s1 = cs[1];
            const float sv[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
            for (int c = 0; c < 8; c++) sc[t][c] = live ? sv[c] : 0.f;
        }
        floatx4 Y[OWN_TILES][2];
#pragma unroll
        for (int t = 0; t < OWN_TILES; t++) Y[t][0] = Y[t][1] = floatx4{0.f, 0.f, 0.f, 0.f};
        const uint4* __restrict__ CA = sA;
#pragma unroll 2
        for (int c = 0; c < 8; c++) {
            bf16x8 A[2][2];
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const uint4 v = CA[((c * 2 + h) * 2 + p) * 64 + lane];
                    typedef unsigned u4 __attribute__((ext_vector_type(4)));
                    const u4 w = {v.x, v.y, v.z, v.w};
                    A[h][p] = __builtin_bit_cast(bf16x8, w);
                }
#pragma unroll
            for (int t = 0; t < OWN_TILES; t++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    floatx4 P = {0.f, 0.f, 0.f, 0.f};
                    P = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[h][1], B[t][0], P, 0, 0, 0);
                    P = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[h][0], B[t][1], P, 0, 0, 0);
                    P = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[h][0], B[t][0], P, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; r++) Y[t][h][r] += sc[t][c] * P[r];
                }
        }
        // lane (n, g) holds rows 4 g .. 4 g + 3 of each half: corner 4 h + g, components 0..2
#pragma unroll
        for (int t = 0; t < OWN_TILES; t++) {
            if (own[t] >= nOwn) continue;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const float v[3] = {Y[t][h][0], Y[t][h][1], Y[t][h][2]};
                __builtin_memcpy(T + 24 * own[t] + 3 * (4 * h + g), v, 12);
            }
        }
    }
};
__global__ __launch_bounds__(OwnRowsMf::NTH) void kOwnRowsMf(DevLevel L, const float* __restrict__ x, const int* stop, int64_t nOwn) {
    OwnRowsMf::run(ownBlock(), L, x, stop, nOwn);
}

__global__ void kOwnVox(const int32_t* __restrict__ words, int64_t nVox, int32_t* __restrict__ ownVox) {
    const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nVox) return;
    const int32_t w = words[v];
    if (w < 0) ownVox[w & 0x7fffffff] = int32_t(v);
}

}  // namespace

void launchOwnVox(const int32_t* words, int64_t nVox, int32_t* ownVox, hipStream_t s) {
    if (nVox > 0) kOwnVox<<<unsigned((nVox + 255) / 256), 256, 0, s>>>(words, nVox, ownVox);
}

// pieces: 2 = BF16 hi + lo of x - c and of Kb (the V-cycle's smoothing), 3 = three of each (y = K x only)
unsigned launchFineEl(const DevLevel& L, int mode, const float* x, float* y, const float* r, const float* zp, double c1, double c2, const int* stop,
                      double* dots, int64_t dotStride, const float* qdot, hipStream_t s, int ninst, int pieces) {
    const int gx = (L.NX + BX - 1) / BX, gy = (L.NY + R - 1) / R;
    // The longest z-chunk (16 or 8 planes) that still gives >= 2048 blocks (8 per CU): large levels are
    // bound by their traffic, which the halo planes of shorter chunks raise. A smaller level is bound
    // by each block's chain of plane steps: its chunk (2..8 planes) minimises the rounds of resident
    // blocks (4 per CU) times the planes a block steps through (its chunk and one halo plane).
    // Measured on the 512x128x128 block: 8 planes on level 0 (4 and 16: 2-11 % slower), 3 on level 1
    // (34.0 us against 40.0 at 2, 37.8 at 4, 43.5 at 5). The rule ignores the instance count, so that
    // an instance's tiles (and its fused dot partials) are those of a lone launch.
    static const int64_t slots = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        return int64_t(4) * cus;
    }();
    int zChunk = 0;
    for (int c : {16, 8})
        if (int64_t(gx) * gy * ((L.NZ + c - 1) / c) >= 2048) {
            zChunk = c;
            break;
        }
    if (zChunk == 0) {
        int64_t best = INT64_MAX;
        for (int c = 2; c <= 8; c++) {
            const int64_t rounds = (int64_t(gx) * gy * ((L.NZ + c - 1) / c) + slots - 1) / slots, cost = rounds * (c + 1);
            if (cost < best) {
                best = cost;
                zChunk = c;
            }
        }
    }
    if (const char* e = std::getenv("PSX_EL_ZCHUNK")) zChunk = std::max(2, std::atoi(e));
    const int gz = (L.NZ + zChunk - 1) / zChunk;
    const int nTiles = gx * gy * gz;
    const dim3 g(unsigned(8 * ((nTiles + 7) / 8)), unsigned(ninst));
    const bool dot = dots && (mode == AM_APPLY || mode == AM_CHEB1 || mode == AM_CHEB2);
    const float f1 = float(c1), f2 = float(c2);
    const bool ownc = L.childForm != 0;
    if (ownc) {
        // the own elements' corner rows of this product's input first
        const int64_t nOwn = L.ownTStride / 24, perBlock = 16 * OWN_TILES * (OwnRowsMf::NTH / 64);
        go<OwnRowsMf>(kOwnRowsMf, dim3(unsigned((nOwn + perBlock - 1) / perBlock), unsigned(ninst)), s, L, x, stop, nOwn);
    }
    auto form = [&](auto m, auto dt, auto pc) {
        constexpr int M = decltype(m)::value, P = decltype(pc)::value;
        constexpr bool DT = decltype(dt)::value;
        if (ownc) go<FineEl<M, DT, P, P, true>>(kFineEl<M, DT, P, P, true>, g, s, L, x, y, r, zp, f1, f2, stop, zChunk, gx, gy, nTiles, dots, dotStride, qdot);
        else go<FineEl<M, DT, P, P, false>>(kFineEl<M, DT, P, P, false>, g, s, L, x, y, r, zp, f1, f2, stop, zChunk, gx, gy, nTiles, dots, dotStride, qdot);
    };
    // three pieces for the outer operator (y = K x); the smoothing steps take two
    auto byPieces = [&](auto m, auto dt) {
        if constexpr (decltype(m)::value == AM_APPLY) {
            if (pieces >= 3) {
                form(m, dt, std::integral_constant<int, 3>{});
                return;
            }
        }
        form(m, dt, std::integral_constant<int, 2>{});
    };
    using yes = std::true_type;
    using no = std::false_type;
    switch (mode) {
        case AM_APPLY:
            if (dot) byPieces(std::integral_constant<int, AM_APPLY>{}, yes{});
            else byPieces(std::integral_constant<int, AM_APPLY>{}, no{});
            break;
        case AM_RAW: byPieces(std::integral_constant<int, AM_RAW>{}, no{}); break;
        case AM_RESID: byPieces(std::integral_constant<int, AM_RESID>{}, no{}); break;
        case AM_CHEB1:
            if (dot) byPieces(std::integral_constant<int, AM_CHEB1>{}, yes{});
            else byPieces(std::integral_constant<int, AM_CHEB1>{}, no{});
            break;
        default:
            if (dot) byPieces(std::integral_constant<int, AM_CHEB2>{}, yes{});
            else byPieces(std::integral_constant<int, AM_CHEB2>{}, no{});
            break;
    }
    return dot ? unsigned(nTiles) : 0u;
}

}  // namespace psx
