// Element-by-element operator on the matrix cores, for FP32 levels whose elements are all s * Kb
// (every finest level, uniform or per-voxel scales): y = K x with the fused epilogues of kFine.
//
// K x = sum over voxels e of s_e Kb x_e, and Kb annihilates rigid translations, so the product is
// taken of x' = x - c, c one node's value per block (the tile's centre): only the spread of x over
// the tile enters the arithmetic, not its magnitude. x' is split into BF16 pieces (hi = bf16(x'),
// lo = bf16(x' - hi), ...), Kb likewise, and each element's 24 x 24 product is a sum of
// v_mfma_f32_16x16x32_bf16 products of the pieces with FP32 accumulation; the piece pairs whose
// weight is below 2^-8 per piece order are left out (two pieces each: hi.hi + hi.lo + lo.hi, an
// error of about 2^-17 |Kb| |x'|; three: six pairs, about 2^-25).
//
// Layout of the work: a block of 256 threads owns the same 32 x 8 column of nodes as kFine and
// marches along z one element layer (the voxels between node planes k and k + 1) per step:
//   1. the layer's 33 x 9 elements, 16 per matrix instruction: columns = elements, K = the 8 corners
//      x 4 components (the fourth zero) read as two adjacent nodes per lane from the LDS node planes
//      (a node's 3 components + a zero, as 4 BF16: one 16-byte read gives the lane its K slice),
//      rows = 4 output corners x 4 components (two instructions per piece pair cover the 8 corners);
//   2. the products go to LDS, [component][corner][element];
//   3. each node sums the corner rows of its 4 elements in the layer, times their voxel scales, in a
//      fixed order: the lower corners complete node k, the upper ones start node k + 1.
// Nothing depends on which voxels differ in scale, so void surfaces and per-voxel densities cost the
// same as solid regions (no pair lists).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <type_traits>

#include "device.hpp"
#include "pack_dev.hpp"
#include "stencil.hpp"

namespace psx {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int TX = 32, TY = 8, NT = TX * TY, NW = NT / 64;
constexpr int PX = TX + 2, PY = TY + 2, NP = PX * PY;  // node plane with halo (34 x 10)
constexpr int EX = TX + 1, EY = TY + 1, NE = EX * EY;  // element layer (33 x 9)
constexpr int NGMAX = (NE + 15) / 16;                  // matrix-instruction groups of 16 elements
constexpr int EPAD = NGMAX * 16;
constexpr int GPW = (NGMAX + NW - 1) / NW;              // groups per wavefront at most (a contiguous run)
constexpr int NVL = (NE + NT - 1) / NT;                 // voxels staged per thread

template <int XP> struct ElShared {
    uint2 xq[2][XP][NP];     // node planes k, k + 1: the pieces of x - c, 4 BF16 per node
    float s[3][NE];          // voxel scales of element layers k - 1 .. k + 1 (0: empty or outside)
    float4 e[8][EPAD];       // the layer's products Kb (x_e - c): [corner][element][component, 0]
};

__device__ __forceinline__ uint32_t bfBits(float v) { return uint32_t(__builtin_bit_cast(uint16_t, (__bf16)v)); }
__device__ __forceinline__ float bfValue(uint32_t b) { return __uint_as_float(b << 16); }

#define PSX_EL_PARAMS                                                                                                                          \
    DevLevel L, const float *__restrict__ x, float *__restrict__ y, const float *__restrict__ r, const float *__restrict__ zp, float c1, float c2, \
        const int *stop, int zChunk, int gx, int gy, int nTiles, double *__restrict__ dots, int64_t dotStride, const float *__restrict__ qdot

template <int MODE, bool DOT, int XP, int AP> __device__ __forceinline__ void elBody(VBlk vb, PSX_EL_PARAMS) {
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
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int i0 = (t % gx) * TX, j0 = ((t / gx) % gy) * TY;
    const int z0 = (t / (gx * gy)) * zChunk, z1 = min(L.NZ, z0 + zChunk);
    // the last tile column of a part (usually one node column) takes the transposed mapping, 8 columns
    // of 8 rows per wavefront, so that whole wavefronts past the part's edge skip the node work
    const bool transposed = L.NX - i0 <= TX / 4;
    const int tx = transposed ? tid / TY : tid % TX, ty = transposed ? tid % TY : tid / TX;
    const int i = i0 + tx, j = j0 + ty;
    const bool mine = i < L.NX && j < L.NY;
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
    constexpr bool NEEDS_R = MODE == AM_RESID || MODE == AM_CHEB1 || MODE == AM_CHEB2;
    constexpr bool NEEDS_D = MODE == AM_CHEB2;
    constexpr bool NEEDS_INV = MODE == AM_CHEB1 || MODE == AM_CHEB2;

    // the block's reference value c: the node at the tile's centre (uniform loads)
    float cref[3];
    {
        const int ic = min(i0 + TX / 2, L.NX - 1), jc = min(j0 + TY / 2, L.NY - 1), zc = min(z0 + zChunk / 2, L.NZ - 1);
        const int64_t n3 = 3 * (ic + NX * jc + NXY * zc);
        cref[0] = x[n3];
        cref[1] = x[n3 + 1];
        cref[2] = x[n3 + 2];
    }

    // the tile's element layer: ncx x ncy elements (local element (ex, ey) = voxel (i0 - 1 + ex,
    // j0 - 1 + ey)), flattened as ex + EX ey, or ey + EY ex in a transposed tile (a part's last tile
    // column: few columns), into groups of 16; the elements past ncx in a row of a tile at the x edge
    // are computed and never gathered
    const int ncx = min(TX, L.NX - i0) + 1, ncy = min(TY, L.NY - j0) + 1;
    const int nEl = transposed ? EY * ncx : EX * ncy, nGroups = (nEl + 15) >> 4;
    auto elemXY = [&](int e, int& ex, int& ey) {
        if (transposed) ex = e / EY, ey = e % EY;
        else ex = e % EX, ey = e / EX;
    };

    // matrix-core operands of Kb (see DevLevel::elA): A[h][piece] for output corners 4 h .. 4 h + 3
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

    // this wavefront's element groups: the contiguous run wid * GPW + u, u < myGroups; for each, the
    // lane's node pair in the node plane (its K slice: corners at y offset (lane / 16) & 1, z offset
    // lane / 32)
    const int myGroups = max(0, min(GPW, nGroups - wid * GPW));
    int bOff[GPW];
#pragma unroll
    for (int u = 0; u < GPW; u++) {
        const int e = 16 * (wid * GPW + u) + (lane & 15);
        int ex, ey;
        elemXY(e < nEl ? e : 0, ex, ey);
        bOff[u] = (ey + ((lane >> 4) & 1)) * PX + ex;
    }
    const int zOff = lane >> 5;  // the lane's corners lie on node plane k + zOff

    // staging: each thread stages its own node (interior position of the plane) and threads
    // 0 .. NP - NT - 1 one halo node each; offsets OUT_OF_RANGE outside the part (loads return 0)
    int lpos[2];
    int32_t poff[2];
    {
        lpos[0] = (ty + 1) * PX + tx + 1;
        const int h = tid;
        int hx, hy;
        if (h < PX) hx = h, hy = 0;
        else if (h < 2 * PX) hx = h - PX, hy = PY - 1;
        else if (h < 2 * PX + TY) hx = 0, hy = h - 2 * PX + 1;
        else hx = PX - 1, hy = h - 2 * PX - TY + 1;
        lpos[1] = h < NP - NT ? hy * PX + hx : -1;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int lp = lpos[q] < 0 ? 0 : lpos[q];
            const int gi = i0 - 1 + lp % PX, gj = j0 - 1 + lp / PX;
            const bool in = lpos[q] >= 0 && gi >= 0 && gj >= 0 && gi < L.NX && gj < L.NY;
            poff[q] = in ? int32_t(12 * (gi + NX * gj)) : OUT_OF_RANGE;
        }
    }
    int32_t voff[NVL];
#pragma unroll
    for (int q = 0; q < NVL; q++) {
        const int e = tid + q * NT;
        const int vi = i0 - 1 + e % EX, vj = j0 - 1 + e / EX;
        const bool in = e < NE && vi >= 0 && vj >= 0 && vi < L.nx && vj < L.ny;
        voff[q] = in ? int32_t(4 * (vi + int64_t(L.nx) * vj)) : OUT_OF_RANGE;
    }
    const uint32_t planeBytes = uint32_t(12 * NXY);
    const int64_t nxy = int64_t(L.nx) * L.ny;
    auto fetchPlane = [&](int z, float (&px)[2][3]) {
        const bool zin = z >= 0 && z < L.NZ;
        const auto rs = planeRsrc(x + 3 * NXY * (zin ? z : 0), zin ? planeBytes : 0u);
#pragma unroll
        for (int q = 0; q < 2; q++) bload3(rs, poff[q], px[q]);
    };
    auto storePlane = [&](int z, const float (&px)[2][3]) {
        const int sl = z & 1;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (lpos[q] < 0) continue;
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
                sh.xq[sl][p][lpos[q]] = uint2{b[0] | (b[1] << 16), b[2]};
            }
        }
    };
    auto fetchVox = [&](int k, float (&pv)[NVL]) {
        const bool kin = k >= 0 && k < L.nz;
        const auto rs = planeRsrc(L.words + nxy * (kin ? k : 0), kin ? uint32_t(4 * nxy) : 0u);
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int32_t w = int32_t(__builtin_amdgcn_raw_buffer_load_b32(rs, voff[q], 0, 0));
            pv[q] = w > 0 ? __int_as_float(w) : 0.f;
        }
    };
    auto storeVox = [&](int k, const float (&pv)[NVL]) {
        const int sl = (k + 3) % 3;
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int e = tid + q * NT;
            if (e < NE) sh.s[sl][e] = pv[q];
        }
    };
    // the epilogue's own operands of node (i, j, z), fetched one step before use
    struct Own {
        float x[3], r[3], d[3], qd[3];
        uint32_t info;
    };
    const int32_t col = mine ? int32_t(i + NX * j) : 0;
    const int32_t noff = mine ? 12 * col : OUT_OF_RANGE;
    const int32_t ioff = mine ? 2 * col : OUT_OF_RANGE;
    auto fetchOwn = [&](int z, Own& o) {
        const bool zin = z >= z0 && z < z1;
        const int64_t pb = NXY * (zin ? z : z0);
        const uint32_t vbytes = zin ? planeBytes : 0u;
        bload3(planeRsrc(x + 3 * pb, vbytes), noff, o.x);
        if constexpr (NEEDS_R) bload3(planeRsrc(r + 3 * pb, vbytes), noff, o.r);
        if constexpr (NEEDS_D) bload3(planeRsrc(zp ? zp + 3 * pb : x, zp ? vbytes : 0u), noff, o.d);
        if constexpr (NEEDS_INV) {
            if constexpr (DOT) bload3(planeRsrc(qdot ? qdot + 3 * pb : x, qdot ? vbytes : 0u), noff, o.qd);
            else o.qd[0] = o.qd[1] = o.qd[2] = 0.f;
        }
        o.info = __builtin_amdgcn_raw_buffer_load_b16(planeRsrc(L.nodeInfo + pb, zin ? uint32_t(2 * NXY) : 0u), ioff, 0, 0);
    };
    // the node's 4 elements in a layer: local element (tx + dx, ty + dy) holds the node as its corner
    // (1 - dx, 1 - dy, .)
    int eIdx[4], sIdx[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int dx = q & 1, dy = q >> 1;
        eIdx[q] = transposed ? (ty + dy) + EY * (tx + dx) : (tx + dx) + EX * (ty + dy);
        sIdx[q] = (tx + dx) + EX * (ty + dy);
    }

    // prologue: node planes z0 - 1, z0 and voxel layer z0 - 1 in LDS; plane z0 + 1 and layer z0 in flight
    float pn[2][3], vn[NVL];
    {
        float pa[2][3], pb[2][3], va[NVL];
        fetchPlane(z0 - 1, pa);
        fetchPlane(z0, pb);
        fetchVox(z0 - 1, va);
        fetchPlane(z0 + 1, pn);
        fetchVox(z0, vn);
        storePlane(z0 - 1, pa);
        storePlane(z0, pb);
        storeVox(z0 - 1, va);
    }
    Own cur{};  // node z0 - 1 has no epilogue; node z0's operands arrive in the first step
    float carry[3] = {0.f, 0.f, 0.f};
    double dot0 = 0, dot1 = 0;
    __syncthreads();

    for (int k = z0 - 1; k < z1; k++) {
        // 1. the layer's element products: the wavefront's groups, every operand read first and the
        //    piece products of all groups interleaved (independent accumulators)
        {
            // node plane k + zOff sits in slot (k + zOff) & 1
            const uint2* __restrict__ xs = &sh.xq[(k + zOff) & 1][0][0];
            auto run = [&](auto cnt) {
                constexpr int CNT = decltype(cnt)::value;
                bf16x8 B[CNT > 0 ? CNT : 1][XP];
#pragma unroll
                for (int u = 0; u < CNT; u++)
#pragma unroll
                    for (int p = 0; p < XP; p++) {
                        const uint2 lo = xs[p * NP + bOff[u]], hi = xs[p * NP + bOff[u] + 1];
                        typedef unsigned u4 __attribute__((ext_vector_type(4)));
                        const u4 w = {lo.x, lo.y, hi.x, hi.y};
                        B[u][p] = __builtin_bit_cast(bf16x8, w);
                    }
                const int g = lane >> 4;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    floatx4 acc[CNT > 0 ? CNT : 1];
#pragma unroll
                    for (int u = 0; u < CNT; u++) acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
                    // the smaller piece products first
#pragma unroll
                    for (int ord = (AP > XP ? AP : XP) - 1; ord >= 0; ord--)
#pragma unroll
                        for (int a = 0; a < AP; a++) {
                            const int b = ord - a;
                            if (b >= 0 && b < XP) {
#pragma unroll
                                for (int u = 0; u < CNT; u++) acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[h][a], B[u][b], acc[u], 0, 0, 0);
                            }
                        }
#pragma unroll
                    for (int u = 0; u < CNT; u++) sh.e[4 * h + g][16 * (wid * GPW + u) + (lane & 15)] = float4{acc[u][0], acc[u][1], acc[u][2], acc[u][3]};
                }
            };
            static_assert(GPW == 5, "group runs of at most 5");
            switch (myGroups) {
                case 5: run(std::integral_constant<int, 5>{}); break;
                case 4: run(std::integral_constant<int, 4>{}); break;
                case 3: run(std::integral_constant<int, 3>{}); break;
                case 2: run(std::integral_constant<int, 2>{}); break;
                case 1: run(std::integral_constant<int, 1>{}); break;
                default: break;
            }
        }
        Own nxt;
        fetchOwn(k + 1, nxt);
        __syncthreads();

        // 2. node k: the upper corners of layer k - 1 (carry) and the lower corners of layer k;
        //    node k + 1 starts with the upper corners of layer k
        if (mine) {
            const int sl = (k + 3) % 3;
            float lower[3] = {0.f, 0.f, 0.f}, upper[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int dx = q & 1, dy = q >> 1;
                const int kl = (1 - dx) + 2 * (1 - dy);  // corner x + 2 y + 4 z of the node in element q
                const float s = sh.s[sl][sIdx[q]];
                const float4 lo = sh.e[kl][eIdx[q]], up = sh.e[kl + 4][eIdx[q]];
                lower[0] += s * lo.x;
                lower[1] += s * lo.y;
                lower[2] += s * lo.z;
                upper[0] += s * up.x;
                upper[1] += s * up.y;
                upper[2] += s * up.z;
            }
            if (k >= z0) {
                float out[3];
#pragma unroll
                for (int c = 0; c < 3; c++) out[c] = carry[c] + lower[c];
                const int64_t pb3 = 3 * NXY * k;
                const int32_t c3 = 3 * col;
                if (L.diagAdd) {
#pragma unroll
                    for (int d = 0; d < 3; d++) out[d] += float((L.diagAdd + pb3)[c3 + d]) * cur.x[d];
                }
                float inv[3] = {0.f, 0.f, 0.f};
                if constexpr (NEEDS_INV) {
                    // inverse diagonal from the voxel scales: diag = Kb_00 sum_v s_v (+ the shift), in
                    // kFine's voxel order
                    const int sm = (k + 2) % 3;
                    float ssum = 0.f;
#pragma unroll
                    for (int v = 0; v < 8; v++) ssum += sh.s[(v >> 2) ? sl : sm][(tx + (v & 1)) + EX * (ty + ((v >> 1) & 1))];
                    const float k00 = L.sym32[0];
                    if (L.diagAdd) {
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const float dg = k00 * ssum + float((L.diagAdd + pb3)[c3 + d]);
                            inv[d] = ((cur.info >> (8 + d)) & 1u) ? 0.f : 1.f / dg;
                        }
                    } else {
                        const float iv = 1.f / (k00 * ssum);
#pragma unroll
                        for (int d = 0; d < 3; d++) inv[d] = ((cur.info >> (8 + d)) & 1u) ? 0.f : iv;
                    }
                }
                float yv[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const bool held = (cur.info >> (8 + d)) & 1u;
                    if constexpr (MODE == AM_APPLY) yv[d] = held ? 0.f : out[d];
                    else if constexpr (MODE == AM_RAW) yv[d] = out[d];
                    else if constexpr (MODE == AM_RESID) yv[d] = held ? 0.f : cur.r[d] - out[d];
                    else if constexpr (MODE == AM_CHEB1) yv[d] = cur.x[d] + c1 * inv[d] * (cur.r[d] - out[d]);
                    else yv[d] = cur.x[d] + c1 * (cur.x[d] - cur.d[d]) + c2 * inv[d] * (cur.r[d] - out[d]);
                    if constexpr (DOT) {
                        if constexpr (MODE == AM_APPLY) dot0 += double(cur.x[d]) * double(yv[d]);
                        if constexpr (NEEDS_INV) {
                            dot0 += double(cur.r[d]) * double(yv[d]);
                            dot1 += double(yv[d]) * double(cur.qd[d]);
                        }
                    }
                }
                bstore3(planeRsrc(y + pb3, planeBytes), noff, yv);
            }
#pragma unroll
            for (int c = 0; c < 3; c++) carry[c] = upper[c];
        }

        // 3. staging: node plane k + 2 into the slot of plane k (read only by this step's products,
        //    before its first barrier) and voxel layer k + 1 into that of layer k - 2, then the next
        //    fetches
        storePlane(k + 2, pn);
        storeVox(k + 1, vn);
        fetchPlane(k + 3, pn);
        fetchVox(k + 2, vn);
        cur = nxt;
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
template <int MODE, bool DOT, int XP, int AP> struct FineEl {
    static constexpr int NTH = NT, WPE = 0;
    __device__ static void run(VBlk vb, PSX_EL_PARAMS) { elBody<MODE, DOT, XP, AP>(vb, PSX_EL_FWD); }
};
template <int MODE, bool DOT, int XP, int AP> __global__ __launch_bounds__(NT) void kFineEl(PSX_EL_PARAMS) {
    elBody<MODE, DOT, XP, AP>(ownBlock(), PSX_EL_FWD);
}

}  // namespace

// pieces: 2 = BF16 hi + lo of x - c and of Kb (the V-cycle's smoothing), 3 = three of each
void launchFineEl(const DevLevel& L, int mode, const float* x, float* y, const float* r, const float* zp, double c1, double c2, const int* stop, int zChunk,
                  double* dots, int64_t dotStride, const float* qdot, hipStream_t s, int ninst, int pieces) {
    const int gx = (L.NX + TX - 1) / TX, gy = (L.NY + TY - 1) / TY, gz = (L.NZ + zChunk - 1) / zChunk;
    const int nTiles = gx * gy * gz;
    const dim3 g(unsigned(8 * ((nTiles + 7) / 8)), unsigned(ninst));
    const bool dot = dots && (mode == AM_APPLY || mode == AM_CHEB1 || mode == AM_CHEB2);
    const float f1 = float(c1), f2 = float(c2);
    auto form = [&](auto m, auto dt, auto pc) {
        constexpr int M = decltype(m)::value, P = decltype(pc)::value;
        constexpr bool DT = decltype(dt)::value;
        go<FineEl<M, DT, P, P>>(kFineEl<M, DT, P, P>, g, s, L, x, y, r, zp, f1, f2, stop, zChunk, gx, gy, nTiles, dots, dotStride, qdot);
    };
    auto byPieces = [&](auto m, auto dt) {
        if (pieces >= 3) form(m, dt, std::integral_constant<int, 3>{});
        else form(m, dt, std::integral_constant<int, 2>{});
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
}

}  // namespace psx
