// Operator kernel for levels whose elements are all s * Kb (every finest level, and the coarse
// levels of uniform parts): y = K x with the fused epilogues of kApply (kernels.hip).
//
// Layout of the work: a block of 256 threads owns a 32 x 8 column of nodes in x-y and marches along
// z. Node planes of x (with a one-node halo) and voxel-scale planes are staged in LDS rings, two
// planes ahead of use. Each plane is read from LDS once per thread: its 9 x-y neighbours feed three
// rolling accumulators, the stencil rows of the nodes above, at and below the plane, so a node's
// 27-point sum is complete one step after its upper neighbour plane arrives.
//
// Nodes whose 8 voxels do not all carry the same scale (the surface of a void, the domain boundary)
// use the difference form K_node x = s_ref S x + sum over differing voxels v of (s_v - s_ref) E_v x,
// where S is the assembled stencil, s_ref the scale of the node's most common voxel value and E_v
// the 3 element rows of the node's corner in voxel v. The differing (node, voxel) pairs of a plane
// (one bit each in a per-node byte precomputed on the host) are packed into a block list and
// computed one pair per lane, each mirrored onto element corner 0 so that every lane uses the same
// compile-time coefficients; the owner then adds its pairs in voxel order (deterministic sums).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include "device.hpp"
#include "pack_dev.hpp"
#include "stencil.hpp"

namespace psx {

namespace {

constexpr int TX = 32, TY = 8, NT = TX * TY;
constexpr int PX = TX + 2, PY = TY + 2;  // node plane with halo
constexpr int VX = TX + 1, VY = TY + 1;  // voxel plane


template <class T> struct FineP;
template <> struct FineP<double> {
    __device__ static const double* invD(const DevLevel& L) { return L.invD64; }
};
template <> struct FineP<float> {
    __device__ static const float* invD(const DevLevel& L) { return L.invD32; }
};

constexpr int NW = NT / 64;    // wavefronts per block
constexpr int WAVE_PAIRS = 512;  // 64 nodes x 8 voxels

template <class T, bool OWNC> struct Ring {
    T x[4][3][PY][PX];                 // node planes k-2 .. k+1, one array per component
    float s[3][VY][VX];                // voxel-scale planes (0 = empty, outside or own matrix)
    int own[OWNC ? 3 : 1][VY][VX];     // own-matrix index of the voxel, -1 for s * Kb
    int4 two[OWNC ? 3 : 1][OWNC ? VY : 1][OWNC ? VX : 1];  // the own voxel's two-scale record (childTwo)
    T pres[3][NT];                     // results of one pass of pairs
    uint16_t list[NW][WAVE_PAIRS];     // pairs: local node << 6 | reference voxel << 3 | voxel
    int waveTot[2][NW];
};

// OWNC: the level has own elements in child-scale form (sum_c s_c M_c, crow = the corner-0 rows
// of the 8 M_c, [c][row d][24]); an own voxel is always a differing voxel of its nodes
#define PSX_FINE_PARAMS                                                                                                                    \
    DevLevel L, Sym<T> C, const T *__restrict__ crow, const T *__restrict__ ctab, const T *__restrict__ x, T *__restrict__ y,                 \
        const T *__restrict__ r, const T *__restrict__ zp, T c1, T c2, const int *stop, int zChunk, int gx, int gy, int nTiles, int xcdOrder, \
        double *__restrict__ dots, int64_t dotStride, const T *__restrict__ qdot, const T *__restrict__ pz, T *__restrict__ pw
// DOT: the fused dot products are accumulated (dots != nullptr). DENSE (levels without own
// elements): every node sums its 8 voxels, s_v E_v (x - x_node) with E_v the node's 3 rows of Kb in
// voxel v (compile-time class coefficients), instead of the stencil and the pair lists; for per-voxel
// densities, where nearly every node would list 7 or 8 pairs.
// H16 (FP32 levels with child-scale own elements, the mixed V-cycle's smoothing): the two-scale
// elements read the subset table in FP16 (ctab16, each row times its power-of-two ctabRowScale), which
// fits the compute unit's cache where the FP32 table does not
template <class T, int MODE, bool OWNC, bool DOT, bool DENSE = false, bool H16 = false>
__device__ __forceinline__ void fineBody(VBlk vb, PSX_FINE_PARAMS) {
    // multi-instance launches: the block's y is the right-hand side (its vectors, stop flag, partials)
    if (stop) stop += int64_t(vb.y) * CG_INTS;
    if (stop && __builtin_nontemporal_load(stop) != 0) return;
    {
        const int64_t io = int64_t(vb.y) * L.nDof;
        x += io;
        y += io;
        if (r) r += io;
        if (zp) zp += io;
        if (qdot) qdot += io;
        if (pz) pz += io;
        if (pw) pw += io;
        if (DOT) dots += int64_t(vb.y) * dotStride;
    }
    // the conjugate gradients' direction update fused into y = K p (pz: the preconditioned residual z,
    // x: the previous direction): the staged planes are p = z + beta x, with beta from the instance's
    // state (the record that holds its stop flag), and the owning threads store p to pw (beside x,
    // never over it: neighbouring tiles still read the previous direction in their halos)
    T beta = T(0);
    if constexpr (MODE == AM_APPLY) {
        if (pz) beta = T(reinterpret_cast<const CgState*>(reinterpret_cast<const char*>(stop) - offsetof(CgState, stop))->beta);
    }
    // XCD-aware tile order: the hardware deals consecutive blocks round-robin to the 8 XCDs, so
    // block b takes tile (b % 8) * per + b / 8 and each XCD walks one contiguous run of tiles (x
    // fastest), whose halos then meet in that XCD's L2
    const int per = int(vb.gx / 8), t = xcdOrder ? int(vb.x % 8) * per + int(vb.x / 8) : int(vb.x);
    if (t >= nTiles) return;
    __shared__ Ring<T, OWNC> sh;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int i0 = (t % gx) * TX, j0 = ((t / gx) % gy) * TY;
    // A wavefront holds two node rows of the tile, so in the last tile row of a part (NY = 2^k + 1
    // leaves one row there) the waves past the part's edge hold no node and skip the arithmetic. The
    // last tile column (usually one node column) takes the transposed mapping, 8 columns of 8 rows
    // per wavefront, for the same effect.
    const bool transposed = L.NX - i0 <= TX / 4;
    const int tx = transposed ? tid / TY : tid % TX, ty = transposed ? tid % TY : tid / TX;
    const int z0 = (t / (gx * gy)) * zChunk, z1 = min(L.NZ, z0 + zChunk);
    const int i = i0 + tx, j = j0 + ty;
    const bool mine = i < L.NX && j < L.NY;
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
    const T* __restrict__ invD = FineP<T>::invD(L);
    constexpr bool NEEDS_R = MODE == AM_RESID || MODE == AM_CHEB1 || MODE == AM_CHEB2;
    constexpr bool NEEDS_D = MODE == AM_CHEB2;

    constexpr int NPL = (PX * PY + NT - 1) / NT, NVL = (VX * VY + NT - 1) / NT;
    // this thread's share of a node plane / voxel plane: byte offsets within the plane, fixed for the
    // march, OUT_OF_RANGE for positions outside the part (their loads return 0); the plane's
    // descriptor carries its block-uniform base
    int32_t poff[NPL], voff[NVL];
#pragma unroll
    for (int q = 0; q < NPL; q++) {
        const int e = tid + q * NT;
        const int gi = i0 - 1 + e % PX, gj = j0 - 1 + e / PX;
        const bool in = e < PX * PY && gi >= 0 && gj >= 0 && gi < L.NX && gj < L.NY;
        poff[q] = in ? int32_t(int64_t(sizeof(T)) * 3 * (gi + NX * gj)) : OUT_OF_RANGE;
    }
#pragma unroll
    for (int q = 0; q < NVL; q++) {
        const int e = tid + q * NT;
        const int vi = i0 - 1 + e % VX, vj = j0 - 1 + e / VX;
        const bool in = e < VX * VY && vi >= 0 && vj >= 0 && vi < L.nx && vj < L.ny;
        voff[q] = in ? int32_t(4 * (vi + int64_t(L.nx) * vj)) : OUT_OF_RANGE;
    }
    const int64_t nxy = int64_t(L.nx) * L.ny;
    const uint32_t planeBytes = uint32_t(3 * NXY * int64_t(sizeof(T)));
    auto fetchPlane = [&](int z, T (&px)[NPL][3]) {
        const bool zin = z >= 0 && z < L.NZ;
        const auto rs = planeRsrc(x + 3 * NXY * (zin ? z : 0), zin ? planeBytes : 0u);
#pragma unroll
        for (int q = 0; q < NPL; q++) bload3(rs, poff[q], px[q]);
        if constexpr (MODE == AM_APPLY) {
            if (pz) {
                const auto rz = planeRsrc(pz + 3 * NXY * (zin ? z : 0), zin ? planeBytes : 0u);
#pragma unroll
                for (int q = 0; q < NPL; q++) {
                    T zv[3];
                    bload3(rz, poff[q], zv);
#pragma unroll
                    for (int c = 0; c < 3; c++) px[q][c] = zv[c] + beta * px[q][c];
                }
            }
        }
    };
    auto storePlane = [&](int slot, const T (&px)[NPL][3]) {
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            const int e = tid + q * NT;
            if (e < PX * PY) {
                sh.x[slot][0][e / PX][e % PX] = px[q][0];
                sh.x[slot][1][e / PX][e % PX] = px[q][1];
                sh.x[slot][2][e / PX][e % PX] = px[q][2];
            }
        }
    };
    // voxel words in flight: the word itself (scale bits, or the own index with the sign bit)
    auto fetchVox = [&](int k, int32_t (&pv)[NVL]) {
        const bool kin = k >= 0 && k < L.nz;
        const auto rs = planeRsrc(L.words + nxy * (kin ? k : 0), kin ? uint32_t(4 * nxy) : 0u);
#pragma unroll
        for (int q = 0; q < NVL; q++) pv[q] = int32_t(__builtin_amdgcn_raw_buffer_load_b32(rs, voff[q], 0, 0));
    };
    auto storeVox = [&](int slot, const int32_t (&pv)[NVL]) {
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int e = tid + q * NT;
            if (e < VX * VY) {
                sh.s[slot][e / VX][e % VX] = pv[q] > 0 ? __int_as_float(pv[q]) : 0.f;
                if constexpr (OWNC) sh.own[slot][e / VX][e % VX] = pv[q] < 0 ? (pv[q] & 0x7fffffff) : -1;
            }
        }
    };
    // the two-scale records of a voxel plane's own voxels, loaded as soon as the plane's words are
    // known and written to LDS just before the step's barrier: the latency hides behind the stencil
    // instead of sitting in every own pair's pass (range-checked: other voxels read 0)
    const auto twoRsrc = planeRsrc(L.childTwo, OWNC && L.childTwo ? 0x7fffffffu : 0u);
    auto fetchTwo = [&](const int32_t (&pv)[NVL], int4 (&tw)[NVL]) {
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int32_t off = pv[q] < 0 ? int32_t(16 * (pv[q] & 0x7fffffff)) : OUT_OF_RANGE;
            const auto u = __builtin_amdgcn_raw_buffer_load_b128(twoRsrc, off, 0, 0);
            tw[q] = int4{int(u[0]), int(u[1]), int(u[2]), int(u[3])};
        }
    };
    auto storeTwo = [&](int slot, const int4 (&tw)[NVL]) {
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int e = tid + q * NT;
            if (e < VX * VY) sh.two[slot][e / VX][e % VX] = tw[q];
        }
    };
    // the epilogue's own operands of node (i, j, z), fetched one step before use; info = the node's
    // differing-voxel bits (0..7) and held-DOF bits (8..10)
    constexpr bool NEEDS_INV = MODE == AM_CHEB1 || MODE == AM_CHEB2;
    struct Own {
        T inv[3], r[3], d[3], qd[3];
        uint32_t info;
    };
    const int32_t col = mine ? int32_t(i + NX * j) : 0;
    const int32_t noff = mine ? int32_t(int64_t(sizeof(T)) * 3 * col) : OUT_OF_RANGE;  // the node's vector entries
    const int32_t ioff = mine ? int32_t(2 * col) : OUT_OF_RANGE;                          // its info word
    auto fetchOwn = [&](int z, Own& o) {
        const bool zin = z >= z0 && z < z1;
        const int64_t pb = NXY * (zin ? z : z0);  // block-uniform plane base
        const uint32_t vb = zin ? planeBytes : 0u;
        if constexpr (NEEDS_INV && OWNC) bload3(planeRsrc(invD + 3 * pb, vb), noff, o.inv);
        if constexpr (NEEDS_R) bload3(planeRsrc(r + 3 * pb, vb), noff, o.r);
        if constexpr (NEEDS_D) bload3(planeRsrc(zp ? zp + 3 * pb : x, zp ? vb : 0u), noff, o.d);
        if constexpr (NEEDS_INV) {
            if constexpr (DOT) bload3(planeRsrc(qdot ? qdot + 3 * pb : x, qdot ? vb : 0u), noff, o.qd);
            else o.qd[0] = o.qd[1] = o.qd[2] = T(0);
        }
        o.info = __builtin_amdgcn_raw_buffer_load_b16(planeRsrc(L.nodeInfo + pb, zin ? uint32_t(2 * NXY) : 0u), ioff, 0, 0);
    };

    // prologue: node plane z0 - 1 in LDS; node plane z0 and voxel plane z0 - 2 in flight
    T pn[NPL][3];
    int32_t vn[NVL];
    {
        T pa[NPL][3];
        fetchPlane(z0 - 1, pa);
        fetchPlane(z0, pn);
        fetchVox(z0 - 2, vn);
        storePlane((z0 - 1) & 3, pa);
    }
    Own cur;
    cur.info = 0;
    T aN[3] = {0, 0, 0}, aC[3] = {0, 0, 0}, aP[3] = {0, 0, 0}, xcPrev[3] = {0, 0, 0}, xcPrev2[3] = {0, 0, 0};
    double dot0 = 0, dot1 = 0;  // fused dot products (DOT)
    __syncthreads();

    for (int k = z0 - 1; k <= z1; k++) {
        // stores of the previous step's fetches, then the next fetches
        storePlane((k + 1) & 3, pn);
        storeVox((k + 2) % 3, vn);  // voxel plane k - 1
        int4 twv[NVL];
        if constexpr (OWNC) fetchTwo(vn, twv);
        fetchPlane(k + 2, pn);
        fetchVox(k, vn);
        Own nxt;
        fetchOwn(k, nxt);

        // stencil rows fed by plane k: node k + 1, node k and node k - 1
        // threads without a node of their own (the tiles past a part's last node column / row) skip
        // the arithmetic: whole waves of them then issue only the plane loads and barriers
        T xm[9][3];
        if (mine) {
            const int sl = k & 3;
#pragma unroll
            for (int q = 0; q < 9; q++)
#pragma unroll
                for (int c = 0; c < 3; c++) xm[q][c] = (!DENSE || q == 4) ? sh.x[sl][c][ty + q / 3][tx + q % 3] : T(0);
        }
        // The stencil in class-sum form: with X_c(ox, oy) the plane's 3 x 3 values of component c and
        // oz = plane - node, every entry is D[|o_d|][nnz of the other two offsets] on the diagonal and
        // E[|o_e|] sgn(o_d o_c) off it (zero unless o_d, o_c != 0). Grouping X by those classes gives
        // the node at the plane (oz = 0) H, the nodes above and below T -+ U; 81 operations per plane
        // instead of 153 multiply-adds.
        // Difference form: the stencil sums to zero, so it is applied to the plane's differences from
        // its centre value, X(ox, oy) - X(0, 0), and the centre column enters once per node as
        // P (x(z - 1) - 2 x(z) + x(z + 1)), P = the diagonal sum of a neighbour plane's blocks. A
        // rigid translation then cancels exactly and a rotation to element scale, which keeps the
        // FP32 operator accurate on the smooth, low-energy modes the conjugate gradients resolve last.
        if (!DENSE && mine) {
            const T D00 = C.v[14], D01 = C.v[15], D02 = C.v[16], D10 = C.v[17], D11 = C.v[18], D12 = C.v[19], E0 = C.v[20], E1 = C.v[21];
            (void)D00;
            // q = (ox + 1) + 3 (oy + 1)
            T xs[3], xd[3], ys[3], yd[3], csum[3], cxy[3], cxs[3], cys[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const T p = xm[8][c], q = xm[0][c], u = xm[2][c], v = xm[6][c], ctr = xm[4][c];
                xs[c] = (xm[5][c] + xm[3][c]) - T(2) * ctr;
                xd[c] = xm[5][c] - xm[3][c];
                ys[c] = (xm[7][c] + xm[1][c]) - T(2) * ctr;
                yd[c] = xm[7][c] - xm[1][c];
                const T s1 = p + q, s2 = u + v;
                csum[c] = (s1 + s2) - T(4) * ctr;
                cxy[c] = s1 - s2;
                cxs[c] = (p + u) - (q + v);  // sum of sgn(ox) X over the corners
                cys[c] = (p + v) - (q + u);  // sum of sgn(oy) X over the corners
            }
            const T G1 = xs[2] + ys[2];
            if (k >= z0 && k < z1) {
                aC[0] += D10 * xs[0] + D01 * ys[0] + D11 * csum[0] + E0 * cxy[1];
                aC[1] += D10 * ys[1] + D01 * xs[1] + D11 * csum[1] + E0 * cxy[0];
                aC[2] += D01 * G1 + D02 * csum[2];
            }
            const bool fN = k + 1 < z1, fP = k - 1 >= z0;
            if (fN || fP) {
                const T Tx = D11 * xs[0] + D02 * ys[0] + D12 * csum[0] + E1 * cxy[1];
                const T Ty = D11 * ys[1] + D02 * xs[1] + D12 * csum[1] + E1 * cxy[0];
                const T Tz = D11 * G1 + D12 * csum[2];
                const T Ux = E0 * xd[2] + E1 * cxs[2];
                const T Uy = E0 * yd[2] + E1 * cys[2];
                const T Uz = E0 * (xd[0] + yd[1]) + E1 * (cxs[0] + cys[1]);
                if (fN) {  // oz = -1
                    aN[0] += Tx - Ux;
                    aN[1] += Ty - Uy;
                    aN[2] += Tz - Uz;
                }
                if (fP) {  // oz = +1
                    aP[0] += Tx + Ux;
                    aP[1] += Ty + Uy;
                    aP[2] += Tz + Uz;
                }
            }
        }

        // node k - 1 is complete: list its differing pairs (per-wave regions, voxel order)
        const bool done = k - 1 >= z0;
        const uint32_t mask = (done && !DENSE) ? (cur.info & 0xffu) : 0u;
        const int cnt = __popc(mask);
        // reference voxel: the lowest clear bit (all 8 differ only when all are own; voxel 0 then
        // reads scale 0)
        const int ref = mask == 0xffu ? 0 : __builtin_ctz(~mask);
        int waveOff = 0, waveTotal = 0;
        // wave-uniform: a wave whose nodes list nothing (solid regions) skips the prefix counts
        if (!DENSE && __ballot(cnt != 0) != 0) {
            const uint64_t lt = (uint64_t(1) << lane) - 1;
            const uint64_t b0 = __ballot(cnt & 1), b1 = __ballot(cnt & 2), b2 = __ballot(cnt & 4), b3 = __ballot(cnt & 8);
            waveOff = __popcll(b0 & lt) + 2 * __popcll(b1 & lt) + 4 * __popcll(b2 & lt) + 8 * __popcll(b3 & lt);
            waveTotal = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2) + 8 * __popcll(b3);
            uint32_t m = mask;
            int at = waveOff;
            while (m) {
                const int v = __builtin_ctz(m);
                m &= m - 1;
                sh.list[wid][at++] = uint16_t(((ty * TX + tx) << 6) | (ref << 3) | v);
            }
        }
        if (!DENSE && lane == 0) sh.waveTot[k & 1][wid] = waveTotal;
        if constexpr (OWNC) storeTwo((k + 2) % 3, twv);
        __syncthreads();

        if (done) {
            int tot[NW], total = 0, mineBase = 0;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                tot[w] = DENSE ? 0 : sh.waveTot[k & 1][w];
                mineBase += w < wid ? tot[w] : 0;
                total += tot[w];
            }
            const int sm = (k + 2) & 3, s0 = (k + 3) & 3, sp = k & 3;  // node planes k - 2, k - 1, k
            const int vm = (k + 1) % 3, v0 = (k + 2) % 3;              // voxel planes k - 2, k - 1
            T out[3] = {T(0), T(0), T(0)};
            if constexpr (DENSE) {
              if (mine) {
                // s_v E_v (x - x_node) over the node's 8 voxels; voxel v = (di, dj, dk) lies at node
                // offset (di - 1, dj - 1, dk - 1) and has the node as its corner a = (1 - di, 1 - dj, 1 - dk)
                T xn[3];
#pragma unroll
                for (int c = 0; c < 3; c++) xn[c] = sh.x[s0][c][ty + 1][tx + 1];
                staticFor<8>([&](auto V) {
                    constexpr int v = decltype(V)::value, di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
                    constexpr int A = cornerOf(1 - di, 1 - dj, 1 - dk);
                    const T sv = T(sh.s[dk ? v0 : vm][ty + dj][tx + di]);
                    T acc[3] = {T(0), T(0), T(0)};
                    staticFor<8>([&](auto Q) {
                        constexpr int q = decltype(Q)::value;
                        if constexpr (q != A) {
                            constexpr int ox = di - 1 + cx(q), oy = dj - 1 + cy(q), oz = dk - 1 + cz(q);
                            const int sl = oz < 0 ? sm : (oz == 0 ? s0 : sp);
                            T dx[3];
#pragma unroll
                            for (int c = 0; c < 3; c++) dx[c] = sh.x[sl][c][ty + 1 + oy][tx + 1 + ox] - xn[c];
                            staticFor<3>([&](auto Dd) {
                                staticFor<3>([&](auto Cc) { acc[Dd] += coefMul<elementClass(A, Dd, q, Cc), T>(C, dx[Cc]); });
                            });
                        }
                    });
#pragma unroll
                    for (int d = 0; d < 3; d++) out[d] += sv * acc[d];
                });
              }
            } else {
                // the centre column's second difference through the neighbour planes' diagonal sums
                const T Pxy = C.v[15] + T(2) * (C.v[18] + C.v[16]) + T(4) * C.v[19], Pz = C.v[17] + T(4) * (C.v[18] + C.v[19]);
                const T sref = T(sh.s[(ref >> 2) ? v0 : vm][ty + ((ref >> 1) & 1)][tx + (ref & 1)]);
                out[0] = sref * (aP[0] + Pxy * ((xcPrev2[0] + xm[4][0]) - T(2) * xcPrev[0]));
                out[1] = sref * (aP[1] + Pxy * ((xcPrev2[1] + xm[4][1]) - T(2) * xcPrev[1]));
                out[2] = sref * (aP[2] + Pz * ((xcPrev2[2] + xm[4][2]) - T(2) * xcPrev[2]));
            }
            const int mine0 = mineBase + waveOff;
            for (int pass = 0; pass * NT < total; pass++) {
                if (pass) __syncthreads();
                if (pass * NT + wid * 64 < total) {
                    const int p = pass * NT + wid * 64 + lane;
                    if (p < total) {
                        int w = 0, bw = 0, acc = tot[0];
#pragma unroll
                        for (int u = 1; u < NW; u++) {
                            if (p >= acc) {
                                w = u;
                                bw = acc;
                            }
                            acc += tot[u];
                        }
                        const int ent = sh.list[w][p - bw];
                        const int o = ent >> 6, rv = (ent >> 3) & 7, v = ent & 7;
                        const int ox0 = o % TX, oy0 = o / TX;
                        const int di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
                        // voxel v mirrored onto corner 0: corner q sits at node offset
                        // ((2 di - 1) qx, (2 dj - 1) qy, (2 dk - 1) qz); components and rows flip
                        // sign along mirrored axes
                        T xr[8][3];
#pragma unroll
                        for (int q = 0; q < 8; q++) {
                            const int qx = cx(q), qy = cy(q), qz = cz(q);
                            const int lx = ox0 + 1 + (di ? qx : -qx), ly = oy0 + 1 + (dj ? qy : -qy);
                            const int sl = qz ? (dk ? sp : sm) : s0;
                            xr[q][0] = sh.x[sl][0][ly][lx];
                            xr[q][1] = sh.x[sl][1][ly][lx];
                            xr[q][2] = sh.x[sl][2][ly][lx];
                        }
                        // element rows sum to zero: apply them to differences from corner 0 (the node)
#pragma unroll
                        for (int q = 1; q < 8; q++)
#pragma unroll
                            for (int c = 0; c < 3; c++) xr[q][c] -= xr[0][c];
                        const T sg[3] = {di ? T(1) : T(-1), dj ? T(1) : T(-1), dk ? T(1) : T(-1)};
                        const int oi = OWNC ? sh.own[dk ? v0 : vm][oy0 + dj][ox0 + di] : -1;
                        // rows of the level's uniform element (compile-time class coefficients)
                        // applied to the mirrored differences, signs restored
                        auto uniformRows = [&](T (&u)[3]) {
                            T Pm[3][3];
                            staticFor<3>([&](auto Dd) {
                                staticFor<3>([&](auto Cc) {
                                    T a = 0;
                                    staticFor<7>([&](auto Q1) { a += coefMul<elementClass(0, Dd, Q1 + 1, Cc), T>(C, xr[Q1 + 1][Cc]); });
                                    Pm[Dd][Cc] = a;
                                });
                            });
#pragma unroll
                            for (int d = 0; d < 3; d++) u[d] = sg[d] * (sg[0] * Pm[d][0] + sg[1] * Pm[d][1] + sg[2] * Pm[d][2]);
                        };
                        const T sref = T(sh.s[(rv >> 2) ? v0 : vm][oy0 + ((rv >> 1) & 1)][ox0 + (rv & 1)]);
                        T res[3];
                        if (OWNC && oi >= 0) {
                            // own element sum_c s_c M_c, mirrored: child c' of the mirrored element
                            // is child c' ^ (7 ^ v) of the original
                            const int mb = 7 ^ v;
                            const int4 tw = sh.two[dk ? v0 : vm][oy0 + dj][ox0 + di];
                            T xs[8][3];
#pragma unroll
                            for (int q = 1; q < 8; q++)
#pragma unroll
                                for (int c = 0; c < 3; c++) xs[q][c] = sg[c] * xr[q][c];
                            if (tw.z >= 0) {
                                // two scales: (a - s_ref) Kb + (b - a) M_B, B the children of scale b
                                const T sa = T(__int_as_float(tw.x)), sb = T(__int_as_float(tw.y));
                                uint32_t Bp = 0;
#pragma unroll
                                for (int c = 0; c < 8; c++) Bp |= ((uint32_t(tw.z) >> (c ^ mb)) & 1u) << c;
                                T u[3];
                                uniformRows(u);
                                if constexpr (H16) {
                                    const _Float16* __restrict__ tab = reinterpret_cast<const _Float16*>(L.ctab16) + 64 * Bp;
                                    const T scale = T(L.ctabRowScale[Bp]);
#pragma unroll
                                    for (int d = 0; d < 3; d++) {
                                        T t = 0;
#pragma unroll
                                        for (int k2 = 0; k2 < 21; k2++) t += T(float(tab[d * 21 + k2])) * xs[1 + k2 / 3][k2 % 3];
                                        res[d] = (sa - sref) * u[d] + (sb - sa) * sg[d] * (t * scale);
                                    }
                                } else {
                                    const T* __restrict__ tab = ctab + 64 * Bp;
#pragma unroll
                                    for (int d = 0; d < 3; d++) {
                                        T t = 0;
#pragma unroll
                                        for (int k2 = 0; k2 < 21; k2++) t += tab[d * 21 + k2] * xs[1 + k2 / 3][k2 % 3];
                                        res[d] = (sa - sref) * u[d] + (sb - sa) * sg[d] * t;
                                    }
                                }
                            } else {
                                const float* cs = L.childS + 8 * int64_t(oi);
                                T o3[3] = {0, 0, 0};
#pragma unroll
                                for (int cc = 0; cc < 8; cc++) {
                                    const T wc = T(cs[cc ^ mb]) - sref;
#pragma unroll
                                    for (int d = 0; d < 3; d++) {
                                        T t = 0;
#pragma unroll
                                        for (int k2 = 3; k2 < 24; k2++) t += crow[(cc * 3 + d) * 24 + k2] * xs[k2 / 3][k2 % 3];
                                        o3[d] += wc * t;
                                    }
                                }
#pragma unroll
                                for (int d = 0; d < 3; d++) res[d] = sg[d] * o3[d];
                            }
                        } else {
                            const T ds = T(sh.s[dk ? v0 : vm][oy0 + dj][ox0 + di]) - sref;
                            T u[3];
                            uniformRows(u);
#pragma unroll
                            for (int d = 0; d < 3; d++) res[d] = ds * u[d];
                        }
#pragma unroll
                        for (int d = 0; d < 3; d++) sh.pres[d][p - pass * NT] = res[d];
                    }
                }
                __syncthreads();
                if (cnt) {
                    const int lo = max(mine0, pass * NT), hi = min(mine0 + cnt, pass * NT + NT);
                    for (int q = lo; q < hi; q++) {
                        out[0] += sh.pres[0][q - pass * NT];
                        out[1] += sh.pres[1][q - pass * NT];
                        out[2] += sh.pres[2][q - pass * NT];
                    }
                }
            }
            // epilogue for node k - 1 by the owning thread
            if (mine) {
                const int64_t pb3 = 3 * NXY * (k - 1);  // block-uniform plane base
                const int32_t c3 = 3 * col;
                if (L.diagAdd) {
#pragma unroll
                    for (int d = 0; d < 3; d++) out[d] += T((L.diagAdd + pb3)[c3 + d]) * xcPrev[d];
                }
                if constexpr (NEEDS_INV && !OWNC) {
                    // inverse diagonal from the voxel scales: every diagonal entry of Kb is the
                    // same class value, so diag = Kb_00 sum_v s_v (+ the diagonal shift)
                    T ssum = 0;
#pragma unroll
                    for (int v = 0; v < 8; v++) ssum += T(sh.s[(v >> 2) ? v0 : vm][ty + ((v >> 1) & 1)][tx + (v & 1)]);
                    if (L.diagAdd) {
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const T dg = C.v[0] * ssum + T((L.diagAdd + pb3)[c3 + d]);
                            cur.inv[d] = ((cur.info >> (8 + d)) & 1u) ? T(0) : T(1) / dg;
                        }
                    } else {
                        // one division for the three components (the same diagonal entry)
                        const T iv = T(1) / (C.v[0] * ssum);
#pragma unroll
                        for (int d = 0; d < 3; d++) cur.inv[d] = ((cur.info >> (8 + d)) & 1u) ? T(0) : iv;
                    }
                }
                T yv[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const bool held = (cur.info >> (8 + d)) & 1u;
                    if constexpr (MODE == AM_APPLY) yv[d] = held ? T(0) : out[d];
                    else if constexpr (MODE == AM_RAW) yv[d] = out[d];
                    else if constexpr (MODE == AM_RESID) yv[d] = held ? T(0) : cur.r[d] - out[d];
                    else if constexpr (MODE == AM_CHEB1) yv[d] = xcPrev[d] + c1 * cur.inv[d] * (cur.r[d] - out[d]);
                    else yv[d] = xcPrev[d] + c1 * (xcPrev[d] - cur.d[d]) + c2 * cur.inv[d] * (cur.r[d] - out[d]);
                    if constexpr (DOT) {
                        if constexpr (MODE == AM_APPLY) dot0 += double(xcPrev[d]) * double(yv[d]);
                        if constexpr (NEEDS_INV) {
                            dot0 += double(cur.r[d]) * double(yv[d]);
                            dot1 += double(yv[d]) * double(cur.qd[d]);
                        }
                    }
                }
                bstore3(planeRsrc(y + pb3, planeBytes), noff, yv);
                if constexpr (MODE == AM_APPLY) {
                    if (pw) bstore3(planeRsrc(pw + pb3, planeBytes), noff, xcPrev);
                }
            }
        }
        // DENSE reads node plane k - 2, which the next step's store replaces
        if constexpr (DENSE) __syncthreads();
        // rotate: accumulators move up one node, the centre of plane k becomes node k's x
#pragma unroll
        for (int d = 0; d < 3; d++) {
            aP[d] = aC[d];
            aC[d] = aN[d];
            aN[d] = T(0);
            xcPrev2[d] = xcPrev[d];
            xcPrev[d] = xm[4][d];
        }
        cur = nxt;
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

#define PSX_FINE_FWD L, C, crow, ctab, x, y, r, zp, c1, c2, stop, zChunk, gx, gy, nTiles, xcdOrder, dots, dotStride, qdot, pz, pw
// the kernel bodies for packed launches: WPE_ as the named kernel's waves-per-SIMD attribute
template <class T, int MODE, bool OWNC, bool DOT, bool DENSE, int WPE_, bool H16 = false> struct Fine {
    static constexpr int NTH = NT, WPE = WPE_;
    __device__ static void run(VBlk vb, PSX_FINE_PARAMS) { fineBody<T, MODE, OWNC, DOT, DENSE, H16>(vb, PSX_FINE_FWD); }
};
template <class T, int MODE, bool OWNC, bool DOT> __global__ __launch_bounds__(NT) void kFine(PSX_FINE_PARAMS) {
    fineBody<T, MODE, OWNC, DOT>(ownBlock(), PSX_FINE_FWD);
}
// FP32 child-scale levels reading the FP16 subset table: held to 168 VGPRs (3 waves per SIMD, as the
// FP32-table kernels reach by themselves; the later Chebyshev step would otherwise take 174 and 2 waves)
template <int MODE, bool DOT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void kFineH(DevLevel L, Sym<float> C, const float* __restrict__ crow,
                                                                                  const float* __restrict__ ctab, const float* __restrict__ x,
                                                                                  float* __restrict__ y, const float* __restrict__ r,
                                                                                  const float* __restrict__ zp, float c1, float c2, const int* stop,
                                                                                  int zChunk, int gx, int gy, int nTiles, int xcdOrder,
                                                                                  double* __restrict__ dots, int64_t dotStride,
                                                                                  const float* __restrict__ qdot, const float* __restrict__ pz,
                                                                                  float* __restrict__ pw) {
    fineBody<float, MODE, true, DOT, false, true>(ownBlock(), PSX_FINE_FWD);
}
// FP32 without own elements: held to 128 VGPRs (4 waves per SIMD instead of 3), which it fits
// without spilling; the outer operator (APPLY with the fused p.q and direction update) to 102, 5 waves
// per SIMD as its LDS allows (103 VGPRs and 4 waves otherwise)
template <class T, int MODE, bool DOT> constexpr int fine4Waves() { return MODE == AM_APPLY && DOT ? 5 : 4; }
template <class T, int MODE, bool DOT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(fine4Waves<T, MODE, DOT>()))) void kFine4(PSX_FINE_PARAMS) {
    fineBody<T, MODE, false, DOT>(ownBlock(), PSX_FINE_FWD);
}
// per-voxel form (levels without own elements)
template <class T, int MODE, bool DOT> __global__ __launch_bounds__(NT) void kFineDense(PSX_FINE_PARAMS) {
    fineBody<T, MODE, false, DOT, true>(ownBlock(), PSX_FINE_FWD);
}
#undef PSX_FINE_FWD

// launches (or records) the operator kernel of one form, mode and dot choice
template <class T, int MODE, bool OWNC, bool DOT, bool DENSE, bool H16 = false>
void fineGo(dim3 g, hipStream_t s, const DevLevel& L, const Sym<T>& C, const T* crow, const T* ctab, const T* x, T* y, const T* r, const T* zp, T c1, T c2,
            const int* stop, int zChunk, int gx, int gy, int nTiles, int xcdOrder, double* dots, int64_t dotStride, const T* qdot, const T* pz,
            T* pw) {
#define PSX_FINE_ARGS L, C, crow, ctab, x, y, r, zp, c1, c2, stop, zChunk, gx, gy, nTiles, xcdOrder, dots, dotStride, qdot, pz, pw
    if constexpr (DENSE) {
        go<Fine<T, MODE, false, DOT, true, 0>>(kFineDense<T, MODE, DOT>, g, s, PSX_FINE_ARGS);
    } else if constexpr (sizeof(T) == 4 && !OWNC) {
        go<Fine<T, MODE, false, DOT, false, fine4Waves<T, MODE, DOT>()>>(kFine4<T, MODE, DOT>, g, s, PSX_FINE_ARGS);
    } else if constexpr (H16) {
        static_assert(sizeof(T) == 4 && OWNC, "FP16 subset table: FP32 child-scale levels");
        go<Fine<float, MODE, true, DOT, false, 3, true>>(kFineH<MODE, DOT>, g, s, PSX_FINE_ARGS);
    } else {
        go<Fine<T, MODE, OWNC, DOT, false, 0>>(kFine<T, MODE, OWNC, DOT>, g, s, PSX_FINE_ARGS);
    }
#undef PSX_FINE_ARGS
}

}  // namespace

template <class T, bool OWNC>
static void launchFineT(const DevLevel& L, int mode, const Sym<T>& C, const T* crow, const T* ctab, const T* x, T* y, const T* r, const T* d, double c1, double c2,
                        const int* stop, int zChunk, double* dots, int64_t dotStride, const T* qdot, hipStream_t s, int ninst,
                        bool half = false, const T* pz = nullptr, T* pw = nullptr) {
    const int gx = (L.NX + TX - 1) / TX, gy = (L.NY + TY - 1) / TY, gz = (L.NZ + zChunk - 1) / zChunk;
    const int nTiles = gx * gy * gz;
    const dim3 g(unsigned(8 * ((nTiles + 7) / 8)), unsigned(ninst));  // whole rounds over the 8 XCDs
    // measured on the 512x128x128 void block: the XCD-contiguous order is 8 % slower than the
    // launch order for this kernel (its halos are served from the shared last-level cache), so it
    // is off unless PSX_XCD_ORDER=1
    static const int xcdOrder = [] {
        const char* e = std::getenv("PSX_XCD_ORDER");
        return e && e[0] == '1' ? 1 : 0;
    }();
    // fused dots exist for APPLY (p.q) and the Chebyshev steps (r.z, z.q) only
    const bool dot = dots && (mode == AM_APPLY || mode == AM_CHEB1 || mode == AM_CHEB2);
    const bool dense = !OWNC && L.dense;
    const bool useH16 = sizeof(T) == 4 && OWNC && half && L.ctab16;
    auto form = [&](auto m, auto dt) {
        constexpr int M = decltype(m)::value;
        constexpr bool DT = decltype(dt)::value;
        const T f1 = T(c1), f2 = T(c2);
        if constexpr (!OWNC) {
            if (dense) {
                fineGo<T, M, false, DT, true>(g, s, L, C, crow, ctab, x, y, r, d, f1, f2, stop, zChunk, gx, gy, nTiles, xcdOrder, dots, dotStride, qdot, pz,
                                              pw);
                return;
            }
        }
        if constexpr (sizeof(T) == 4 && OWNC) {
            if (useH16) {
                fineGo<T, M, OWNC, DT, false, true>(g, s, L, C, crow, ctab, x, y, r, d, f1, f2, stop, zChunk, gx, gy, nTiles, xcdOrder, dots, dotStride,
                                                    qdot, pz, pw);
                return;
            }
        }
        fineGo<T, M, OWNC, DT, false>(g, s, L, C, crow, ctab, x, y, r, d, f1, f2, stop, zChunk, gx, gy, nTiles, xcdOrder, dots, dotStride, qdot, pz, pw);
    };
    using yes = std::true_type;
    using no = std::false_type;
    switch (mode) {
        case AM_APPLY:
            if (dot) form(std::integral_constant<int, AM_APPLY>{}, yes{});
            else form(std::integral_constant<int, AM_APPLY>{}, no{});
            break;
        case AM_RAW: form(std::integral_constant<int, AM_RAW>{}, no{}); break;
        case AM_RESID: form(std::integral_constant<int, AM_RESID>{}, no{}); break;
        case AM_CHEB1:
            if (dot) form(std::integral_constant<int, AM_CHEB1>{}, yes{});
            else form(std::integral_constant<int, AM_CHEB1>{}, no{});
            break;
        default:
            if (dot) form(std::integral_constant<int, AM_CHEB2>{}, yes{});
            else form(std::integral_constant<int, AM_CHEB2>{}, no{});
            break;
    }
}

template <class T>
unsigned launchFineDot(const DevLevel& L, int mode, const T* x, T* y, const T* r, const T* d, double c1, double c2, const int* stop, const DotFusion& fd,
                       hipStream_t s, int ninst, int el, bool half) {
    const bool fused = (fd.kind == FD_PQ && mode == AM_APPLY) || (fd.kind == FD_RZ && (mode == AM_CHEB1 || mode == AM_CHEB2));
    double* dots = fused ? fd.red.partials : nullptr;
    const T* qdot = fused && fd.kind == FD_RZ ? static_cast<const T*>(fd.q) : nullptr;
    // the direction update p = z + beta p_old fused into the outer operator (FD_PQ with z and p given)
    const T* pz = fused && fd.kind == FD_PQ ? static_cast<const T*>(fd.pz) : nullptr;
    T* pw = fused && fd.kind == FD_PQ ? static_cast<T*>(fd.pw) : nullptr;
    if ((pz != nullptr) != (pw != nullptr)) throw std::invalid_argument("fused direction update: z and the new direction go together");
    if constexpr (sizeof(T) == 4) {
        // the element-by-element product on the matrix cores (fine_el.hip), on its own tiles
        // (the child-scale level with its own elements' corner-row pass)
        if (el > 0 && (!L.childForm || L.ownVox)) {
            if (pz) throw std::invalid_argument("the fused direction update needs the VALU operator");
            return launchFineEl(L, mode, x, y, r, d, c1, c2, stop, dots, fd.red.stride, qdot, s, ninst, el);
        }
    }
    Sym<T> C;
    for (int q = 0; q < 22; q++) C.v[q] = sizeof(T) == 8 ? T(L.sym64[q]) : T(L.sym32[q]);
    // the longest z-chunk (16, 8 or 4 planes) that still gives >= 4096 blocks (16 per CU); shorter
    // chunks repeat two halo planes per chunk but keep small levels' CUs busy. The rule ignores the
    // instance count so that an instance's tiles, and so its fused dot partials, are those of a lone
    // launch (bitwise the same results)
    const int gx = (L.NX + TX - 1) / TX, gy = (L.NY + TY - 1) / TY;
    int zChunk = 4;
    for (int c : {16, 8})
        if (int64_t(gx) * gy * ((L.NZ + c - 1) / c) >= 4096) {
            zChunk = c;
            break;
        }
    // a level that 4-plane chunks give fewer than 2048 blocks is latency-bound: 2-plane chunks
    // (measured on the 256x64x64 child-scale level of the 512x128x128 block: 6 % faster than 4)
    if (zChunk == 4 && int64_t(gx) * gy * ((L.NZ + 3) / 4) < 2048) zChunk = 2;
    if (const char* e = std::getenv("PSX_ZCHUNK")) zChunk = std::max(2, std::atoi(e));
    // (measured: near-equal chunks of 8-9 planes, which drop the one-plane last chunk of NZ = 2^k + 1,
    // are 4 % slower on the 512x128x128 block than this rule's 8-plane chunks)
    const int gz = (L.NZ + zChunk - 1) / zChunk;
    if (L.childForm) {
        const T *crow = nullptr, *ctab = nullptr;
        if constexpr (sizeof(T) == 8) crow = L.crow64, ctab = L.ctab64;
        else crow = L.crow32, ctab = L.ctab32;
        launchFineT<T, true>(L, mode, C, crow, ctab, x, y, r, d, c1, c2, stop, zChunk, dots, fd.red.stride, qdot, s, ninst, half, pz, pw);
    } else {
        launchFineT<T, false>(L, mode, C, nullptr, nullptr, x, y, r, d, c1, c2, stop, zChunk, dots, fd.red.stride, qdot, s, ninst, false, pz, pw);
    }
    return fused ? unsigned(gx * gy * gz) : 0u;
}

template <class T>
void launchFine(const DevLevel& L, int mode, const T* x, T* y, const T* r, const T* d, double c1, double c2, const int* stop, hipStream_t s, int ninst,
                int el, bool half) {
    launchFineDot<T>(L, mode, x, y, r, d, c1, c2, stop, DotFusion{}, s, ninst, el, half);
}

template unsigned launchFineDot<double>(const DevLevel&, int, const double*, double*, const double*, const double*, double, double, const int*,
                                        const DotFusion&, hipStream_t, int, int, bool);
template unsigned launchFineDot<float>(const DevLevel&, int, const float*, float*, const float*, const float*, double, double, const int*,
                                       const DotFusion&, hipStream_t, int, int, bool);
template void launchFine<double>(const DevLevel&, int, const double*, double*, const double*, const double*, double, double, const int*, hipStream_t, int,
                                 int, bool);
template void launchFine<float>(const DevLevel&, int, const float*, float*, const float*, const float*, double, double, const int*, hipStream_t, int, int,
                                bool);

}  // namespace psx
