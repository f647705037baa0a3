// Operator kernel for levels whose elements are all s * Kb (every finest level, and the coarse
// levels of uniform parts): y = K x with the fused epilogues of kApply (kernels.hip).
//
// Layout of the work: a block of 256 threads owns a 32 x 8 column of nodes in x-y and marches along z.
// Node planes of x (with a one-node halo) and voxel-scale planes are staged in LDS rings; the next
// plane is fetched from HBM while the current one is computed.
//
// Divergence: 94 % of 64-node wavefronts of a void part contain a node whose 8 voxels differ (the
// surface of a void). Such a node needs the per-voxel sum (8 x 72 multiply-adds) instead of the
// assembled stencil (153). Rather than letting those lanes diverge inside every wavefront, each
// z-step runs in two phases: (A) every uniform node applies the stencil; irregular nodes append
// themselves to a block list (wavefront ballots, deterministic order); (B) the first threads of
// the block take the listed nodes and do the per-voxel sums on packed lanes, with the element
// matrix rows as wave-uniform (scalar) operands. Results go back to the owning thread through LDS.
#include <hip/hip_runtime.h>

#include <utility>

#include "device.hpp"

namespace psx {

namespace {

constexpr int TX = 32, TY = 8, NT = TX * TY;
constexpr int PX = TX + 2, PY = TY + 2;  // node plane with halo
constexpr int VX = TX + 1, VY = TY + 1;  // voxel plane

__device__ constexpr int corner(int x, int y, int z) { return (y ? (x ? 2 : 3) : (x ? 1 : 0)) + 4 * z; }
__device__ constexpr int cx(int b) { return (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0; }
__device__ constexpr int cy(int b) { return (b == 2 || b == 3 || b == 6 || b == 7) ? 1 : 0; }
__device__ constexpr int cz(int b) { return b >= 4 ? 1 : 0; }
// Symmetry classes (the host's elementClass / stencilClass, hierarchy.cpp): every coefficient is
// +-one of 22 values, at an index known at compile time once the loops are unrolled, so the values
// stay in scalar registers for the whole kernel.
__device__ constexpr int elementClass(int a, int d, int b, int c) {
    const int ca[3] = {cx(a), cy(a), cz(a)};
    int r[3] = {cx(a) ^ cx(b), cy(a) ^ cy(b), cz(a) ^ cz(b)};
    const bool neg = ((ca[d] + ca[c]) & 1) != 0;
    int t = r[0];
    r[0] = r[d];
    r[d] = t;
    int c2 = c == 0 ? d : (c == d ? 0 : c);
    if (c2 == 2) {
        t = r[1];
        r[1] = r[2];
        r[2] = t;
        c2 = 1;
    } else if (c2 == 0 && r[1] < r[2]) {
        t = r[1];
        r[1] = r[2];
        r[2] = t;
    }
    const int idx = c2 == 0 ? r[0] * 3 + r[1] + r[2] : 6 + r[0] + 2 * r[1] + 4 * r[2];
    return idx | (neg ? 0x100 : 0);
}
__device__ constexpr int stencilClass(int m, int d, int c) {
    const int o[3] = {m % 3 - 1, (m / 3) % 3 - 1, m / 9 - 1};
    if (d == c) {
        int other = 0;
        for (int e = 0; e < 3; e++)
            if (e != d) other += o[e] != 0;
        return 14 + (o[d] != 0) * 3 + other;
    }
    if (o[d] == 0 || o[c] == 0) return -1;
    const int e = 3 - d - c;
    return (20 + (o[e] != 0)) | (o[d] * o[c] < 0 ? 0x100 : 0);
}
template <class T> struct Sym {
    T v[22];
};
// coefficient times x, class and sign fixed at compile time
template <int CLS, class T> __device__ __forceinline__ T coefMul(const Sym<T>& C, T x) {
    if constexpr ((CLS & 0x100) != 0) return -C.v[CLS & 0xff] * x;
    else return C.v[CLS & 0xff] * x;
}
// loop with a compile-time index: f(std::integral_constant<int, i>) for i in [0, N)
template <class F, int... I> __device__ __forceinline__ void staticForImpl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void staticFor(F&& f) { staticForImpl(f, std::make_integer_sequence<int, N>{}); }

template <class T> struct FineP;
template <> struct FineP<double> {
    __device__ static const double* invD(const DevLevel& L) { return L.invD64; }
};
template <> struct FineP<float> {
    __device__ static const float* invD(const DevLevel& L) { return L.invD32; }
};

template <class T> struct Planes {
    T x[4][3][PY][PX];      // ring of node planes, one array per component
    float s[3][VY][VX];     // ring of voxel-scale planes (0 = empty or outside)
    T res[3][NT];           // phase-B results by local node
    int list[NT];
    int waveCount[NT / 64];
};

template <class T, int MODE>
__global__ __launch_bounds__(NT) void kFine(DevLevel L, Sym<T> C, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ r, T* __restrict__ dv,
                                            T c1, T c2, const int* stop, int zChunk) {
    if (stop && __builtin_nontemporal_load(stop) != 0) return;
    __shared__ Planes<T> sh;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX, lane = tid & 63, wid = tid >> 6;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const int z0 = blockIdx.z * zChunk, z1 = min(L.NZ, z0 + zChunk);
    const int i = i0 + tx, j = j0 + ty;
    const bool mine = i < L.NX && j < L.NY;
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
    const T* __restrict__ invD = FineP<T>::invD(L);
    constexpr bool NEEDS_R = MODE == AM_RESID || MODE == AM_CHEB1 || MODE == AM_CHEB2;
    constexpr bool NEEDS_D = MODE == AM_CHEB2;

    constexpr int NPL = (PX * PY + NT - 1) / NT, NVL = (VX * VY + NT - 1) / NT;
    // global -> register fetches of a node plane / voxel plane (this thread's share)
    auto fetchPlane = [&](int z, T (&px)[NPL][3]) {
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            const int e = tid + q * NT;
            const int gi = i0 - 1 + e % PX, gj = j0 - 1 + e / PX;
            const bool in = e < PX * PY && gi >= 0 && gj >= 0 && gi < L.NX && gj < L.NY && z >= 0 && z < L.NZ;
            const T* p = x + 3 * (in ? gi + NX * gj + NXY * z : 0);
            px[q][0] = in ? p[0] : T(0);
            px[q][1] = in ? p[1] : T(0);
            px[q][2] = in ? p[2] : T(0);
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
    auto fetchVox = [&](int k, float (&pv)[NVL]) {
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int e = tid + q * NT;
            const int vi = i0 - 1 + e % VX, vj = j0 - 1 + e / VX;
            const bool in = e < VX * VY && vi >= 0 && vj >= 0 && vi < L.nx && vj < L.ny && k >= 0 && k < L.nz;
            const int32_t w = L.words[in ? vi + int64_t(L.nx) * (vj + int64_t(L.ny) * k) : 0];
            pv[q] = (in && w > 0) ? __int_as_float(w) : 0.f;
        }
    };
    auto storeVox = [&](int slot, const float (&pv)[NVL]) {
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int e = tid + q * NT;
            if (e < VX * VY) sh.s[slot][e / VX][e % VX] = pv[q];
        }
    };
    // the epilogue's own operands of node (i, j, z): fetched one step ahead
    struct Own {
        T inv[3], r[3], d[3];
    };
    auto fetchOwn = [&](int z, Own& o) {
        const bool in = mine && z < z1;
        const int64_t q0 = 3 * (in ? i + NX * j + NXY * z : 0);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            o.inv[d] = in ? invD[q0 + d] : T(0);
            if constexpr (NEEDS_R) o.r[d] = in ? r[q0 + d] : T(0);
            if constexpr (NEEDS_D) o.d[d] = in ? dv[q0 + d] : T(0);
        }
    };

    // prologue: all loads first, then the LDS stores
    {
        T pa[NPL][3], pb[NPL][3], pc[NPL][3];
        float va[NVL], vb[NVL];
        fetchPlane(z0 - 1, pa);
        fetchPlane(z0, pb);
        fetchPlane(z0 + 1, pc);
        fetchVox(z0 - 1, va);
        fetchVox(z0, vb);
        storePlane((z0 + 3) & 3, pa);
        storePlane(z0 & 3, pb);
        storePlane((z0 + 1) & 3, pc);
        storeVox((z0 + 2) % 3, va);
        storeVox(z0 % 3, vb);
    }
    Own cur;
    fetchOwn(z0, cur);
    __syncthreads();

    for (int z = z0; z < z1; z++) {
        // prefetch what the next step needs: node plane z+2, voxel plane z+1, own operands of z+1
        T pn[NPL][3];
        float vn[NVL];
        Own nxt;
        fetchPlane(z + 2, pn);
        fetchVox(z + 1, vn);
        fetchOwn(z + 1, nxt);
        const int sm = (z + 3) & 3, s0 = z & 3, sp = (z + 1) & 3;  // node planes z-1, z, z+1
        const int vm = (z + 2) % 3, v0 = z % 3;                   // voxel planes z-1, z
        float sv[8];
#pragma unroll
        for (int v = 0; v < 8; v++) sv[v] = sh.s[(v >> 2) ? v0 : vm][ty + ((v >> 1) & 1)][tx + (v & 1)];
        bool uniform = sv[0] > 0.f;
#pragma unroll
        for (int v = 1; v < 8; v++) uniform = uniform && sv[v] == sv[0];
        uniform = uniform && i > 0 && j > 0 && z > 0 && i < L.nx && j < L.ny && z < L.nz;
        const bool irregular = mine && !uniform;
        T acc[3] = {0, 0, 0};
        // ---- phase A: uniform nodes, assembled stencil, one neighbour plane at a time
        if (mine && uniform) {
            staticFor<3>([&](auto OZ) {
                const int sl = OZ == 0 ? sm : (OZ == 1 ? s0 : sp);
                T xm[9][3];
#pragma unroll
                for (int k = 0; k < 9; k++)
#pragma unroll
                    for (int c = 0; c < 3; c++) xm[k][c] = sh.x[sl][c][ty + k / 3][tx + k % 3];
                staticFor<9>([&](auto K) {
                    constexpr int m = K + 9 * OZ;
                    staticFor<3>([&](auto Dd) {
                        staticFor<3>([&](auto Cc) {
                            constexpr int cls = stencilClass(m, Dd, Cc);
                            if constexpr (cls >= 0) acc[Dd] += coefMul<cls, T>(C, xm[K][Cc]);
                        });
                    });
                });
            });
            const T s = T(sv[0]);
            acc[0] *= s;
            acc[1] *= s;
            acc[2] *= s;
        }
        // irregular nodes -> block list (wavefront ballots, waves in order)
        const uint64_t mask = __ballot(irregular);
        if (lane == 0) sh.waveCount[wid] = __popcll(mask);
        __syncthreads();
        int base = 0, nIrr = 0;
#pragma unroll
        for (int w = 0; w < NT / 64; w++) {
            base += w < wid ? sh.waveCount[w] : 0;
            nIrr += sh.waveCount[w];
        }
        if (irregular) sh.list[base + __popcll(mask & ((uint64_t(1) << lane) - 1))] = tid;
        __syncthreads();
        // ---- phase B: 8 lanes per listed node, one voxel per lane. Voxel v is mirrored onto
        // corner 0 (neighbour offsets (2 di - 1, 2 dj - 1, 2 dk - 1) * corner bits, components
        // and rows sign-flipped along mirrored axes), so every lane uses the same element rows.
        {
            const int v = lane & 7, di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
            const T sgx = di ? T(1) : T(-1), sgy = dj ? T(1) : T(-1), sgz = dk ? T(1) : T(-1);
            for (int pass = 0; pass * (NT / 8) < nIrr; pass++) {
                const int slot = pass * (NT / 8) + wid * 8 + (lane >> 3);
                const bool act = slot < nIrr;
                const int o = act ? sh.list[slot] : 0;
                const int ox0 = o % TX, oy0 = o / TX;
                const T s = act ? T(sh.s[dk ? v0 : vm][oy0 + dj][ox0 + di]) : T(0);
                T xq[8][3];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int qx = (q == 1 || q == 2 || q == 5 || q == 6), qy = (q == 2 || q == 3 || q == 6 || q == 7), qz = q >= 4;
                    const int lx = ox0 + 1 + (di ? qx : -qx), ly = oy0 + 1 + (dj ? qy : -qy);
                    const int sl = qz ? (dk ? sp : sm) : s0;
                    xq[q][0] = sgx * sh.x[sl][0][ly][lx];
                    xq[q][1] = sgy * sh.x[sl][1][ly][lx];
                    xq[q][2] = sgz * sh.x[sl][2][ly][lx];
                }
                T e[3] = {0, 0, 0};
                staticFor<3>([&](auto Dd) {
                    staticFor<8>([&](auto Q) {
                        staticFor<3>([&](auto Cc) { e[Dd] += coefMul<elementClass(0, Dd, Q, Cc), T>(C, xq[Q][Cc]); });
                    });
                });
                T b0 = s * sgx * e[0], b1 = s * sgy * e[1], b2 = s * sgz * e[2];
#pragma unroll
                for (int off = 4; off >= 1; off >>= 1) {
                    b0 += __shfl_xor(b0, off, 64);
                    b1 += __shfl_xor(b1, off, 64);
                    b2 += __shfl_xor(b2, off, 64);
                }
                if (act && v == 0) {
                    sh.res[0][o] = b0;
                    sh.res[1][o] = b1;
                    sh.res[2][o] = b2;
                }
            }
        }
        __syncthreads();
        // ---- epilogue by the owning thread
        if (mine) {
            if (irregular) {
                acc[0] = sh.res[0][tid];
                acc[1] = sh.res[1][tid];
                acc[2] = sh.res[2][tid];
            }
            const int64_t n = i + NX * j + NXY * z;
            if (L.diagAdd) {
#pragma unroll
                for (int d = 0; d < 3; d++) acc[d] += T(L.diagAdd[3 * n + d]) * sh.x[s0][d][ty + 1][tx + 1];
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int64_t q = 3 * n + d;
                const T id = cur.inv[d];
                if constexpr (MODE == AM_APPLY) y[q] = id != T(0) ? acc[d] : T(0);
                else if constexpr (MODE == AM_RAW) y[q] = acc[d];
                else if constexpr (MODE == AM_RESID) y[q] = id != T(0) ? cur.r[d] - acc[d] : T(0);
                else if constexpr (MODE == AM_CHEB1) {
                    const T dn = c1 * id * (cur.r[d] - acc[d]);
                    dv[q] = dn;
                    y[q] = sh.x[s0][d][ty + 1][tx + 1] + dn;
                } else {
                    const T dn = c1 * cur.d[d] + c2 * id * (cur.r[d] - acc[d]);
                    dv[q] = dn;
                    y[q] = sh.x[s0][d][ty + 1][tx + 1] + dn;
                }
            }
        }
        // rotate: plane z+2 replaces z-2, voxel plane z+1 replaces z-2
        storePlane((z + 2) & 3, pn);
        storeVox((z + 1) % 3, vn);
        cur = nxt;
        __syncthreads();
    }
}

}  // namespace

template <class T> void launchFine(const DevLevel& L, int mode, const T* x, T* y, const T* r, T* d, double c1, double c2, const int* stop, hipStream_t s) {
    const int zChunk = 8;
    Sym<T> C;
    for (int q = 0; q < 22; q++) C.v[q] = sizeof(T) == 8 ? T(L.sym64[q]) : T(L.sym32[q]);
    const dim3 g((L.NX + TX - 1) / TX, (L.NY + TY - 1) / TY, (L.NZ + zChunk - 1) / zChunk), b(NT);
    switch (mode) {
        case AM_APPLY: kFine<T, AM_APPLY><<<g, b, 0, s>>>(L, C, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
        case AM_RAW: kFine<T, AM_RAW><<<g, b, 0, s>>>(L, C, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
        case AM_RESID: kFine<T, AM_RESID><<<g, b, 0, s>>>(L, C, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
        case AM_CHEB1: kFine<T, AM_CHEB1><<<g, b, 0, s>>>(L, C, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
        default: kFine<T, AM_CHEB2><<<g, b, 0, s>>>(L, C, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
    }
}

template void launchFine<double>(const DevLevel&, int, const double*, double*, const double*, double*, double, double, const int*, hipStream_t);
template void launchFine<float>(const DevLevel&, int, const float*, float*, const float*, float*, double, double, const int*, hipStream_t);

}  // namespace psx
