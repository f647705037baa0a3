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
// stencil entry [m][d][c] vanishes by the cube's reflection symmetry when d != c and the offset is
// zero along d or c
__device__ constexpr bool stencilZero(int m, int d, int c) {
    const int o[3] = {m % 3 - 1, (m / 3) % 3 - 1, m / 9 - 1};
    return d != c && (o[d] == 0 || o[c] == 0);
}

template <class T> struct FineP;
template <> struct FineP<double> {
    __device__ static const double* Kb(const DevLevel& L) { return L.Kb64; }
    __device__ static const double* St(const DevLevel& L) { return L.St64; }
    __device__ static const double* invD(const DevLevel& L) { return L.invD64; }
};
template <> struct FineP<float> {
    __device__ static const float* Kb(const DevLevel& L) { return L.Kb32; }
    __device__ static const float* St(const DevLevel& L) { return L.St32; }
    __device__ static const float* invD(const DevLevel& L) { return L.invD32; }
};

template <class T> struct Planes {
    T x[4][3][PY][PX];      // ring of node planes, one array per component
    float s[3][VY][VX];     // ring of voxel-scale planes (0 = empty or outside)
    T res[3][NT];           // phase-B results by local node
    int list[NT];
    int count;
};

template <class T, int MODE>
__global__ __launch_bounds__(NT) void kFine(DevLevel L, const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ r, T* __restrict__ dv, T c1,
                                            T c2, const int* stop, int zChunk) {
    if (stop && __builtin_nontemporal_load(stop) != 0) return;
    __shared__ Planes<T> sh;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const int z0 = blockIdx.z * zChunk, z1 = min(L.NZ, z0 + zChunk);
    const int i = i0 + tx, j = j0 + ty;
    const bool mine = i < L.NX && j < L.NY;
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;

    // fetch of node plane z (this thread's share) and voxel plane k into registers
    constexpr int NPL = (PX * PY + NT - 1) / NT, NVL = (VX * VY + NT - 1) / NT;
    T px[NPL][3];
    float pv[NVL];
    auto fetchPlane = [&](int z) {
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            const int e = tid + q * NT;
            const int gi = i0 - 1 + e % PX, gj = j0 - 1 + e / PX;
            const bool in = e < PX * PY && gi >= 0 && gj >= 0 && gi < L.NX && gj < L.NY && z >= 0 && z < L.NZ;
            const T* p = x + 3 * (gi + NX * gj + NXY * z);
            px[q][0] = in ? p[0] : T(0);
            px[q][1] = in ? p[1] : T(0);
            px[q][2] = in ? p[2] : T(0);
        }
    };
    auto storePlane = [&](int slot) {
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
    auto fetchVox = [&](int k) {
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int e = tid + q * NT;
            const int vi = i0 - 1 + e % VX, vj = j0 - 1 + e / VX;
            const bool in = e < VX * VY && vi >= 0 && vj >= 0 && vi < L.nx && vj < L.ny && k >= 0 && k < L.nz;
            const int32_t w = in ? L.words[vi + int64_t(L.nx) * (vj + int64_t(L.ny) * k)] : 0;
            pv[q] = w > 0 ? __int_as_float(w) : 0.f;
        }
    };
    auto storeVox = [&](int slot) {
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int e = tid + q * NT;
            if (e < VX * VY) sh.s[slot][e / VX][e % VX] = pv[q];
        }
    };

    // prologue: node planes z0-1, z0, z0+1 and voxel planes z0-1, z0
    for (int z = z0 - 1; z <= z0 + 1; z++) {
        fetchPlane(z);
        storePlane((z + 4) & 3);
    }
    for (int k = z0 - 1; k <= z0; k++) {
        fetchVox(k);
        storeVox((k + 3) % 3);
    }
    __syncthreads();

    const T* __restrict__ S = FineP<T>::St(L);
    const T* __restrict__ Kb = FineP<T>::Kb(L);
    const T* __restrict__ invD = FineP<T>::invD(L);
    const int lane = tid & 63, wid = tid >> 6;

    for (int z = z0; z < z1; z++) {
        // prefetch the planes the next step needs (latency hidden behind this step's compute)
        fetchPlane(z + 2);
        fetchVox(z + 1);
        const int sm = (z + 3) & 3, s0 = z & 3, sp = (z + 1) & 3;  // node planes z-1, z, z+1
        const int vm = (z + 2) % 3, v0 = z % 3;                   // voxel planes z-1, z
        // scales of the 8 voxels around the node: v = di + 2 dj + 4 dk at (i-1+di, j-1+dj, z-1+dk)
        float sv[8];
#pragma unroll
        for (int v = 0; v < 8; v++) sv[v] = sh.s[(v >> 2) ? v0 : vm][ty + ((v >> 1) & 1)][tx + (v & 1)];
        bool uniform = sv[0] > 0.f;
#pragma unroll
        for (int v = 1; v < 8; v++) uniform = uniform && sv[v] == sv[0];
        // the stencil needs all 27 neighbours inside the grid; a uniform node is never on the box surface
        uniform = uniform && i > 0 && j > 0 && z > 0 && i < L.nx && j < L.ny && z < L.nz;
        const bool irregular = mine && !uniform;
        T acc[3] = {0, 0, 0};
        // ---- phase A: uniform nodes, assembled stencil
        if (mine && uniform) {
#pragma unroll
            for (int m = 0; m < 27; m++) {
                const int ox = m % 3, oy = (m / 3) % 3, oz = m / 9;
                const int sl = oz == 0 ? sm : (oz == 1 ? s0 : sp);
                const T xm[3] = {sh.x[sl][0][ty + oy][tx + ox], sh.x[sl][1][ty + oy][tx + ox], sh.x[sl][2][ty + oy][tx + ox]};
#pragma unroll
                for (int d = 0; d < 3; d++)
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        if (!stencilZero(m, d, c)) acc[d] += S[m * 9 + d * 3 + c] * xm[c];
            }
            const T s = T(sv[0]);
            acc[0] *= s;
            acc[1] *= s;
            acc[2] *= s;
        }
        // irregular nodes -> block list (wavefront ballot + prefix, waves in order)
        const uint64_t mask = __ballot(irregular);
        __shared__ int waveBase[NT / 64];
        if (lane == 0) waveBase[wid] = __popcll(mask);
        __syncthreads();
        int base = 0;
        for (int w = 0; w < wid; w++) base += waveBase[w];
        if (irregular) sh.list[base + __popcll(mask & ((uint64_t(1) << lane) - 1))] = tid;
        if (tid == NT - 1) sh.count = base + waveBase[wid];
        __syncthreads();
        // ---- phase B: packed per-voxel sums for the listed nodes (uniform voxel loop -> scalar K rows)
        const int nIrr = sh.count;
        if (tid < nIrr) {
            const int o = sh.list[tid];
            const int ox0 = o % TX, oy0 = o / TX;
            T xn[27][3];
#pragma unroll
            for (int m = 0; m < 27; m++) {
                const int ox = m % 3, oy = (m / 3) % 3, oz = m / 9;
                const int sl = oz == 0 ? sm : (oz == 1 ? s0 : sp);
                xn[m][0] = sh.x[sl][0][oy0 + oy][ox0 + ox];
                xn[m][1] = sh.x[sl][1][oy0 + oy][ox0 + ox];
                xn[m][2] = sh.x[sl][2][oy0 + oy][ox0 + ox];
            }
            T b[3] = {0, 0, 0};
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const int di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
                const T s = T(sh.s[dk ? v0 : vm][oy0 + dj][ox0 + di]);
                const int a = corner(1 - di, 1 - dj, 1 - dk);
                T e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int m = (di + cx(q)) + 3 * (dj + cy(q)) + 9 * (dk + cz(q));
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        e0 += Kb[(3 * a) * 24 + 3 * q + c] * xn[m][c];
                        e1 += Kb[(3 * a + 1) * 24 + 3 * q + c] * xn[m][c];
                        e2 += Kb[(3 * a + 2) * 24 + 3 * q + c] * xn[m][c];
                    }
                }
                b[0] += s * e0;
                b[1] += s * e1;
                b[2] += s * e2;
            }
            sh.res[0][o] = b[0];
            sh.res[1][o] = b[1];
            sh.res[2][o] = b[2];
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
                const T id = invD[q];
                if constexpr (MODE == AM_APPLY) y[q] = id != T(0) ? acc[d] : T(0);
                else if constexpr (MODE == AM_RAW) y[q] = acc[d];
                else if constexpr (MODE == AM_RESID) y[q] = id != T(0) ? r[q] - acc[d] : T(0);
                else if constexpr (MODE == AM_CHEB1) {
                    const T dn = c1 * id * (r[q] - acc[d]);
                    dv[q] = dn;
                    y[q] = sh.x[s0][d][ty + 1][tx + 1] + dn;
                } else {
                    const T dn = c1 * dv[q] + c2 * id * (r[q] - acc[d]);
                    dv[q] = dn;
                    y[q] = sh.x[s0][d][ty + 1][tx + 1] + dn;
                }
            }
        }
        // rotate: plane z+2 replaces z-2, voxel plane z+1 replaces z-2
        storePlane((z + 2) & 3);
        storeVox((z + 1) % 3);
        __syncthreads();
    }
}

}  // namespace

template <class T> void launchFine(const DevLevel& L, int mode, const T* x, T* y, const T* r, T* d, double c1, double c2, const int* stop, hipStream_t s) {
    const int zChunk = 8;
    const dim3 g((L.NX + TX - 1) / TX, (L.NY + TY - 1) / TY, (L.NZ + zChunk - 1) / zChunk), b(NT);
    switch (mode) {
        case AM_APPLY: kFine<T, AM_APPLY><<<g, b, 0, s>>>(L, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
        case AM_RAW: kFine<T, AM_RAW><<<g, b, 0, s>>>(L, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
        case AM_RESID: kFine<T, AM_RESID><<<g, b, 0, s>>>(L, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
        case AM_CHEB1: kFine<T, AM_CHEB1><<<g, b, 0, s>>>(L, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
        default: kFine<T, AM_CHEB2><<<g, b, 0, s>>>(L, x, y, r, d, T(c1), T(c2), stop, zChunk); break;
    }
}

template void launchFine<double>(const DevLevel&, int, const double*, double*, const double*, double*, double, double, const int*, hipStream_t);
template void launchFine<float>(const DevLevel&, int, const float*, float*, const float*, float*, double, double, const int*, hipStream_t);

}  // namespace psx
