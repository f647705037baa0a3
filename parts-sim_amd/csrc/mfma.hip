// BF16 matrix-core operator for the FP32 multigrid smoother (levels without own matrices).
//
// y = K x is assembled element by element: each voxel's 24 DOF values become one column of a
// 24 x N matrix U, and the uniform element matrix multiplies all columns at once on the matrix
// cores, K0 (24 x 24, padded to 32 x 32) times U in BF16 with FP32 accumulation
// (v_mfma_f32_16x16x32_bf16: 16 voxels per instruction pair). Each node then sums the rows of its
// 8 voxels' products from LDS. Every node takes the same path, so a void surface costs nothing
// extra.
//
// Accuracy: a column is s_v (u_e - u_ref), the voxel's scale times its displacements relative to
// its corner 0. K0 annihilates rigid translations, so the product is unchanged, while BF16 now
// rounds displacement differences instead of displacements (a smoother only needs an
// approximate operator; the outer iteration uses FP32 / FP64 operators).
//
// Work layout as kFine (fine.hip): 32 x 8 node columns per block marching along z, node planes and
// voxel scales in LDS rings; the products of one voxel plane are computed once and used by the two
// node planes that touch it.
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace psx {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int TX = 32, TY = 8, NT = TX * TY;
constexpr int PX = TX + 2, PY = TY + 2;
constexpr int VX = TX + 1, VY = TY + 1, NV = VX * VY;
constexpr int GROUPS = (NV + 15) / 16;
constexpr int PSTRIDE = 25;  // 24 products per voxel, padded to an odd stride

struct Smem {
    float x[4][3][PY][PX];
    float s[3][VY][VX];
    float prod[2][NV * PSTRIDE];
};

__device__ constexpr int corner(int x, int y, int z) { return (y ? (x ? 2 : 3) : (x ? 1 : 0)) + 4 * z; }

template <int MODE>
__global__ __launch_bounds__(NT) void kFineMfma(DevLevel L, const bf16x8* __restrict__ Aop, const float* __restrict__ x, float* __restrict__ y,
                                                const float* __restrict__ r, const float* __restrict__ zp, float c1, float c2, const int* stop, int zChunk) {
    if (stop && __builtin_nontemporal_load(stop) != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Smem& sh = *reinterpret_cast<Smem*>(smem_raw);
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX, lane = tid & 63, wid = tid >> 6;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const int z0 = blockIdx.z * zChunk, z1 = min(L.NZ, z0 + zChunk);
    const int i = i0 + tx, j = j0 + ty;
    const bool mine = i < L.NX && j < L.NY;
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
    // A operands: K0 rows 0-15 and 16-31 (zero padded), this lane's K slice
    const bf16x8 A0 = Aop[lane], A1 = Aop[64 + lane];

    constexpr int NPL = (PX * PY + NT - 1) / NT, NVL = (NV + NT - 1) / NT;
    constexpr bool NEEDS_R = MODE == AM_RESID || MODE == AM_CHEB1 || MODE == AM_CHEB2;
    constexpr bool NEEDS_D = MODE == AM_CHEB2;
    const float* __restrict__ invD = L.invD32;
    // the epilogue's own operands of node (i, j, z), fetched one step ahead
    struct Own {
        float inv[3], r[3], d[3];
    };
    auto fetchOwn = [&](int z, Own& o) {
        const bool in = mine && z < z1;
        const int64_t q0 = 3 * (in ? i + NX * j + NXY * z : 0);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            o.inv[d] = in ? invD[q0 + d] : 0.f;
            if constexpr (NEEDS_R) o.r[d] = in ? r[q0 + d] : 0.f;
            if constexpr (NEEDS_D) o.d[d] = in && zp ? zp[q0 + d] : 0.f;
        }
    };
    float px[NPL][3], pv[NVL];
    auto fetchPlane = [&](int z) {
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            const int e = tid + q * NT;
            const int gi = i0 - 1 + e % PX, gj = j0 - 1 + e / PX;
            const bool in = e < PX * PY && gi >= 0 && gj >= 0 && gi < L.NX && gj < L.NY && z >= 0 && z < L.NZ;
            const float* p = x + 3 * (in ? gi + NX * gj + NXY * z : 0);
            px[q][0] = in ? p[0] : 0.f;
            px[q][1] = in ? p[1] : 0.f;
            px[q][2] = in ? p[2] : 0.f;
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
            const bool in = e < NV && vi >= 0 && vj >= 0 && vi < L.nx && vj < L.ny && k >= 0 && k < L.nz;
            const int32_t w = L.words[in ? vi + int64_t(L.nx) * (vj + int64_t(L.ny) * k) : 0];
            pv[q] = (in && w > 0) ? __int_as_float(w) : 0.f;
        }
    };
    auto storeVox = [&](int slot) {
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int e = tid + q * NT;
            if (e < NV) sh.s[slot][e / VX][e % VX] = pv[q];
        }
    };
    // products of voxel plane k (between node planes k and k+1) into prod[k & 1]
    auto products = [&](int k) {
        const int lo = k & 3, hi = (k + 1) & 3, vs = (k + 3) % 3;
        float* P = sh.prod[k & 1];
        const int q = lane >> 4;  // this lane's K slice: DOFs 8q .. 8q+7 (q = 3: padding)
        for (int g = wid; g < GROUPS; g += NT / 64) {
            const int e = 16 * g + (lane & 15);
            bf16x8 b;
            if (e < NV && q < 3) {
                const int vi = e % VX, vj = e / VX;
                const float s = sh.s[vs][vj][vi];
                const float ref[3] = {sh.x[lo][0][vj][vi], sh.x[lo][1][vj][vi], sh.x[lo][2][vj][vi]};
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    const int dof = 8 * q + t, c = dof % 3, bc = dof / 3;
                    const int ox = ((bc + 1) >> 1) & 1, oy = (bc >> 1) & 1, oz = bc >> 2;
                    b[t] = (__bf16)(s * (sh.x[oz ? hi : lo][c][vj + oy][vi + ox] - ref[c]));
                }
            } else {
#pragma unroll
                for (int t = 0; t < 8; t++) b[t] = (__bf16)0.f;
            }
            const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
            const floatx4 d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, b, zero, 0, 0, 0);
            const floatx4 d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, b, zero, 0, 0, 0);
            if (e < NV) {
                // result rows 4q .. 4q+3 (first product) and 16+4q .. (second, rows < 24)
#pragma unroll
                for (int v = 0; v < 4; v++) P[e * PSTRIDE + 4 * q + v] = d0[v];
                if (q < 2)
#pragma unroll
                    for (int v = 0; v < 4; v++) P[e * PSTRIDE + 16 + 4 * q + v] = d1[v];
            }
        }
    };

    // prologue: planes z0-1 .. z0+1 and voxel planes z0-1, z0; loads of the next fetch are issued
    // before the previous one is stored so two round trips overlap
    fetchPlane(z0 - 1);
    fetchVox(z0 - 1);
    {
        float qx[NPL][3], qv[NVL];
#pragma unroll
        for (int q = 0; q < NPL; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) qx[q][c] = px[q][c];
#pragma unroll
        for (int q = 0; q < NVL; q++) qv[q] = pv[q];
        fetchPlane(z0);
        fetchVox(z0);
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            const int e = tid + q * NT;
            if (e < PX * PY)
#pragma unroll
                for (int c = 0; c < 3; c++) sh.x[(z0 + 3) & 3][c][e / PX][e % PX] = qx[q][c];
        }
#pragma unroll
        for (int q = 0; q < NVL; q++) {
            const int e = tid + q * NT;
            if (e < NV) sh.s[(z0 + 2) % 3][e / VX][e % VX] = qv[q];
        }
    }
    storePlane(z0 & 3);
    storeVox(z0 % 3);
    fetchPlane(z0 + 1);
    storePlane((z0 + 1) & 3);
    Own cur;
    fetchOwn(z0, cur);
    __syncthreads();
    products(z0 - 1);

    for (int z = z0; z < z1; z++) {
        fetchPlane(z + 2);
        fetchVox(z + 1);
        Own nxt;
        fetchOwn(z + 1, nxt);
        products(z);
        __syncthreads();
        if (mine) {
            float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const int di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
                const int a = corner(1 - di, 1 - dj, 1 - dk);
                const float* P = sh.prod[(z - 1 + dk) & 1] + ((ty + dj) * VX + tx + di) * PSTRIDE + 3 * a;
                acc[0] += P[0];
                acc[1] += P[1];
                acc[2] += P[2];
            }
            const int64_t n = i + NX * j + NXY * z;
            const int s0 = z & 3;
            if (L.diagAdd) {
#pragma unroll
                for (int d = 0; d < 3; d++) acc[d] += float(L.diagAdd[3 * n + d]) * sh.x[s0][d][ty + 1][tx + 1];
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int64_t q = 3 * n + d;
                const float id = cur.inv[d];
                if constexpr (MODE == AM_APPLY) y[q] = id != 0.f ? acc[d] : 0.f;
                else if constexpr (MODE == AM_RAW) y[q] = acc[d];
                else if constexpr (MODE == AM_RESID) y[q] = id != 0.f ? cur.r[d] - acc[d] : 0.f;
                else if constexpr (MODE == AM_CHEB1) y[q] = sh.x[s0][d][ty + 1][tx + 1] + c1 * id * (cur.r[d] - acc[d]);
                else {
                    const float xc = sh.x[s0][d][ty + 1][tx + 1];
                    y[q] = xc + c1 * (xc - cur.d[d]) + c2 * id * (cur.r[d] - acc[d]);
                }
            }
        }
        storePlane((z + 2) & 3);
        storeVox((z + 1) % 3);
        cur = nxt;
        __syncthreads();
    }
}

template <int MODE> void launchMode(const DevLevel& L, const void* Aop, const float* x, float* y, const float* r, const float* d, float c1, float c2, const int* stop,
                                    hipStream_t s) {
    static bool attr = [] {
        return hipFuncSetAttribute(reinterpret_cast<const void*>(&kFineMfma<MODE>), hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(Smem))) ==
               hipSuccess;
    }();
    (void)attr;
    const int zChunk = 8;
    const dim3 g((L.NX + TX - 1) / TX, (L.NY + TY - 1) / TY, (L.NZ + zChunk - 1) / zChunk);
    kFineMfma<MODE><<<g, NT, sizeof(Smem), s>>>(L, static_cast<const bf16x8*>(Aop), x, y, r, d, c1, c2, stop, zChunk);
}

}  // namespace

size_t mfmaOperandBytes() { return 2 * 64 * 8 * 2; }

// the A operands of the two MFMA instructions for a level's uniform element matrix: lane l of
// instruction m holds K[16 m + l % 16][8 (l / 16) .. +8] in BF16 (rows and columns >= 24 zero)
void mfmaOperands(const double* Kb, uint16_t* out) {
    for (int m = 0; m < 2; m++)
        for (int l = 0; l < 64; l++)
            for (int t = 0; t < 8; t++) {
                const int row = 16 * m + l % 16, col = 8 * (l / 16) + t;
                const float v = (row < 24 && col < 24) ? float(Kb[row * 24 + col]) : 0.f;
                uint32_t u;
                __builtin_memcpy(&u, &v, 4);
                // round to nearest even
                const uint32_t rounded = u + 0x7fffu + ((u >> 16) & 1u);
                out[(m * 64 + l) * 8 + t] = uint16_t(rounded >> 16);
            }
}

void launchFineMfma(const DevLevel& L, const void* Aop, int mode, const float* x, float* y, const float* r, const float* d, double c1, double c2, const int* stop,
                    hipStream_t s) {
    switch (mode) {
        case AM_APPLY: launchMode<AM_APPLY>(L, Aop, x, y, r, d, float(c1), float(c2), stop, s); break;
        case AM_RAW: launchMode<AM_RAW>(L, Aop, x, y, r, d, float(c1), float(c2), stop, s); break;
        case AM_RESID: launchMode<AM_RESID>(L, Aop, x, y, r, d, float(c1), float(c2), stop, s); break;
        case AM_CHEB1: launchMode<AM_CHEB1>(L, Aop, x, y, r, d, float(c1), float(c2), stop, s); break;
        default: launchMode<AM_CHEB2>(L, Aop, x, y, r, d, float(c1), float(c2), stop, s); break;
    }
}

}  // namespace psx
