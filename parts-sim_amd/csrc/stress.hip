// Nodal stresses on the device (VoxelFEA::stresses / stressTensors, voxel_fea.cpp:968-1084): at
// each active node, the density-weighted average over its (up to 8) voxels of the voxel's corner
// stress sigma = (E / h) S_a u_e, with S_a the 6 x 24 matrix D Bbar at the node's corner a of the
// voxel and u_e the voxel's 24 displacements. Outputs: von Mises (and the first and third principal
// stresses) averaged per node, or the averaged 6-component tensor. FP64 arithmetic, FP32 output.
//
// One thread per node; the corner matrices are wave-uniform (the voxel loop is unrolled, so the
// corner a and every row/column index are compile-time constants), read through the scalar cache.
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace psx {

namespace {

constexpr int BLOCK = 256;

__device__ constexpr int corner(int x, int y, int z) { return (y ? (x ? 2 : 3) : (x ? 1 : 0)) + 4 * z; }
__device__ constexpr int cx(int b) { return (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0; }
__device__ constexpr int cy(int b) { return (b == 2 || b == 3 || b == 6 || b == 7) ? 1 : 0; }
__device__ constexpr int cz(int b) { return b >= 4 ? 1 : 0; }

__device__ inline double vonMises(const double* s) {
    return sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) + (s[2] - s[0]) * (s[2] - s[0])) +
                3 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
}

// principal stresses, largest first: the closed-form eigenvalues of a symmetric 3 x 3 (the
// trigonometric solution of the characteristic cubic), as voxel_fea.cpp:948-966
__device__ inline void principal(const double* s, double& p1, double& p3) {
    const double sx = s[0], sy = s[1], sz = s[2], txy = s[3], tyz = s[4], tzx = s[5];
    const double q1 = txy * txy + tyz * tyz + tzx * tzx;
    const double q = (sx + sy + sz) / 3;
    const double scale = fabs(sx) + fabs(sy) + fabs(sz) + sqrt(q1);
    if (q1 <= 1e-24 * scale * scale) {
        p1 = fmax(sx, fmax(sy, sz));
        p3 = fmin(sx, fmin(sy, sz));
        return;
    }
    const double q2 = (sx - q) * (sx - q) + (sy - q) * (sy - q) + (sz - q) * (sz - q) + 2 * q1;
    const double p = sqrt(q2 / 6);
    const double b11 = (sx - q) / p, b22 = (sy - q) / p, b33 = (sz - q) / p, b12 = txy / p, b23 = tyz / p, b13 = tzx / p;
    const double det = b11 * (b22 * b33 - b23 * b23) - b12 * (b12 * b33 - b23 * b13) + b13 * (b12 * b23 - b22 * b13);
    const double r = fmin(1.0, fmax(-1.0, det / 2));
    const double phi = acos(r) / 3;
    p1 = q + 2 * p * cos(phi);
    p3 = q + 2 * p * cos(phi + 2 * M_PI / 3);
}

enum { SM_VM = 0, SM_PRINCIPAL = 1, SM_TENSOR = 2 };

// S: 8 corner matrices [a][6][24], already scaled by E / h
template <int MODE>
__global__ __launch_bounds__(BLOCK) void kStresses(DevLevel L, const double* __restrict__ S, const double* __restrict__ u, float* __restrict__ o0,
                                                   float* __restrict__ o1, float* __restrict__ o2) {
    const int64_t n = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (n >= L.nNodes) return;
    const int64_t NX = L.NX, NXY = int64_t(L.NX) * L.NY;
    const int i = int(n % NX), j = int((n / NX) % L.NY), k = int(n / NXY);
    double vm = 0, p1 = 0, p3 = 0, w = 0, acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 8; v++) {
        const int di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
        const int ei = i + di - 1, ej = j + dj - 1, ek = k + dk - 1;
        if (ei < 0 || ej < 0 || ek < 0 || ei >= L.nx || ej >= L.ny || ek >= L.nz) continue;
        const int32_t word = L.words[ei + int64_t(L.nx) * (ej + int64_t(L.ny) * ek)];
        if (word <= 0) continue;  // empty (the finest level has no own matrices)
        const double rho = double(__int_as_float(word));
        // the node is corner a = (1 - di, 1 - dj, 1 - dk) of the voxel
        const int a = corner(1 - di, 1 - dj, 1 - dk);
        const int64_t base = ei + NX * ej + NXY * ek;
        double sig[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const double* ub = u + 3 * (base + cx(b) + NX * cy(b) + NXY * cz(b));
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double uv = ub[c];
#pragma unroll
                for (int rr = 0; rr < 6; rr++) sig[rr] += S[a * 144 + rr * 24 + 3 * b + c] * uv;
            }
        }
        if constexpr (MODE == SM_TENSOR) {
#pragma unroll
            for (int rr = 0; rr < 6; rr++) acc[rr] += rho * sig[rr];
        } else {
            vm += rho * vonMises(sig);
            if constexpr (MODE == SM_PRINCIPAL) {
                double a1, a3;
                principal(sig, a1, a3);
                p1 += rho * a1;
                p3 += rho * a3;
            }
        }
        w += rho;
    }
    if constexpr (MODE == SM_TENSOR) {
#pragma unroll
        for (int rr = 0; rr < 6; rr++) o0[6 * n + rr] = w > 0 ? float(acc[rr] / w) : 0.f;
    } else {
        o0[n] = w > 0 ? float(vm / w) : 0.f;
        if constexpr (MODE == SM_PRINCIPAL) {
            o1[n] = w > 0 ? float(p1 / w) : 0.f;
            o2[n] = w > 0 ? float(p3 / w) : 0.f;
        }
    }
}

}  // namespace

void launchStresses(const DevLevel& L, const double* S, const double* u, int principals, float* vm, float* p1, float* p3, hipStream_t s) {
    const unsigned g = unsigned((L.nNodes + BLOCK - 1) / BLOCK);
    if (principals) kStresses<SM_PRINCIPAL><<<g, BLOCK, 0, s>>>(L, S, u, vm, p1, p3);
    else kStresses<SM_VM><<<g, BLOCK, 0, s>>>(L, S, u, vm, nullptr, nullptr);
}

void launchStressTensors(const DevLevel& L, const double* S, const double* u, float* out, hipStream_t s) {
    const unsigned g = unsigned((L.nNodes + BLOCK - 1) / BLOCK);
    kStresses<SM_TENSOR><<<g, BLOCK, 0, s>>>(L, S, u, out, nullptr, nullptr);
}

}  // namespace psx
