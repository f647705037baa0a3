// Voxelization of a triangle mesh on the device: the filled fraction of every voxel from sub^3
// samples (voxelize.cpp:25-128 is the reference form, whose results these kernels reproduce bit for
// bit). Along each axis a, rays run through the sample centres of the other two axes (offset by a
// tiny jitter that keeps them off shared edges); every triangle crossing a ray contributes the depth
// of the crossing; a ray's sorted depths pair up into inside spans (even-odd rule) and each sample in
// a span gets one vote. A sample is solid when at least two of its three rays agree, which tolerates
// small holes and flipped triangles; a voxel's fraction is its solid samples over sub^3.
//
// Device layout: per axis, a count pass (one thread per triangle and 8-row band of rays, atomic
// counts per ray), an exclusive scan of the counts, a fill pass that stores every crossing depth in
// its ray's slice, then one thread per ray sorts its slice and casts its votes (each sample of the
// axis belongs to exactly one ray, so the votes need no atomics); a last kernel turns the votes into
// fractions. Arithmetic is FP64 in the reference's order with no fused multiply-adds, so the
// crossings, their FP32 depths and hence the fractions are identical to the reference's.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/psx.h"
#include "abi.hpp"

#pragma clang fp contract(off)

namespace psx {

namespace {

constexpr int BLOCK = 256;
constexpr int ROWS = 8;  // ray rows per thread of the raster passes

struct Axis {
    int a, u, v, Nu, Nv, Na;
    double ou, ov, oa, ds;
    int64_t strideU, strideV, strideA;
};

__device__ __forceinline__ double min3(double x, double y, double z) { return x < y ? (x < z ? x : z) : (y < z ? y : z); }
__device__ __forceinline__ double max3(double x, double y, double z) { return x > y ? (x > z ? x : z) : (y > z ? y : z); }

// FILL = false: count the crossings per ray; true: store their depths at cnt[ray]++, within the ray's
// slice [.., end[ray + 1]) that the count pass sized (a crossing the count pass did not see is not
// stored but counted in overflow, which the host turns into an error)
template <bool FILL>
__global__ __launch_bounds__(BLOCK) void kRaster(const float* __restrict__ pos, const uint32_t* __restrict__ index, int64_t nTri, Axis X,
                                                 uint32_t* __restrict__ cnt, float* __restrict__ hits, const uint32_t* __restrict__ end,
                                                 uint32_t* __restrict__ overflow) {
    const int64_t t = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (t >= nTri) return;
    const int rowLo = int(blockIdx.y) * ROWS, rowHi = min(X.Nv - 1, rowLo + ROWS - 1);
    const int64_t i0 = 3 * int64_t(index[3 * t]), i1 = 3 * int64_t(index[3 * t + 1]), i2 = 3 * int64_t(index[3 * t + 2]);
    const double av = pos[i0 + X.v], bv = pos[i1 + X.v], cv = pos[i2 + X.v];
    const int sv0 = max(rowLo, int(ceil((min3(av, bv, cv) - X.ov) / X.ds - 0.5)));
    const int sv1 = min(rowHi, int(floor((max3(av, bv, cv) - X.ov) / X.ds - 0.5)));
    if (sv0 > sv1) return;
    const double au = pos[i0 + X.u], aa = pos[i0 + X.a];
    const double bu = pos[i1 + X.u], ba = pos[i1 + X.a];
    const double cu = pos[i2 + X.u], ca = pos[i2 + X.a];
    const double det = (bv - cv) * (au - cu) + (cu - bu) * (av - cv);
    if (det == 0) return;
    const int su0 = max(0, int(ceil((min3(au, bu, cu) - X.ou) / X.ds - 0.5)));
    const int su1 = min(X.Nu - 1, int(floor((max3(au, bu, cu) - X.ou) / X.ds - 0.5)));
    if (su0 > su1) return;
    const double inv = 1 / det;
    for (int sv = sv0; sv <= sv1; sv++) {
        const double pv = X.ov + (sv + 0.5) * X.ds;
        for (int su = su0; su <= su1; su++) {
            const double pu = X.ou + (su + 0.5) * X.ds;
            const double l1 = ((bv - cv) * (pu - cu) + (cu - bu) * (pv - cv)) * inv;
            if (l1 < 0) continue;
            const double l2 = ((cv - av) * (pu - cu) + (au - cu) * (pv - cv)) * inv;
            if (l2 < 0) continue;
            const double l3 = 1 - l1 - l2;
            if (l3 < 0) continue;
            const int64_t ray = int64_t(su) + int64_t(X.Nu) * sv;
            if constexpr (FILL) {
                const uint32_t slot = atomicAdd(&cnt[ray], 1u);
                if (slot < end[ray + 1]) hits[slot] = float(l1 * aa + l2 * ba + l3 * ca);
                else atomicAdd(overflow, 1u);
            } else {
                atomicAdd(&cnt[ray], 1u);
            }
        }
    }
}

// one thread per ray: sort its depths, then one vote per sample inside each span
__global__ __launch_bounds__(BLOCK) void kSpans(const uint32_t* __restrict__ off, float* __restrict__ hits, Axis X, uint8_t* __restrict__ votes) {
    const int64_t ray = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (ray >= int64_t(X.Nu) * X.Nv) return;
    const uint32_t start = off[ray], end = off[ray + 1];
    if (end - start < 2) return;
    float* h = hits + start;
    const int n = int(end - start);
    for (int q = 1; q < n; q++) {  // insertion sort (a ray crosses few triangles)
        const float key = h[q];
        int p = q - 1;
        while (p >= 0 && h[p] > key) {
            h[p + 1] = h[p];
            p--;
        }
        h[p + 1] = key;
    }
    const int su = int(ray % X.Nu), sv = int(ray / X.Nu);
    const int64_t base = su * X.strideU + sv * X.strideV;
    for (int q = 0; q + 1 < n; q += 2) {
        const int s0 = max(0, int(ceil((double(h[q]) - X.oa) / X.ds - 0.5)));
        const int s1 = min(X.Na - 1, int(floor((double(h[q + 1]) - X.oa) / X.ds - 0.5)));
        for (int s = s0; s <= s1; s++) votes[base + s * X.strideA]++;
    }
}

__global__ __launch_bounds__(BLOCK) void kFractions(const uint8_t* __restrict__ votes, int nx, int ny, int nz, int sub, int S0, int S1,
                                                    float* __restrict__ frac) {
    const int64_t e = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= int64_t(nx) * ny * nz) return;
    const int i = int(e % nx), j = int((e / nx) % ny);
    const int64_t k = e / (int64_t(nx) * ny);
    int c = 0;
    for (int dk = 0; dk < sub; dk++)
        for (int dj = 0; dj < sub; dj++) {
            const int64_t row = int64_t(i) * sub + S0 * (int64_t(j) * sub + dj + int64_t(S1) * (k * sub + dk));
            for (int di = 0; di < sub; di++) c += votes[row + di] >= 2;
        }
    const double inv = 1.0 / (sub * sub * sub);
    frac[e] = float(c * inv);
}

// device buffers released on every exit path, in stream order
struct Buffers {
    hipStream_t s = nullptr;
    std::vector<void*> p;
    template <class T> T* get(size_t count) {
        void* q = nullptr;
        HIPCHK(hipMallocAsync(&q, std::max<size_t>(count, 1) * sizeof(T), s));
        p.push_back(q);
        return static_cast<T*>(q);
    }
    ~Buffers() {
        for (void* q : p) (void)hipFreeAsync(q, s);
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

}  // namespace

}  // namespace psx

extern "C" {

int psx_grid_for_box(const double mn[3], const double mx[3], int n, int pad, double origin[3], double* h, int dims[3]) {
    return psx::abi::guarded([&] {
        if (!mn || !mx || !origin || !h || !dims || n < 1 || pad < 0) throw std::invalid_argument("bad arguments");
        const double size[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
        *h = std::max({size[0], size[1], size[2]}) / n;
        if (!(*h > 0)) throw std::invalid_argument("Empty bounding box.");
        for (int d = 0; d < 3; d++) {
            dims[d] = std::max(1, int(std::ceil(size[d] / *h - 1e-6))) + 2 * pad;
            origin[d] = mn[d] - (dims[d] * *h - size[d]) / 2;
        }
        return PSX_OK;
    });
}

int psx_voxelize(const float* pos, int64_t n_vertices, const uint32_t* index, int64_t n_tris, const double origin[3], double h, const int dims[3], int sub,
                 float* frac) {
    using namespace psx;
    return abi::guarded([&] {
        if (!pos || !index || !origin || !dims || !frac || n_vertices < 0 || n_tris < 0 || sub < 1 || !(h > 0))
            throw std::invalid_argument("bad arguments");
        for (int d = 0; d < 3; d++)
            if (dims[d] < 1) throw std::invalid_argument("bad grid dimensions");
        for (int64_t q = 0; q < 3 * n_tris; q++)
            if (int64_t(index[q]) >= n_vertices) throw std::invalid_argument("triangle index out of range");
        const int S[3] = {dims[0] * sub, dims[1] * sub, dims[2] * sub};
        const int64_t nSamples = int64_t(S[0]) * S[1] * S[2], nVox = int64_t(dims[0]) * dims[1] * dims[2];
        const double ds = h / sub;
        const double jitter[3] = {0.000137 * ds, 0.000291 * ds, 0.000213 * ds};
        const int64_t strides[3] = {1, S[0], int64_t(S[0]) * S[1]};
        Buffers B;
        HIPCHK(hipStreamCreateWithFlags(&B.s, hipStreamNonBlocking));
        float* dPos = B.get<float>(size_t(3 * n_vertices));
        uint32_t* dIdx = B.get<uint32_t>(size_t(3 * n_tris));
        uint8_t* votes = B.get<uint8_t>(size_t(nSamples));
        float* dFrac = B.get<float>(size_t(nVox));
        if (n_vertices) HIPCHK(hipMemcpyAsync(dPos, pos, size_t(3 * n_vertices) * sizeof(float), hipMemcpyHostToDevice, B.s));
        if (n_tris) HIPCHK(hipMemcpyAsync(dIdx, index, size_t(3 * n_tris) * sizeof(uint32_t), hipMemcpyHostToDevice, B.s));
        HIPCHK(hipMemsetAsync(votes, 0, size_t(nSamples), B.s));
        int64_t maxRays = 0;
        for (int a = 0; a < 3; a++) maxRays = std::max(maxRays, int64_t(S[(a + 1) % 3]) * S[(a + 2) % 3]);
        uint32_t* cnt = B.get<uint32_t>(size_t(maxRays + 1));
        uint32_t* off = B.get<uint32_t>(size_t(maxRays + 1));
        size_t scanBytes = 0;
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, cnt, off, int(maxRays + 1), B.s));
        void* scanTmp = B.get<uint8_t>(scanBytes);
        uint32_t* overflow = B.get<uint32_t>(1);
        HIPCHK(hipMemsetAsync(overflow, 0, sizeof(uint32_t), B.s));
        uint32_t* total = nullptr;
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&total), sizeof(uint32_t), hipHostMallocDefault));
        struct Pinned {
            uint32_t* p;
            ~Pinned() { (void)hipHostFree(p); }
        } pinned{total};
        for (int a = 0; a < 3; a++) {
            Axis X;
            X.a = a;
            X.u = (a + 1) % 3;
            X.v = (a + 2) % 3;
            X.Nu = S[X.u];
            X.Nv = S[X.v];
            X.Na = S[a];
            X.ds = ds;
            X.ou = origin[X.u] + jitter[X.u];
            X.ov = origin[X.v] + jitter[X.v];
            X.oa = origin[a];
            X.strideU = strides[X.u];
            X.strideV = strides[X.v];
            X.strideA = strides[a];
            const int64_t nRays = int64_t(X.Nu) * X.Nv;
            if (nRays + 1 > int64_t(UINT32_MAX) / 2 || (X.Nv + ROWS - 1) / ROWS > 65535) throw std::invalid_argument("sample grid too large");
            HIPCHK(hipMemsetAsync(cnt, 0, size_t(nRays + 1) * sizeof(uint32_t), B.s));
            const dim3 g(unsigned((n_tris + BLOCK - 1) / BLOCK), unsigned((X.Nv + ROWS - 1) / ROWS));
            if (n_tris) kRaster<false><<<g, BLOCK, 0, B.s>>>(dPos, dIdx, n_tris, X, cnt, nullptr, nullptr, nullptr);
            HIPCHK(hipcub::DeviceScan::ExclusiveSum(scanTmp, scanBytes, cnt, off, int(nRays + 1), B.s));
            HIPCHK(hipMemcpyAsync(total, off + nRays, sizeof(uint32_t), hipMemcpyDeviceToHost, B.s));
            HIPCHK(hipStreamSynchronize(B.s));
            float* hits = B.get<float>(size_t(*total));
            HIPCHK(hipMemcpyAsync(cnt, off, size_t(nRays + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, B.s));
            if (n_tris) kRaster<true><<<g, BLOCK, 0, B.s>>>(dPos, dIdx, n_tris, X, cnt, hits, off, overflow);
            kSpans<<<unsigned((nRays + BLOCK - 1) / BLOCK), BLOCK, 0, B.s>>>(off, hits, X, votes);
            HIPCHK(hipGetLastError());
        }
        kFractions<<<unsigned((nVox + BLOCK - 1) / BLOCK), BLOCK, 0, B.s>>>(votes, dims[0], dims[1], dims[2], sub, S[0], S[1], dFrac);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(total, overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, B.s));
        HIPCHK(hipStreamSynchronize(B.s));
        if (*total != 0) throw DeviceError("voxelizer: the fill pass found " + std::to_string(*total) + " ray crossings the count pass did not");
        HIPCHK(hipMemcpyAsync(frac, dFrac, size_t(nVox) * sizeof(float), hipMemcpyDeviceToHost, B.s));
        HIPCHK(hipStreamSynchronize(B.s));
        return PSX_OK;
    });
}

}  // extern "C"
