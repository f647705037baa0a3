// Hierarchy build on the device (gfx950): the multigrid levels of a part, their exact Galerkin
// coarse elements and the coarsest level's dense inverse, computed by HIP kernels from the part's
// density and held DOFs instead of on the host CPU (native/src/fea/voxel_fea.cpp:138-386, 612-661 is
// the reference's CPU build; hierarchy.cpp is this library's host build of the same).
//
// Every level array is the host build's bit for bit: each value is computed by one thread with the
// host's operations in the host's order (this file is compiled without floating-point contraction,
// so a product followed by a sum stays two roundings, as in the host code), and every index that
// depends on an ordering (element numbers, own-matrix numbers, the coarsest free-DOF numbering) comes
// from a deterministic scan in voxel order. The one exception is the coarsest inverse: the host
// factors A = L L^T and forms L^-T L^-1; here A^-1 comes from m Gauss-Jordan sweeps of the symmetric
// matrix (one launch per two pivots, the whole matrix updated per launch), with the host's pivot
// regularisation (a pivot at most 1e-10 of the largest diagonal entry is replaced by that entry,
// voxel_fea.cpp:644-650). Both are the same matrix to rounding.
//
// Per level, in launch order:
//   kWords0 / kClassify + kCountFlags + kAssignOwn   voxel words (s bits, own index), element count
//   kChildForm                                       child-scale form (finer level all s * Kb)
//   kGalerkin                                        own matrices: sum over children of s M_c or
//                                                    P_c^T K_child P_c, one workgroup per element
//   kCoarseBc, kCoarseDiagAdd                        held DOFs, lumped diagonal shift
//   kNodes                                           inverse diagonal, node info, node scale, counts
// The host reads back 2-3 counters per level (element, own-matrix and free-DOF counts: the sizes of
// the next allocations and the coarsening stop rule of voxel_fea.cpp:174-180).
#include <hip/hip_runtime.h>

#include "model.hpp"

namespace psx {

namespace {

constexpr int BB = 256;
constexpr int SCAN = 1024;  // voxels per block of the own-index scan (4 per thread)
typedef unsigned long long u64;

inline unsigned grid1(int64_t n) { return unsigned(std::max<int64_t>(1, (n + BB - 1) / BB)); }

__device__ constexpr int dCorner(int x, int y, int z) { return (y ? (x ? 2 : 3) : (x ? 1 : 0)) + 4 * z; }
__device__ constexpr int dCX(int b) { return (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0; }
__device__ constexpr int dCY(int b) { return (b == 2 || b == 3 || b == 6 || b == 7) ? 1 : 0; }
__device__ constexpr int dCZ(int b) { return b >= 4 ? 1 : 0; }

// The counting kernels run on a fixed grid (grid-stride loops) and add one block sum per counter
// with one atomic: a few thousand same-address atomics instead of one per wavefront.
constexpr unsigned COUNT_GRID = 2048;
inline unsigned gridCount(int64_t n) { return std::min(grid1(n), COUNT_GRID); }
// adds the block's sum of v to *c (every thread of the block calls it)
__device__ __forceinline__ void blockAdd(u64* c, u64 v) {
    __shared__ u64 sh[BB / 64];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const u64 t = sh[0] + sh[1] + sh[2] + sh[3];
        if (t) atomicAdd(c, t);
    }
}

// ------------------------------------------------------------------------------------------------
// level 0: density check and voxel words

__global__ void kWords0(const float* dens, int64_t n, int32_t* words, u64* cnt) {
    u64 e = 0, bad = 0;
    for (int64_t v = int64_t(blockIdx.x) * BB + threadIdx.x; v < n; v += int64_t(gridDim.x) * BB) {
        const float d = dens[v];
        if (!(isfinite(d) && d >= 0.f && d <= 1.f)) bad++;
        else if (d > 0.f) {
            words[v] = __float_as_int(d);
            e++;
        }
    }
    blockAdd(&cnt[0], e);
    blockAdd(&cnt[1], bad);
}

// ------------------------------------------------------------------------------------------------
// coarsening F -> C

struct Grid {
    int nx, ny, nz, NX, NY, NZ;
};

// per coarse voxel: empty (no child), uniform (8 children, all s * Kb with one s: the word is s),
// or own (flag 1; the word is set by kAssignOwn)
__global__ void kClassify(Grid F, Grid C, const int32_t* fw, int32_t* cw, uint8_t* ownFlag, u64* cnt) {
    const int64_t nV = int64_t(C.nx) * C.ny * C.nz;
    u64 e = 0;
    for (int64_t V = int64_t(blockIdx.x) * BB + threadIdx.x; V < nV; V += int64_t(gridDim.x) * BB) {
        const int I = int(V % C.nx), J = int((V / C.nx) % C.ny), K = int(V / (int64_t(C.nx) * C.ny));
        int count = 0;
        bool uniform = true, first = true;
        float s = 0.f;
        for (int c = 0; c < 8; c++) {
            const int i = 2 * I + (c & 1), j = 2 * J + ((c >> 1) & 1), k = 2 * K + (c >> 2);
            if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
            const int32_t w = fw[i + int64_t(F.nx) * (j + int64_t(F.ny) * k)];
            if (w == 0) continue;
            count++;
            if (w < 0) uniform = false;
            else {
                const float sc = __int_as_float(w);
                if (first) s = sc, first = false;
                else if (sc != s) uniform = false;
            }
        }
        uniform = uniform && count == 8;
        ownFlag[V] = count > 0 && !uniform ? 1 : 0;
        cw[V] = count > 0 && uniform ? __float_as_int(s) : 0;
        e += count > 0 ? 1 : 0;
    }
    blockAdd(&cnt[0], e);
}

// own flags per block of SCAN voxels
__global__ void kCountFlags(const uint8_t* flag, int64_t n, int32_t* blockCount) {
    const int64_t base = int64_t(blockIdx.x) * SCAN + 4 * threadIdx.x;
    int c = 0;
    for (int t = 0; t < 4; t++) c += base + t < n ? flag[base + t] : 0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    __shared__ int sh[BB / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) blockCount[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// own index = number of own voxels before it in voxel order; word = 0x80000000 | index
__global__ void kAssignOwn(const uint8_t* flag, int64_t n, const int32_t* blockOffset, int32_t* cw, int32_t* ownVox) {
    const int64_t base = int64_t(blockIdx.x) * SCAN + 4 * threadIdx.x;
    int f[4], c = 0;
    for (int t = 0; t < 4; t++) {
        f[t] = base + t < n ? flag[base + t] : 0;
        c += f[t];
    }
    // inclusive prefix over the wave, then over the block's waves
    const int lane = threadIdx.x & 63;
    int inc = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    __shared__ int sh[BB / 64];
    if (lane == 63) sh[threadIdx.x >> 6] = inc;
    __syncthreads();
    int before = blockOffset[blockIdx.x] + inc - c;
    for (int w = 0; w < int(threadIdx.x >> 6); w++) before += sh[w];
    for (int t = 0; t < 4; t++)
        if (f[t]) {
            cw[base + t] = int32_t(0x80000000u | uint32_t(before));
            ownVox[before] = int32_t(base + t);
            before++;
        }
}

// child-scale form (the finer level has no own elements): the 8 child scales of each own element
// (0 for an absent child) and its two-scale record [a, b, mask of the b children or -1, 0]
__global__ void kChildForm(Grid F, Grid C, const int32_t* fw, const int32_t* ownVox, int64_t nOwn, float* childS, int32_t* childTwo, u64* cnt) {
    u64 two = 0;
    for (int64_t e = int64_t(blockIdx.x) * BB + threadIdx.x; e < nOwn; e += int64_t(gridDim.x) * BB) {
        const int64_t V = ownVox[e];
        const int I = int(V % C.nx), J = int((V / C.nx) % C.ny), K = int(V / (int64_t(C.nx) * C.ny));
        float cs[8];
        for (int c = 0; c < 8; c++) {
            const int i = 2 * I + (c & 1), j = 2 * J + ((c >> 1) & 1), k = 2 * K + (c >> 2);
            const int32_t w = (i < F.nx && j < F.ny && k < F.nz) ? fw[i + int64_t(F.nx) * (j + int64_t(F.ny) * k)] : 0;
            cs[c] = w > 0 ? __int_as_float(w) : 0.f;
            childS[8 * e + c] = cs[c];
        }
        const float a = cs[0];
        float b = a;
        int mask = 0;
        bool isTwo = true;
        for (int c = 1; c < 8; c++) {
            if (cs[c] == a) continue;
            if (mask == 0) b = cs[c];
            if (cs[c] == b) mask |= 1 << c;
            else isTwo = false;
        }
        const int t2 = isTwo && mask ? mask : -1;
        childTwo[4 * e + 0] = __float_as_int(a);
        childTwo[4 * e + 1] = __float_as_int(b);
        childTwo[4 * e + 2] = t2;
        childTwo[4 * e + 3] = 0;
        two += t2 >= 0 ? 1 : 0;
    }
    blockAdd(&cnt[0], two);
}

// own coarse element = sum over its children in child order of s M_c (child s * Kb) or
// P_c^T K_child P_c (child with an own matrix), in FP64, stored FP32. One workgroup per element;
// W[c][a][b] = trilinear weight of coarse corner b at fine corner a of child c
constexpr int GB = 192;
__global__ __launch_bounds__(GB) void kGalerkin(Grid F, Grid C, const int32_t* fw, const float* fK, const int32_t* ownVox, const double* M,
                                                const double* W, float* cK) {
    __shared__ double out[576], Kf[576], T[576];
    const int64_t e = blockIdx.x;
    const int64_t V = ownVox[e];
    const int I = int(V % C.nx), J = int((V / C.nx) % C.ny), K = int(V / (int64_t(C.nx) * C.ny));
    for (int u = threadIdx.x; u < 576; u += GB) out[u] = 0.0;
    for (int c = 0; c < 8; c++) {
        const int i = 2 * I + (c & 1), j = 2 * J + ((c >> 1) & 1), k = 2 * K + (c >> 2);
        if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
        const int32_t w = fw[i + int64_t(F.nx) * (j + int64_t(F.ny) * k)];
        if (w == 0) continue;
        const double* Wc = W + c * 64;
        if (w > 0) {
            const double s = double(__int_as_float(w));
            for (int u = threadIdx.x; u < 576; u += GB) out[u] += s * M[c * 576 + u];
            continue;
        }
        const float* Ko = fK + int64_t(w & 0x7fffffff) * 576;
        for (int u = threadIdx.x; u < 576; u += GB) Kf[u] = double(Ko[u]);
        __syncthreads();
        // T = K P_c: T[r][3 b + d] = sum_a K[r][3 a + d] W[a][b]
        for (int u = threadIdx.x; u < 576; u += GB) {
            const int r = u / 24, col = u % 24, b = col / 3, d = col % 3;
            double s = 0;
            for (int a = 0; a < 8; a++) {
                const double wab = Wc[a * 8 + b];
                if (wab != 0) s += Kf[r * 24 + 3 * a + d] * wab;
            }
            T[u] = s;
        }
        __syncthreads();
        // out[3 b + d][col] += sum_a W[a][b] T[3 a + d][col]
        for (int u = threadIdx.x; u < 576; u += GB) {
            const int row = u / 24, col = u % 24, b = row / 3, d = row % 3;
            double s = 0;
            for (int a = 0; a < 8; a++) {
                const double wab = Wc[a * 8 + b];
                if (wab != 0) s += wab * T[(3 * a + d) * 24 + col];
            }
            out[u] += s;
        }
        __syncthreads();
    }
    __syncthreads();
    for (int u = threadIdx.x; u < 576; u += GB) cK[e * 576 + u] = float(out[u]);
}

// a coarse DOF is held when it interpolates onto a held fine DOF: OR over the fine nodes
// 2I - 1 .. 2I + 1 per axis
__global__ void kCoarseBc(Grid F, Grid C, const uint8_t* fbc, uint8_t* cbc) {
    const int64_t N = int64_t(blockIdx.x) * BB + threadIdx.x;
    if (N >= int64_t(C.NX) * C.NY * C.NZ) return;
    const int I = int(N % C.NX), J = int((N / C.NX) % C.NY), K = int(N / (int64_t(C.NX) * C.NY));
    uint8_t h[3] = {0, 0, 0};
    for (int k = 2 * K - 1; k <= 2 * K + 1; k++) {
        if (k < 0 || k >= F.NZ) continue;
        for (int j = 2 * J - 1; j <= 2 * J + 1; j++) {
            if (j < 0 || j >= F.NY) continue;
            for (int i = 2 * I - 1; i <= 2 * I + 1; i++) {
                if (i < 0 || i >= F.NX) continue;
                const int64_t fn = 3 * (i + int64_t(F.NX) * (j + int64_t(F.NY) * k));
                h[0] |= fbc[fn];
                h[1] |= fbc[fn + 1];
                h[2] |= fbc[fn + 2];
            }
        }
    }
    cbc[3 * N] = h[0];
    cbc[3 * N + 1] = h[1];
    cbc[3 * N + 2] = h[2];
}

// lumped Galerkin product of a diagonal shift: P^T (D 1), full weighting
__global__ void kCoarseDiagAdd(Grid F, Grid C, const double* fda, double* cda) {
    const int64_t N = int64_t(blockIdx.x) * BB + threadIdx.x;
    if (N >= int64_t(C.NX) * C.NY * C.NZ) return;
    const int I = int(N % C.NX), J = int((N / C.NX) % C.NY), K = int(N / (int64_t(C.NX) * C.NY));
    double s[3] = {0, 0, 0};
    for (int dk = -1; dk <= 1; dk++) {
        const int kf = 2 * K + dk;
        if (kf < 0 || kf >= F.NZ) continue;
        for (int dj = -1; dj <= 1; dj++) {
            const int jf = 2 * J + dj;
            if (jf < 0 || jf >= F.NY) continue;
            for (int di = -1; di <= 1; di++) {
                const int fi = 2 * I + di;
                if (fi < 0 || fi >= F.NX) continue;
                const double w = (di ? 0.5 : 1.0) * (dj ? 0.5 : 1.0) * (dk ? 0.5 : 1.0);
                const int64_t fn = 3 * (fi + int64_t(F.NX) * (jf + int64_t(F.NY) * kf));
                for (int d = 0; d < 3; d++) s[d] += w * fda[fn + d];
            }
        }
    }
    for (int d = 0; d < 3; d++) cda[3 * N + d] = s[d];
}

// ------------------------------------------------------------------------------------------------
// per node: diagonal gathered over the node's voxels in voxel order (s Kb or the own matrix), held /
// inactive DOFs, inverse diagonal, the voxels whose scale differs from the node's reference voxel,
// the uniform-node scale; counts free DOFs and irregular nodes

__global__ void kNodes(Grid G, const int32_t* words, const float* Kown, const double* Kb, const uint8_t* bc, const double* diagAdd, double* invD64,
                       float* invD32, uint16_t* info, float* nodeScale, u64* cnt) {
    const int64_t nN = int64_t(G.NX) * G.NY * G.NZ;
    u64 nfree = 0, irregular = 0;
    for (int64_t n = int64_t(blockIdx.x) * BB + threadIdx.x; n < nN; n += int64_t(gridDim.x) * BB) {
        const int i = int(n % G.NX), j = int((n / G.NX) % G.NY), k = int(n / (int64_t(G.NX) * G.NY));
        double diag[3] = {0, 0, 0};
        bool act = false;
        float sv[8];
        for (int dk = 0; dk < 2; dk++)
            for (int dj = 0; dj < 2; dj++)
                for (int di = 0; di < 2; di++) {
                    const int v = di + 2 * dj + 4 * dk;
                    const int vi = i - 1 + di, vj = j - 1 + dj, vk = k - 1 + dk;
                    sv[v] = 0.f;
                    if (vi < 0 || vj < 0 || vk < 0 || vi >= G.nx || vj >= G.ny || vk >= G.nz) continue;
                    const int32_t w = words[vi + int64_t(G.nx) * (vj + int64_t(G.ny) * vk)];
                    if (w == 0) continue;
                    act = true;
                    const int a = dCorner(1 - di, 1 - dj, 1 - dk);
                    if (w < 0) {
                        sv[v] = -1.f;
                        const float* Ko = Kown + int64_t(w & 0x7fffffff) * 576;
                        for (int d = 0; d < 3; d++) diag[d] += double(Ko[(3 * a + d) * 25]);
                    } else {
                        const float s = __int_as_float(w);
                        sv[v] = s;
                        for (int d = 0; d < 3; d++) diag[d] += double(s) * Kb[(3 * a + d) * 25];
                    }
                }
        uint16_t word = 0;
        for (int d = 0; d < 3; d++) {
            const int64_t q = 3 * n + d;
            const double dg = diagAdd ? diag[d] + diagAdd[q] : diag[d];
            double inv = 0;
            if (!(!act || bc[q] || !(dg > 0))) {
                inv = 1 / dg;
                nfree++;
            } else word |= uint16_t(1u << (8 + d));
            invD64[q] = inv;
            invD32[q] = float(inv);
        }
        int ref = 0, best = 0;
        for (int v = 0; v < 8; v++) {
            int c = 0;
            for (int u = 0; u < 8; u++) c += sv[u] == sv[v] && sv[v] >= 0;
            if (c > best) best = c, ref = v;
        }
        uint16_t m = 0;
        for (int v = 0; v < 8; v++)
            if (!(sv[v] == sv[ref] && sv[v] >= 0)) m |= uint16_t(1u << v);
        info[n] = uint16_t(word | m);
        irregular += m ? 1 : 0;
        float ns = 0.f;
        if (i >= 1 && j >= 1 && k >= 1 && i < G.nx && j < G.ny && k < G.nz && sv[0] > 0) {
            bool same = true;
            for (int v = 1; v < 8; v++) same = same && sv[v] == sv[0];
            if (same) ns = sv[0];
        }
        nodeScale[n] = ns;
    }
    blockAdd(&cnt[0], nfree);
    blockAdd(&cnt[1], irregular);
}

// ------------------------------------------------------------------------------------------------
// coarsest level

// row r of the free-DOF matrix, gathered over the row node's voxels in voxel order (the host's
// element-order scatter gives every entry the same sum), plus the diagonal shift
__global__ void kAssemble(Grid G, const int32_t* words, const float* Kown, const double* Kb, const int32_t* freeIdx, const int32_t* dense, int64_t m,
                          const double* diagAdd, double* A) {
    const int64_t r = int64_t(blockIdx.x) * BB + threadIdx.x;
    if (r >= m) return;
    const int64_t q = freeIdx[r];
    const int64_t n = q / 3;
    const int d = int(q % 3);
    const int i = int(n % G.NX), j = int((n / G.NX) % G.NY), k = int(n / (int64_t(G.NX) * G.NY));
    double* row = A + r * m;
    for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                const int vi = i - 1 + di, vj = j - 1 + dj, vk = k - 1 + dk;
                if (vi < 0 || vj < 0 || vk < 0 || vi >= G.nx || vj >= G.ny || vk >= G.nz) continue;
                const int32_t w = words[vi + int64_t(G.nx) * (vj + int64_t(G.ny) * vk)];
                if (w == 0) continue;
                const int rr = 3 * dCorner(1 - di, 1 - dj, 1 - dk) + d;
                const float* Ko = w < 0 ? Kown + int64_t(w & 0x7fffffff) * 576 : nullptr;
                const double s = w < 0 ? 0.0 : double(__int_as_float(w));
                for (int b = 0; b < 8; b++) {
                    const int64_t nb = (vi + dCX(b)) + int64_t(G.NX) * ((vj + dCY(b)) + int64_t(G.NY) * (vk + dCZ(b)));
                    for (int dd = 0; dd < 3; dd++) {
                        const int col = dense[3 * nb + dd];
                        if (col < 0) continue;
                        const int cc = 3 * b + dd;
                        row[col] += Ko ? double(Ko[rr * 24 + cc]) : s * Kb[rr * 24 + cc];
                    }
                }
            }
    if (diagAdd) row[r] += diagAdd[q];
}

__global__ void kMaxDiag(const double* A, int64_t m, double* out) {
    double v = 0;
    for (int64_t r = threadIdx.x; r < m; r += BB) v = fmax(v, A[r * m + r]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double sh[BB / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) *out = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// one Gauss-Jordan sweep of the symmetric matrix on pivot k: B = sweep_k(A). After all m sweeps
// the matrix is -A^-1. The pivot is regularised as the host's Cholesky regularises it.
__global__ void kSweep(const double* A, double* B, int64_t m, int64_t k, const double* maxDiag) {
    const int64_t t = int64_t(blockIdx.x) * BB + threadIdx.x;
    if (t >= m * m) return;
    const int64_t i = t / m, j = t % m;
    double d = A[k * m + k];
    const double md = *maxDiag;
    if (!(d > 1e-10 * md)) d = md;
    const double aik = A[i * m + k], akj = A[k * m + j];
    double v;
    if (i == k && j == k) v = -1 / d;
    else if (i == k) v = akj / d;
    else if (j == k) v = aik / d;
    else v = A[t] - aik * akj / d;
    B[t] = v;
}

// two sweeps in one launch (pivots k and k + 1): every thread evaluates the entries of sweep_k(A)
// it needs with kSweep's expressions, then sweep_(k+1) of those; bitwise the two launches' result
__global__ void kSweep2(const double* A, double* B, int64_t m, int64_t k, const double* maxDiag) {
    const int64_t t = int64_t(blockIdx.x) * BB + threadIdx.x;
    if (t >= m * m) return;
    const int64_t i = t / m, j = t % m, k1 = k + 1;
    const double md = *maxDiag;
    double d0 = A[k * m + k];
    if (!(d0 > 1e-10 * md)) d0 = md;
    auto s1 = [&](int64_t a, int64_t b) -> double {
        if (a == k && b == k) return -1 / d0;
        if (a == k) return A[k * m + b] / d0;
        if (b == k) return A[a * m + k] / d0;
        return A[a * m + b] - A[a * m + k] * A[k * m + b] / d0;
    };
    double d1 = s1(k1, k1);
    if (!(d1 > 1e-10 * md)) d1 = md;
    double v;
    if (i == k1 && j == k1) v = -1 / d1;
    else if (i == k1) v = s1(k1, j) / d1;
    else if (j == k1) v = s1(i, k1) / d1;
    else v = s1(i, j) - s1(i, k1) * s1(k1, j) / d1;
    B[t] = v;
}

__global__ void kNegate(const double* A, double* out, int64_t n) {
    const int64_t t = int64_t(blockIdx.x) * BB + threadIdx.x;
    if (t < n) out[t] = -A[t];
}

// ------------------------------------------------------------------------------------------------
// Lanczos start vector: the LCG x <- (1103515245 x + 12345) mod 2^31 from 12345, DOF q taking the
// state after q + 1 steps (computed by repeated squaring of the affine map), held DOFs 0

__device__ uint32_t lcgAfter(uint64_t steps) {
    uint32_t A = 1, C = 0, a = 1103515245u, c = 12345u;
    while (steps) {
        if (steps & 1) {
            C = a * C + c;
            A = a * A;
        }
        c = a * c + c;
        a = a * a;
        steps >>= 1;
    }
    return (A * 12345u + C) & 0x7fffffffu;
}

__global__ void kLzStart(const double* invD, int64_t n, double* v) {
    const int64_t q = int64_t(blockIdx.x) * BB + threadIdx.x;
    if (q >= n) return;
    v[q] = invD[q] == 0 ? 0.0 : double(lcgAfter(uint64_t(q) + 1)) / 0x7fffffff - 0.5;
}

__global__ void kLzScale(double* v, int64_t n, const double* nn) {
    const int64_t q = int64_t(blockIdx.x) * BB + threadIdx.x;
    if (q >= n) return;
    const double s = *nn > 0 ? 1 / sqrt(*nn) : 1;
    v[q] *= s;
}

// ------------------------------------------------------------------------------------------------
// host side

Grid gridOf(const Level& L) { return Grid{L.nx, L.ny, L.nz, L.NX, L.NY, L.NZ}; }

Level levelShape(int nx, int ny, int nz) {
    Level L;
    L.nx = nx; L.ny = ny; L.nz = nz;
    L.NX = nx + 1; L.NY = ny + 1; L.NZ = nz + 1;
    L.nNodes = i64(L.NX) * L.NY * L.NZ;
    L.nDof = 3 * L.nNodes;
    L.nVox = i64(nx) * ny * nz;
    return L;
}

// temporary device buffers of one build, freed stream-ordered at the end
struct Scratch {
    hipStream_t s;
    std::vector<void*> bufs;
    template <class T> T* get(int64_t count) {
        void* p = nullptr;
        HIPCHK(hipMallocAsync(&p, size_t(std::max<int64_t>(count, 1)) * sizeof(T), s));
        bufs.push_back(p);
        return static_cast<T*>(p);
    }
    ~Scratch() {
        for (void* p : bufs) (void)hipFreeAsync(p, s);
    }
};

}  // namespace

void launchLanczosStart(const double* invD, int64_t n, double* v, double* tmp, Reduce red, hipStream_t s) {
    kLzStart<<<grid1(n), BB, 0, s>>>(invD, n, v);
    launchNormFree64(v, invD, n, tmp, red, s);
    kLzScale<<<grid1(n), BB, 0, s>>>(v, n, tmp);
}

void buildOnDevice(Model& M, const psx_model_desc& d) {
    checkGridLimits(d.nx, d.ny, d.nz);
    if (!d.density || !d.bc) throw std::invalid_argument("density and bc are required");
    const int maxLevels = d.max_levels <= 0 || d.max_levels > 12 ? 12 : d.max_levels;
    const i64 coarsestMaxDof = d.coarsest_max_dof > 0 ? d.coarsest_max_dof : 1100;
    hipStream_t s = M.stream;
    Hierarchy& H = M.H;
    H.nu = d.nu;
    H.el = hexElement(d.nu);
    H.levels.clear();
    M.deviceBuilt = true;
    Scratch tmp{s, {}};
    u64* cnt = tmp.get<u64>(4);
    u64 hc[4];
    auto counters = [&](int k) {
        HIPCHK(hipMemcpyAsync(hc, cnt, size_t(k) * sizeof(u64), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipMemsetAsync(cnt, 0, 4 * sizeof(u64), s));
    };
    HIPCHK(hipMemsetAsync(cnt, 0, 4 * sizeof(u64), s));
    // interpolation weights W[c][a][b]
    double Wt[8 * 64];
    for (int c = 0; c < 8; c++)
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++) Wt[c * 64 + a * 8 + b] = childWeight(c, a, b);
    double* dW = tmp.get<double>(8 * 64);
    HIPCHK(hipMemcpyAsync(dW, Wt, sizeof Wt, hipMemcpyHostToDevice, s));

    // level 0
    H.levels.push_back(levelShape(d.nx, d.ny, d.nz));
    M.lv.resize(1);
    {
        Level& L = H.levels[0];
        std::memcpy(L.Kb, H.el.K, sizeof L.Kb);
        DevLevel& D = M.lv[0].dl;
        float* dens = tmp.get<float>(L.nVox);
        HIPCHK(hipMemcpyAsync(dens, d.density, size_t(L.nVox) * 4, hipMemcpyHostToDevice, s));
        int32_t* words = M.alloc<int32_t>(L.nVox);
        kWords0<<<gridCount(L.nVox), BB, 0, s>>>(dens, L.nVox, words, cnt);
        D.words = words;
        counters(2);
        if (hc[1]) throw std::invalid_argument("Voxel density must be between zero and one.");
        L.nElem = i64(hc[0]);
        M.bc0 = M.upload(d.bc, L.nDof);
        D.diagAdd = d.diag_add ? M.upload(d.diag_add, L.nDof) : nullptr;
        D.Kown = M.alloc<float>(0);
    }
    const uint8_t* bcL = M.bc0;
    for (size_t l = 0;; l++) {
        // finish level l: per-node data
        {
            Level& L = H.levels[l];
            DevLevel& D = M.lv[l].dl;
            D.nx = L.nx; D.ny = L.ny; D.nz = L.nz; D.NX = L.NX; D.NY = L.NY; D.NZ = L.NZ;
            D.nNodes = L.nNodes;
            D.nDof = L.nDof;
            M.uploadConstants(l);
            double* invD64 = M.alloc<double>(L.nDof);
            float* invD32 = M.alloc<float>(L.nDof);
            uint16_t* info = M.alloc<uint16_t>(L.nNodes);
            float* nodeScale = M.alloc<float>(L.nNodes);
            kNodes<<<gridCount(L.nNodes), BB, 0, s>>>(gridOf(L), D.words, D.Kown, D.Kb64, bcL, D.diagAdd, invD64, invD32, info, nodeScale, cnt);
            D.invD64 = invD64;
            D.invD32 = invD32;
            D.nodeInfo = info;
            D.nodeScale = nodeScale;
            counters(2);
            L.freeDof = i64(hc[0]);
            const i64 irregular = i64(hc[1]);
            const char* e = std::getenv("PSX_DENSE");
            D.dense = L.nOwn == 0 && (e && e[0] ? e[0] == '1' : 2 * irregular > L.nNodes);
        }
        const Level& F = H.levels[l];
        if (F.freeDof <= coarsestMaxDof || int(H.levels.size()) >= maxLevels) break;
        if (F.nx <= 1 && F.ny <= 1 && F.nz <= 1) break;
        // coarsen l -> l + 1
        H.levels.push_back(levelShape((F.nx + 1) >> 1, (F.ny + 1) >> 1, (F.nz + 1) >> 1));
        M.lv.resize(H.levels.size());
        Level& Fl = H.levels[l];
        Level& C = H.levels[l + 1];
        const DevLevel& DF = M.lv[l].dl;
        DevLevel& DC = M.lv[l + 1].dl;
        std::vector<double> Mc(8 * 576);
        childProducts(Fl.Kb, Mc.data(), C.Kb);
        double* dM = tmp.get<double>(8 * 576);
        HIPCHK(hipMemcpyAsync(dM, Mc.data(), Mc.size() * 8, hipMemcpyHostToDevice, s));
        int32_t* cw = M.alloc<int32_t>(C.nVox);
        uint8_t* flag = tmp.get<uint8_t>(C.nVox);
        kClassify<<<gridCount(C.nVox), BB, 0, s>>>(gridOf(Fl), gridOf(C), DF.words, cw, flag, cnt);
        const int64_t nb = (C.nVox + SCAN - 1) / SCAN;
        int32_t* bcount = tmp.get<int32_t>(nb);
        kCountFlags<<<unsigned(nb), BB, 0, s>>>(flag, C.nVox, bcount);
        std::vector<int32_t> hb(static_cast<size_t>(nb));
        HIPCHK(hipMemcpyAsync(hb.data(), bcount, size_t(nb) * 4, hipMemcpyDeviceToHost, s));
        counters(1);
        C.nElem = i64(hc[0]);
        int64_t nOwn = 0;
        for (auto& v : hb) {
            const int32_t c = v;
            v = int32_t(nOwn);
            nOwn += c;
        }
        C.nOwn = nOwn;
        int32_t* boff = tmp.get<int32_t>(nb);
        HIPCHK(hipMemcpyAsync(boff, hb.data(), size_t(nb) * 4, hipMemcpyHostToDevice, s));
        int32_t* ownVox = tmp.get<int32_t>(nOwn);
        kAssignOwn<<<unsigned(nb), BB, 0, s>>>(flag, C.nVox, boff, cw, ownVox);
        DC.words = cw;
        float* cK = M.alloc<float>(nOwn * 576);
        if (nOwn > 0) kGalerkin<<<unsigned(nOwn), GB, 0, s>>>(gridOf(Fl), gridOf(C), DF.words, DF.Kown, ownVox, dM, dW, cK);
        DC.Kown = cK;
        if (Fl.nOwn == 0) {
            C.childForm = true;
            C.childM = Mc;
            C.childTab.assign(256 * 64, 0.0);
            childTable(Mc.data(), C.childTab.data());
            DC.childForm = nOwn > 0;
            if (nOwn > 0) {
                float* cs = M.alloc<float>(nOwn * 8);
                int32_t* two = M.alloc<int32_t>(nOwn * 4);
                kChildForm<<<gridCount(nOwn), BB, 0, s>>>(gridOf(Fl), gridOf(C), DF.words, ownVox, nOwn, cs, two, cnt);
                DC.childS = cs;
                DC.childTwo = two;
                M.uploadChildConstants(l + 1);
                counters(1);
                M.lv[l + 1].childCheap = 2 * i64(hc[0]) >= nOwn;
            }
        }
        uint8_t* cbc = tmp.get<uint8_t>(C.nDof);
        kCoarseBc<<<grid1(C.nNodes), BB, 0, s>>>(gridOf(Fl), gridOf(C), bcL, cbc);
        bcL = cbc;
        if (DF.diagAdd) {
            double* cda = M.alloc<double>(C.nDof);
            kCoarseDiagAdd<<<grid1(C.nNodes), BB, 0, s>>>(gridOf(Fl), gridOf(C), DF.diagAdd, cda);
            DC.diagAdd = cda;
        } else DC.diagAdd = nullptr;
    }
    HIPCHK(hipGetLastError());

    // coarsest level: free-DOF numbering, dense matrix, inverse
    {
        const Level& L = H.levels.back();
        const DevLevel& D = M.lv.back().dl;
        std::vector<double> inv(size_t(L.nDof));
        HIPCHK(hipMemcpyAsync(inv.data(), D.invD64, size_t(L.nDof) * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        H.freeIdx.clear();
        std::vector<int32_t> dense(size_t(L.nDof), -1);
        for (i64 q = 0; q < L.nDof; q++)
            if (inv[size_t(q)] != 0) {
                dense[size_t(q)] = int32_t(H.freeIdx.size());
                H.freeIdx.push_back(int32_t(q));
            }
        const i64 m = i64(H.freeIdx.size());
        H.m = m;
        if (m == 0 || m > 3000) throw std::runtime_error("the coarsest level has no direct solver");
        M.freeIdx = M.upload(H.freeIdx.data(), m);
        M.denseIdx = M.upload(dense.data(), L.nDof);
        double* A = tmp.get<double>(m * m);
        double* B = tmp.get<double>(m * m);
        double* md = tmp.get<double>(1);
        HIPCHK(hipMemsetAsync(A, 0, size_t(m * m) * 8, s));
        kAssemble<<<grid1(m), BB, 0, s>>>(gridOf(L), D.words, D.Kown, D.Kb64, M.freeIdx, M.denseIdx, m, D.diagAdd, A);
        kMaxDiag<<<1, BB, 0, s>>>(A, m, md);
        for (i64 k = 0; k < m; k += 2) {
            if (k + 1 < m) kSweep2<<<grid1(m * m), BB, 0, s>>>(A, B, m, k, md);
            else kSweep<<<grid1(m * m), BB, 0, s>>>(A, B, m, k, md);
            std::swap(A, B);
        }
        double* Ainv = M.alloc<double>(m * m);
        kNegate<<<grid1(m * m), BB, 0, s>>>(A, Ainv, m * m);
        M.Ainv = Ainv;
        HIPCHK(hipGetLastError());
    }
    M.finishSetup();
}

}  // namespace psx
