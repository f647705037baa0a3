// Topology optimization on the device (topology.cpp:21-183 is the reference form): SIMP densities
// x in [xmin, 1] with stiffness x^p, compliance minimized for a volume budget by the optimality-
// criteria update with a density filter of radius rmin voxels. Every per-voxel step runs as a
// kernel on HBM-resident arrays: the element energies u_e^T K0 u_e of the solution (a gather over
// the voxel's 8 nodes, FP64), the filter and its transpose (a gather over the ball of neighbour
// offsets in the reference's order, so the sums are the reference's), the sensitivities, the
// bisection on the volume multiplier (one update + filter + volume reduction per step, the host
// only halves the interval) and the change. Each iteration solves the static problem with this
// library's multigrid solver, warm-started from the previous solution in HBM; the hierarchy of the
// new densities is built on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../../include/psx.h"
#include "abi.hpp"

#pragma clang fp contract(off)

namespace psx {

namespace {

constexpr int BLOCK = 256;
constexpr double TOPO_XMIN = 1e-3;  // topology.hpp:23

struct Off {
    int i, j, k;
    double w;
};

// u_e^T K0 u_e per element (voxel order), FP64 in the reference's order
__global__ __launch_bounds__(BLOCK) void kEnergies(const int32_t* __restrict__ elemVox, int64_t nE, int nx, int ny, const double* __restrict__ K,
                                                   const double* __restrict__ u, double* __restrict__ out) {
    const int64_t e = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (e >= nE) return;
    const int v = elemVox[e];
    const int i = v % nx, j = (v / nx) % ny, k = v / (nx * ny);
    const int64_t NX = nx + 1, NXY = NX * (ny + 1);
    double ue[24];
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const int cx = (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0, cy = (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0, cz = a >= 4 ? 1 : 0;
        const int64_t n = 3 * ((i + cx) + NX * (j + cy) + NXY * (k + cz));
        ue[3 * a] = u[n];
        ue[3 * a + 1] = u[n + 1];
        ue[3 * a + 2] = u[n + 2];
    }
    double s = 0;
    for (int r = 0; r < 24; r++) {
        double t = 0;
#pragma unroll
        for (int c = 0; c < 24; c++) t += K[r * 24 + c] * ue[c];
        s += ue[r] * t;
    }
    out[e] = s;
}

struct FilterGrid {
    int nx, ny, nz, nOff;
    const Off* off;
    const int32_t* solid;  // solid index -> voxel
    const int32_t* pos;    // voxel -> solid index, -1 outside
};

// MODE 0: out[q] = sum w x[p] / W[q]; 1: out[q] = sum w g[p] / W[p]; 2: out[q] = sum w (W itself)
template <int MODE>
__global__ __launch_bounds__(BLOCK) void kFilter(FilterGrid G, int64_t m, const double* __restrict__ W, const double* __restrict__ x, double* __restrict__ out) {
    const int64_t q = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (q >= m) return;
    const int e = G.solid[q];
    const int i = e % G.nx, j = (e / G.nx) % G.ny, k = e / (G.nx * G.ny);
    double s = 0;
    for (int t = 0; t < G.nOff; t++) {
        const Off o = G.off[t];
        const int a = i + o.i, b = j + o.j, c = k + o.k;
        if (a < 0 || b < 0 || c < 0 || a >= G.nx || b >= G.ny || c >= G.nz) continue;
        const int p = G.pos[a + G.nx * (b + G.ny * c)];
        if (p < 0) continue;
        if constexpr (MODE == 0) s += o.w * x[p];
        else if constexpr (MODE == 1) s += o.w * x[p] / W[p];
        else s += o.w;
    }
    out[q] = MODE == 0 ? s / W[q] : s;
}

// xPhys of the kept voxels = 1; density of the solve
__global__ __launch_bounds__(BLOCK) void kDensity(int64_t m, const int32_t* __restrict__ solid, const float* __restrict__ fill, const uint8_t* __restrict__ isFree,
                                                  double* __restrict__ xPhys, double penal, float* __restrict__ density) {
    const int64_t q = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (q >= m) return;
    if (!isFree[q]) xPhys[q] = 1;
    const int v = solid[q];
    density[v] = float(fmin(1.0, fill[v] * (TOPO_XMIN + (1 - TOPO_XMIN) * pow(xPhys[q], penal))));
}

// dc = -p (1 - xmin) fill xPhys^(p - 1) energy of the voxel's element
__global__ __launch_bounds__(BLOCK) void kSensitivity(int64_t m, const int32_t* __restrict__ solid, const int32_t* __restrict__ elemOf,
                                                      const float* __restrict__ fill, const double* __restrict__ xPhys, const double* __restrict__ energy,
                                                      double penal, double* __restrict__ dc) {
    const int64_t q = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (q >= m) return;
    const int v = solid[q];
    const int e = elemOf[v];
    const double en = e >= 0 ? energy[e] : 0.0;
    dc[q] = -penal * (1 - TOPO_XMIN) * fill[v] * pow(xPhys[q], penal - 1) * en;
}

// the filtered densities per voxel for the iteration callback (the voxels outside stay 0)
__global__ __launch_bounds__(BLOCK) void kScatter(int64_t m, const int32_t* __restrict__ solid, const double* __restrict__ xPhys, float* __restrict__ out) {
    const int64_t q = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (q < m) out[solid[q]] = float(xPhys[q]);
}

// the optimality-criteria candidate for a multiplier lm
__global__ __launch_bounds__(BLOCK) void kOc(int64_t m, const uint8_t* __restrict__ isFree, const double* __restrict__ x, const double* __restrict__ dcx,
                                             const double* __restrict__ dvx, double lm, double move, double* __restrict__ xNew) {
    const int64_t q = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (q >= m) return;
    if (!isFree[q]) {
        xNew[q] = 1;
        return;
    }
    const double B = sqrt(fmax(0.0, -dcx[q]) / (lm * dvx[q] + 1e-300));
    xNew[q] = fmax(0.0, fmax(x[q] - move, fmin(1.0, fmin(x[q] + move, x[q] * B))));
}

// deterministic reductions: per-block partial sums (fixed blocks), then one block in order
template <int KIND>  // 0: sum a b (b null: sum a), 1: sum vol (isFree ? a : 1), 2: max |a - b|
__global__ __launch_bounds__(BLOCK) void kReduce(int64_t n, const double* __restrict__ a, const double* __restrict__ b, const uint8_t* __restrict__ isFree,
                                                 const double* __restrict__ vol, double* __restrict__ part) {
    double s = 0;
    for (int64_t q = int64_t(blockIdx.x) * BLOCK + threadIdx.x; q < n; q += int64_t(gridDim.x) * BLOCK) {
        if constexpr (KIND == 0) s += b ? a[q] * b[q] : a[q];
        else if constexpr (KIND == 1) s += vol[q] * (isFree[q] ? a[q] : 1.0);
        else s = fmax(s, fabs(a[q] - b[q]));
    }
    __shared__ double sh[BLOCK];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) sh[threadIdx.x] = KIND == 2 ? fmax(sh[threadIdx.x], sh[threadIdx.x + w]) : sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
template <int KIND> __global__ void kReduceFinish(int nb, const double* __restrict__ part, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0;
    for (int b = 0; b < nb; b++) s = KIND == 2 ? fmax(s, part[b]) : s + part[b];
    *out = s;
}

struct Dev {
    hipStream_t s = nullptr;
    std::vector<void*> p;
    template <class T> T* get(size_t count) {
        void* q = nullptr;
        HIPCHK(hipMallocAsync(&q, std::max<size_t>(count, 1) * sizeof(T), s));
        HIPCHK(hipMemsetAsync(q, 0, std::max<size_t>(count, 1) * sizeof(T), s));
        p.push_back(q);
        return static_cast<T*>(q);
    }
    template <class T> T* put(const T* src, size_t count) {
        T* d = get<T>(count);
        if (count) HIPCHK(hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, s));
        return d;
    }
    ~Dev() {
        for (void* q : p) (void)hipFreeAsync(q, s);
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

constexpr int RED_BLOCKS = 256;

// offsets of the filter ball in the order k, j, i with weights rmin - distance > 0
// (topology.cpp:21-37)
std::vector<Off> ballOffsets(double rmin) {
    std::vector<Off> offs;
    const int r = int(std::ceil(rmin)) - 1;
    for (int k = -r - 1; k <= r + 1; k++)
        for (int j = -r - 1; j <= r + 1; j++)
            for (int i = -r - 1; i <= r + 1; i++) {
                const double w = rmin - std::sqrt(double(i * i + j * j + k * k));
                if (w > 0) offs.push_back({i, j, k, w});
            }
    return offs;
}

template <int KIND> double reduce(Dev& D, int64_t n, const double* a, const double* b, const uint8_t* isFree, const double* vol, double* part, double* res) {
    kReduce<KIND><<<RED_BLOCKS, BLOCK, 0, D.s>>>(n, a, b, isFree, vol, part);
    kReduceFinish<KIND><<<1, 64, 0, D.s>>>(RED_BLOCKS, part, res);
    double h = 0;
    HIPCHK(hipMemcpyAsync(&h, res, sizeof(double), hipMemcpyDeviceToHost, D.s));
    HIPCHK(hipStreamSynchronize(D.s));
    return h;
}

inline unsigned grid(int64_t n) { return unsigned(std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK)); }

}  // namespace

}  // namespace psx

extern "C" {

int psx_element_energies(psx_model m, const double* u, double* out) {
    using namespace psx;
    return abi::guarded([&] {
        if (!m || !u || !out) throw std::invalid_argument("bad arguments");
        Model& M = *m->m;
        const Level& L = M.hostLevel0();
        const int64_t nE = int64_t(L.elemVox.size());
        Dev D;
        HIPCHK(hipStreamCreateWithFlags(&D.s, hipStreamNonBlocking));
        const int32_t* ev = D.put(L.elemVox.data(), size_t(nE));
        const double* K = D.put(M.H.el.K, 576);
        const double* du = D.put(u, size_t(L.nDof));
        double* dout = D.get<double>(size_t(nE));
        kEnergies<<<grid(nE), BLOCK, 0, D.s>>>(ev, nE, L.nx, L.ny, K, du, dout);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dout, size_t(nE) * sizeof(double), hipMemcpyDeviceToHost, D.s));
        HIPCHK(hipStreamSynchronize(D.s));
        return PSX_OK;
    });
}

int psx_topology_filter(const int dims[3], const float* fill, double rmin, int transpose, const double* x, double* out) {
    using namespace psx;
    return abi::guarded([&] {
        if (!dims || !fill || !x || !out || dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || !(rmin > 0)) throw std::invalid_argument("bad arguments");
        const int64_t nV = int64_t(dims[0]) * dims[1] * dims[2];
        std::vector<int32_t> solid, pos(size_t(nV), -1);
        for (int64_t e = 0; e < nV; e++)
            if (fill[e] > 0) {
                pos[size_t(e)] = int32_t(solid.size());
                solid.push_back(int32_t(e));
            }
        const int64_t m = int64_t(solid.size());
        const std::vector<Off> offs = ballOffsets(rmin);
        Dev D;
        HIPCHK(hipStreamCreateWithFlags(&D.s, hipStreamNonBlocking));
        FilterGrid G{dims[0], dims[1], dims[2], int(offs.size()), D.put(offs.data(), offs.size()), D.put(solid.data(), solid.size()), D.put(pos.data(), pos.size())};
        double* W = D.get<double>(size_t(m));
        double* dx = D.put(x, size_t(m));
        double* dout = D.get<double>(size_t(m));
        kFilter<2><<<grid(m), BLOCK, 0, D.s>>>(G, m, nullptr, nullptr, W);
        if (transpose) kFilter<1><<<grid(m), BLOCK, 0, D.s>>>(G, m, W, dx, dout);
        else kFilter<0><<<grid(m), BLOCK, 0, D.s>>>(G, m, W, dx, dout);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dout, size_t(m) * sizeof(double), hipMemcpyDeviceToHost, D.s));
        HIPCHK(hipStreamSynchronize(D.s));
        return PSX_OK;
    });
}

int psx_optimize_topology(const int dims[3], const float* fill, const uint8_t* keep, const uint8_t* bc, const double* f, double nu,
                          const psx_topology_opts* opts, float* density, double* history, int* iterations, double* kept_fraction) {
    using namespace psx;
    return abi::guarded([&] {
        if (!dims || !fill || !keep || !bc || !f || !density || !history || !iterations) throw std::invalid_argument("bad arguments");
        const int nx = dims[0], ny = dims[1], nz = dims[2];
        if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("dims must be positive");
        psx_topology_opts o{};
        if (opts) o = *opts;
        if (o.vol_frac <= 0) o.vol_frac = 0.4;
        if (o.penal <= 0) o.penal = 3;
        if (o.rmin <= 0) o.rmin = 1.5;
        if (o.move <= 0) o.move = 0.2;
        if (o.tol <= 0) o.tol = 0.01;
        if (o.max_iter <= 0) o.max_iter = 40;
        if (o.solve_tol <= 0) o.solve_tol = 1e-8;
        if (o.coarsest_max_dof <= 0) o.coarsest_max_dof = 1100;
        if (o.solve_max_iter <= 0) o.solve_max_iter = 1000;
        if (o.precision < PSX_PREC_FP64 || o.precision > PSX_PREC_MIXED) throw std::invalid_argument("unknown precision");
        const int64_t nV = int64_t(nx) * ny * nz, nDof = 3 * int64_t(nx + 1) * (ny + 1) * (nz + 1);
        abi::checkFinite(f, nDof, "Loads");
        // the design voxels (topology.cpp:104-117): fill > 0; kept voxels stay solid
        std::vector<int32_t> solid, pos(size_t(nV), -1);
        for (int64_t e = 0; e < nV; e++)
            if (fill[e] > 0) {
                pos[size_t(e)] = int32_t(solid.size());
                solid.push_back(int32_t(e));
            }
        const int64_t m = int64_t(solid.size());
        if (m == 0) throw std::invalid_argument("no voxel has fill > 0");
        std::vector<double> vol(static_cast<size_t>(m));
        std::vector<uint8_t> isFree(size_t(m), 0);
        double totalVol = 0, keptVol = 0;
        for (int64_t q = 0; q < m; q++) {
            vol[size_t(q)] = fill[solid[size_t(q)]];
            totalVol += vol[size_t(q)];
            if (keep[solid[size_t(q)]]) keptVol += vol[size_t(q)];
            else isFree[size_t(q)] = 1;
        }
        const double target = std::max(o.vol_frac * totalVol, keptVol * 1.02);
        const double start = std::min(1.0, std::max(TOPO_XMIN, (target - keptVol) / std::max(1e-12, totalVol - keptVol)));
        std::vector<double> x0(static_cast<size_t>(m));
        for (int64_t q = 0; q < m; q++) x0[size_t(q)] = isFree[size_t(q)] ? start : 1;
        const std::vector<Off> offs = ballOffsets(o.rmin);
        Dev D;
        HIPCHK(hipStreamCreateWithFlags(&D.s, hipStreamNonBlocking));
        FilterGrid G{nx, ny, nz, int(offs.size()), D.put(offs.data(), offs.size()), D.put(solid.data(), solid.size()), D.put(pos.data(), pos.size())};
        const float* dFill = D.put(fill, size_t(nV));
        const uint8_t* dFree = D.put(isFree.data(), isFree.size());
        const double* dVol = D.put(vol.data(), vol.size());
        double* x = D.put(x0.data(), x0.size());
        double *W = D.get<double>(size_t(m)), *xPhys = D.get<double>(size_t(m)), *dc = D.get<double>(size_t(m)), *dcx = D.get<double>(size_t(m));
        double *dvx = D.get<double>(size_t(m)), *xNew = D.get<double>(size_t(m)), *phys = D.get<double>(size_t(m));
        float* dDens = D.get<float>(size_t(nV));
        const double* dF = D.put(f, size_t(nDof));
        double* u = D.get<double>(size_t(nDof));
        double* part = D.get<double>(RED_BLOCKS);
        double* res = D.get<double>(1);
        int32_t* elemOf = D.get<int32_t>(size_t(nV));
        int32_t* ev = D.get<int32_t>(size_t(nV));
        double* energy = D.get<double>(size_t(nV));
        double* K0 = D.get<double>(576);
        float* dShow = o.on_iter ? D.get<float>(size_t(nV)) : nullptr;
        std::vector<float> hShow(o.on_iter ? size_t(nV) : 0);
        kFilter<2><<<grid(m), BLOCK, 0, D.s>>>(G, m, nullptr, nullptr, W);
        // dv = vol in every iteration, so its filtered transpose is computed once
        kFilter<1><<<grid(m), BLOCK, 0, D.s>>>(G, m, W, dVol, dvx);
        HIPCHK(hipGetLastError());
        std::vector<float> hDens(size_t(nV), 0.f);
        std::vector<int32_t> hElemOf(static_cast<size_t>(nV));
        double change = 1;
        int iters = 0;
        bool warm = false;
        for (int it = 0; it < o.max_iter; it++) {
            kFilter<0><<<grid(m), BLOCK, 0, D.s>>>(G, m, W, x, xPhys);
            kDensity<<<grid(m), BLOCK, 0, D.s>>>(m, G.solid, dFill, dFree, xPhys, o.penal, dDens);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hDens.data(), dDens, size_t(nV) * sizeof(float), hipMemcpyDeviceToHost, D.s));
            HIPCHK(hipStreamSynchronize(D.s));
            // one static solve with these densities, warm-started from the previous solution
            psx_model_desc desc{};
            desc.nx = nx;
            desc.ny = ny;
            desc.nz = nz;
            desc.density = hDens.data();
            desc.nu = nu;
            desc.bc = bc;
            desc.coarsest_max_dof = o.coarsest_max_dof;
            desc.build_on_device = 1;
            Model M(desc);
            const Level& L0 = M.hostLevel0();
            const int64_t nE = int64_t(L0.elemVox.size());
            std::fill(hElemOf.begin(), hElemOf.end(), -1);
            for (int64_t e = 0; e < nE; e++) hElemOf[size_t(L0.elemVox[size_t(e)])] = int32_t(e);
            HIPCHK(hipMemcpyAsync(elemOf, hElemOf.data(), size_t(nV) * sizeof(int32_t), hipMemcpyHostToDevice, D.s));
            HIPCHK(hipMemcpyAsync(ev, L0.elemVox.data(), size_t(nE) * sizeof(int32_t), hipMemcpyHostToDevice, D.s));
            HIPCHK(hipMemcpyAsync(K0, M.H.el.K, 576 * sizeof(double), hipMemcpyHostToDevice, D.s));
            HIPCHK(hipStreamSynchronize(D.s));
            psx_solve_opts so{};
            so.tol = o.solve_tol;
            so.max_iter = o.solve_max_iter;
            so.precision = o.precision;
            psx_solve_info info{};
            M.solveDevice(dF, warm ? u : nullptr, so, u, info);
            warm = true;
            HIPCHK(hipStreamSynchronize(M.stream));
            kEnergies<<<grid(nE), BLOCK, 0, D.s>>>(ev, nE, nx, ny, K0, u, energy);
            kSensitivity<<<grid(m), BLOCK, 0, D.s>>>(m, G.solid, elemOf, dFill, xPhys, energy, o.penal, dc);
            HIPCHK(hipGetLastError());
            const double c = reduce<0>(D, nDof, dF, u, nullptr, nullptr, part, res);
            const double used = reduce<0>(D, m, dVol, xPhys, nullptr, nullptr, part, res);
            history[4 * it] = it;
            history[4 * it + 1] = c;
            history[4 * it + 2] = used / totalVol;
            history[4 * it + 3] = change;
            iters = it + 1;
            if (o.on_iter) {
                kScatter<<<grid(m), BLOCK, 0, D.s>>>(m, G.solid, xPhys, dShow);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(hShow.data(), dShow, size_t(nV) * sizeof(float), hipMemcpyDeviceToHost, D.s));
                HIPCHK(hipStreamSynchronize(D.s));
                if (o.on_iter(it, c, used / totalVol, change, hShow.data(), o.user)) break;
            }
            if (it > 4 && change < o.tol) break;
            kFilter<1><<<grid(m), BLOCK, 0, D.s>>>(G, m, W, dc, dcx);
            // optimality criteria: bisection on the volume multiplier, one candidate per step
            double l1 = 0, l2 = 1e9;
            while ((l2 - l1) / (l1 + l2 + 1e-30) > 1e-4) {
                const double lm = 0.5 * (l1 + l2);
                kOc<<<grid(m), BLOCK, 0, D.s>>>(m, dFree, x, dcx, dvx, lm, o.move, xNew);
                kFilter<0><<<grid(m), BLOCK, 0, D.s>>>(G, m, W, xNew, phys);
                const double v = reduce<1>(D, m, phys, nullptr, dFree, dVol, part, res);
                if (v > target) l1 = lm;
                else l2 = lm;
            }
            change = reduce<2>(D, m, xNew, x, nullptr, nullptr, part, res);
            HIPCHK(hipMemcpyAsync(x, xNew, size_t(m) * sizeof(double), hipMemcpyDeviceToDevice, D.s));
        }
        kFilter<0><<<grid(m), BLOCK, 0, D.s>>>(G, m, W, x, xPhys);
        kDensity<<<grid(m), BLOCK, 0, D.s>>>(m, G.solid, dFill, dFree, xPhys, o.penal, dDens);
        HIPCHK(hipGetLastError());
        std::vector<double> hx(static_cast<size_t>(m));
        HIPCHK(hipMemcpyAsync(hx.data(), xPhys, size_t(m) * sizeof(double), hipMemcpyDeviceToHost, D.s));
        HIPCHK(hipStreamSynchronize(D.s));
        for (int64_t e = 0; e < nV; e++) density[e] = 0.f;
        for (int64_t q = 0; q < m; q++) density[solid[size_t(q)]] = float(hx[size_t(q)]);
        *iterations = iters;
        if (kept_fraction) *kept_fraction = keptVol / totalVol;
        return PSX_OK;
    });
}

}  // extern "C"
