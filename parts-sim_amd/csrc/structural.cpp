// Input contract of a CAD part (native/src/fea/structural.cpp:19-209, native/src/core/voxelize.cpp:130-276,
// native/src/core/mesh.cpp:315-330, 540-568, native/src/core/bvh.cpp): the triangle data of a surface
// mesh, its typical wall thickness, the density of thin walls, the voxel densities of the structural
// model (the fill fractions come from the device voxelizer, psx_voxelize) and the assembly of
// fixtures and loads onto grid nodes. These results are graded bit for bit (fixture masks, nodal
// loads, densities), so every step keeps the reference's operations and their order; the work is
// small (per triangle, per sample point) next to a solve and runs on the host threads.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/psx.h"
#include "abi.hpp"
#include "psx_core.hpp"

namespace psx {

namespace {

// fn(lo, hi) over [0, n) in contiguous chunks on the host threads (order-independent work only)
template <class F> void parallelChunks(int64_t n, int64_t minChunk, F&& fn) {
    const int T = hostThreads();
    const int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(T, (n + minChunk - 1) / std::max<int64_t>(1, minChunk)));
    if (chunks <= 1 || n <= minChunk) {
        fn(int64_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    for (int64_t c = 0; c < chunks; c++) th.emplace_back([&, c] { fn(n * c / chunks, n * (c + 1) / chunks); });
    for (auto& t : th) t.join();
}

struct Mesh {
    const float* V = nullptr;
    int64_t nV = 0;
    const uint32_t* T = nullptr;
    int nT = 0;
    std::vector<float> normal, area;  // per triangle (mesh.cpp:315-330)
    double bmin[3], bmax[3], diag = 0;  // bounding box (mesh.cpp:427-439)
};

Mesh makeMesh(const psx_mesh* m) {
    if (!m || !m->vertices || !m->tris || m->n_vertices < 0 || m->n_tris < 0 || m->n_tris > INT32_MAX) throw std::invalid_argument("bad mesh");
    Mesh M;
    M.V = m->vertices;
    M.nV = m->n_vertices;
    M.T = m->tris;
    M.nT = int(m->n_tris);
    for (int64_t q = 0; q < 3 * m->n_tris; q++)
        if (int64_t(M.T[q]) >= M.nV) throw std::invalid_argument("triangle index out of range");
    M.normal.assign(3 * size_t(M.nT), 0.f);
    M.area.assign(size_t(M.nT), 0.f);
    for (int t = 0; t < M.nT; t++) {
        const uint32_t a = M.T[3 * t], b = M.T[3 * t + 1], c = M.T[3 * t + 2];
        const double ax = M.V[3 * a], ay = M.V[3 * a + 1], az = M.V[3 * a + 2];
        const double ux = M.V[3 * b] - ax, uy = M.V[3 * b + 1] - ay, uz = M.V[3 * b + 2] - az;
        const double vx = M.V[3 * c] - ax, vy = M.V[3 * c + 1] - ay, vz = M.V[3 * c + 2] - az;
        const double n[3] = {uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx};
        double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        const double ar = 0.5 * l;
        if (!(l > 0)) l = 1;
        for (int d = 0; d < 3; d++) M.normal[3 * size_t(t) + d] = float(n[d] / l);
        M.area[size_t(t)] = float(ar);
    }
    for (int d = 0; d < 3; d++) M.bmin[d] = 1e300, M.bmax[d] = -1e300;
    for (int64_t i = 0; i < M.nV; i++)
        for (int d = 0; d < 3; d++) {
            M.bmin[d] = std::min(M.bmin[d], double(M.V[3 * i + d]));
            M.bmax[d] = std::max(M.bmax[d], double(M.V[3 * i + d]));
        }
    const double s[3] = {M.bmax[0] - M.bmin[0], M.bmax[1] - M.bmin[1], M.bmax[2] - M.bmin[2]};
    M.diag = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    return M;
}

// ------------------------------------------------------------------------------------------------
// bounding-volume hierarchy with binned SAH splits and nearest-hit ray casts (bvh.cpp). Nearest hits
// at exactly equal distance keep the triangle visited last, so the tree is built with the same splits
// and partition calls as the reference's, which makes the visit order (and such ties) the same.

class Bvh {
public:
    explicit Bvh(const Mesh& m) : M(m) {
        const int n = m.nT;
        order.resize(size_t(n));
        cen.resize(3 * size_t(n));
        lo.resize(3 * size_t(n));
        hi.resize(3 * size_t(n));
        for (int t = 0; t < n; t++) {
            order[size_t(t)] = t;
            for (int d = 0; d < 3; d++) {
                const float a = m.V[3 * m.T[3 * t] + d], b = m.V[3 * m.T[3 * t + 1] + d], c = m.V[3 * m.T[3 * t + 2] + d];
                lo[3 * size_t(t) + d] = std::min({a, b, c});
                hi[3 * size_t(t) + d] = std::max({a, b, c});
                cen[3 * size_t(t) + d] = (a + b + c) / 3;
            }
        }
        nodes.reserve(size_t(n) + 1);
        if (n) build(0, n, 0);
        cen.clear();
    }

    bool raycast(const double o[3], const double d[3], double tmin, double tmax, int& hitTri, double& hitDist) const {
        if (nodes.empty()) return false;
        double inv[3];
        for (int k = 0; k < 3; k++) inv[k] = d[k] != 0 ? 1 / d[k] : std::copysign(1e300, d[k] == 0 ? 1.0 : d[k]);
        bool found = false;
        double best = tmax;
        int stack[128], sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const Node& n = nodes[size_t(stack[--sp])];
            double t0 = tmin, t1 = best;
            bool miss = false;
            for (int k = 0; k < 3 && !miss; k++) {
                double a = (n.min[k] - o[k]) * inv[k], b = (n.max[k] - o[k]) * inv[k];
                if (a > b) std::swap(a, b);
                t0 = std::max(t0, a);
                t1 = std::min(t1, b);
                miss = t0 > t1 * (1 + 1e-12) + 1e-300;
            }
            if (miss) continue;
            if (n.left < 0) {
                for (int i = n.start; i < n.start + n.count; i++) {
                    const int t = order[size_t(i)];
                    const float* A = &M.V[3 * M.T[3 * t]];
                    const float* B = &M.V[3 * M.T[3 * t + 1]];
                    const float* C = &M.V[3 * M.T[3 * t + 2]];
                    // Moller-Trumbore in FP64
                    const double e1[3] = {double(B[0]) - A[0], double(B[1]) - A[1], double(B[2]) - A[2]};
                    const double e2[3] = {double(C[0]) - A[0], double(C[1]) - A[1], double(C[2]) - A[2]};
                    const double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
                    const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
                    if (std::abs(det) < 1e-300) continue;
                    const double idet = 1 / det;
                    const double s[3] = {o[0] - A[0], o[1] - A[1], o[2] - A[2]};
                    const double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * idet;
                    if (u < 0 || u > 1) continue;
                    const double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
                    const double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * idet;
                    if (v < 0 || u + v > 1) continue;
                    const double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * idet;
                    if (tt < tmin || tt > best) continue;
                    best = tt;
                    hitTri = t;
                    hitDist = tt;
                    found = true;
                }
            } else if (sp + 2 <= 128) {
                stack[sp++] = n.left;
                stack[sp++] = n.right;
            }
        }
        return found;
    }

private:
    struct Node {
        float min[3], max[3];
        int left = -1, right = -1;
        int start = 0, count = 0;
    };
    const Mesh& M;
    std::vector<Node> nodes;
    std::vector<int> order;
    std::vector<float> cen, lo, hi;

    int build(int start, int count, int depth) {
        const int id = int(nodes.size());
        nodes.push_back({});
        Node node;
        for (int d = 0; d < 3; d++) node.min[d] = std::numeric_limits<float>::max(), node.max[d] = -std::numeric_limits<float>::max();
        float cmin[3] = {1e30f, 1e30f, 1e30f}, cmax[3] = {-1e30f, -1e30f, -1e30f};
        for (int i = start; i < start + count; i++) {
            const size_t t = size_t(order[size_t(i)]);
            for (int d = 0; d < 3; d++) {
                node.min[d] = std::min(node.min[d], lo[3 * t + d]);
                node.max[d] = std::max(node.max[d], hi[3 * t + d]);
                cmin[d] = std::min(cmin[d], cen[3 * t + d]);
                cmax[d] = std::max(cmax[d], cen[3 * t + d]);
            }
        }
        if (count <= 4 || depth > 60) {
            node.start = start;
            node.count = count;
            nodes[size_t(id)] = node;
            return id;
        }
        int axis = 0;
        for (int d = 1; d < 3; d++)
            if (cmax[d] - cmin[d] > cmax[axis] - cmin[axis]) axis = d;
        const float extent = cmax[axis] - cmin[axis];
        int mid = start + count / 2;
        if (extent > 0) {
            constexpr int NB = 16;
            struct Bin {
                float lo[3], hi[3];
                int n;
            } bins[NB];
            for (auto& b : bins) {
                b.n = 0;
                for (int d = 0; d < 3; d++) b.lo[d] = 1e30f, b.hi[d] = -1e30f;
            }
            const float scale = NB / extent * 0.99999f;
            auto binOf = [&](int t) { return std::min(NB - 1, int((cen[3 * size_t(t) + axis] - cmin[axis]) * scale)); };
            for (int i = start; i < start + count; i++) {
                const int t = order[size_t(i)];
                Bin& b = bins[binOf(t)];
                b.n++;
                for (int d = 0; d < 3; d++) {
                    b.lo[d] = std::min(b.lo[d], lo[3 * size_t(t) + d]);
                    b.hi[d] = std::max(b.hi[d], hi[3 * size_t(t) + d]);
                }
            }
            auto surface = [](const float* l, const float* h) {
                const float x = h[0] - l[0], y = h[1] - l[1], z = h[2] - l[2];
                return x < 0 ? 0.f : 2 * (x * y + y * z + z * x);
            };
            float leftA[NB], rightA[NB];
            int leftN[NB], rightN[NB];
            float l3[3] = {1e30f, 1e30f, 1e30f}, h3[3] = {-1e30f, -1e30f, -1e30f};
            int acc = 0;
            for (int b = 0; b < NB; b++) {
                acc += bins[b].n;
                for (int d = 0; d < 3; d++) l3[d] = std::min(l3[d], bins[b].lo[d]), h3[d] = std::max(h3[d], bins[b].hi[d]);
                leftA[b] = surface(l3, h3);
                leftN[b] = acc;
            }
            for (int d = 0; d < 3; d++) l3[d] = 1e30f, h3[d] = -1e30f;
            acc = 0;
            for (int b = NB - 1; b >= 0; b--) {
                acc += bins[b].n;
                for (int d = 0; d < 3; d++) l3[d] = std::min(l3[d], bins[b].lo[d]), h3[d] = std::max(h3[d], bins[b].hi[d]);
                rightA[b] = surface(l3, h3);
                rightN[b] = acc;
            }
            float best = std::numeric_limits<float>::max();
            int split = -1;
            for (int b = 0; b + 1 < NB; b++) {
                if (!leftN[b] || !rightN[b + 1]) continue;
                const float cost = leftA[b] * leftN[b] + rightA[b + 1] * rightN[b + 1];
                if (cost < best) best = cost, split = b;
            }
            if (split >= 0) {
                auto it = std::partition(order.begin() + start, order.begin() + start + count, [&](int t) { return binOf(t) <= split; });
                mid = int(it - order.begin());
            }
        }
        if (mid <= start || mid >= start + count) {
            mid = start + count / 2;
            std::nth_element(order.begin() + start, order.begin() + mid, order.begin() + start + count,
                             [&](int a, int b) { return cen[3 * size_t(a) + axis] < cen[3 * size_t(b) + axis]; });
        }
        node.left = build(start, mid - start, depth + 1);
        node.right = build(mid, start + count - mid, depth + 1);
        nodes[size_t(id)] = node;
        return id;
    }
};

// ------------------------------------------------------------------------------------------------
// walls (voxelize.cpp:130-276)

// material thickness behind triangle t, along the inward normal from just inside its centroid to a
// facing-away triangle, or +inf (no hit within far, or the hit faces the same way)
double wallThicknessAt(const Mesh& M, const Bvh& bvh, int t, double far) {
    const double eps = M.diag * 1e-6;
    double c[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++)
        for (int d = 0; d < 3; d++) c[d] += M.V[3 * M.T[3 * t + k] + d] / 3.0;
    const double n[3] = {M.normal[3 * size_t(t)], M.normal[3 * size_t(t) + 1], M.normal[3 * size_t(t) + 2]};
    const double o[3] = {c[0] - n[0] * eps, c[1] - n[1] * eps, c[2] - n[2] * eps};
    const double dir[3] = {-n[0], -n[1], -n[2]};
    int q = -1;
    double dist = 0;
    if (!bvh.raycast(o, dir, 0, far, q, dist) || q == t) return std::numeric_limits<double>::infinity();
    const float* N = M.normal.data();
    if (N[3 * q] * n[0] + N[3 * q + 1] * n[1] + N[3 * q + 2] * n[2] > -0.3) return std::numeric_limits<double>::infinity();
    return dist + eps;
}

// separating-axis test of a triangle against an axis-aligned cube (centre c, half size hs)
bool triBoxOverlap(const double c[3], double hs, const double a[3], const double b[3], const double d[3]) {
    double v0[3], v1[3], v2[3], e[3][3];
    for (int k = 0; k < 3; k++) {
        v0[k] = a[k] - c[k];
        v1[k] = b[k] - c[k];
        v2[k] = d[k] - c[k];
    }
    for (int k = 0; k < 3; k++) {
        e[0][k] = v1[k] - v0[k];
        e[1][k] = v2[k] - v1[k];
        e[2][k] = v0[k] - v2[k];
    }
    for (auto& ed : e)
        for (int k = 0; k < 3; k++) {
            // edge x unit axis k
            double ax[3];
            if (k == 0) ax[0] = 0, ax[1] = -ed[2], ax[2] = ed[1];
            else if (k == 1) ax[0] = ed[2], ax[1] = 0, ax[2] = -ed[0];
            else ax[0] = -ed[1], ax[1] = ed[0], ax[2] = 0;
            const double p0 = ax[0] * v0[0] + ax[1] * v0[1] + ax[2] * v0[2];
            const double p1 = ax[0] * v1[0] + ax[1] * v1[1] + ax[2] * v1[2];
            const double p2 = ax[0] * v2[0] + ax[1] * v2[1] + ax[2] * v2[2];
            const double r = hs * (std::abs(ax[0]) + std::abs(ax[1]) + std::abs(ax[2]));
            if (std::min({p0, p1, p2}) > r || std::max({p0, p1, p2}) < -r) return false;
        }
    for (int k = 0; k < 3; k++)
        if (std::min({v0[k], v1[k], v2[k]}) > hs || std::max({v0[k], v1[k], v2[k]}) < -hs) return false;
    const double n[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2], e[0][0] * e[1][1] - e[0][1] * e[1][0]};
    const double dist = n[0] * v0[0] + n[1] * v0[1] + n[2] * v0[2];
    return std::abs(dist) <= hs * (std::abs(n[0]) + std::abs(n[1]) + std::abs(n[2]));
}

// 20th percentile of the wall thickness at `samples` area-weighted triangles (LCG seed 7), or +inf
// when fewer than 15 % of them see a wall
double typicalWallThickness(const Mesh& M, const Bvh& bvh, int samples) {
    const double far = M.diag * 0.15;
    std::vector<double> cdf(static_cast<size_t>(M.nT));
    double acc = 0;
    for (int t = 0; t < M.nT; t++) cdf[size_t(t)] = acc += M.area[size_t(t)];
    std::vector<int> picks(static_cast<size_t>(samples));
    uint32_t seed = 7;
    for (int s = 0; s < samples; s++) {
        seed = (seed * 1103515245u + 12345u) & 0x7fffffffu;
        const double r = double(seed) / 0x7fffffff * acc;
        picks[size_t(s)] = int(std::lower_bound(cdf.begin(), cdf.end(), r) - cdf.begin());
        if (picks[size_t(s)] >= M.nT) picks[size_t(s)] = M.nT - 1;
    }
    std::vector<double> thick(static_cast<size_t>(samples));
    parallelChunks(samples, 64, [&](int64_t lo, int64_t hi) {
        for (int64_t s = lo; s < hi; s++) thick[size_t(s)] = wallThicknessAt(M, bvh, picks[size_t(s)], far);
    });
    std::sort(thick.begin(), thick.end());
    int finite = 0;
    for (double t : thick) finite += std::isfinite(t);
    return finite > samples * 0.15 ? thick[size_t(int(samples * 0.2))] : std::numeric_limits<double>::infinity();
}

// per voxel, the volume thin walls (thinner than maxThickness) put into it — each wall triangle's
// area x thickness / 2 shared among the voxels it overlaps, in triangle order — averaged over the
// non-empty voxels of its 3 x 3 x 3 neighbourhood, as a fraction of a voxel, capped at 1
std::vector<float> thinWallDensity(const Mesh& M, const Bvh& bvh, const double origin[3], double h, const int dims[3], double maxThickness) {
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    std::vector<double> thick(static_cast<size_t>(M.nT));
    parallelChunks(M.nT, 256, [&](int64_t lo, int64_t hi) {
        for (int64_t t = lo; t < hi; t++) thick[size_t(t)] = wallThicknessAt(M, bvh, int(t), maxThickness);
    });
    std::vector<float> vol(size_t(nx) * ny * nz, 0.f);
    std::vector<int64_t> hit;
    for (int t = 0; t < M.nT; t++) {
        if (!std::isfinite(thick[size_t(t)])) continue;
        double a[3], b[3], d[3], c[3];
        for (int k = 0; k < 3; k++) {
            a[k] = M.V[3 * M.T[3 * t] + k];
            b[k] = M.V[3 * M.T[3 * t + 1] + k];
            d[k] = M.V[3 * M.T[3 * t + 2] + k];
        }
        int lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            lo[k] = std::max(0, int(std::floor((std::min({a[k], b[k], d[k]}) - origin[k]) / h)));
            hi[k] = std::min(dims[k] - 1, int(std::floor((std::max({a[k], b[k], d[k]}) - origin[k]) / h)));
        }
        hit.clear();
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    c[0] = origin[0] + (i + 0.5) * h;
                    c[1] = origin[1] + (j + 0.5) * h;
                    c[2] = origin[2] + (k + 0.5) * h;
                    if (triBoxOverlap(c, h / 2, a, b, d)) hit.push_back(i + int64_t(nx) * (j + int64_t(ny) * k));
                }
        const double share = M.area[size_t(t)] * thick[size_t(t)] / 2 / std::max<size_t>(1, hit.size());
        for (int64_t e : hit) vol[size_t(e)] += float(share);
    }
    std::vector<float> out(vol.size(), 0.f);
    const double h3 = h * h * h;
    parallelChunks(nz, 1, [&](int64_t klo, int64_t khi) {
        for (int64_t k = klo; k < khi; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    const size_t e = i + size_t(nx) * (j + size_t(ny) * k);
                    if (!vol[e]) continue;
                    double sum = 0;
                    int n = 0;
                    for (int dk = -1; dk <= 1; dk++)
                        for (int dj = -1; dj <= 1; dj++)
                            for (int di = -1; di <= 1; di++) {
                                const int ii = i + di, jj = j + dj;
                                const int64_t kk = k + dk;
                                if (ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) continue;
                                const float v = vol[ii + size_t(nx) * (jj + size_t(ny) * kk)];
                                if (v) sum += v, n++;
                            }
                    out[e] = float(std::min(1.0, sum / (n * h3)));
                }
    });
    return out;
}

// ------------------------------------------------------------------------------------------------
// the structural model's nodes and the assembly of fixtures and loads (structural.cpp:47-209)

struct Model3 {
    double origin[3], h;
    int dims[3], NX, NY, NZ;
    int64_t nNodes;
    std::vector<uint8_t> surface;
    int64_t node(int i, int j, int k) const { return i + int64_t(NX) * (j + int64_t(NY) * k); }
};

// active nodes (a corner of a voxel with density > 0) and surface nodes (active, fewer than 8 solid
// voxels around)
Model3 nodesOf(const double origin[3], double h, const int dims[3], const float* density, uint8_t* activeOut) {
    Model3 S;
    for (int d = 0; d < 3; d++) S.origin[d] = origin[d], S.dims[d] = dims[d];
    S.h = h;
    S.NX = dims[0] + 1, S.NY = dims[1] + 1, S.NZ = dims[2] + 1;
    S.nNodes = int64_t(S.NX) * S.NY * S.NZ;
    std::vector<uint8_t> active(size_t(S.nNodes), 0), full(size_t(S.nNodes), 0);
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                if (!(density[i + size_t(nx) * (j + size_t(ny) * k)] > 0)) continue;
                for (int c = 0; c < 8; c++) {
                    const size_t n = size_t(S.node(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)));
                    active[n] = 1;
                    full[n]++;
                }
            }
    S.surface.assign(size_t(S.nNodes), 0);
    for (size_t n = 0; n < size_t(S.nNodes); n++)
        if (active[n] && full[n] < 8) S.surface[n] = 1;
    if (activeOut) std::copy(active.begin(), active.end(), activeOut);
    return S;
}

int64_t nearestSurfaceNode(const Model3& S, double x, double y, double z, double maxCells = 2.5) {
    const double gx = (x - S.origin[0]) / S.h, gy = (y - S.origin[1]) / S.h, gz = (z - S.origin[2]) / S.h;
    const int r = int(std::ceil(maxCells));
    const int ci = int(std::lround(gx)), cj = int(std::lround(gy)), ck = int(std::lround(gz));
    int64_t best = -1;
    double bestD = maxCells * maxCells;
    for (int k = std::max(0, ck - r); k <= std::min(S.NZ - 1, ck + r); k++)
        for (int j = std::max(0, cj - r); j <= std::min(S.NY - 1, cj + r); j++)
            for (int i = std::max(0, ci - r); i <= std::min(S.NX - 1, ci + r); i++) {
                const int64_t n = S.node(i, j, k);
                if (!S.surface[size_t(n)]) continue;
                const double d = (i - gx) * (i - gx) + (j - gy) * (j - gy) + (k - gz) * (k - gz);
                if (d < bestD) bestD = d, best = n;
            }
    return best;
}

template <class F> void surfaceNodesNear(const Model3& S, double x, double y, double z, double radius, F&& fn) {
    const double gx = (x - S.origin[0]) / S.h, gy = (y - S.origin[1]) / S.h, gz = (z - S.origin[2]) / S.h;
    const double r2 = radius * radius;
    for (int k = std::max(0, int(std::ceil(gz - radius))); k <= std::min(S.NZ - 1, int(std::floor(gz + radius))); k++)
        for (int j = std::max(0, int(std::ceil(gy - radius))); j <= std::min(S.NY - 1, int(std::floor(gy + radius))); j++)
            for (int i = std::max(0, int(std::ceil(gx - radius))); i <= std::min(S.NX - 1, int(std::floor(gx + radius))); i++) {
                const int64_t n = S.node(i, j, k);
                if (S.surface[size_t(n)] && (i - gx) * (i - gx) + (j - gy) * (j - gy) + (k - gz) * (k - gz) <= r2) fn(n);
            }
}

struct Samples {
    std::vector<float> points, weights;
    std::vector<int32_t> tris;
};

// about one point per spacing^2 on each listed triangle (m x m sub-triangle centroids), each with its
// area share, optionally clipped to a sphere (mesh.cpp:540-568)
void sampleTriangles(const Mesh& M, const psx_patch& p, double spacing, Samples& s) {
    const double r2 = p.has_clip ? p.clip_radius * p.clip_radius : 0;
    for (int64_t q = 0; q < p.n_tris; q++) {
        const int32_t t = p.tris[q];
        if (t < 0 || t >= M.nT) throw std::invalid_argument("patch triangle out of range");
        const size_t a = 3 * size_t(M.T[3 * t]), b = 3 * size_t(M.T[3 * t + 1]), c = 3 * size_t(M.T[3 * t + 2]);
        const float area = M.area[size_t(t)];
        const int m = int(std::max(1.0, std::min(40.0, std::ceil(std::sqrt(area / (spacing * spacing))))));
        const double w = area / double(m * m);
        auto push = [&](double su, double u) {
            const double r = 1 - su - u;
            const double x = r * M.V[a] + su * M.V[b] + u * M.V[c];
            const double y = r * M.V[a + 1] + su * M.V[b + 1] + u * M.V[c + 1];
            const double z = r * M.V[a + 2] + su * M.V[b + 2] + u * M.V[c + 2];
            if (p.has_clip) {
                const double dx = x - p.clip_center[0], dy = y - p.clip_center[1], dz = z - p.clip_center[2];
                if (dx * dx + dy * dy + dz * dz > r2) return;
            }
            s.points.insert(s.points.end(), {float(x), float(y), float(z)});
            s.weights.push_back(float(w));
            s.tris.push_back(t);
        };
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m - i; j++) {
                push((i + 1.0 / 3) / m, (j + 1.0 / 3) / m);
                if (i + j < m - 1) push((i + 2.0 / 3) / m, (j + 2.0 / 3) / m);
            }
    }
}

// the union of a selection's patches: every (triangle, point) counted once (structural.cpp:89-104)
Samples samplePatches(const Mesh& M, const Model3& S, const psx_patch* patches, int n) {
    Samples out;
    std::set<std::tuple<int32_t, float, float, float>> seen;
    for (int q = 0; q < n; q++) {
        Samples s;
        sampleTriangles(M, patches[q], S.h * 0.4, s);
        for (size_t i = 0; i < s.weights.size(); i++) {
            const float x = s.points[3 * i], y = s.points[3 * i + 1], z = s.points[3 * i + 2];
            if (!seen.emplace(s.tris[i], x, y, z).second) continue;
            out.points.insert(out.points.end(), {x, y, z});
            out.weights.push_back(s.weights[i]);
            out.tris.push_back(s.tris[i]);
        }
    }
    return out;
}

void validateMaterial(const psx_material& m) {
    const std::pair<double, const char*> positive[] = {{m.E, "Young's modulus"}, {m.yield, "Yield strength"}, {m.uts, "Tensile strength"}, {m.density, "Density"}};
    for (auto& [v, name] : positive)
        if (!std::isfinite(v) || v <= 0) throw std::invalid_argument(std::string(name) + " must be a positive finite number.");
    if (!std::isfinite(m.nu) || m.nu <= -1 || m.nu >= 0.5) throw std::invalid_argument("Poisson's ratio must be greater than -1 and less than 0.5.");
    if (!m.brittle && m.uts < m.yield) throw std::invalid_argument("Tensile strength must be at least the yield strength for a ductile material.");
}

}  // namespace

}  // namespace psx

extern "C" {

int psx_wall_thickness(const psx_mesh* mesh, int samples, double* out) {
    return psx::abi::guarded([&] {
        if (!out || samples < 1) throw std::invalid_argument("bad arguments");
        const psx::Mesh M = psx::makeMesh(mesh);
        if (M.nT == 0) throw std::invalid_argument("empty mesh");
        const psx::Bvh bvh(M);
        *out = psx::typicalWallThickness(M, bvh, samples);
        return PSX_OK;
    });
}

int psx_thin_wall_density(const psx_mesh* mesh, const double origin[3], double h, const int dims[3], double max_thickness, float* out) {
    return psx::abi::guarded([&] {
        if (!origin || !dims || !out || !(h > 0) || dims[0] < 1 || dims[1] < 1 || dims[2] < 1) throw std::invalid_argument("bad arguments");
        const psx::Mesh M = psx::makeMesh(mesh);
        const psx::Bvh bvh(M);
        const auto v = psx::thinWallDensity(M, bvh, origin, h, dims, max_thickness);
        std::copy(v.begin(), v.end(), out);
        return PSX_OK;
    });
}

int psx_structural_model(const psx_mesh* mesh, const double origin[3], double h, const int dims[3], float* density, uint8_t* active_node,
                         uint8_t* surface_node, psx_structural_info* info) {
    return psx::abi::guarded([&] {
        if (!origin || !dims || !density || !(h > 0) || dims[0] < 1 || dims[1] < 1 || dims[2] < 1) throw std::invalid_argument("bad arguments");
        const psx::Mesh M = psx::makeMesh(mesh);
        if (M.nT == 0) throw std::invalid_argument("empty mesh");
        const psx::Bvh bvh(M);
        const int nx = dims[0], ny = dims[1], nz = dims[2];
        // sampling fine enough for thin walls, within the sample-grid budget
        const double wall = psx::typicalWallThickness(M, bvh, 1500);
        int sub = 5;
        if (std::isfinite(wall)) sub = std::max(sub, std::min(6, int(std::ceil(3 * h / wall))));
        while (sub > 2 && double(nx) * ny * nz * sub * sub * sub > 48e6) sub--;
        const size_t nVox = size_t(nx) * ny * nz;
        std::vector<float> frac(nVox);
        const int vrc = psx_voxelize(M.V, M.nV, M.T, M.nT, origin, h, dims, sub, frac.data());
        if (vrc == PSX_ERR_ARGUMENT) throw std::invalid_argument(psx_last_error());
        if (vrc == PSX_ERR_NO_DEVICE) throw psx::NoDevice(psx_last_error());
        if (vrc != PSX_OK) throw psx::DeviceError(psx_last_error());
        // walls under ~1.5 voxels thick get a connected layer with their true cross-section
        std::vector<float> shell;
        if (wall < 2.5 * h) shell = psx::thinWallDensity(M, bvh, origin, h, dims, 1.5 * h);
        int voxels = 0, thin = 0;
        for (size_t e = 0; e < nVox; e++) {
            const float f = frac[e], w = shell.empty() ? 0.f : shell[e];
            float rho = 0;
            if (w > 0) {
                rho = f >= 0.9f ? f : std::max(w, 0.02f);
                if (f < 0.9f) thin++;
            } else if (f >= 0.3f) rho = f;
            density[e] = rho > 0 ? std::min(1.f, rho) : 0.f;
            if (rho > 0) voxels++;
        }
        const psx::Model3 S = psx::nodesOf(origin, h, dims, density, active_node);
        if (surface_node) std::copy(S.surface.begin(), S.surface.end(), surface_node);
        if (info) {
            info->wall_thickness = wall;
            info->sub = sub;
            info->voxel_count = voxels;
            info->thin_voxels = thin;
            info->used_shell = shell.empty() ? 0 : 1;
        }
        return PSX_OK;
    });
}

int psx_assemble(const psx_mesh* mesh, const double origin[3], double h, const int dims[3], const float* density, const psx_fixture* fixtures,
                 int n_fixtures, const psx_load* loads, int n_loads, int gravity, const psx_material* material, double to_meters, uint8_t* bc,
                 double* f, psx_assembly_info* info, char* warnings, size_t warnings_len) {
    return psx::abi::guarded([&] {
        if (!origin || !dims || !density || !bc || !f || !material || !(h > 0) || n_fixtures < 0 || n_loads < 0 || (n_fixtures && !fixtures) ||
            (n_loads && !loads))
            throw std::invalid_argument("bad arguments");
        psx::validateMaterial(*material);
        if (!std::isfinite(to_meters) || to_meters <= 0) throw std::invalid_argument("Invalid model units.");
        const psx::Mesh M = psx::makeMesh(mesh);
        const psx::Model3 S = psx::nodesOf(origin, h, dims, density, nullptr);
        const size_t nDof = 3 * size_t(S.nNodes);
        std::fill(bc, bc + nDof, uint8_t(0));
        std::fill(f, f + nDof, 0.0);
        int fixedNodes = 0;
        double total[3] = {0, 0, 0};
        std::vector<std::string> warn;
        auto hold = [&](int64_t n) {
            if (!bc[3 * n]) fixedNodes++;
            bc[3 * n] = bc[3 * n + 1] = bc[3 * n + 2] = 1;
        };
        for (int q = 0; q < n_fixtures; q++) {
            const psx_fixture& fx = fixtures[q];
            if (fx.n_patches && !fx.patches) throw std::invalid_argument("bad fixture");
            const psx::Samples s = psx::samplePatches(M, S, fx.patches, fx.n_patches);
            int hit = 0;
            for (size_t i = 0; i < s.weights.size(); i++) {
                const double x = s.points[3 * i], y = s.points[3 * i + 1], z = s.points[3 * i + 2];
                const int64_t n = psx::nearestSurfaceNode(S, x, y, z);
                if (n < 0) continue;
                hit++;
                hold(n);
                // the grid can overhang the selected face by part of a voxel: hold every surface node
                // within 3/4 voxel of a sample
                psx::surfaceNodesNear(S, x, y, z, 0.75, hold);
            }
            if (!hit) warn.push_back(std::string(fx.name ? fx.name : "") + ": no solid voxels under the selection - try a finer mesh.");
        }
        for (int q = 0; q < n_loads; q++) {
            const psx_load& ld = loads[q];
            const std::string name = ld.name ? ld.name : "";
            if (ld.type == PSX_LOAD_WIND) {
                if (!ld.forces || ld.n_forces != int64_t(M.nT) * 3) throw std::invalid_argument(name + ": invalid airflow forces.");
                int lost = 0;
                for (int t = 0; t < M.nT; t++) {
                    const double fx = ld.forces[3 * t], fy = ld.forces[3 * t + 1], fz = ld.forces[3 * t + 2];
                    if (fx == 0 && fy == 0 && fz == 0) continue;
                    double c[3] = {0, 0, 0};
                    for (int k = 0; k < 3; k++)
                        for (int d = 0; d < 3; d++) c[d] += M.V[3 * M.T[3 * t + k] + d] / 3.0;
                    const int64_t n = psx::nearestSurfaceNode(S, c[0], c[1], c[2], 3.5);
                    if (n < 0) {
                        lost++;
                        continue;
                    }
                    f[3 * n] += fx;
                    f[3 * n + 1] += fy;
                    f[3 * n + 2] += fz;
                    total[0] += fx;
                    total[1] += fy;
                    total[2] += fz;
                }
                if (lost) warn.push_back(name + ": " + std::to_string(lost) + " surface triangles had no nearby voxel.");
                continue;
            }
            if (ld.type != PSX_LOAD_FORCE && ld.type != PSX_LOAD_PRESSURE) throw std::invalid_argument(name + ": unknown load type.");
            if (!std::isfinite(ld.magnitude) || ld.magnitude < 0) throw std::invalid_argument(name + ": magnitude must be a finite, nonnegative number.");
            double direction[3] = {0, 0, 0};
            if (ld.type == PSX_LOAD_FORCE) {
                const double len = std::sqrt(ld.dir[0] * ld.dir[0] + ld.dir[1] * ld.dir[1] + ld.dir[2] * ld.dir[2]);
                if (!(len > 0) || !std::isfinite(len)) throw std::invalid_argument(name + ": choose a nonzero force direction.");
                for (int d = 0; d < 3; d++) direction[d] = ld.dir[d] / len;
            }
            if (ld.n_patches && !ld.patches) throw std::invalid_argument("bad load");
            const psx::Samples s = psx::samplePatches(M, S, ld.patches, ld.n_patches);
            // per node: [sample weight, pressure force direction x area (3)], in node order
            std::map<int64_t, std::array<double, 4>> nodes;
            double W = 0;
            for (size_t i = 0; i < s.weights.size(); i++) {
                const int64_t n = psx::nearestSurfaceNode(S, s.points[3 * i], s.points[3 * i + 1], s.points[3 * i + 2]);
                if (n < 0) continue;
                const double w = s.weights[i];
                auto& e = nodes[n];
                e[0] += w;
                if (ld.type == PSX_LOAD_PRESSURE) {
                    const int t = s.tris[i];
                    for (int d = 0; d < 3; d++) e[1 + d] -= M.normal[3 * size_t(t) + d] * w;
                }
                W += w;
            }
            if (nodes.empty()) {
                warn.push_back(name + ": no solid voxels under the selection - try a finer mesh.");
                continue;
            }
            for (const auto& [n, e] : nodes) {
                double v[3];
                if (ld.type == PSX_LOAD_PRESSURE) {
                    const double k = ld.magnitude * 1e6 * to_meters * to_meters;  // MPa x model-unit^2 -> N
                    for (int d = 0; d < 3; d++) v[d] = e[1 + d] * k;
                } else {
                    const double s2 = ld.magnitude * e[0] / W;
                    for (int d = 0; d < 3; d++) v[d] = direction[d] * s2;
                }
                for (int d = 0; d < 3; d++) {
                    f[3 * n + d] += v[d];
                    total[d] += v[d];
                }
            }
        }
        if (gravity) {
            const int nx = dims[0], ny = dims[1];
            const double hm = h * to_meters;
            const size_t nVox = size_t(dims[0]) * dims[1] * dims[2];
            double mass = 0;
            for (size_t e = 0; e < nVox; e++) {
                const double d = density[e];
                if (!(d > 0)) continue;
                const double m = material->density * d * hm * hm * hm;
                mass += m;
                const int i = int(e % size_t(nx)), j = int((e / size_t(nx)) % size_t(ny)), k = int(e / (size_t(nx) * ny));
                for (int c = 0; c < 8; c++) f[3 * S.node(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)) + 1] -= m * 9.81 / 8;
            }
            total[1] -= mass * 9.81;
        }
        if (fixedNodes > 0 && fixedNodes < 3) warn.push_back("Very small fixture - the part may pivot. Select a larger area.");
        if (info) {
            info->fixed_nodes = fixedNodes;
            for (int d = 0; d < 3; d++) info->total[d] = total[d];
            info->n_warnings = int(warn.size());
        }
        if (warnings && warnings_len) {
            std::string all;
            for (const auto& w : warn) all += w + "\n";
            const size_t n = std::min(all.size(), warnings_len - 1);
            std::memcpy(warnings, all.data(), n);
            warnings[n] = 0;
        }
        return PSX_OK;
    });
}

}  // extern "C"
