// Host-side hierarchy build: element matrix, levels, exact Galerkin coarsening and the coarsest
// level's dense inverse. Runs once per model on the CPU cores; everything per-iteration runs on the
// GPU (solver.cpp, kernels.hip).
#include "psx_core.hpp"

#include <algorithm>
#include <numeric>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>

namespace psx {

// host threads of one hierarchy build: PSX_HOST_THREADS, else OMP_NUM_THREADS (the process's CPU
// share where a scheduler sets it), else the hardware's, at most 64
int hostThreads() {
    static const int n = [] {
        for (const char* name : {"PSX_HOST_THREADS", "OMP_NUM_THREADS"})
            if (const char* e = std::getenv(name)) {
                const int v = std::atoi(e);
                if (v > 0) return std::min(v, 64);
            }
        return int(std::max(1u, std::min(64u, std::thread::hardware_concurrency())));
    }();
    return n;
}

// fn(lo, hi) over [0, n) in contiguous chunks on the host threads
template <class F> static void parallelRange(i64 n, F&& fn, i64 minChunk = 256) {
    const int T = int(std::min<i64>(hostThreads(), std::max<i64>(1, n / minChunk)));
    if (T <= 1) {
        fn(i64(0), n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T - 1);
    for (int t = 1; t < T; t++) th.emplace_back([&, t] { fn(n * t / T, n * (t + 1) / T); });
    fn(0, n / T);
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------------------------
// hex8 element (native/src/fea/hex8.cpp:8-155)

// isotropic D for E = 1 (Voigt xx yy zz xy yz zx, engineering shear)
static void elasticityD(double nu, double D[6][6]) {
    if (!std::isfinite(nu) || nu <= -1 || nu >= 0.5) throw std::invalid_argument("Poisson's ratio must be greater than -1 and less than 0.5.");
    const double lam = nu / ((1 + nu) * (1 - 2 * nu)), mu = 1 / (2 * (1 + nu));
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) D[i][j] = 0;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) D[i][j] = lam;
        D[i][i] = lam + 2 * mu;
        D[i + 3][i + 3] = mu;
    }
}

// strain rows of a displacement triple with gradient g: exx eyy ezz gxy gyz gzx
static void strainOf(double (*B)[24], int col, const double g[3]) {
    B[0][col] = g[0];
    B[1][col + 1] = g[1];
    B[2][col + 2] = g[2];
    B[3][col] = g[1];
    B[3][col + 1] = g[0];
    B[4][col + 1] = g[2];
    B[4][col + 2] = g[1];
    B[5][col] = g[2];
    B[5][col + 2] = g[0];
}

// B (6 x 24) of the trilinear shape functions and G (6 x 9) of the three bubble modes 1 - xi^2,
// 1 - eta^2, 1 - zeta^2 at natural coordinates p in [-1, 1]^3 (unit cube: d xi / dx = 2)
static void shapeGradients(const double p[3], double B[6][24], double G[6][24]) {
    for (int r = 0; r < 6; r++)
        for (int c = 0; c < 24; c++) B[r][c] = G[r][c] = 0;
    for (int a = 0; a < 8; a++) {
        const double s[3] = {2.0 * CX[a] - 1, 2.0 * CY[a] - 1, 2.0 * CZ[a] - 1};
        double g[3];
        for (int ax = 0; ax < 3; ax++) {
            double v = 2 * s[ax] / 8;
            for (int o = 0; o < 3; o++)
                if (o != ax) v *= 1 + s[o] * p[o];
            g[ax] = v;
        }
        strainOf(B, 3 * a, g);
    }
    for (int m = 0; m < 3; m++) {
        double g[3] = {0, 0, 0};
        g[m] = -4 * p[m];
        strainOf(G, 3 * m, g);
    }
}

Element hexElement(double nu) {
    double D[6][6];
    elasticityD(nu, D);
    double Kuu[24][24] = {}, Kua[24][9] = {}, Kaa[9][9] = {};
    const double gp = 1 / std::sqrt(3.0);
    double B[6][24], G[6][24], DB[6][24], DG[6][24];
    for (int q = 0; q < 8; q++) {
        const double p[3] = {(2.0 * CX[q] - 1) * gp, (2.0 * CY[q] - 1) * gp, (2.0 * CZ[q] - 1) * gp};
        shapeGradients(p, B, G);
        for (int r = 0; r < 6; r++)
            for (int c = 0; c < 24; c++) {
                double b = 0, g = 0;
                for (int t = 0; t < 6; t++) {
                    b += D[r][t] * B[t][c];
                    g += D[r][t] * G[t][c];
                }
                DB[r][c] = b;
                DG[r][c] = g;
            }
        // Gauss weight 1, det J = 1/8
        for (int i = 0; i < 24; i++) {
            for (int j = 0; j < 24; j++) {
                double s = 0;
                for (int t = 0; t < 6; t++) s += B[t][i] * DB[t][j];
                Kuu[i][j] += s / 8;
            }
            for (int j = 0; j < 9; j++) {
                double s = 0;
                for (int t = 0; t < 6; t++) s += B[t][i] * DG[t][j];
                Kua[i][j] += s / 8;
            }
        }
        for (int i = 0; i < 9; i++)
            for (int j = 0; j < 9; j++) {
                double s = 0;
                for (int t = 0; t < 6; t++) s += G[t][i] * DG[t][j];
                Kaa[i][j] += s / 8;
            }
    }
    // static condensation of the bubble amplitudes: alpha = A u with A = -Kaa^-1 Kau (Cholesky of Kaa)
    double Lc[9][9] = {};
    for (int j = 0; j < 9; j++) {
        double s = Kaa[j][j];
        for (int k = 0; k < j; k++) s -= Lc[j][k] * Lc[j][k];
        Lc[j][j] = std::sqrt(s);
        for (int i = j + 1; i < 9; i++) {
            double t = Kaa[i][j];
            for (int k = 0; k < j; k++) t -= Lc[i][k] * Lc[j][k];
            Lc[i][j] = t / Lc[j][j];
        }
    }
    double A[9][24];
    for (int c = 0; c < 24; c++) {
        double y[9];
        for (int i = 0; i < 9; i++) {
            double t = -Kua[c][i];
            for (int k = 0; k < i; k++) t -= Lc[i][k] * y[k];
            y[i] = t / Lc[i][i];
        }
        for (int i = 8; i >= 0; i--) {
            double t = y[i];
            for (int k = i + 1; k < 9; k++) t -= Lc[k][i] * y[k];
            y[i] = t / Lc[i][i];
        }
        for (int i = 0; i < 9; i++) A[i][c] = y[i];
    }
    Element el;
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) {
            double s = Kuu[i][j];
            for (int t = 0; t < 9; t++) s += Kua[i][t] * A[t][j];
            el.K[i * 24 + j] = s;
        }
    for (int i = 0; i < 24; i++)
        for (int j = i + 1; j < 24; j++) el.K[i * 24 + j] = el.K[j * 24 + i] = 0.5 * (el.K[i * 24 + j] + el.K[j * 24 + i]);
    // corner stresses D (B + G A) at the element corners
    for (int a = 0; a < 8; a++) {
        const double p[3] = {2.0 * CX[a] - 1, 2.0 * CY[a] - 1, 2.0 * CZ[a] - 1};
        shapeGradients(p, B, G);
        double Be[6][24];
        for (int r = 0; r < 6; r++)
            for (int c = 0; c < 24; c++) {
                double s = B[r][c];
                for (int t = 0; t < 9; t++) s += G[r][t] * A[t][c];
                Be[r][c] = s;
            }
        for (int r = 0; r < 6; r++)
            for (int c = 0; c < 24; c++) {
                double s = 0;
                for (int t = 0; t < 6; t++) s += D[r][t] * Be[t][c];
                el.corner[a][r * 24 + c] = s;
            }
    }
    return el;
}

void stencil(const double* K, double* S) {
    std::fill(S, S + 243, 0.0);
    for (int v = 0; v < 8; v++) {
        const int di = v & 1, dj = (v >> 1) & 1, dk = v >> 2;
        const int a = corner(1 - di, 1 - dj, 1 - dk);
        for (int b = 0; b < 8; b++) {
            const int m = (di + CX[b]) + 3 * (dj + CY[b]) + 9 * (dk + CZ[b]);
            for (int d = 0; d < 3; d++)
                for (int c = 0; c < 3; c++) S[m * 9 + d * 3 + c] += K[(3 * a + d) * 24 + 3 * b + c];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// cube symmetry of the element and its stencil

int elementClass(int a, int d, int b, int c) {
    // reflect corner a onto corner 0 (flips the sign of each DOF component along a reflected axis)
    const int ca[3] = {CX[a], CY[a], CZ[a]};
    int r[3] = {CX[a] ^ CX[b], CY[a] ^ CY[b], CZ[a] ^ CZ[b]};
    const bool neg = ((ca[d] + ca[c]) & 1) != 0;
    // permute axis d onto axis 0 (row d of corner 0 becomes row 0)
    std::swap(r[0], r[d]);
    int c2 = c == 0 ? d : (c == d ? 0 : c);
    // y <-> z mirror to a canonical form
    if (c2 == 2) {
        std::swap(r[1], r[2]);
        c2 = 1;
    } else if (c2 == 0 && r[1] < r[2]) std::swap(r[1], r[2]);
    const int idx = c2 == 0 ? r[0] * 3 + r[1] + r[2] : 6 + r[0] + 2 * r[1] + 4 * r[2];
    return idx | (neg ? 0x100 : 0);
}

int stencilClass(int m, int d, int c) {
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

double symmetryClasses(const double* K, double* out) {
    double sum[22] = {}, lo[22], hi[22];
    int cnt[22] = {};
    for (int q = 0; q < 22; q++) lo[q] = INFINITY, hi[q] = -INFINITY;
    auto add = [&](int cls, double v) {
        const int q = cls & 0xff;
        const double s = (cls & 0x100) ? -v : v;
        sum[q] += s;
        cnt[q]++;
        lo[q] = std::min(lo[q], s);
        hi[q] = std::max(hi[q], s);
    };
    double kmax = 0;
    for (int a = 0; a < 8; a++)
        for (int d = 0; d < 3; d++)
            for (int b = 0; b < 8; b++)
                for (int c = 0; c < 3; c++) {
                    const double v = K[(3 * a + d) * 24 + 3 * b + c];
                    kmax = std::max(kmax, std::abs(v));
                    add(elementClass(a, d, b, c), v);
                }
    double S[243];
    stencil(K, S);
    double dev = 0;
    for (int m = 0; m < 27; m++)
        for (int d = 0; d < 3; d++)
            for (int c = 0; c < 3; c++) {
                const int cls = stencilClass(m, d, c);
                if (cls < 0) dev = std::max(dev, std::abs(S[m * 9 + d * 3 + c]));
                else add(cls, S[m * 9 + d * 3 + c]);
            }
    for (int q = 0; q < 22; q++) {
        out[q] = cnt[q] ? sum[q] / cnt[q] : 0;
        if (cnt[q]) dev = std::max(dev, hi[q] - lo[q]);
    }
    return kmax > 0 ? dev / kmax : 0;
}

// ---------------------------------------------------------------------------------------------
// levels

static Level emptyLevel(int nx, int ny, int nz) {
    Level L;
    L.nx = nx; L.ny = ny; L.nz = nz;
    L.NX = nx + 1; L.NY = ny + 1; L.NZ = nz + 1;
    L.nNodes = i64(L.NX) * L.NY * L.NZ;
    L.nDof = 3 * L.nNodes;
    L.nVox = i64(nx) * ny * nz;
    return L;
}

// diagonal, held DOFs, stencil scales and device words of a level whose elements are set
// (voxel_fea.cpp:208-266)
static void finishLevel(Level& L) {
    const i64 nE = i64(L.elemVox.size());
    L.nElem = nE;
    L.active.assign(L.nNodes, 0);
    L.fixed.assign(L.nDof, 0);
    L.invDiag.assign(L.nDof, 0.0);
    // per node, gathered over its voxels in increasing voxel (= element) order: the same sums as
    // the reference's scatter over elements
    std::vector<i64> freeRows(i64(L.NZ) * L.NY, 0);
    parallelRange(i64(L.NZ) * L.NY, [&](i64 lo, i64 hi) {
        for (i64 row = lo; row < hi; row++) {
            const int j = int(row % L.NY), k = int(row / L.NY);
            i64 nfree = 0;
            for (int i = 0; i < L.NX; i++) {
                const i64 n = L.node(i, j, k);
                double diag[3] = {0, 0, 0};
                bool act = false;
                for (int dk = 0; dk < 2; dk++)
                    for (int dj = 0; dj < 2; dj++)
                        for (int di = 0; di < 2; di++) {
                            const int vi = i - 1 + di, vj = j - 1 + dj, vk = k - 1 + dk;
                            if (vi < 0 || vj < 0 || vk < 0 || vi >= L.nx || vj >= L.ny || vk >= L.nz) continue;
                            const int e = L.emap[vi + i64(L.nx) * (vj + i64(L.ny) * vk)];
                            if (e < 0) continue;
                            act = true;
                            const int a = corner(1 - di, 1 - dj, 1 - dk);
                            const float* Ko = L.own[e] >= 0 ? L.K.data() + i64(L.own[e]) * 576 : nullptr;
                            for (int d = 0; d < 3; d++) {
                                const int r = 3 * a + d;
                                diag[d] += Ko ? double(Ko[r * 25]) : L.rho[e] * L.Kb[r * 25];
                            }
                        }
                L.active[n] = act ? 1 : 0;
                for (int d = 0; d < 3; d++) {
                    const i64 q = 3 * n + d;
                    const double dg = L.diagAdd.empty() ? diag[d] : diag[d] + L.diagAdd[q];
                    if (!act || L.bc[q] || !(dg > 0)) L.fixed[q] = 1;
                    else {
                        L.invDiag[q] = 1 / dg;
                        nfree++;
                    }
                }
            }
            freeRows[row] = nfree;
        }
    }, 16);
    L.freeDof = 0;
    for (i64 c : freeRows) L.freeDof += c;
    L.words.assign(L.nVox, 0);
    std::vector<float> vs(L.nVox, 0.f);
    parallelRange(nE, [&](i64 lo, i64 hi) {
        for (i64 e = lo; e < hi; e++) {
            const int v = L.elemVox[e];
            if (L.own[e] >= 0) {
                L.words[v] = int32_t(0x80000000u | uint32_t(L.own[e]));
                vs[v] = -1.f;
            } else {
                const float s = float(L.rho[e]);
                std::memcpy(&L.words[v], &s, 4);
                vs[v] = s;
            }
        }
    });
    L.nodeScale.assign(L.nNodes, 0.f);
    parallelRange(std::max(0, L.nz - 1), [&](i64 lo, i64 hi) {
        for (int k = int(lo) + 1; k < int(hi) + 1; k++)
            for (int j = 1; j < L.ny; j++)
                for (int i = 1; i < L.nx; i++) {
                    const i64 v0 = (i - 1) + i64(L.nx) * ((j - 1) + i64(L.ny) * (k - 1));
                    const float s0 = vs[v0];
                    if (!(s0 > 0)) continue;
                    bool same = true;
                    for (int v = 1; v < 8 && same; v++) same = vs[v0 + (v & 1) + i64(L.nx) * (((v >> 1) & 1) + i64(L.ny) * (v >> 2))] == s0;
                    if (same) L.nodeScale[L.node(i, j, k)] = s0;
                }
    }, 1);
    // per node: the voxels (v = di + 2 dj + 4 dk at (i - 1 + di, j - 1 + dj, k - 1 + dk)) whose
    // scale differs from the node's reference voxel, the lowest-index voxel of the most common
    // scale (outside voxels count as 0, own-matrix voxels as distinct)
    L.diffMask.assign(L.nNodes, 0);
    parallelRange(i64(L.NZ) * L.NY, [&](i64 lo, i64 hi) {
        for (i64 row = lo; row < hi; row++) {
            const int j = int(row % L.NY), k = int(row / L.NY);
            for (int i = 0; i < L.NX; i++) {
                float sv[8];
                for (int v = 0; v < 8; v++) {
                    const int vi = i - 1 + (v & 1), vj = j - 1 + ((v >> 1) & 1), vk = k - 1 + (v >> 2);
                    const bool in = vi >= 0 && vj >= 0 && vk >= 0 && vi < L.nx && vj < L.ny && vk < L.nz;
                    sv[v] = in ? vs[vi + i64(L.nx) * (vj + i64(L.ny) * vk)] : 0.f;
                }
                int ref = 0, best = 0;
                for (int v = 0; v < 8; v++) {
                    int c = 0;
                    for (int u = 0; u < 8; u++) c += sv[u] == sv[v] && sv[v] >= 0;
                    if (c > best) best = c, ref = v;
                }
                uint8_t m = 0;
                for (int v = 0; v < 8; v++)
                    if (!(sv[v] == sv[ref] && sv[v] >= 0)) m |= uint8_t(1u << v);
                L.diffMask[L.node(i, j, k)] = m;
            }
        }
    }, 16);
}

// interpolation weight of coarse corner b at fine corner a of child c (trilinear)
double childWeight(int c, int a, int b) {
    const int cb[3] = {c & 1, (c >> 1) & 1, c >> 2};
    const int ab[3] = {CX[a], CY[a], CZ[a]}, bb[3] = {CX[b], CY[b], CZ[b]};
    double w = 1;
    for (int ax = 0; ax < 3; ax++) {
        const double s = (cb[ax] + ab[ax]) / 2.0;
        w *= bb[ax] ? s : 1 - s;
    }
    return w;
}

// out += P_c^T K P_c for child position c (P_c: 24 fine DOFs of the child <- 24 coarse DOFs)
static void galerkinChild(double* out, const double* K, int c) {
    double W[8][8];
    for (int a = 0; a < 8; a++)
        for (int b = 0; b < 8; b++) W[a][b] = childWeight(c, a, b);
    double T[24][24];  // K P
    for (int r = 0; r < 24; r++)
        for (int b = 0; b < 8; b++)
            for (int d = 0; d < 3; d++) {
                double s = 0;
                for (int a = 0; a < 8; a++)
                    if (W[a][b] != 0) s += K[r * 24 + 3 * a + d] * W[a][b];
                T[r][3 * b + d] = s;
            }
    for (int b = 0; b < 8; b++)
        for (int d = 0; d < 3; d++)
            for (int col = 0; col < 24; col++) {
                double s = 0;
                for (int a = 0; a < 8; a++)
                    if (W[a][b] != 0) s += W[a][b] * T[3 * a + d][col];
                out[(3 * b + d) * 24 + col] += s;
            }
}

void childProducts(const double* Kb, double* M, double* KbCoarse) {
    std::fill(M, M + 8 * 576, 0.0);
    for (int c = 0; c < 8; c++) galerkinChild(M + c * 576, Kb, c);
    for (int u = 0; u < 576; u++) {
        double s = 0;
        for (int c = 0; c < 8; c++) s += M[c * 576 + u];
        KbCoarse[u] = s;
    }
}

void childTable(const double* M, double* tab) {
    // corner-0 rows of M_B for every subset B, columns of corners 1..7 (the rows sum to zero over
    // the corners, so corner 0's columns are not needed): [256][3][21], padded to 64
    std::fill(tab, tab + 256 * 64, 0.0);
    for (int B = 1; B < 256; B++)
        for (int c = 0; c < 8; c++) {
            if (!((B >> c) & 1)) continue;
            for (int d = 0; d < 3; d++)
                for (int u = 0; u < 21; u++) tab[B * 64 + d * 21 + u] += M[size_t(c) * 576 + size_t(d) * 24 + 3 + u];
        }
}

// fine level F -> coarse level (voxel_fea.cpp:268-386): children of each coarse voxel, exact
// Galerkin element matrices, held coarse DOFs and the lumped diagonal shift
static Level coarsen(const Level& F) {
    Level C = emptyLevel((F.nx + 1) >> 1, (F.ny + 1) >> 1, (F.nz + 1) >> 1);
    const i64 nF = i64(F.elemVox.size());
    std::vector<int32_t> count(C.nVox, 0);
    std::vector<int32_t> parent(nF);
    std::vector<uint8_t> pos(nF);
    for (i64 q = 0; q < nF; q++) {
        const int v = F.elemVox[q];
        const int i = v % F.nx, j = (v / F.nx) % F.ny, k = v / (F.nx * F.ny);
        parent[q] = (i >> 1) + C.nx * ((j >> 1) + C.ny * (k >> 1));
        pos[q] = uint8_t((i & 1) | ((j & 1) << 1) | ((k & 1) << 2));
        count[parent[q]]++;
    }
    C.emap.assign(C.nVox, -1);
    for (i64 v = 0; v < C.nVox; v++)
        if (count[v]) {
            C.emap[v] = int32_t(C.elemVox.size());
            C.elemVox.push_back(int32_t(v));
        }
    const i64 nE = i64(C.elemVox.size());
    std::vector<int32_t> start(nE + 1, 0), kids(nF);
    for (i64 q = 0; q < nF; q++) start[C.emap[parent[q]] + 1]++;
    for (i64 e = 0; e < nE; e++) start[e + 1] += start[e];
    {
        std::vector<int32_t> fill(start.begin(), start.end() - 1);
        for (i64 q = 0; q < nF; q++) kids[fill[C.emap[parent[q]]]++] = int32_t(q);
    }
    // the fine uniform element's Galerkin product per child position; their sum is the coarse Kb
    std::vector<double> M(8 * 576);
    childProducts(F.Kb, M.data(), C.Kb);
    C.rho.assign(nE, 0.0);
    C.own.assign(nE, -1);
    int nOwn = 0;
    for (i64 e = 0; e < nE; e++) {
        const int b = start[e], t = start[e + 1];
        bool uniform = t - b == 8 && F.own[kids[b]] < 0;
        const double s = uniform ? F.rho[kids[b]] : 0;
        for (int q = b; q < t && uniform; q++) uniform = F.own[kids[q]] < 0 && F.rho[kids[q]] == s;
        if (uniform) C.rho[e] = s;
        else C.own[e] = nOwn++;
    }
    C.nOwn = nOwn;
    C.K.assign(i64(nOwn) * 576, 0.f);
    parallelRange(nE, [&](i64 lo, i64 hi) {
        std::vector<double> out(576), Kf(576);
        for (i64 e = lo; e < hi; e++) {
            if (C.own[e] < 0) continue;
            std::fill(out.begin(), out.end(), 0.0);
            for (int q = start[e]; q < start[e + 1]; q++) {
                const int f = kids[q];
                const int c = pos[f];
                if (F.own[f] < 0) {
                    const double s = F.rho[f];
                    const double* Mc = M.data() + c * 576;
                    for (int u = 0; u < 576; u++) out[u] += s * Mc[u];
                } else {
                    const float* Ko = F.K.data() + i64(F.own[f]) * 576;
                    for (int u = 0; u < 576; u++) Kf[u] = Ko[u];
                    galerkinChild(out.data(), Kf.data(), c);
                }
            }
            float* dst = C.K.data() + i64(C.own[e]) * 576;
            for (int u = 0; u < 576; u++) dst[u] = float(out[u]);
        }
    }, 64);
    // child-scale form: when every fine element is s * Kb, an own coarse element is exactly
    // sum_c s_c M_c (M_c the Galerkin product of the fine Kb at child position c, s_c = 0 for an
    // absent child), so 8 scales replace its 576 stored entries
    if (F.nOwn == 0) {
        C.childForm = true;
        C.childS.assign(i64(nOwn) * 8, 0.f);
        for (i64 e = 0; e < nE; e++) {
            if (C.own[e] < 0) continue;
            for (int q = start[e]; q < start[e + 1]; q++) C.childS[i64(C.own[e]) * 8 + pos[kids[q]]] = float(F.rho[kids[q]]);
        }
        C.childM.assign(M.begin(), M.end());
        // two-valued children (void parts: solid and void scales, or a scale and absent children):
        // sum_c s_c M_c = a Kb + (b - a) M_B with M_B the sum over the children of scale b
        C.childTwo.assign(i64(nOwn) * 4, 0);
        for (i64 e = 0; e < nE; e++) {
            if (C.own[e] < 0) continue;
            const float* cs = C.childS.data() + i64(C.own[e]) * 8;
            const float a = cs[0];
            float b = a;
            int mask = 0;
            bool two = true;
            for (int c = 1; c < 8; c++) {
                if (cs[c] == a) continue;
                if (mask == 0) b = cs[c];
                if (cs[c] == b) mask |= 1 << c;
                else two = false;
            }
            int32_t* t = C.childTwo.data() + i64(C.own[e]) * 4;
            std::memcpy(&t[0], &a, 4);
            std::memcpy(&t[1], &b, 4);
            t[2] = two && mask ? mask : -1;
        }
        C.childTab.assign(256 * 64, 0.0);
        childTable(M.data(), C.childTab.data());
    }
    // a coarse DOF is held when it interpolates onto a held fine DOF
    C.bc.assign(C.nDof, 0);
    for (int k = 0; k < F.NZ; k++)
        for (int j = 0; j < F.NY; j++)
            for (int i = 0; i < F.NX; i++) {
                const i64 fn = 3 * F.node(i, j, k);
                const uint8_t h0 = F.bc[fn], h1 = F.bc[fn + 1], h2 = F.bc[fn + 2];
                if (!(h0 | h1 | h2)) continue;
                const int Ks[2] = {k / 2, (k & 1) ? (k + 1) / 2 : -1}, Js[2] = {j / 2, (j & 1) ? (j + 1) / 2 : -1},
                          Is[2] = {i / 2, (i & 1) ? (i + 1) / 2 : -1};
                for (int K : Ks)
                    for (int J : Js)
                        for (int I : Is) {
                            if (K < 0 || J < 0 || I < 0) continue;
                            const i64 cn = 3 * C.node(I, J, K);
                            C.bc[cn] |= h0;
                            C.bc[cn + 1] |= h1;
                            C.bc[cn + 2] |= h2;
                        }
            }
    if (!F.diagAdd.empty()) {
        // lumped Galerkin product of a diagonal: P^T (D 1), full weighting over all coarse nodes
        C.diagAdd.assign(C.nDof, 0.0);
        for (int K = 0; K < C.NZ; K++)
            for (int J = 0; J < C.NY; J++)
                for (int I = 0; I < C.NX; I++) {
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
                                const i64 fn = 3 * F.node(fi, jf, kf);
                                for (int d = 0; d < 3; d++) s[d] += w * F.diagAdd[fn + d];
                            }
                        }
                    }
                    const i64 cn = 3 * C.node(I, J, K);
                    for (int d = 0; d < 3; d++) C.diagAdd[cn + d] = s[d];
                }
    }
    finishLevel(C);
    return C;
}

// coarsest free-DOF matrix, its Cholesky factor (pivots regularised like voxel_fea.cpp:639-659)
// and the explicit inverse L^-T L^-1 that the device applies as one matrix-vector product
static void coarseInverse(Hierarchy& H) {
    const Level& L = H.levels.back();
    std::vector<i64> map(L.nDof, -1);
    H.freeIdx.clear();
    for (i64 q = 0; q < L.nDof; q++)
        if (!L.fixed[q]) {
            map[q] = i64(H.freeIdx.size());
            H.freeIdx.push_back(int32_t(q));
        }
    const i64 m = i64(H.freeIdx.size());
    H.m = m;
    if (m == 0 || m > 3000) throw std::runtime_error("the coarsest level has no direct solver");
    std::vector<double> A(size_t(m * m), 0.0);
    i64 dofs[24];
    for (size_t e = 0; e < L.elemVox.size(); e++) {
        const int v = L.elemVox[e];
        const int i = v % L.nx, j = (v / L.nx) % L.ny, k = v / (L.nx * L.ny);
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) dofs[3 * a + d] = map[3 * L.node(i + CX[a], j + CY[a], k + CZ[a]) + d];
        const float* Ko = L.own[e] >= 0 ? L.K.data() + i64(L.own[e]) * 576 : nullptr;
        for (int r = 0; r < 24; r++) {
            if (dofs[r] < 0) continue;
            for (int c = 0; c < 24; c++)
                if (dofs[c] >= 0) A[dofs[r] * m + dofs[c]] += Ko ? double(Ko[r * 24 + c]) : L.rho[e] * L.Kb[r * 24 + c];
        }
    }
    if (!L.diagAdd.empty())
        for (i64 q = 0; q < L.nDof; q++)
            if (map[q] >= 0) A[map[q] * m + map[q]] += L.diagAdd[q];
    double maxDiag = 0;
    for (i64 q = 0; q < m; q++) maxDiag = std::max(maxDiag, A[q * m + q]);
    // right-looking Cholesky, lower triangle, trailing rows in parallel
    for (i64 j = 0; j < m; j++) {
        double s = A[j * m + j];
        if (!(s > 1e-10 * maxDiag)) s = maxDiag;
        const double ljj = std::sqrt(s), inv = 1 / ljj;
        A[j * m + j] = ljj;
        for (i64 q = j + 1; q < m; q++) A[q * m + j] *= inv;
        const i64 rest = m - j - 1;
        auto upd = [&](i64 lo, i64 hi) {
            for (i64 q = j + 1 + lo; q < j + 1 + hi; q++) {
                const double lqj = A[q * m + j];
                if (lqj == 0) continue;
                double* row = A.data() + q * m;
                for (i64 t = j + 1; t <= q; t++) row[t] -= lqj * A[t * m + j];
            }
        };
        if (rest > 256) parallelRange(rest, upd, 64);
        else upd(0, rest);
    }
    // W = L^-1 (lower), column by column
    std::vector<double> W(size_t(m * m), 0.0);
    parallelRange(m, [&](i64 lo, i64 hi) {
        std::vector<double> y(m);
        for (i64 c = lo; c < hi; c++) {
            std::fill(y.begin(), y.end(), 0.0);
            y[c] = 1 / A[c * m + c];
            for (i64 q = c + 1; q < m; q++) {
                double s = 0;
                const double* row = A.data() + q * m;
                for (i64 t = c; t < q; t++) s += row[t] * y[t];
                y[q] = -s / row[q];
            }
            for (i64 q = c; q < m; q++) W[q * m + c] = y[q];
        }
    }, 8);
    // A^-1 = W^T W (symmetric)
    H.Ainv.assign(size_t(m * m), 0.0);
    parallelRange(m, [&](i64 lo, i64 hi) {
        for (i64 r = lo; r < hi; r++)
            for (i64 c = 0; c <= r; c++) {
                double s = 0;
                for (i64 t = r; t < m; t++) s += W[t * m + r] * W[t * m + c];
                H.Ainv[r * m + c] = s;
                H.Ainv[c * m + r] = s;
            }
    }, 8);
}

void checkGridLimits(int nx, int ny, int nz) {
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("The structural grid must have three positive dimensions.");
    const i64 NX = i64(nx) + 1, NY = i64(ny) + 1, NZ = i64(nz) + 1, lim = INT32_MAX;
    // a node plane's FP64 components (24 bytes a node) and a voxel plane's words (4 bytes) stay within a
    // 32-bit byte offset: the operator kernels address a plane through one buffer descriptor
    if (24 * NX * NY > lim || 4 * i64(nx) * ny > lim || NX * NY * NZ > lim - 2 || i64(nx) * ny * nz > lim - 2)
        throw std::invalid_argument("The structural grid of " + std::to_string(nx) + " x " + std::to_string(ny) + " x " + std::to_string(nz) +
                                    " voxels exceeds the 32-bit node and plane indices of the device kernels.");
}

void buildHierarchy(Hierarchy& H, int nx, int ny, int nz, const float* density, double nu, const uint8_t* bc, i64 coarsestMaxDof,
                    const double* diagAdd, int maxLevels) {
    checkGridLimits(nx, ny, nz);
    if (!density || !bc) throw std::invalid_argument("density and bc are required");
    const i64 nVox = i64(nx) * ny * nz;
    for (i64 e = 0; e < nVox; e++)
        if (!std::isfinite(density[e]) || density[e] < 0 || density[e] > 1) throw std::invalid_argument("Voxel density must be between zero and one.");
    if (maxLevels <= 0 || maxLevels > 12) maxLevels = 12;
    H.nu = nu;
    H.el = hexElement(nu);
    H.levels.clear();
    Level L0 = emptyLevel(nx, ny, nz);
    std::memcpy(L0.Kb, H.el.K, sizeof L0.Kb);
    L0.emap.assign(nVox, -1);
    {
        // element numbering in voxel order: per-chunk counts, then offsets, then the fill
        const int T = hostThreads();
        std::vector<i64> cnt(T + 1, 0);
        parallelRange(T, [&](i64 lo, i64 hi) {
            for (i64 t = lo; t < hi; t++)
                for (i64 v = nVox * t / T; v < nVox * (t + 1) / T; v++) cnt[t + 1] += density[v] > 0 ? 1 : 0;
        }, 1);
        for (int t = 0; t < T; t++) cnt[t + 1] += cnt[t];
        L0.elemVox.resize(cnt[T]);
        L0.rho.resize(cnt[T]);
        parallelRange(T, [&](i64 lo, i64 hi) {
            for (i64 t = lo; t < hi; t++) {
                i64 at = cnt[t];
                for (i64 v = nVox * t / T; v < nVox * (t + 1) / T; v++)
                    if (density[v] > 0) {
                        L0.emap[v] = int32_t(at);
                        L0.elemVox[at] = int32_t(v);
                        L0.rho[at] = density[v];
                        at++;
                    }
            }
        }, 1);
    }
    L0.own.assign(L0.elemVox.size(), -1);
    L0.bc.assign(bc, bc + L0.nDof);
    if (diagAdd) L0.diagAdd.assign(diagAdd, diagAdd + L0.nDof);
    finishLevel(L0);
    H.levels.push_back(std::move(L0));
    for (;;) {
        const Level& L = H.levels.back();
        if (L.freeDof <= coarsestMaxDof || int(H.levels.size()) >= maxLevels) break;
        if (L.nx <= 1 && L.ny <= 1 && L.nz <= 1) break;
        Level C = coarsen(L);
        H.levels.push_back(std::move(C));
    }
    coarseInverse(H);
}

std::vector<double> lanczosStart(const Level& L) {
    std::vector<double> v(L.nDof);
    uint32_t seed = 12345;
    double nn = 0;
    for (i64 q = 0; q < L.nDof; q++) {
        seed = (seed * 1103515245u + 12345u) & 0x7fffffffu;
        v[q] = L.fixed[q] ? 0 : double(seed) / 0x7fffffff - 0.5;
        nn += v[q] * v[q];
    }
    const double s = nn > 0 ? 1 / std::sqrt(nn) : 1;
    for (auto& x : v) x *= s;
    return v;
}

double tridiagonalMax(const std::vector<double>& a, const std::vector<double>& b) {
    const int k = int(a.size());
    if (k == 0) return 0;
    double lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < k; i++) {
        const double r = (i > 0 ? std::abs(b[i - 1]) : 0) + (i + 1 < k ? std::abs(b[i]) : 0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    // number of eigenvalues above x from the signs of the LDL^T pivots of T - x I
    auto above = [&](double x) {
        int n = 0;
        double q = 1;
        for (int i = 0; i < k; i++) {
            q = a[i] - x - (i > 0 ? b[i - 1] * b[i - 1] / q : 0);
            if (q == 0) q = 1e-300;
            if (q > 0) n++;
        }
        return n;
    };
    for (int it = 0; it < 100 && hi - lo > 1e-12 * std::abs(hi); it++) {
        const double mid = 0.5 * (lo + hi);
        if (above(mid) > 0) lo = mid;
        else hi = mid;
    }
    return hi;
}

// voxels not face-connected to a held node, removed (density 0): the input rule applied before
// every break-step re-solve (voxel_fea.cpp:1086-1119). Seeds are the solid voxels with a held corner;
// the fill walks the six face neighbours. Returns the number of voxels removed.
std::vector<int32_t> crackSelect(const float* measure, size_t n, double smax, size_t cap) {
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return measure[a] > measure[b]; });
    std::vector<int32_t> out;
    for (int q : order) {
        if (out.size() >= cap || measure[q] < 0.95 * smax) break;
        out.push_back(q);
    }
    return out;
}

int pruneFloating(int nx, int ny, int nz, const float* density, const uint8_t* heldNode, float* out) {
    const i64 NX = nx + 1, NY = ny + 1, nE = i64(nx) * ny * nz;
    std::vector<uint8_t> reached(nE, 0);
    std::vector<i64> frontier;
    auto solid = [&](i64 e) { return density[e] > 0; };
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const i64 e = i + i64(nx) * (j + i64(ny) * k);
                if (!solid(e)) continue;
                bool held = false;
                for (int c = 0; c < 8 && !held; c++) held = heldNode[(i + CX[c]) + NX * ((j + CY[c]) + NY * (k + CZ[c]))] != 0;
                if (held) {
                    reached[e] = 1;
                    frontier.push_back(e);
                }
            }
    while (!frontier.empty()) {
        const i64 e = frontier.back();
        frontier.pop_back();
        const int i = int(e % nx), j = int((e / nx) % ny), k = int(e / (i64(nx) * ny));
        const i64 step[3] = {1, nx, i64(nx) * ny};
        const int at[3] = {i, j, k}, len[3] = {nx, ny, nz};
        for (int ax = 0; ax < 3; ax++)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                const int c = at[ax] + sgn;
                if (c < 0 || c >= len[ax]) continue;
                const i64 m = e + sgn * step[ax];
                if (!reached[m] && solid(m)) {
                    reached[m] = 1;
                    frontier.push_back(m);
                }
            }
    }
    int removed = 0;
    for (i64 e = 0; e < nE; e++) {
        const bool keep = solid(e) && reached[e];
        removed += solid(e) && !reached[e];
        out[e] = keep ? density[e] : 0.f;
    }
    return removed;
}

}  // namespace psx
