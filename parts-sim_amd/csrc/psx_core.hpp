// Host-side model of the voxel hierarchy: the level data the device kernels read, built on the CPU
// from the solver's inputs (dims, density, nu, held DOFs, optional diagonal shift).
//
// Level contents follow the reference's Level (native/src/fea/voxel_fea.hpp:37-78): element list,
// uniform element matrix Kb (K0 on the finest level, then the exact Galerkin product of 8 uniform
// children), own 24x24 matrices for the coarse elements that are not s * Kb, held DOFs, inverse
// diagonal, and the per-node stencil scale. The device form packs an element's kind into one
// 32-bit word per voxel: 0 empty, the float bits of s for s * Kb, or 0x80000000 | own index.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace psx {

using i64 = int64_t;

// element corner at unit offset (x, y, z): bottom face counter-clockwise, then the top face
// (native/src/fea/hex8.hpp:14, voxel_fea.cpp:52)
constexpr int corner(int x, int y, int z) { return (y ? (x ? 2 : 3) : (x ? 1 : 0)) + 4 * z; }
constexpr int CX[8] = {0, 1, 1, 0, 0, 1, 1, 0}, CY[8] = {0, 0, 1, 1, 0, 0, 1, 1}, CZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

struct Element {
    double K[576];          // condensed 24x24 stiffness of the unit cube, E = 1
    double corner[8][144];  // D * Bbar at each corner (6 x 24), for stresses
};
// Wilson incompatible-mode hex8 on the unit cube (native/src/fea/hex8.cpp:104-155); throws for nu
// outside (-1, 0.5)
Element hexElement(double nu);

struct Level {
    int nx = 0, ny = 0, nz = 0, NX = 0, NY = 0, NZ = 0;
    i64 nNodes = 0, nDof = 0, nVox = 0, freeDof = 0, nOwn = 0;
    std::vector<int32_t> emap;      // voxel -> element, -1 empty
    std::vector<int32_t> elemVox;   // element -> voxel (increasing)
    std::vector<double> rho;        // element -> s for s * Kb (0 for own-matrix elements)
    std::vector<int32_t> own;       // element -> own matrix index, -1 for s * Kb
    std::vector<float> K;           // own matrices, 576 each (row-major)
    double Kb[576]{};
    std::vector<uint8_t> bc, fixed, active;
    std::vector<double> invDiag, diagAdd;
    std::vector<float> nodeScale;   // s when the node's 8 voxels are all s * Kb, else 0
    std::vector<uint8_t> diffMask;  // per node: voxels whose scale differs from the reference voxel
    // child-scale form of the own matrices (levels whose finer level has no own matrices): own
    // element e = sum_c childS[8 e + c] * childM[c] (childM: 8 x 576, the Galerkin products of the
    // finer Kb per child position c = cx + 2 cy + 4 cz)
    bool childForm = false;
    std::vector<float> childS;
    std::vector<double> childM;
    // own elements whose children carry at most two scales a (child 0's) and b: [a bits, b bits,
    // mask of the b children or -1 when more than two scales, 0]; childTab = corner-0 rows of
    // sum_{c in B} M_c for every subset B (columns of corners 1..7), [256][64]
    std::vector<int32_t> childTwo;
    std::vector<double> childTab;
    std::vector<int32_t> words;     // device voxel words
    i64 nElem = 0;                  // elements (non-empty voxels)
    double lmax = 0;
    i64 node(int i, int j, int k) const { return i + i64(NX) * (j + i64(NY) * k); }
};

struct Hierarchy {
    double nu = 0.3;
    Element el;
    std::vector<Level> levels;
    i64 m = 0;                       // free DOFs of the coarsest level
    std::vector<int32_t> freeIdx;    // dense index -> coarsest DOF
    std::vector<double> Ainv;        // m x m inverse of the coarsest free-DOF matrix (row-major)
};

// Index widths the device kernels rely on: a node plane's FP64 components (24 NX NY bytes) and a voxel
// plane's words (4 nx ny bytes) fit a signed 32-bit byte offset from a 64-bit plane base, and node and
// voxel indices fit signed 32-bit values (queue entries, voxel words, decoded node indices). Throws
// std::invalid_argument for a grid beyond them, before any array is read.
void checkGridLimits(int nx, int ny, int nz);

// Builds the hierarchy (voxel_fea.cpp:138-183): coarsens until the free DOFs are at most
// coarsestMaxDof, 12 levels, or a single voxel; then factors the coarsest level. Throws
// std::invalid_argument on bad input and std::runtime_error when the coarsest level has no direct
// solve (0 or more than 3000 free DOFs).
void buildHierarchy(Hierarchy& H, int nx, int ny, int nz, const float* density, double nu, const uint8_t* bc, i64 coarsestMaxDof,
                    const double* diagAdd, int maxLevels);

// Galerkin products of a level's uniform element Kb per child position c = cx + 2 cy + 4 cz
// (M: 8 x 576, M_c = P_c^T Kb P_c) and their sum, the coarser level's uniform element
void childProducts(const double* Kb, double* M, double* KbCoarse);
// corner-0 rows of sum_{c in B} M_c for every subset B of the children, columns of corners 1..7:
// [256][64]
void childTable(const double* M, double* tab);
// trilinear weight of coarse corner b at fine corner a of child position c
double childWeight(int c, int a, int b);

// The 27-point stencil of a uniform node for element matrix K: S[m][d][c] couples DOF d of the node
// with DOF c of neighbour m = ox + 3 oy + 9 oz (offsets -1..1 shifted to 0..2).
void stencil(const double* K, double* S243);

// Symmetry-reduced coefficients of a level's uniform element: out[0..14) element classes,
// [14..20) stencil diagonal classes, [20..22) stencil off-diagonal classes (class means). Returns the
// largest deviation of any entry from its class value relative to max |K| (0 for an exactly
// cube-symmetric element).
double symmetryClasses(const double* K, double* out22);
// class of element entry K[3a+d][3b+c] (index | 0x100 when negated) and of stencil entry S[m][d][c]
// (-1: structural zero)
int elementClass(int a, int d, int b, int c);
int stencilClass(int m, int d, int c);

// Lanczos start vector of a level (an LCG with seed 12345, fixed DOFs 0; voxel_fea.cpp:584-588).
std::vector<double> lanczosStart(const Level& L);
// largest eigenvalue of a symmetric tridiagonal (Sturm-count bisection; voxel_fea.cpp:506-530)
double tridiagonalMax(const std::vector<double>& a, const std::vector<double>& b);

// The crack selection of one break step (static_study.cpp:245-258): the element indices by decreasing
// measure as std::sort orders all n of them (equal measures in that sort's order), the first cap of
// those whose measure is at least 0.95 smax.
std::vector<int32_t> crackSelect(const float* measure, size_t n, double smax, size_t cap);

// solid voxels not face-connected to a held node set to 0 (voxel_fea.cpp:1086-1119); heldNode per
// node; returns the count removed
int pruneFloating(int nx, int ny, int nz, const float* density, const uint8_t* heldNode, float* out);

// number of worker threads for host-side builds
int hostThreads();

}  // namespace psx
