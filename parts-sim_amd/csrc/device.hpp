// Device-side view of one multigrid level and the launchers of the HIP kernels (kernels.hip).
// Everything here is plain data and pointers into HBM; the host code in solver.cpp owns the buffers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace psx {

// One level as the kernels see it. Vectors are AoS [node][3] over the full (NX, NY, NZ) node box,
// node n = i + NX (j + NY k); voxels v = i + nx (j + ny k). A DOF is held (or inactive) exactly
// when its inverse diagonal is 0, so no separate mask array is read.
struct DevLevel {
    int nx, ny, nz, NX, NY, NZ;
    int64_t nNodes, nDof;
    const int32_t* words;      // per voxel: 0 empty, >0 float bits of s (s * Kb), <0 0x80000000 | own index
    const float* Kown;         // own element matrices, 576 floats each
    // the same in FP16 for the mixed V-cycle's smoothing, each element divided by a power of two
    // (ownScale) that puts its largest entry in [2^13, 2^14) (null on levels without own matrices)
    const uint16_t* KownH;
    const float* ownScale;
    const double* Kb64;        // the level's uniform element matrix
    const float* Kb32;
    const double* St64;        // 27-point stencil of Kb: [27 neighbours][3 rows][3 cols]
    const float* St32;
    // Kb as the matrix-core A operands of fine_el.hip, [half][piece][lane] x 8 BF16: Kb split into three
    // BF16 pieces (hi, mid, lo); rows 4 component + output corner (x + 2 y, at z offset half), input
    // corner pairs in the K slices
    const uint4* elA;
    const float* nodeScale;    // s for nodes whose 8 voxels are s * Kb, else 0
    const uint16_t* nodeInfo;  // per node: bit v (0..7) set when voxel v's scale differs from the
                               // reference voxel's (the lowest clear bit); bit 8 + d when DOF d is
                               // held or inactive (inverse diagonal 0)
    // child-scale form of the own matrices (levels whose finer level is all s * Kb): childS[8 e + c]
    // scales the finer Kb's Galerkin product at child position c; crow = the corner-0 rows of those
    // 8 products, [c][row][24]. kFine uses it when childForm is set.
    const float* childS;
    const double* crow64;
    const float* crow32;
    // two-scale own elements: [a, b, mask of the b children or -1, 0] per own element and the
    // subset table [256][64] (corner-0 rows of sum_{c in B} M_c over corners 1..7)
    const int32_t* childTwo;  // 4 per own element, 16-byte aligned
    const double* ctab64;
    const float* ctab32;
    // the subset table in FP16 for the mixed V-cycle's smoothing: row B (64 entries) is ctab's row B over
    // ctabRowScale[B], the power of two that puts the row's largest entry in (2^13, 2^14] (null on levels
    // without the child-scale form)
    const uint16_t* ctab16;
    const float* ctabRowScale;
    int childForm;
    // the child-scale level's own elements for the matrix-core operator: the voxel of each own index,
    // and a work buffer of their corner rows ([instance][own][corner x + 2 y + 4 z][component], FP32;
    // ownTStride floats per instance), written before each product (null elsewhere)
    const int32_t* ownVox;
    const uint4* childA;  // the 8 child products M_c as matrix-core A operands ([c][half][piece][lane])
    float* ownT;
    int64_t ownTStride;
    // kFine sums the 8 voxels of every node directly (no stencil, no pair lists): set on levels
    // without own elements where most nodes have differing voxels (per-voxel densities)
    int dense;
    const double* invD64;      // inverse diagonal, 0 on held / inactive DOFs
    const float* invD32;
    const double* diagAdd;     // optional diagonal shift (nullptr when absent)
    // The cube's symmetry leaves 22 distinct values in Kb and its stencil (kFine reads them from
    // here, as wave-uniform kernel arguments): [0, 14) the element entries, [14, 20) the stencil
    // diagonal classes, [20, 22) its off-diagonal classes. Valid when symOk.
    double sym64[22];
    float sym32[22];
    int symOk;
};

// Scalars of one conjugate-gradient run, resident on the device. Every kernel of an iteration
// returns at once when `stop` is set, so the host can queue iterations ahead and read the state
// only now and then.
struct CgState {
    double rz, pq, alpha, beta, rr, rzNew, zq, bnorm, tol, rel;
    double rmax, scale, peak, trueRel, spare[6];
    int stop, iters, maxIter, broken, cancel, folds, sinceFold, foldEvery;
    int noFold, pad[3];
    double lz[32];             // Lanczos alpha[0..9], beta at [16..25]
};

// partial sums of the reductions: one slot per block (blocks own RED_TILE elements), up to 4
// values per kernel
constexpr int RED_TILE = 2048;
struct Reduce {
    double* partials;          // per instance: 4 * (nDof / RED_TILE + 2) or more
    unsigned* counter;
    int64_t stride;            // doubles between consecutive instances' partials
};

// Multi-instance launches (several right-hand sides of one level, one launch): blockIdx.y is the
// instance; vectors are instance-major with the level's nDof per instance, and the conjugate-
// gradient state is CgState[instance]. Flags and scalars that point into CgState step by these.
constexpr int CG_INTS = int(sizeof(CgState) / sizeof(int));
constexpr int CG_DBLS = int(sizeof(CgState) / sizeof(double));
static_assert(sizeof(CgState) % sizeof(double) == 0, "CgState must tile in doubles");

enum ApplyMode { AM_APPLY = 0, AM_RAW = 1, AM_RESID = 2, AM_CHEB1 = 3, AM_CHEB2 = 4 };
// operator forms (bit flags): own elements in child-scale form; BF16 matrix-core operator
// OF_CHILD: every level that has the child-scale form uses it; OF_AUTO: only levels where it is the
// cheaper form (most own elements two-valued; a level of general own elements streams its stored
// matrices faster than it evaluates 8 children per element)
// OF_MFMA_OWN: FP32 levels with stored own matrices take the matrix-core element product
// OF_MFMA: FP32 levels without own elements take the element-by-element product on the matrix cores
// with two BF16 pieces of x - c and of Kb (fine_el.hip); OF_EL3: the same with three pieces each
enum OpForm { OF_CHILD = 1, OF_MFMA = 2, OF_AUTO = 4, OF_MFMA_OWN = 8, OF_EL3 = 16 };

// y = K x on a level with the epilogue of `mode` (T = double or float):
//   AM_APPLY  y = K x, held DOFs 0               AM_RAW   y = K x
//   AM_RESID  y = r - K x, held DOFs 0
//   AM_CHEB1  y = x + c1 D^-1 (r - K x)                        (first step with a nonzero start)
//   AM_CHEB2  y = x + c1 (x - zp) + c2 D^-1 (r - K x)          (later steps; zp = the previous
//             iterate, nullptr for a zero one: the step d = y - x needs no vector of its own)
// y must not alias x or zp.
template <class T> void launchApply(const DevLevel& L, int mode, const T* x, T* y, const T* r, const T* zp, double c1, double c2, const int* stop,
                                    hipStream_t s, int ninst = 1);
// the same with 8 lanes per node (one per voxel): levels with own matrices
// half: read the FP16 own matrices (FP32 vectors only)
template <class T> void launchApplyOwn(const DevLevel& L, int mode, const T* x, T* y, const T* r, const T* zp, double c1, double c2, const int* stop,
                                       hipStream_t s, int ninst = 1, bool half = false);
// FP16 copies of n own matrices with their per-element power-of-two scales
void launchOwnToHalf(const float* Kown, int64_t n, uint16_t* half, float* scale, hipStream_t s);
// the same product on the matrix cores (FP32 vectors, BF16 operands on the element-corner differences,
// FP32 accumulation; one wavefront per node, 16 instances per pass): the mixed V-cycle's smoothing on
// levels with own matrices when asked (OF_MFMA_OWN)
void launchApplyOwnMf(const DevLevel& L, int mode, const float* x, float* y, const float* r, const float* zp, double c1, double c2, const int* stop,
                      hipStream_t s, int ninst = 1);
// the same for levels whose elements are s * Kb, or own elements in child-scale form: z-marching
// LDS planes, class-sum stencil in difference form + one lane per differing voxel (fine.hip)
// el: FP32 levels without own elements take the element-by-element product on the matrix cores with
// el (2 or 3) BF16 pieces of the operands (fine_el.hip); 0 = the VALU forms
// half: FP32 levels with child-scale own elements read the FP16 subset table (the mixed V-cycle's
// smoothing)
template <class T> void launchFine(const DevLevel& L, int mode, const T* x, T* y, const T* r, const T* zp, double c1, double c2, const int* stop,
                                   hipStream_t s, int ninst = 1, int el = 0, bool half = false);
// ownVox[own index] = the voxel of each own element (words < 0)
void launchOwnVox(const int32_t* words, int64_t nVox, int32_t* ownVox, hipStream_t s);
// the element-by-element matrix-core operator itself (fine_el.hip); dots: fused dot partials, one per
// tile (the kinds of launchFineDot); returns the number of partials written (0: no fused dot)
unsigned launchFineEl(const DevLevel& L, int mode, const float* x, float* y, const float* r, const float* zp, double c1, double c2, const int* stop,
                      double* dots, int64_t dotStride, const float* qdot, hipStream_t s, int ninst, int pieces);
// Dot products fused into an operator launch (conjugate gradients), block partials into red:
//   FD_PQ  (AM_APPLY)  p.q with p = x, q = y                         -> launchFinishPQ
//   FD_RZ  (AM_CHEB1/2) r.z and z.q with z = y, q = qdot             -> launchFinishRZ
enum FuseDot { FD_NONE = 0, FD_PQ = 1, FD_RZ = 2 };
struct DotFusion {
    int kind = FD_NONE;
    const void* q = nullptr;  // FD_RZ: the previous search direction's product K p (type T)
    Reduce red{};
    // FD_PQ on the VALU operator (type T): the direction update p = z + beta x fused into y = K p,
    // z the preconditioned residual, the new p stored to pw (not over x); both null: x is p
    const void* pz = nullptr;
    void* pw = nullptr;
};
// launchFine with a fused dot; returns the number of block partials written (0: not fused)
template <class T> unsigned launchFineDot(const DevLevel& L, int mode, const T* x, T* y, const T* r, const T* zp, double c1, double c2, const int* stop,
                                          const DotFusion& fd, hipStream_t s, int ninst = 1, int el = 0, bool half = false);
void launchFinishPQ(Reduce red, unsigned nb, CgState* st, hipStream_t s, int ninst = 1);
void launchFinishRZ(Reduce red, unsigned nb, CgState* st, hipStream_t s, int ninst = 1);
// z = c0 D^-1 r (first Chebyshev step from zero)
template <class T> void launchChebFirst(const DevLevel& L, const T* r, T* z, double c0, const int* stop, hipStream_t s, int ninst = 1);
// rc = P^T rf on the coarse level's free DOFs; z0 (when not null) = c0 D_C^-1 rc, the coarse level's
// first Chebyshev step
template <class T>
void launchRestrict(const DevLevel& F, const DevLevel& C, const T* rf, T* rc, const int* stop, hipStream_t s, int ninst = 1, T* z0 = nullptr,
                    double c0 = 0);
// zf += P zc on the fine level's free DOFs
template <class T> void launchProlong(const DevLevel& F, const DevLevel& C, const T* zc, T* zf, const int* stop, hipStream_t s, int ninst = 1);
// coarsest level: z[freeIdx] = Ainv r[freeIdx], z = 0 on held DOFs (dense[dof] = dense index or -1)
template <class T>
void launchCoarse(const int32_t* freeIdx, const int32_t* dense, int64_t m, int64_t nDof, const double* Ainv, const T* r, T* z, const int* stop, hipStream_t s, int ninst = 1);

// conjugate-gradient steps (all accumulate in FP64; T is the vector type)
// pq = p.q; alpha = rz / pq (or broken + stop when pq <= 0)
template <class T> void launchDotPQ(const T* p, const T* q, int64_t n, CgState* st, Reduce red, hipStream_t s, int ninst = 1);
// x += alpha p; r -= alpha q; rr = r.r; iteration count, relative residual, stop test
template <class T> void launchUpdateXR(T* x, const T* p, T* r, const T* q, int64_t n, CgState* st, Reduce red, hipStream_t s, int ninst = 1);
// rzNew = r.z, zq = z.q; beta = max(0, -alpha zq / rz) (flexible); rz = rzNew
template <class T> void launchDotsRZ(const T* r, const T* z, const T* q, int64_t n, CgState* st, Reduce red, hipStream_t s, int ninst = 1);
// p = z + beta p
template <class T> void launchUpdateP(const T* z, T* p, int64_t n, CgState* st, hipStream_t s, int ninst = 1);

// fold, part 1: u += x * rmax, x = 0 (only when a fold is pending)
// nodal stresses (stress.hip): S = the 8 corner stress matrices [a][6][24] scaled by E / h
void launchStresses(const DevLevel& L, const double* S, const double* u, int principals, float* vm, float* p1, float* p3, hipStream_t s);
void launchStressTensors(const DevLevel& L, const double* S, const double* u, float* out, hipStream_t s);
// mixed with FP32 recurrences: x32 += alpha p, r32 -= alpha q (fold decision); at a fold u += rmax x32
// z0 (may be null): also z0 = c0 D^-1 r from the new r (invD32: the level's FP32 inverse diagonal),
// the next V-cycle's first Chebyshev step; launchFoldR rewrites it when it replaces r32
void launchUpdateXR32Fold(float* x, const float* p, float* r, const float* q, int64_t n, CgState* st, Reduce red, hipStream_t s, int ninst = 1,
                          float* z0 = nullptr, const float* invD32 = nullptr, double c0 = 0);
void launchFoldX(double* u, float* x, int64_t n, CgState* st, hipStream_t s, int ninst = 1);
// fold, part 2 (after Kx = K u in FP64): r64 = f - Kx (r64 may be null: not stored), r32 = r64 / rmax,
// true residual, stop test; held DOFs from invD (FP64) or, when it is null, invD32
void launchFoldR(const double* f, const double* Kx, const double* invD, double* r64, float* r32, int64_t n, CgState* st, Reduce red, hipStream_t s, int ninst = 1,
                 float* z0 = nullptr, const float* invD32 = nullptr, double c0 = 0);

// r = f - Kx on free DOFs (Kx already masked), out[0] += |r|^2 over the grid (FP64)
void launchResidual64(const double* f, const double* Kx, const double* invD, double* r, int64_t n, double* out, Reduce red, hipStream_t s, int ninst = 1);
// out = sum over free DOFs of f^2
void launchNormFree64(const double* f, const double* invD, int64_t n, double* out, Reduce red, hipStream_t s, int ninst = 1);
// x[q] = held ? 0 : x0[q] (or 0 when x0 is null)
void launchMaskCopy64(const double* x0, const double* invD, double* x, int64_t n, hipStream_t s, int ninst = 1);
// out = max |v| over DOFs (FP64 vector)
void launchMaxAbs64(const double* v, int64_t n, double* out, Reduce red, hipStream_t s, int ninst = 1);
// y32 = float(x64 * (1 / *scaleFrom)) (the V-cycle's order-one input), z64 = double(z32) * *scaleFrom
void launchToF32(const double* x, float* y, int64_t n, const double* scaleFrom, hipStream_t s, int ninst = 1);
void launchFromF32(const float* z, double* y, int64_t n, const double* scaleFrom, hipStream_t s, int ninst = 1);
// Lanczos steps for lmax(D^-1 K) (voxel_fea.cpp:579-610)
void launchLzX(const double* v, const double* invD, double* x, int64_t n, hipStream_t s);
void launchLzW(double* w, const double* invD, const double* vPrev, const double* v, int64_t n, CgState* st, int j, Reduce red, hipStream_t s);
void launchLzAxpy(double* w, const double* v, int64_t n, CgState* st, int j, Reduce red, hipStream_t s);
void launchLzNext(const double* w, double* v, double* vPrev, int64_t n, CgState* st, int j, hipStream_t s);

}  // namespace psx
