/*
 * psx.h — C ABI of the MI355X-native voxel FEA linear-static solve.
 *
 * The boundary the reference's GPU solver class sits on: ps::GpuFeaSolver
 * (native/src/gpu/gpu_fea.hpp:17-35) over ps::VoxelFEA (native/src/fea/voxel_fea.hpp:130-189).
 * A model owns the multigrid hierarchy and every device buffer of one part; the C++ shim in
 * parts-sim_amd/host/gpu_fea.hpp re-exposes the reference's class over these calls so the bend,
 * break, nonlinear-ramp and topology loops call it unchanged.
 *
 * Conventions
 *  - Units: the normalised system of the reference (E = 1, voxel size = 1; voxel_fea.hpp:17-18):
 *    forces in N, u_physical = u / (E h).
 *  - Layout: density[e], e = i + nx (j + ny k); DOF vectors [node][3], node n = i + NX (j + NY k),
 *    NX = nx + 1 (structural.hpp:67, voxel_fea.hpp:132-135).
 *  - Every call returns PSX_OK (0) or an error code; psx_last_error() gives the message of the
 *    calling thread's last failure. No exception crosses this ABI.
 *  - Host-pointer calls copy in and out; *_device calls take HBM pointers on the model's device
 *    and run on the model's stream.
 *  - A model is used by one thread at a time (the reference locks its GPU context per call,
 *    gpu_fea.cpp:256, 433, 459); different models may run concurrently on their own streams.
 */
#ifndef PSX_H
#define PSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psx_model_s* psx_model;

enum psx_status {
    PSX_OK = 0,
    PSX_ERR_ARGUMENT = 1,   /* bad sizes / values: VoxelFEA ctor and solve() throw invalid_argument (voxel_fea.cpp:141-146, 866-868) */
    PSX_ERR_NO_DEVICE = 2,  /* no usable GPU: GpuFeaSolver ctor throws runtime_error (gpu_fea.cpp:254) */
    PSX_ERR_NO_COARSE = 3,  /* coarsest level without a direct solve, m = 0 or m > 3000 (gpu_fea.cpp:255) */
    PSX_ERR_DEVICE = 4,     /* HIP runtime or kernel failure (the reference's error scopes, gpu.cpp:220-246) */
    PSX_ERR_INTERNAL = 5
};

/* Arithmetic of a solve.
 *  PSX_PREC_FP64   plain FP64 MG-PCG: FP64 CG, FP64 operator, FP64 V-cycle (the comparator)
 *  PSX_PREC_REF    the reference CPU algorithm (voxel_fea.cpp:863-927): FP64 CG, FP32 V-cycle with r
 *                  scaled to order one (voxel_fea.cpp:835-861)
 *  PSX_PREC_MIXED  FP32 CG and operator with FP64 dots and alpha/beta, FP32 V-cycle, FP64 true
 *                  residual and residual-corrected refinement (reliable updates), FP64 coarsest solve */
enum psx_precision { PSX_PREC_FP64 = 0, PSX_PREC_REF = 1, PSX_PREC_MIXED = 2 };

/* Replaces VoxelFEA(dims, density, nu, bc, coarsestMaxDof, diagAdd) (voxel_fea.hpp:136-137). */
typedef struct {
    int nx, ny, nz;
    const float* density;       /* [nx*ny*nz], values in [0, 1], 0 = empty */
    double nu;                  /* Poisson's ratio in (-1, 0.5) */
    const uint8_t* bc;          /* [3*(nx+1)*(ny+1)*(nz+1)], 1 = held at zero */
    const double* diag_add;     /* optional [3*nodes] shift added to K's diagonal, or NULL */
    int64_t coarsest_max_dof;   /* coarsen until free DOFs <= this (default 1100, gpu_fea.hpp:15) */
    int max_levels;             /* 0 = the reference's cap of 12 (voxel_fea.cpp:176) */
    int build_on_device;        /* 1: build the hierarchy (levels, Galerkin coarse elements, coarsest
                                   inverse) with device kernels instead of on the host CPU; the level
                                   data are the host build's bit for bit, the coarsest inverse agrees
                                   to rounding. No counterpart in the reference (its VoxelFEA builds on
                                   the CPU, voxel_fea.cpp:138-183). */
} psx_model_desc;

/* Replaces SolveOptions (voxel_fea.hpp:112-117). */
typedef int (*psx_progress_fn)(int iteration, double rel_residual, void* user); /* nonzero cancels */
typedef struct {
    double tol;                 /* relative residual ||f - K u|| / ||f_free|| (default 1e-6) */
    int max_iter;               /* default 500 */
    int precision;              /* enum psx_precision */
    int check_every;            /* iterations queued between host reads of the state (0 = 8) */
    int fold_every;             /* PSX_PREC_MIXED: true residual at least every N iterations (0 = 20) */
    psx_progress_fn on_progress;
    void* user;
    int cheb_degree;            /* Chebyshev smoother degree 1..6 (0 = 2, the reference's, voxel_fea.cpp:692-725) */
    int mfma_smoother;          /* PSX_PREC_MIXED, bit flags (0 = FP32 VALU everywhere): 1 = the V-cycle's levels
                                   without own element matrices as the element-by-element product on the
                                   matrix cores (two BF16 pieces of x - c and of Kb, FP32 accumulation),
                                   and the child-scale level likewise, its own elements' products
                                   sum_c s_c M_c (x_e - x_e0) over the 8 child matrices in a matrix-core
                                   pass of their own;
                                   4 = the outer FP32 operator as the same product with three pieces each;
                                   2 = levels with stored own element matrices apply them on the matrix
                                   cores (one v_mfma_f32_16x16x32_bf16 per node and voxel over 16 load
                                   cases: BF16 element rows and element-corner differences, FP32
                                   accumulation) */
    int eager;                  /* 1: launch every kernel from the host; 0: replay a captured HIP graph of
                                   check_every iterations (the same kernels and arguments) */
    int pre_degree;             /* Chebyshev degree of the pre-smoothing steps (0 = cheb_degree); the
                                   post-smoothing keeps cheb_degree */
} psx_solve_opts;

/* Replaces SolveResult's scalars (voxel_fea.hpp:119-124). */
typedef struct {
    int iterations;
    double residual;            /* final true relative residual, FP64 */
    int converged;              /* residual <= 10 tol and not cancelled (voxel_fea.cpp:925) */
    int cancelled;
    int broken;                 /* p.q <= 0: loss of positive definiteness (voxel_fea.cpp:904) */
    int folds;                  /* PSX_PREC_MIXED: FP64 true-residual refinements */
    double device_ms;           /* wall time of the device solve */
} psx_solve_info;

typedef struct {
    int nx, ny, nz;
    int64_t n_nodes, n_dof, free_dof, n_elems, n_own;
    double lmax;                /* largest eigenvalue estimate of D^-1 K (0 = not estimated yet) */
} psx_level_info;

const char* psx_last_error(void);
/* 1 when a GPU is usable (the reference's gpuAvailable(), gpu.cpp:250-257). */
int psx_available(void);
/* Name of the current HIP device (the reference's gpuName(), gpu.hpp:26). */
int psx_device_name(char* buf, size_t len);

/* Builds the hierarchy and uploads it (VoxelFEA ctor voxel_fea.cpp:138-183 + GpuFeaSolver ctor
 * gpu_fea.cpp:252-423). */
int psx_create(const psx_model_desc* desc, psx_model* out);
void psx_destroy(psx_model m);

int psx_num_levels(psx_model m);
int psx_get_level_info(psx_model m, int level, psx_level_info* out);
/* The level data a model's kernels read, downloaded from the device (any pointer may be NULL):
 * held-or-inactive mask per DOF (inverse diagonal 0), inverse diagonal [n_dof], node scale and node
 * info word [n_nodes], voxel words [nx*ny*nz], uniform element [576], own matrices [n_own*576];
 * child-scale form (levels with it): child scales [n_own*8] and two-scale records [n_own*4]; the
 * diagonal shift [n_dof] (levels that have one). The coarsest level's free-DOF map [m] and dense
 * inverse [m*m]. No counterpart in the reference: they let a test compare a device-built hierarchy
 * with the host build (VoxelFEA::Level, voxel_fea.hpp:37-78, is the reference's form of the data). */
int psx_model_level(psx_model m, int level, uint8_t* fixed, double* inv_diag, float* node_scale, int32_t* words, double* kb, float* own,
                    uint16_t* node_info);
int psx_model_level_child(psx_model m, int level, float* child_s, int32_t* child_two, double* diag_add);
int psx_model_coarse(psx_model m, int64_t* mdense, int32_t* free_idx, double* ainv);
/* Smoother eigenvalue estimates per level (the reference stores them in Level::lmax,
 * gpu_fea.cpp:412-417); set_lmax overrides one (e.g. the CPU model's value). */
int psx_estimate_smoothers(psx_model m);
int psx_set_lmax(psx_model m, int level, double lmax);

/* y = K x on a level, held DOFs zeroed unless raw (VoxelFEA::apply, voxel_fea.cpp:388). */
int psx_apply(psx_model m, int level, const double* x, double* y, int raw);
/* The same in FP32 (VoxelFEA::applyT<float>, voxel_fea.cpp:393). */
int psx_apply32(psx_model m, int level, const float* x, float* y, int raw);

/* y = K x as the FP64 and mixed solves apply it: own element matrices of a level whose finer level
 * has none in child-scale form, sum_c s_c P_c^T Kb P_c over its 8 children evaluated in the
 * working precision, instead of the reference's FP32-rounded stored entries (voxel_fea.cpp:362-375).
 * fp32 selects float x and y. No reference counterpart (the reference always streams its matrices). */
int psx_apply_child(psx_model m, int level, int fp32, const void* x, void* y, int raw);

/* y = K x with the matrix-core operator of the mixed-precision solve (FP32 vectors; levels without own
 * element matrices and the child-scale level): the element-by-element product sum_e s_e Kb (x_e - c), c
 * one node's value per 60 x 8 x zChunk tile, as v_mfma_f32_16x16x32_bf16 products of `pieces` BF16
 * pieces of x - c and of Kb with FP32 accumulation. pieces = 2 (hi + lo: the V-cycle's smoothing,
 * opts.mfma_smoother bit 0) or 3 (the outer operator, bit 2). On the child-scale level the own elements
 * add sum_c s_c M_c (x_e - x_e0), two-piece products over the 8 child matrices, in either form. Levels
 * with stored own matrices are rejected. No reference counterpart: the reference's operator is FP32
 * (voxel_fea.cpp:393, fea.wgsl:173-206). */
int psx_apply_mfma(psx_model m, int level, int pieces, const float* x, float* y);

/* One V-cycle z ~ K^-1 r, FP32 with r scaled to order one (GpuFeaSolver::precondition,
 * gpu_fea.cpp:431-450; VoxelFEA::precondition voxel_fea.cpp:830-861). */
int psx_precondition(psx_model m, const double* r, double* z);
int psx_precondition_device(psx_model m, const double* d_r, double* d_z);
/* nvec V-cycles in one set of launches (a block preconditioner for LOBPCG, whose FullPreconditioner
 * takes a block of residuals, eigen.cpp:508-553; the reference runs one V-cycle per vector): r and z
 * hold nvec x n_dof values, vector-major, and each z equals psx_precondition of its r alone. */
int psx_precondition_multi(psx_model m, int nvec, const double* r, double* z);
int psx_precondition_multi_device(psx_model m, int nvec, const double* d_r, double* d_z);

/* K u = f (GpuFeaSolver::solve gpu_fea.cpp:452-574, VoxelFEA::solve voxel_fea.cpp:863-927).
 * x0 may be NULL (zero start). opts may be NULL (defaults). */
int psx_solve(psx_model m, const double* f, const double* x0, const psx_solve_opts* opts, double* u, psx_solve_info* info);
int psx_solve_device(psx_model m, const double* d_f, const double* d_x0, const psx_solve_opts* opts, double* d_u, psx_solve_info* info);

/* nrhs independent load cases on one model in one set of launches: every kernel of the V-cycle and
 * the CG runs all load cases (one launch per level and step), and each case converges and stops on
 * its own state, so each solution equals the one psx_solve gives for it alone. f, u (and x0 when not
 * NULL) are nrhs x n_dof, load-case major; info receives nrhs entries. The reference solves its load
 * cases one after another (GpuFeaSolver::solve per case, gpu_fea.cpp:452-574; the topology and
 * eigen loops call it per vector, eigen.cpp:508-553). */
int psx_solve_multi(psx_model m, int nrhs, const double* f, const double* x0, const psx_solve_opts* opts, double* u, psx_solve_info* info);
int psx_solve_multi_device(psx_model m, int nrhs, const double* d_f, const double* d_x0, const psx_solve_opts* opts, double* d_u,
                           psx_solve_info* info);

/* Load cases of several models in one set of launches: model i solves nrhs[i] load cases (d_f[i],
 * d_u[i] and, when d_x0 and d_x0[i] are not NULL, d_x0[i]: nrhs[i] x its n_dof, device memory), and
 * each step of an iteration is one launch for every model whose run takes it (the grids of the
 * models' own launches side by side). Each model's solutions, iterations and residuals are bit for
 * bit those of psx_solve_multi_device on it alone; runs leave the pack as they stop. info receives
 * sum(nrhs) entries, model-major. At most 64 distinct models; opts->on_progress is not called. No
 * reference counterpart: the reference solves one part at a time (gpu_fea.cpp:452-574); for the
 * many small parts of a serving stream, whose kernels cannot fill the device, this pays one launch
 * per step for the whole pack. */
int psx_solve_packed(int n_models, const psx_model* models, const int* nrhs, const double* const* d_f, const double* const* d_x0,
                     const psx_solve_opts* opts, double* const* d_u, psx_solve_info* info);

/* Grid enclosing a box with n voxels along its longest side plus pad voxels each side (gridForBox,
 * voxelize.cpp:14-23): origin, voxel size h and dims. */
int psx_grid_for_box(const double mn[3], const double mx[3], int n, int pad, double origin[3], double* h, int dims[3]);
/* Filled fraction per voxel of a triangle mesh (voxelize, voxelize.cpp:25-128; the same fractions bit
 * for bit): pos = 3 floats per vertex, index = 3 vertex indices per triangle, sub^3 samples per voxel,
 * each solid when at least two of its three axis rays find it inside (even-odd). frac has
 * dims[0] * dims[1] * dims[2] entries, index i + nx (j + ny k). Runs on the device. */
int psx_voxelize(const float* pos, int64_t n_vertices, const uint32_t* index, int64_t n_tris, const double origin[3], double h, const int dims[3], int sub,
                 float* frac);

/* ---- topology optimization (optimizeTopology, topology.cpp:104-183) ----
 * SIMP densities on the voxels with fill > 0 (kept voxels stay 1), compliance minimised for a volume
 * budget by optimality-criteria updates with a density filter of radius rmin voxels. Each iteration
 * builds the part's hierarchy from fill * (xmin + (1 - xmin) x^penal), xmin = 1e-3, solves on the device
 * (warm-started from the previous solution) and runs the energies, filter, sensitivities and the
 * volume bisection as kernels on HBM-resident arrays. Zero fields take the defaults. */
/* nonzero stops; density = the iteration's filtered densities per voxel (TopologyIter::xPhys scattered to
 * the voxels, 0 outside the part), valid during the call */
typedef int (*psx_topology_fn)(int it, double compliance, double volume, double change, const float* density, void* user);
typedef struct {
    double vol_frac;            /* volume budget as a fraction of the solid volume (0 = 0.4) */
    double penal;               /* SIMP exponent (0 = 3) */
    double rmin;                /* filter radius in voxels (0 = 1.5) */
    double move;                /* largest change of a density per iteration (0 = 0.2) */
    double tol;                 /* stop when the largest change is below this, after iteration 5 (0 = 0.01) */
    int max_iter;               /* 0 = 40 */
    int precision;              /* enum psx_precision of the static solves */
    double solve_tol;           /* relative residual of the static solves (0 = 1e-8) */
    int64_t coarsest_max_dof;   /* 0 = 1100 */
    psx_topology_fn on_iter;    /* TopologyOptions::onIter (topology.hpp:48), or NULL */
    void* user;
    int solve_max_iter;         /* CG iteration cap of the static solves (0 = 1000) */
} psx_topology_opts;
/* dims, fill [nx*ny*nz], keep [nx*ny*nz] (1 = stays solid), bc [3*nodes], f [3*nodes], nu as for
 * psx_create. density receives the final filtered densities per voxel (0 outside the part); history
 * receives 4 doubles per iteration (iteration, compliance f.u, volume fraction, change) and must hold
 * 4 * max_iter; iterations receives the count; kept_fraction (may be NULL) the kept share of the
 * solid volume. */
int psx_optimize_topology(const int dims[3], const float* fill, const uint8_t* keep, const uint8_t* bc, const double* f, double nu,
                          const psx_topology_opts* opts, float* density, double* history, int* iterations, double* kept_fraction);
/* Strain energies u_e^T K0 u_e of the finest level's elements (elementEnergies, topology.cpp:92-102),
 * elements in voxel order, n_elems values, FP64. */
int psx_element_energies(psx_model m, const double* u, double* out);
/* The density filter (applyFilter, topology.cpp:54-63) or its transpose (applyFilterT, topology.cpp:
 * 65-79) over the voxels with fill > 0 in voxel order: x and out hold one value per such voxel. */
int psx_topology_filter(const int dims[3], const float* fill, double rmin, int transpose, const double* x, double* out);

/* ---- break test (breakTest, static_study.cpp:197-279) ----
 * Repeats: build and solve (warm-started from the last step's solution), nodal stresses, the failure
 * measure of every voxel (its worst corner stress, von Mises or first principal), the load factor
 * lambda = strength / largest measure, the crack (the worst voxels, at most round(0.02 n^(2/3)) of
 * the n voxels, each within 95 % of the largest) and the voxels the crack detaches from the held
 * nodes. Solves, stresses, measures, the largest displacement and the prune (a breadth-first search
 * over face neighbours from the held voxels) run on the device; the hierarchy is built on the host. */
enum psx_break_reason {
    PSX_BREAK_STEP_LIMIT = 0,     /* "step limit" */
    PSX_BREAK_SEPARATED = 1,      /* "separated" */
    PSX_BREAK_UNSUPPORTED = 2,    /* "unsupported load" */
    PSX_BREAK_NOT_CONVERGED = 3,  /* "solver did not converge" */
    PSX_BREAK_NO_STRESS = 4,      /* "no stress" */
    PSX_BREAK_STOPPED = 5         /* on_step returned nonzero (no reference counterpart) */
};
/* One step (BreakStep, static_study.hpp:52-60); the arrays are valid during the on_step call. */
typedef struct {
    int step;
    double lambda;              /* strength / the step's largest measure */
    double max_disp;            /* largest nodal displacement magnitude at lambda */
    int voxels;                 /* solid voxels of the step's model */
    const int32_t* cracked;     /* voxels cracked, worst first */
    int n_cracked;
    const int32_t* detached;    /* voxels cut off by the crack, increasing */
    int n_detached;
    const float* u;             /* 3 * nodes physical displacements at lambda (fields = 1, else NULL) */
    const float* node_vm;       /* nodes, von Mises at lambda (fields = 1, else NULL) */
    const uint8_t* active_node; /* nodes (fields = 1, else NULL) */
} psx_break_step;
typedef int (*psx_break_fn)(const psx_break_step* step, void* user); /* nonzero stops */
typedef struct {
    double strength;            /* failure stress, > 0 */
    int principal;              /* 1: first principal stress, 0: von Mises */
    int max_steps;
    double E, h;                /* Young's modulus and voxel size (u_physical = u / (E h)) */
    double tol;                 /* relative residual of the solves (0 = 1e-6, static_study.cpp:111) */
    int max_iter;               /* 0 = 3000 (the GPU engine's, static_study.cpp:125) */
    int precision;              /* enum psx_precision */
    int tie_order;              /* 0: equal measures in the reference's std::sort order (the measures of
                                   all voxels are sorted on the host, as the reference does); 1: equal
                                   measures by increasing voxel index (only the voxels within 95 % of
                                   the largest leave the device) */
    int fields;                 /* 1: fill psx_break_step's u, node_vm and active_node */
    psx_break_fn on_step;
    void* user;
} psx_break_opts;
typedef struct {
    int reason;                 /* enum psx_break_reason */
    int steps;
    int removed;                /* voxels the first prune removed */
    double peak;                /* largest lambda */
} psx_break_result;
/* part: dims, density, nu, bc and coarsest_max_dof as for psx_create (diag_add must be NULL); f the
 * 3 * nodes loads. */
int psx_break_test(const psx_model_desc* part, const double* f, const psx_break_opts* opts, psx_break_result* out);
/* pruneFloating (voxel_fea.cpp:1086-1119) on the device: d_density, d_out nx*ny*nz floats in HBM,
 * d_held one byte per node; removed receives the count of solid voxels set to 0. */
int psx_prune_floating_device(int nx, int ny, int nz, const float* d_density, const uint8_t* d_held, float* d_out, int* removed);
/* The same prune without allocating: no counterpart in the reference. It is for callers that cannot
 * allocate device or pinned memory (code recording a stream into a graph, tight serving loops,
 * harnesses that forbid it). Its search scratch is a static device array, one copy per GPU held until
 * the process ends (about 34 MB), so parts of at most 8388608 (2^23) voxels are accepted and larger
 * ones fail with PSX_ERR_ARGUMENT. Work runs on the null stream and the call returns when it is done.
 * Calls that overlap on one GPU share its scratch and race. The result equals
 * psx_prune_floating_device's bit for bit. */
int psx_prune_floating_device_noalloc(int nx, int ny, int nz, const float* d_density, const uint8_t* d_held, float* d_out, int* removed);

/* Support reactions: sum over held active DOFs of (K u)_raw - f (VoxelFEA::reactions,
 * voxel_fea.cpp:929-940). f may be NULL. */
int psx_reactions(psx_model m, const double* u, const double* f, double out[3]);

/* Nodal stresses of a solution u (normalised units, nDof values): the density-weighted average over
 * each active node's voxels of the corner stress (E / h) D Bbar u_e; von Mises always, the first and
 * third principal stresses when principals != 0 (p1, p3 may be NULL otherwise); nNodes floats each,
 * 0 at inactive nodes (VoxelFEA::stresses, voxel_fea.cpp:968-1037). */
int psx_stresses(psx_model m, const double* u, double E, double h, int principals, float* vm, float* p1, float* p3);
/* The same with HBM-resident u and outputs. */
int psx_stresses_device(psx_model m, const double* d_u, double E, double h, int principals, float* d_vm, float* d_p1, float* d_p3);
/* The averaged 6-component stress tensor per node, [node][xx yy zz xy yz zx] (VoxelFEA::stressTensors,
 * voxel_fea.cpp:1039-1084). */
int psx_stress_tensors(psx_model m, const double* u, double E, double h, float* out);

/* Per-level kernel timings, one line each (GpuFeaSolver::profile, gpu_fea.cpp:576-652). */
int psx_profile(psx_model m, int reps, char* buf, size_t len);

/* Device memory of the model's buffers in bytes (hierarchy + work vectors). */
int64_t psx_device_bytes(psx_model m);

/* ---- serving: a device-resident request loop with continuous batching ----
 * No reference counterpart: the reference solves one part at a time behind a global GPU lock
 * (gpu.hpp:43, gpu_fea.cpp:256-578). A server admits queued solve requests in arrival order as
 * slots (max_inflight) and device memory (hbm_budget) free up; each admitted request builds its
 * part's hierarchy (on the device unless asked otherwise). Parts of at most pack_max_dof DOFs join
 * the pack, whose runs advance together with one shared launch per step for all of them (as
 * psx_solve_packed); larger parts solve on their own HIP stream, overlapping with the rest. Every
 * result is the same as the request's solve alone (per-request deterministic reductions, no shared
 * device state). */
typedef struct psx_server_s* psx_server;
typedef struct {
    int max_inflight;           /* admission workers: requests built or solved alone at once (0 = 4) */
    int64_t hbm_budget;         /* device bytes requests may hold at once (0 = 80 % of the free memory) */
    int pack_max;               /* requests in the pack at once (0 = 16, 1 = no packing, at most 64) */
    int64_t pack_max_dof;       /* largest part that joins the pack, in DOFs (0 = 100 000) */
} psx_server_opts;
typedef struct {
    psx_model_desc part;        /* copied at submit */
    int nrhs;                   /* load cases (>= 1) */
    const double* f;            /* nrhs x n_dof, copied at submit */
    const double* x0;           /* NULL, n_dof (every load case) or nrhs x n_dof (x0_per_rhs) */
    int x0_per_rhs;
    int64_t part_id;            /* >= 0: keep the part's model resident for later requests with the same
                                   id, which must then carry the same part (ids < 0: build afresh) */
    psx_solve_opts opts;        /* on_progress is not called */
} psx_request;
typedef struct {
    int status;                 /* PSX_OK or the error of the request */
    int iterations;             /* summed over the load cases */
    int max_iterations;
    int converged;              /* every load case */
    double residual;            /* largest final true relative residual */
    double queue_ms;            /* submit -> admitted */
    double build_ms;            /* hierarchy build + upload + smoother estimates (0 when resident) */
    double solve_ms;            /* all load cases incl. transfers */
    double total_ms;            /* submit -> done */
    int64_t device_bytes;       /* the model's device memory */
    int reused_part;            /* the model was resident */
    int packed;                 /* solved in the pack (shared launches) */
} psx_result;
int psx_server_create(const psx_server_opts* opts, psx_server* out);
/* Queues a request (non-blocking); ticket identifies it for psx_server_wait. */
int psx_server_submit(psx_server s, const psx_request* req, int64_t* ticket);
/* Blocks until the request is done; u (nrhs x n_dof, may be NULL) receives the solutions. The
 * return value is the request's status; the ticket is released. */
int psx_server_wait(psx_server s, int64_t ticket, double* u, psx_result* out);
/* Requests queued or running. */
int psx_server_pending(psx_server s);
/* Finishes the running requests, fails the queued ones, frees everything. */
void psx_server_destroy(psx_server s);

/* ---- host-side inspection (no device needed): the hierarchy exactly as it is uploaded ---- */
typedef struct psx_hierarchy_s* psx_hierarchy;
/* Element matrix K (576) and corner stresses D*Bbar (8 x 6 x 24) of the unit cube for E = 1
 * (hexElement, native/src/fea/hex8.cpp:104-155). */
int psx_hex_element(double nu, double* K, double* corner_stress);
/* Input rule before a re-solve of a damaged part: solid voxels (density > 0) not face-connected to a
 * voxel touching a held node are set to 0 in out (may alias density). held: per node (any DOF held,
 * static_study.cpp:17-21). Returns the number removed, or a negative status (pruneFloating,
 * voxel_fea.cpp:1086-1119). */
int psx_prune_floating(int nx, int ny, int nz, const float* density, const uint8_t* held, float* out);

/* ---- Input contract of a CAD part (host; the fill fractions come from psx_voxelize) ----
 * A surface mesh as the reference's Part holds it (native/src/core/mesh.hpp:37-57): float xyz per
 * vertex and 3 vertex indices per outward triangle; triangle normals and areas are derived as
 * computeTriangleData does (mesh.cpp:315-330), the bounding box as boundingBox (mesh.cpp:427-439). */
typedef struct {
    const float* vertices;      /* 3 * n_vertices */
    int64_t n_vertices;
    const uint32_t* tris;       /* 3 * n_tris */
    int64_t n_tris;
} psx_mesh;
/* Replaces typicalWallThickness(part, samples) (native/src/core/voxelize.cpp:190-217): the 20th
 * percentile of the wall thickness at `samples` area-weighted triangles, or +inf (1500 = default). */
int psx_wall_thickness(const psx_mesh* mesh, int samples, double* out);
/* Replaces thinWallDensity(part, grid, maxThickness) (voxelize.cpp:219-276): [nx*ny*nz]. */
int psx_thin_wall_density(const psx_mesh* mesh, const double origin[3], double h, const int dims[3], double max_thickness, float* out);
typedef struct {
    double wall_thickness;      /* typicalWallThickness of the part */
    int sub;                    /* voxelizer samples per voxel edge */
    int voxel_count, thin_voxels;
    int used_shell;             /* the thin-wall layer was added (wall < 2.5 h) */
} psx_structural_info;
/* Replaces StructuralModel(part, resolution) (native/src/fea/structural.cpp:19-64) on the grid
 * psx_grid_for_box(bbox, resolution) gives: voxel densities [nx*ny*nz] (fill fraction >= 0.3, thin
 * walls as a connected layer), active and surface nodes [(nx+1)(ny+1)(nz+1)] (either may be NULL).
 * Needs the device (the voxelizer). */
int psx_structural_model(const psx_mesh* mesh, const double origin[3], double h, const int dims[3], float* density, uint8_t* active_node,
                         uint8_t* surface_node, psx_structural_info* info);
/* A selected surface area (structural.hpp:18-21): triangles, optionally clipped to a sphere. */
typedef struct {
    const int32_t* tris;
    int64_t n_tris;
    int has_clip;
    double clip_center[3];
    double clip_radius;
} psx_patch;
typedef struct {                /* Fixture (structural.hpp:23-26) */
    const char* name;
    const psx_patch* patches;
    int n_patches;
} psx_fixture;
enum psx_load_type { PSX_LOAD_FORCE = 0, PSX_LOAD_PRESSURE = 1, PSX_LOAD_WIND = 2 };
typedef struct {                /* Load (structural.hpp:28-35) */
    int type;                   /* enum psx_load_type */
    const char* name;
    double magnitude;           /* N (force), MPa (pressure) */
    double dir[3];              /* force direction (normalised here) */
    const psx_patch* patches;
    int n_patches;
    const float* forces;        /* wind: N per part triangle (xyz), n_forces = 3 * n_tris */
    int64_t n_forces;
} psx_load;
typedef struct {                /* the Material fields assemble validates and uses (materials.hpp) */
    double E, nu, yield, uts, density;
    int brittle;
} psx_material;
typedef struct {
    int fixed_nodes;
    double total[3];            /* applied load, N */
    int n_warnings;
} psx_assembly_info;
/* Replaces StructuralModel::assemble(fixtures, loads, gravity, material, toMeters)
 * (structural.cpp:107-209) for the model of `density` on the grid: held DOFs bc and nodal forces f
 * [3*(nx+1)(ny+1)(nz+1)]; warnings newline-separated into the buffer (may be NULL). Errors as the
 * reference's invalid_argument (material, units, magnitudes, directions, wind force sizes). */
int psx_assemble(const psx_mesh* mesh, const double origin[3], double h, const int dims[3], const float* density, const psx_fixture* fixtures,
                 int n_fixtures, const psx_load* loads, int n_loads, int gravity, const psx_material* material, double to_meters, uint8_t* bc,
                 double* f, psx_assembly_info* info, char* warnings, size_t warnings_len);
/* Builds the hierarchy on the host only (VoxelFEA ctor, voxel_fea.cpp:138-183). */
int psx_hierarchy_create(const psx_model_desc* desc, psx_hierarchy* out);
void psx_hierarchy_destroy(psx_hierarchy h);
int psx_hierarchy_num_levels(psx_hierarchy h);
int psx_hierarchy_level_info(psx_hierarchy h, int level, psx_level_info* out);
/* Any pointer may be NULL. fixed/inv_diag: [n_dof]; node_scale: [n_nodes]; words: [voxels];
 * kb: [576]; own: [n_own * 576] (row-major per element, elements in voxel order). */
int psx_hierarchy_level(psx_hierarchy h, int level, uint8_t* fixed, double* inv_diag, float* node_scale, int32_t* words, double* kb, float* own);
/* Coarsest free DOFs: free_idx [m] (level DOF of each dense index) and the dense inverse [m * m]. */
int psx_hierarchy_coarse(psx_hierarchy h, int64_t* m, int32_t* free_idx, double* ainv);

#ifdef __cplusplus
}
#endif

#endif /* PSX_H */
