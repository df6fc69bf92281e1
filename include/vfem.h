/*
 * vfem.h -- C-ABI of the MI355X matrix-free multigrid-preconditioned CG solver
 * for voxel linear elasticity.
 *
 * Drop-in boundary for the GPU solver path of hirademirkol/VulkanFEM.  Every
 * entry point names the reference interface it replaces (file:line).  Plain
 * pointers and sizes only; all buffers are caller-owned host memory unless the
 * name says "device".  Every call is synchronous and returns a vfem_status
 * (0 = success); nothing throws across this boundary.  One vfem_system per host
 * thread; the library owns all device memory it allocates.
 *
 * Conventions (identical to the reference):
 *   - voxels: int32, value 1 = solid, x fastest: voxels[x + nx*(y + ny*z)]
 *     (src/inout.cpp:16-36, include/utils.hpp:85-86).
 *   - compact node ids: corners of solid voxels numbered in ascending grid-linear
 *     order x + y*(nx+1) + z*(nx+1)*(ny+1) (src/FEM.cpp:88-99,179-184).
 *   - DOF vectors: 3 per compact node, dof = 3*node + component
 *     (src/FEM.cpp:580-582).  A fixed node pins all three DOFs
 *     (include/MatrixFreeSparse.hpp:210-213).
 *   - ke: 24x24 row-major element stiffness; only its upper triangle is read
 *     (src/FEM.cpp:200-210).  Local DOF 3*L + c, local node L = x + 2y + 4z.
 */
#ifndef VFEM_H
#define VFEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vfem_system vfem_system;
typedef struct vfem_unet vfem_unet;

enum vfem_status {
    VFEM_OK = 0,
    VFEM_ERR_ARGUMENT = 1,      /* invalid argument (null pointer, size, option) */
    VFEM_ERR_EMPTY = 2,         /* the voxel model has no solid voxel */
    VFEM_ERR_DEVICE = 3,        /* a HIP call failed (see vfem_last_error) */
    VFEM_ERR_SINGULAR = 4,      /* the coarsest-level matrix is not positive definite */
    VFEM_ERR_NOT_CONVERGED = 5, /* max_iter reached before the tolerance */
    VFEM_ERR_BREAKDOWN = 6,     /* CG breakdown (p.Ap <= 0 or a non-finite scalar) */
    VFEM_ERR_UNSUPPORTED = 7    /* option combination not implemented */
};

/* arithmetic of the multigrid cycle and of the CG matvec */
enum vfem_precision {
    VFEM_FP64 = 0,  /* everything FP64 ("plain FP64 MGPCG") */
    VFEM_FP32 = 1,  /* V-cycle and CG matvec in FP32; CG dots, scalars, x/r updates FP64 */
    VFEM_BF16 = 2,  /* as VFEM_FP32 with the V-cycle's level-0 matvec on BF16 MFMA */
    VFEM_MIXED = 3  /* FP64 vectors, CG and cycle; every level-0 matvec of the cycle and of the CG
                       operator in FP32 arithmetic on the matrix cores (BF16 splits of Ke and of the
                       FP64 element differences x - x_centre); FP64 true residual every
                       true_res_every iterations; flexible CG */
};

/* form of the level-0 operator (vfem_set_level0_kernel) */
enum vfem_level0 {
    VFEM_L0_AUTO = 0,     /* the stencil form where it is available and faster (default) */
    VFEM_L0_ELEMENT = 1,  /* element form: Ke times each element's 24 values, on the matrix cores */
    VFEM_L0_STENCIL = 2,  /* assembled 27-point stencil plus a recomputation of the surface nodes;
                             sum-factorised (one wavefront per node row) when the stencil is a
                             tensor-product form */
    VFEM_L0_STENCIL_DIRECT = 3, /* as VFEM_L0_STENCIL with the structured 27-point kernel (no sum
                                   factorisation), for comparison */
    VFEM_L0_STENCIL_PLANES = 4  /* as VFEM_L0_STENCIL with the sum-factorised kernel that stages
                                   16 x 18-node planes in LDS (the row kernel's predecessor), for
                                   comparison */
};

/* multigrid cycle */
enum vfem_cycle {
    VFEM_CYCLE_REFERENCE = 0, /* additive cycle of MultigridPreconditioner.hpp:158-193 (omega = 1) */
    VFEM_CYCLE_DAMPED = 1     /* multiplicative V(1,1) cycle with damped Jacobi (omega) */
};

typedef struct {
    double tol;              /* relative residual |r| < tol |f| (FEMGPU.cpp:416,492); default 1e-8 */
    int32_t max_iter;        /* default 30000 (FEMGPU.cpp:407) */
    int32_t precision;       /* enum vfem_precision */
    int32_t cycle;           /* enum vfem_cycle */
    double omega;            /* damped Jacobi weight (VFEM_CYCLE_DAMPED), default 0.5: the V(1,1)
                                cycle is a positive definite preconditioner only while
                                omega * lambda_max(D^-1 A) < 2 on every level, and the trilinear
                                elasticity element has lambda_max ~ 3.8 around surface nodes */
    int32_t true_res_every;  /* mixed precision: FP64 true residual every k iterations, default 20 */
    int32_t check_every;     /* iterations between host convergence polls, default 8 */
    int32_t use_graph;       /* capture one CG iteration in a hipGraph (default 1) */
    int32_t warm_start;      /* 1: start from the u passed in instead of zero */
    int32_t op_passes;       /* reduced precision: BF16 products of the CG operator matvec, 3 or 6; 0 = 6 */
    int32_t cycle_passes;    /* reduced precision: BF16 products of the level-0 cycle matvec, 1, 3 or 6;
                                0 = 6 (VFEM_FP32) or 1 (VFEM_BF16) */
    int32_t learned;         /* 1: damped cycle with the attached learned coarse correction (flexible CG) */
    double stall_ratio;      /* learned: fall back to the plain cycle for the rest of the solve when the
                                residual norm falls by less than this factor per iteration over a
                                polling interval; default 0.9 */
    int32_t op_fp64;         /* VFEM_FP32 / VFEM_BF16: 1 = the CG operator (A p) in FP64 with FP64 p, the
                                cycle in reduced precision; default 0 (operator in the cycle's precision) */
    int32_t coarse_fp64;     /* VFEM_FP32 / VFEM_BF16 with the damped cycle: 1 = only level 0 of the cycle
                                in reduced precision, the coarser levels in FP64; default 0 */
    int32_t cheb_level;      /* damped cycle: levels >= cheb_level (>= 1; the coarsest level keeps its
                                exact solve) smooth with cheb_degree Chebyshev steps before and after
                                the coarse correction instead of one damped Jacobi step, fitted to
                                [lmax / cheb_ratio, lmax] of D^-1 A (vfem_level_lmax); the cycle stays
                                symmetric positive definite.  0 = damped Jacobi on every level (default) */
    int32_t cheb_degree;     /* Chebyshev steps, 1..8; default 2 */
    double cheb_ratio;       /* > 1; default 60 */
} vfem_options;

typedef struct {
    int32_t iterations;      /* completed CG iterations, the converged step excluded (FEMGPU.cpp:471,494) */
    int32_t converged;
    double rel_residual;     /* true FP64 |f - A u| / |f| at exit */
    double rec_residual;     /* recursively updated |r| / |f| at exit */
    double solve_ms;         /* device time of the solve (hipEvents) */
    double setup_ms;         /* host + device setup time of vfem_assemble */
    int32_t true_res_count;  /* FP64 true-residual recomputations done */
    int32_t status;
    int32_t learned_iters;   /* iterations preconditioned with the learned correction */
    int32_t fell_back;       /* 1 if the solve fell back to the plain cycle */
} vfem_stats;

/* Fill *o with the defaults listed above. */
void vfem_default_options(vfem_options *o);

/* Replaces assembleSystemMatrix<double>(voxelModel, dims, Ke, fixedNodes, numLevels,
 * skipLevels) (include/FEM.hpp:33-34, src/FEM.cpp:157-517): numbers the nodes and
 * builds the multigrid hierarchy (levels may stop early, FEM.cpp:465-470) on the
 * current HIP device -- element and node masks, compact numbering, fixed-node
 * propagation and inverse diagonals by grid kernels, the coarsest-level inverses
 * by blocked Gauss-Jordan elimination.  fixed_nodes: compact node ids.
 * skip_levels: 0 or 1 (FEM.cpp:266-272). */
int vfem_assemble(const int32_t *voxels, int32_t nx, int32_t ny, int32_t nz, const double *ke,
                  const int64_t *fixed_nodes, int64_t n_fixed, int32_t num_levels, int32_t skip_levels,
                  vfem_system **out);

/* Releases everything vfem_assemble allocated. */
void vfem_destroy(vfem_system *sys);

/* MatrixFreeSparse<double>::rows() (include/MatrixFreeSparse.hpp:57): 3 * #nodes. */
int64_t vfem_num_dofs(const vfem_system *sys);

/* MatrixFreeSparse::numLevels after the early stop (include/MatrixFreeSparse.hpp:86). */
int32_t vfem_num_levels(const vfem_system *sys);

/* Node-grid extents of level `level` (no reference counterpart: the reference's
 * level dims are only known inside FEM.cpp:253-259). */
int vfem_level_extent(const vfem_system *sys, int32_t level, int32_t *nx, int32_t *ny, int32_t *nz);

/* Per-level sizes: elements, used nodes, fixed nodes of level `level`. */
int vfem_level_info(const vfem_system *sys, int32_t level, int64_t *n_elems, int64_t *n_nodes, int64_t *n_fixed);

/* Upper end of the Chebyshev smoother's interval on level `level` (no reference
 * counterpart: the reference smooths with D^-1 only): a bound of lambda_max(D^-1 A)
 * that vfem_assemble forms by 30 FP64 power steps from a fixed start, times 1.1.
 * 0 on level 0 and on the coarsest level, which never use it. */
int vfem_level_lmax(const vfem_system *sys, int32_t level, double *out);

/* Replaces applyBoundaryConditions (include/FEM.hpp:63-64, src/FEM.cpp:568-584,
 * MATRIX_FREE branch): f[3*node + c] = loads[3*k + c] for node = loaded_nodes[k]. */
int vfem_apply_boundary_conditions(double *f, int64_t n_dofs, const int64_t *loaded_nodes, const double *loads,
                                   int64_t n_loaded);

/* Replaces solveWithKompute<double>(systemMatrix, f, u) (include/FEMGPU.hpp:11-12,
 * src/FEMGPU.cpp:32-554) with a device-resident solve: f, u are host arrays of
 * vfem_num_dofs doubles.  opts may be NULL (defaults); stats may be NULL. */
int vfem_solve(vfem_system *sys, const double *f, double *u, const vfem_options *opts, vfem_stats *stats);

/* Repeated solves on a changing geometry (no reference counterpart; the reference
 * is the solver of a topology-optimisation loop, Readme.md:76-77): sets the
 * resident solution of `to` (its vfem_solve_resident start vector under
 * warm_start = 1) from the resident solution of `from`, node by node at the same
 * grid position: nodes both systems use keep their displacement, nodes only `to`
 * uses and `to`'s fixed nodes start at zero.  Both systems must be assembled on
 * the same device from voxel grids of the same extents. */
int vfem_map_solution(const vfem_system *from, vfem_system *to);

/* Multi-RHS block (no reference counterpart; the reference solves one load case,
 * include/FEMGPU.hpp:11-12): nrhs load cases on one geometry, f and u are
 * nrhs x ndof row-major.  Up to 256 columns run as one block -- every kernel of
 * an iteration covers all columns, each column keeps its own FP64 CG scalars and
 * stops on its own tolerance; stats receives nrhs entries.  Each column's
 * solution equals the single-RHS solve of that column to the solver tolerance. */
int vfem_solve_batched(vfem_system *sys, const double *f, double *u, int32_t nrhs, const vfem_options *opts,
                       vfem_stats *stats);

/* Resident-data split of vfem_solve_batched for one block of 1..256 columns:
 * upload nrhs right-hand sides (nrhs x ndof row-major), solve them on the device,
 * download the solutions. */
int vfem_set_rhs_block(vfem_system *sys, const double *f, int32_t nrhs);
int vfem_solve_resident_block(vfem_system *sys, int32_t nrhs, const vfem_options *opts, vfem_stats *stats);
int vfem_get_solution_block(vfem_system *sys, double *u, int32_t nrhs);

/* Resident-data split of vfem_solve for timing without PCIe transfers:
 * vfem_set_rhs uploads f, vfem_solve_resident solves on the device and keeps u
 * there, vfem_get_solution downloads it. */
int vfem_set_rhs(vfem_system *sys, const double *f);
int vfem_solve_resident(vfem_system *sys, const vfem_options *opts, vfem_stats *stats);
int vfem_get_solution(vfem_system *sys, double *u);

/* Replaces the matrix-free product y = A x with fixed rows zeroed
 * (include/MatrixFreeSparse.hpp:191-215; GPU: shaders/MatxVec.comp +
 * shaders/FixedNodes.comp).  precision selects the arithmetic (vfem_precision):
 * FP64, FP32, BF16 (1 BF16 product on the matrix cores), MIXED (6 BF16 products
 * on FP64 input differences).  FP64 and FP32 use the level-0 form chosen with
 * vfem_set_level0_kernel. */
int vfem_matvec(vfem_system *sys, const double *x, double *y, int32_t precision);

/* Operator form (enum vfem_level0; no reference counterpart, whose MatxVec.comp
 * is the element form).
 *   VFEM_L0_STENCIL: the assembled stencil on every level whose full-element
 *     matrix has its structure (level 0: Ke; a damped-cycle coarse level: the
 *     Galerkin matrix of the full child pattern), sum-factorised where it is a
 *     tensor-product form.
 *   VFEM_L0_STENCIL_DIRECT: as VFEM_L0_STENCIL with the structured 27-point
 *     kernel everywhere (no sum factorisation), for comparison.
 *   VFEM_L0_STENCIL_PLANES: as VFEM_L0_STENCIL with the sum-factorised kernel
 *     that stages each plane of a 14 x 16-node tile in LDS, for comparison.
 *   VFEM_L0_AUTO (default): the stencil form on level 0 and on coarse levels of
 *     at least 100000 nodes; smaller levels are latency-bound and faster in
 *     element form.
 *   VFEM_L0_ELEMENT: the element form everywhere.
 * The damped cycle's restriction and prolongation between div-2 levels follow
 * the same choice: plane-tiled kernels (both stencil kinds, or VFEM_L0_AUTO on
 * fine levels of at least 500000 nodes) or row-segment kernels.  The stencil
 * kinds fail with VFEM_ERR_UNSUPPORTED when level 0 has no stencil form.
 * vfem_level0_kernel returns the form in effect on level 0 and
 * vfem_level_operator_form on `level` (VFEM_L0_ELEMENT, VFEM_L0_STENCIL for the
 * sum-factorised row kernel, VFEM_L0_STENCIL_PLANES for the plane-staged one or
 * VFEM_L0_STENCIL_DIRECT for the structured one), or -1
 * for a null system or a level out of range. */
int vfem_set_level0_kernel(vfem_system *sys, int32_t kind);
int32_t vfem_level0_kernel(const vfem_system *sys);
int32_t vfem_level_operator_form(const vfem_system *sys, int32_t level);

/* Multi-RHS form of vfem_matvec (no reference counterpart): y_k = A x_k for nrhs
 * columns (x, y nrhs x ndof row-major).  FP32/BF16 use the multi-RHS matrix-core
 * kernel (16 columns per MFMA N tile) with the tile shape last chosen by
 * vfem_autotune_block or vfem_set_block_config. */
int vfem_matvec_block(vfem_system *sys, const double *x, double *y, int32_t nrhs, int32_t precision);

/* Tile shapes of the multi-RHS kernel: elements per layer tex x tey, node planes
 * per block mz.  The last index (vfem_block_config_count() - 1) is not a shape
 * of that kernel but the stencil form of the level-0 operator (tex = 0, tey = 1
 * column per block row, mz planes per block), the other candidate of a
 * full-accuracy FP32 batch; vfem_set_block_config with it selects the stencil
 * form for the FP32 batch, with any other index the matrix-core shape. */
int32_t vfem_block_config_count(void);
int vfem_block_config(int32_t index, int32_t *tex, int32_t *tey, int32_t *mz);
int vfem_set_block_config(vfem_system *sys, int32_t index);

/* Autotuner of the multi-RHS level-0 operator (no reference counterpart): times
 * every tile shape of the matrix-core kernel for nrhs columns and `passes` BF16
 * products (1, 3 or 6) over `reps` launches on a deterministic pseudo-random
 * block, rejects any shape whose product differs from shape 0's; at 6 products
 * (FP32 accuracy) it also times the stencil form (last index), accepted when its
 * product agrees with shape 0's to 1e-5 of the largest entry.  The fastest is
 * kept for this system -- at 6 products it is the operator the FP32 batched solve
 * runs -- and returned in *best; ms_out (vfem_block_config_count() entries, may
 * be null) receives the mean time of each candidate (-1: rejected or not timed). */
int vfem_autotune_block(vfem_system *sys, int32_t nrhs, int32_t passes, int32_t reps, double *ms_out,
                        int32_t *best);

/* Replaces MultigridPreconditioner::solve (include/MultigridPreconditioner.hpp:158-193;
 * GPU: FEMGPU.cpp:500-534): z = M r with the cycle/precision of opts. */
int vfem_precondition(vfem_system *sys, const double *r, double *z, const vfem_options *opts);

/* Replaces shaders/VecDotVec.comp: *out = a . b over vfem_num_dofs entries (FP64). */
int vfem_dot(vfem_system *sys, const double *a, const double *b, double *out);

/* Allocation-free variant of vfem_dot (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  *out = a . b over n
 * entries, a and b DEVICE pointers to n doubles on the current device, out a
 * host pointer; no vfem_system is needed.  The same kernels as vfem_dot (same
 * grid for the same n, so the same result bit for bit), run on the current
 * device's null stream; synchronous.  It calls no hipMalloc, hipFree or
 * allocating routine: the block partials and the finalizer's state live in a
 * fixed per-device scratch array of 32.9 KB (2 x 2048 partials + one state)
 * that stays allocated until the process ends.  The grid is capped at 2048
 * blocks, so any n >= 0 fits it.  Every call on a device shares that device's
 * scratch: calls that overlap on one device (other host threads) race. */
int vfem_dot_noalloc(const double *a, const double *b, int64_t n, double *out);

/* Kernel timing hook (no reference counterpart): runs `reps` launches of one
 * kernel on the library's stream between two hipEvents and returns the mean
 * duration.  which: 0 = level-0 matvec of the CG operator in the form chosen by
 * vfem_set_level0_kernel (stencil form: the stencil kernel and the surface pass;
 * element form: FP64 exact products and FP32 three-way BF16 splits, 6 products,
 * on the matrix cores; BF16: one BF16 product; MIXED: matrix cores on FP64
 * input), 1 = one damped V(1,1) cycle (default omega), 2 = one CG iteration with
 * that cycle, 3 = level-0 element-form matvec with the two-way BF16 split (3
 * products), 4 = plain copy of one FP32 level-0 vector (12 bytes read and 12
 * written per node: the reference for HBM byte counters), 5 = copy of one FP64
 * level-0 vector, 6 = FP64 dot product of two level-0 vectors, 7 = the level-0
 * stencil kernel alone, without the surface pass (FP32 or FP64; fails with
 * VFEM_ERR_UNSUPPORTED when level 0 has no stencil form), 8 = the same kernel as
 * it runs inside CG iterations: `reps` iterations as in 2 (after one untimed),
 * launched eagerly, with each launch of the level-0 stencil kernel without a
 * fused smoothing epilogue (the CG operator and the cycle's first; the cycle's
 * second carries the post-smoothing where the stencil form allows it) bracketed
 * by its own events; the mean over those launches.  Kind 7 repeats one launch on the same vectors,
 * which the last-level cache can partly hold; kind 8 sees the vectors as the
 * solve leaves them; VFEM_ERR_UNSUPPORTED when the plan's iterations launch
 * no level-0 stencil kernel (element form at level 0).  reps is at most 4096. */
int vfem_time_kernel(vfem_system *sys, int32_t which, int32_t precision, int32_t reps, double *avg_ms);

/* Host-only inspection of the hierarchy vfem_assemble builds (no device call; no
 * reference counterpart -- a test hook for the setup logic of FEM.cpp:157-517):
 * *n_levels_out = levels after the early stop; for `level`, the used-node count,
 * the fixed-node count and, when the pointers are non-null, the inverse diagonal
 * (3 per node, compact order, FEM.cpp:104-151,463) and the fixed-node mask
 * (1 byte per node).  free_dofs_out (non-null, level = last) receives the count
 * of free coarsest DOFs (FEM.cpp:473-496). */
int vfem_inspect_hierarchy(const int32_t *voxels, int32_t nx, int32_t ny, int32_t nz, const double *ke,
                           const int64_t *fixed_nodes, int64_t n_fixed, int32_t num_levels, int32_t skip_levels,
                           int32_t level, int32_t *n_levels_out, int64_t *n_nodes, int64_t *n_fixed_out,
                           double *invdiag, uint8_t *fixed_mask, int64_t *free_dofs_out);

/* The hierarchy vfem_assemble built on the device (test hook; no reference
 * counterpart): for `level`, the inverse diagonal (3 per node, compact order)
 * and the fixed-node mask (1 byte per node), copied from the device when the
 * pointers are non-null, and the free coarsest DOFs.  Equal, bit for bit, to
 * what vfem_inspect_hierarchy builds on the host for the same input. */
int vfem_level_data(vfem_system *sys, int32_t level, double *invdiag, uint8_t *fixed_mask, int64_t *free_dofs_out);

/* ---- learned coarse-grid correction (no reference counterpart) --------------
 * A 3-D conv U-Net (4 resolutions, 16-64 channels, 274,800 weights, BF16
 * arithmetic with FP32 accumulation, FP32 master weights) that maps a coarse
 * level's residual and occupancy to the correction the damped cycle would get
 * from the levels below; trained and run by the library's own HIP kernels.
 * Tensors are [batch][channel][z][y][x] float32.  Input channels: the level
 * residual's three components divided by their RMS over the free DOFs, and the
 * node code 1 (free), -1 (fixed), 0 (unused); output: the error estimate in
 * the same scaling.  Weights are layer by layer, each [cout][cin][k][k][k]. */
int64_t vfem_unet_num_params(void);
int vfem_unet_create(uint64_t seed, vfem_unet **out);
void vfem_unet_destroy(vfem_unet *net);
int vfem_unet_get_weights(vfem_unet *net, float *w);
int vfem_unet_set_weights(vfem_unet *net, const float *w);
/* test hooks: forward of a host batch, then the weight gradients (and the input
 * gradient when gx is non-null) for the upstream gradient gout of that batch */
int vfem_unet_forward(vfem_unet *net, const float *x, int32_t batch, int32_t d, int32_t h, int32_t w, float *out);
int vfem_unet_backward(vfem_unet *net, const float *gout, float *gw, float *gx);
/* Adam training on (x, y) pairs with a node mask [n][d][h][w]: loss per sample
 * |mask (N(x) - y)|^2 / |mask y|^2; cosine-decayed learning rate; epoch_loss
 * receives the mean loss of each epoch. */
int vfem_unet_train(vfem_unet *net, const float *x, const float *y, const float *mask, int64_t n_samples, int32_t d,
                    int32_t h, int32_t w, int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss);

/* Training pairs produced by the solver: runs up to n_iter FP64 PCG iterations
 * of the damped cycle on right-hand side f and, after every iteration, takes
 * the residual r_l that the cycle restricted to `level` (1 <= level < levels-1)
 * and writes, on that level's node grid (d, h, w = level node extents):
 *   x[k]   = [r_l / s, node code]          (4 channels, s = RMS of r_l on free DOFs)
 *   y[k]   = A_l^-1 r_l / s                (3 channels; exact, dense Cholesky of
 *                                           the level operator on its free DOFs)
 *   yvc[k] = (the cycle's own correction from level l down) / s   (3 channels)
 *   mask[k] = 1 on free used nodes
 * *count receives the number of pairs (<= n_iter); grid receives d, h, w. */
int vfem_collect_coarse_pairs(vfem_system *sys, const double *f, int32_t level, int32_t n_iter, float *x, float *y,
                              float *yvc, float *mask, int32_t *count, int32_t *grid);

/* Use `net` as the coarse-grid correction of the damped cycle at `level`
 * (options.learned = 1 turns it on per solve); net = NULL detaches. */
int vfem_attach_unet(vfem_system *sys, vfem_unet *net, int32_t level);

/* Name of a status code, and the last error message of the calling thread. */
const char *vfem_status_string(int status);
const char *vfem_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VFEM_H */
