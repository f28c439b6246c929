/*
 * ehip.h -- C ABI of the MI355X-native assemble-then-iterate library
 * (libehip.so).  Plain pointers and sizes only; callable from C, from
 * Fortran through ISO_C_BINDING (elmerfem_amd/fortran/ehip.F90) and from
 * Python through ctypes.
 *
 * Conventions shared with the reference's Matrix_t (fem/src/Types.src:155-228):
 *  - CRS arrays are 1-based as Elmer holds them: rows[n+1], cols[nnz],
 *    diag[n] (int32), values[nnz] and rhs[n] (double).
 *  - Unknowns are interleaved per node: row = dofs*(perm(node)-1)+dof
 *    (fem/src/CRSMatrix.src:553-583).
 *  - Status codes: every entry point returns EHIP_OK (0) or a negative
 *    EHIP_E* code and never aborts; ehip_last_error() describes the failure.
 *    Solver outcomes come back in ehip_report.info with HUTI's codes
 *    (hutiter/src/huti_fdefs.h:19-56): 1 converged, 2 max iterations,
 *    3 diverged, 20 CG rho = 0, 35/36/37 BiCGStab rho/s-norm/omega.
 *  - Threading: a system is used by one host thread at a time; each system
 *    owns one HIP stream.  Calls on different systems may overlap.
 */
#ifndef EHIP_H
#define EHIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define EHIP_OK 0
#define EHIP_EARG (-1)     /* invalid argument */
#define EHIP_EHIP (-2)     /* HIP runtime error */
#define EHIP_ENOMEM (-3)   /* device or host allocation failed */
#define EHIP_EUNSUP (-4)   /* unsupported element type / option */
#define EHIP_ECAP (-5)     /* request exceeds a capacity limit */
#define EHIP_ESTATE (-6)   /* call out of order (e.g. solve before values) */

/* Linear System Iterative Method (IterSolve.src:221-253) */
#define EHIP_CG 0
#define EHIP_BICGSTAB 1
/* Linear System Preconditioning (IterSolve.src:372-411) */
#define EHIP_PRECOND_NONE 0
#define EHIP_PRECOND_DIAGONAL 1
/* ILU(k), k = 0..9 ("ILU0".."ILU9", or "ILU" with "Linear System ILU Order"):
 * EHIP_PRECOND_ILU0 + k.  CRS_IncompleteLU's non-Cholesky factors of the
 * scaled matrix (CRSMatrix.src:2157-2500), applied as CRS_LUPrecondition
 * (:3259-3268, 3293-3397).  FP64 only; the Krylov loop runs in HUTI's order
 * whatever ehip_params.mode says. */
#define EHIP_PRECOND_ILU0 2
/* precision forms */
#define EHIP_FP64 0
#define EHIP_MIXED_IR 1 /* FP32 inner Krylov inside FP64 iterative refinement */
/* iteration modes */
#define EHIP_FAITHFUL 0 /* HUTI loop order: true residual every iteration */
#define EHIP_FAST 1     /* fused kernels, recurrence residual, true-residual exit check */

/* HUTI_INFO codes */
#define EHIP_HUTI_CONVERGENCE 1
#define EHIP_HUTI_MAXITER 2
#define EHIP_HUTI_DIVERGENCE 3
#define EHIP_HUTI_CG_RHO 20
#define EHIP_HUTI_BICGSTAB_RHO 35
#define EHIP_HUTI_BICGSTAB_SNORM 36
#define EHIP_HUTI_BICGSTAB_OMEGA 37

typedef struct ehip_system ehip_system;

/* Solver parameters: the Linear System keywords IterSolver and
 * SolveLinearSystem read (IterSolve.src:221-366, SolverUtils.src:7256). */
typedef struct {
  int method;        /* EHIP_CG | EHIP_BICGSTAB ("Linear System Iterative Method") */
  int precond;       /* EHIP_PRECOND_* ("Linear System Preconditioning") */
  double tol;        /* "Linear System Convergence Tolerance" */
  int maxit;         /* "Linear System Max Iterations" */
  int scaling;       /* "Linear System Scaling", default 1 */
  int precision;     /* EHIP_FP64 | EHIP_MIXED_IR */
  int mode;          /* EHIP_FAITHFUL | EHIP_FAST */
  double inner_tol;  /* mixed form: residual reduction per refinement step.  CG: drop of the
                        FP32 residual between two FP64 residual replacements (0 = 1e-2);
                        BiCGStab: relative tolerance of each restarted inner solve (0 = 1e-2) */
  int inner_maxit;   /* mixed BiCGStab: inner iteration cap per refinement step (0 = 1000) */
  int check_every;   /* host polls convergence every k iterations (0 = default) */
  double divtol;     /* "Linear System Divergence Tolerance" (0 = none) */
  int record_history;/* 1: keep the residual of every iteration */
  int profile_every; /* > 0: time every k-th SpMV launch of the Krylov loop with HIP events */
  int true_residual_stop; /* 0: HUTI's stop test alone, ||D(b - A x)|| / ||D b|| < tol on the
                        scaled system (TRESID_SCALED_BYB, IterSolve.src:318-340).  1: convergence
                        also needs the unscaled FP64 residual ||b - A x|| / ||b|| < tol, evaluated
                        on the device from the scaled residual as ||r_s / D|| / ||b_s / D||.  The
                        loop order stays HUTI's in FAITHFUL mode: iterations continue until both
                        hold.  No counterpart in the reference. */
} ehip_params;

typedef struct {
  int info;             /* HUTI_INFO */
  int iters;            /* HUTI_ITERS (sum of inner iterations for mixed IR) */
  int outer_iters;      /* refinement steps (mixed IR), 0 otherwise */
  int spmv_count;       /* SpMV launches in the solve */
  double res_scaled;    /* final ||D(b - A x)|| / ||D b|| (the stop criterion) */
  double res_true;      /* final ||b - A x|| / ||b|| on the unscaled system, FP64 */
  double norm;          /* ComputeNorm: sqrt(sum x^2 / n) */
  double t_total_ms;    /* wall time of the solve call on the device stream */
  double t_krylov_ms;   /* time inside the Krylov iteration */
  double spmv_us;       /* mean duration of the sampled SpMV launches (profile_every > 0) */
  int spmv_samples;     /* number of sampled launches */
  int spmv_precision;   /* 0: the sampled launches were FP64, 1: FP32 */
} ehip_report;

/* ---- library ---- */
int ehip_init(int device);
/* Pinned host memory (hipHostMalloc) for arrays a caller uploads often --
 * per-request coefficient fields, say: copies from it are DMA transfers
 * that do not hold the calling thread.  No counterpart in the reference. */
int ehip_host_alloc(int64_t bytes, void **p);
int ehip_host_free(void *p);
const char *ehip_last_error(void);
int ehip_version(void);

/* ---- exact "Matrix Vector Proc" ABI ----
 * Replaces CRS_MatrixVectorProd's plugin branch (CRSMatrix.src:3549-3552,
 * reached through matvecsubr, fem/src/Load.c:734-747): v = A u with host
 * arrays.  *spmv is Matrix_t%SpMV (Types.src:210): the library keeps its
 * device copy of the matrix there and re-uploads Values when reinit != 0 or
 * when their checksum changes. */
void ehip_matvec(void **spmv, int *n, int *rows, int *cols, double *vals, double *u, double *v,
                 int *reinit);
/* The same entry under the name ElmerSolver's loader looks up when it was
 * built with the lower-case + trailing-underscore Fortran link type
 * (fortranMangle, fem/src/Load.c:183-208): "Matrix Vector Proc =
 * "libehip" "ehip_matvec"" in a .sif resolves to this symbol there. */
void ehip_matvec_(void **spmv, int *n, int *rows, int *cols, double *vals, double *u, double *v,
                  int *reinit);
/* Releases what ehip_matvec cached under *spmv. */
void ehip_matvec_free(void **spmv);

/* ---- device-resident systems ---- */
/* From an existing 1-based CRS pattern (Matrix_t Rows/Cols/Diag). */
int ehip_system_create(int n, int64_t nnz, const int *rows, const int *cols, const int *diag,
                       ehip_system **out);
/* From a mesh: the CreateMatrix chain (ElementUtils.src:698-935) with
 * InitialPermutation identity, MakeListMatrix, OptimizeBandwidth
 * (BandwidthOptimize.src:108-355) when optimize != 0, CRS sort/Diag, plus
 * the element-to-CRS-slot map used by ehip_assemble.  elements: nelem x npe
 * 1-based node ids; bnd_elements (may be NULL) take part in the pattern as
 * the reference's boundary elements do. */
int ehip_system_from_mesh(int nnodes, int nelem, int npe, const int *elements, int nbnd,
                          int npe_bnd, const int *bnd_elements, int dofs, int optimize,
                          ehip_system **out);
int ehip_system_free(ehip_system *s);
int ehip_system_info(const ehip_system *s, int *n, int64_t *nnz, int *dofs, int *bw_before,
                     int *bw_after, int *bw_accepted);
/* Copies of the host-side pattern (1-based); any pointer may be NULL. */
int ehip_system_get_pattern(const ehip_system *s, int *perm, int *rows, int *cols, int *diag);
int ehip_system_set_values(ehip_system *s, const double *values, const double *rhs);
int ehip_system_get_values(ehip_system *s, double *values, double *rhs);
/* Device memory the system holds, in bytes. */
int64_t ehip_system_bytes(const ehip_system *s);

/* ---- mesh and result files ----
 * An ElmerGrid mesh directory as LoadMesh reads it (MeshUtils.src:313-1009
 * through EIO): mesh.header, mesh.nodes ("tag partition x y z"),
 * mesh.elements ("tag body type nodes..."), mesh.boundary ("tag bc parent1
 * parent2 type nodes...").  Nodes are renumbered 1..nnodes in file order and
 * element node references mapped through the node tags (LocalPerm,
 * :493-499).  Element nodes per type = type mod 100. */
typedef struct ehip_mesh ehip_mesh;
int ehip_mesh_read(const char *dir, ehip_mesh **out);
/* max_npe_*: nodes of the largest bulk / boundary element type (row widths below) */
int ehip_mesh_sizes(const ehip_mesh *m, int *nnodes, int *nbulk, int *nbnd, int *max_npe_bulk, int *max_npe_bnd);
/* coords: nnodes x 3; bulk_nodes: nbulk x max_npe_bulk and bnd_nodes: nbnd x
 * max_npe_bnd, 1-based node numbers, rows of smaller types 0-padded;
 * bnd_parents: nbnd x 2 (0: none).  Any pointer may be NULL. */
int ehip_mesh_get(const ehip_mesh *m, double *coords, int *bulk_types, int *bulk_bodies, int *bulk_nodes,
                  int *bnd_types, int *bnd_tags, int *bnd_parents, int *bnd_nodes);
int ehip_mesh_free(ehip_mesh *m);
/* Writes the four files of a serial ElmerGrid mesh (partition -1) with one
 * bulk and one boundary element type; bulk_bodies, bnd_tags and bnd_parents
 * may be NULL (body 1, tag 1, no parents). */
int ehip_mesh_write(const char *dir, int nnodes, const double *coords, int nbulk, int bulk_type,
                    const int *bulk_nodes, const int *bulk_bodies, int nbnd, int bnd_type, const int *bnd_nodes,
                    const int *bnd_tags, const int *bnd_parents);
/* SaveResult (ModelDescription.src:2442-2865), ASCII format 3, one scalar
 * nodal variable: the header when append = 0, then the "Time:" line, the
 * variable name, its Perm (perm[nnodes], 1-based, 0 = no value; NULL writes
 * "Perm: NULL" and values in node order) and the values. */
int ehip_result_write(const char *path, const char *var, const char *equation, int nnodes, const int *perm,
                      int dofs, const double *values, int append, int step, double time);

/* ---- assembly ---- */
#define EHIP_PHYS_MODELPDE 0   /* ModelPDE: diffusion + convection + source (ModelPDE.src:75-166) */
#define EHIP_PHYS_ELASTIC 1    /* StressSolve isotropic elasticity (Stress.src:59-521) */
/* Bulk element loop + DefaultUpdateEquations glue, in element order, as a
 * deterministic row-owner gather (no float atomics): every slot receives
 * its element contributions in element order, from zero.
 * coords: nnodes x 3 (row-major), or NULL for the coordinates of the
 * previous call (kept on the device).  Per-element arrays (length nelem, or
 * 3*nelem for velo) may be NULL (= 0).  elem_code: 303 (cdim 2), 404
 * (cdim 2), 504, 510, 808, 827 (cdim 3); elasticity takes 808, 827, 510.
 * Values and RHS are overwritten (DefaultInitialize).  t_ms non-NULL: the
 * call waits and returns the device time; NULL: the work is only enqueued
 * on the system's stream (no allocation after the first call), and
 * degenerate elements are reported (EHIP_EARG) by the next call that waits
 * on the system (ehip_solve, ehip_system_get_values, a batch or server
 * admission, or ehip_assemble again). */
int ehip_assemble(ehip_system *s, int physics, int elem_code, int cdim, const double *coords,
                  const double *elem_coef, const double *elem_conv, const double *elem_velo,
                  const double *elem_load, double poisson, double *t_ms);
/* Boundary traction (StressBoundary, Stress.src:814-1020) added to RHS in
 * boundary-element order: bnd_code 404 or 409, belems nb x npe 1-based. */
int ehip_assemble_traction(ehip_system *s, int bnd_code, const double *coords, int nb, int npe,
                           const int *belems, const double *load3);
/* DefaultDirichletBCs for nodal Dirichlet rows (DefUtils.src:3530-4030):
 * scale, ZeroRow + unit diagonal + b = val/D (or CRS_SetSymmDirichlet when
 * symmetric != 0), back-scale.  bc_rows are 1-based matrix rows.  The rows
 * and values stay on the device: a call with the same ones as the last only
 * enqueues the kernels and returns without waiting. */
int ehip_dirichlet(ehip_system *s, int nbc, const int *bc_rows, const double *bc_vals,
                   int symmetric, int scaling);

/* ---- solve ----
 * SolveLinearSystem's iterative branch (SolverUtils.src:7151-7525):
 * zero-RHS exit, ScaleLinearSystem, x = 1e-8 when x = 0 (IterSolve.src:358),
 * HUTI CG/BiCGStab with TRESID_SCALED_BYB, back-scale, ComputeNorm.  x is the
 * host initial guess on entry and the solution on exit (may be NULL: zero
 * guess, solution left on the device).  hist (may be NULL, length maxit)
 * receives the per-iteration residual when params->record_history. */
int ehip_solve(ehip_system *s, const ehip_params *p, double *x, ehip_report *r, double *hist);
/* Solution currently on the device. */
int ehip_get_solution(ehip_system *s, double *x);
/* One-shot drop-in for a Before-Linsolve hook: host CRS + b in, x out. */
int ehip_solve_csr(int n, const int *rows, const int *cols, const int *diag, const double *values,
                   const double *b, double *x, const ehip_params *p, ehip_report *r);

/* ---- no-allocation solve ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * solve as ehip_solve_csr (same arguments and meaning), with every device
 * buffer -- the staged CRS and RHS, the solution, D, the Krylov workspace,
 * partial sums and scalars -- carved from one static device array of
 * EHIP_NOALLOC_SCRATCH_BYTES that the library owns on each GPU it runs on
 * and holds until the process ends.  It calls no hipMalloc/hipFree and
 * creates no stream or event; it runs on the null stream of the current
 * device and times the solve with the host clock.
 * Limits: n <= EHIP_NOALLOC_MAX_N and nnz <= EHIP_NOALLOC_MAX_NNZ (else
 * EHIP_ECAP); precision EHIP_FP64 with record_history = 0,
 * profile_every = 0 and precond EHIP_PRECOND_NONE or _DIAGONAL (else
 * EHIP_EUNSUP).  The SpMV is the CSR stream layout,
 * whose FP64 row sums equal every other layout's.
 * All calls on one GPU share its array and leave their data in it: calls
 * that overlap on one GPU (several host threads) race, and the caller must
 * serialise them. */
#define EHIP_NOALLOC_SCRATCH_BYTES (96ll << 20)
#define EHIP_NOALLOC_MAX_N 262144
#define EHIP_NOALLOC_MAX_NNZ 4194304
int ehip_solve_csr_noalloc(int n, const int *rows, const int *cols, const int *diag, const double *values,
                           const double *b, double *x, const ehip_params *p, ehip_report *r);

/* ---- ILU(k) factors ----
 * CRS_IncompleteLU (CRSMatrix.src:2157-2500), non-Cholesky, of the system's
 * current Values with fill order k = 0..9 (the pattern of InitializeILU1's
 * k rounds, :2387-2498; the matrix pattern for k = 0).  Copies out the ILU
 * pattern (1-based, irows[n+1], icols and idiag) and the factors (unit lower
 * part below idiag, U above it, U(i,i) stored inverted as the reference's
 * LU solve expects); every pointer may be NULL, *ilu_nnz receives the
 * pattern size.  ehip_solve factorizes the scaled matrix itself when
 * params->precond selects ILU(k); this entry exposes what it uses. */
int ehip_ilu_factor(ehip_system *s, int order, int64_t *ilu_nnz, int *irows, int *icols, int *idiag,
                    double *ilu);
/* CRS_LUPrecondition (CRSMatrix.src:3259-3268): x = (LU)^-1 b with the last
 * factors of ehip_ilu_factor or of an ILU-preconditioned ehip_solve. */
int ehip_lu_solve(ehip_system *s, const double *b, double *x);
/* The factors' shape: fill order (-1: none yet), pattern size and the
 * number of dependency levels of the forward and backward sweeps (one
 * launch each per LU solve).  Any pointer may be NULL.  No counterpart in
 * the reference (its LU solve runs row by row on one core). */
int ehip_ilu_info(const ehip_system *s, int *order, int64_t *ilu_nnz, int *levels_forward, int *levels_backward);

/* ---- measurement helpers ---- */
/* v = A u on the device-resident system (host vectors), precision 0 = FP64
 * values, 1 = FP32 copy of the values. */
int ehip_spmv(ehip_system *s, int precision, const double *u, double *v);
/* Times `reps` SpMV launches on device-resident vectors; returns the mean
 * kernel time in microseconds (HIP events on the system's stream). */
int ehip_spmv_bench(ehip_system *s, int precision, int reps, double *us_per_call);
/* Streams a device buffer of `bytes` through a plain copy kernel `reps`
 * times; returns the achieved GB/s (read + write) -- the measured HBM
 * ceiling, and the calibration kernel for hardware byte counters. */
int ehip_copy_bench(int64_t bytes, int reps, double *gbps, double *us_per_call);
/* SpMV storage formats (the CRS arrays stay the system's reference copy):
 *  EHIP_SPMV_CSR : CRS streamed in chunks of nonzeros, products staged in LDS;
 *  EHIP_SPMV_SELL: SELL-64 copy (slices of 64 rows, column-major, padded to
 *                  the slice's longest row), one lane per row -- built when
 *                  padding adds at most 25% to the nonzeros;
 *  EHIP_SPMV_SELL_SHARED: SELL-64 whose slices share one list of column
 *                  offsets from the row index (entry k of every row of a
 *                  slice at column row + offset k, absent entries holding
 *                  0, a per-row bit mask naming the row's nonzeros), so no
 *                  column indices are streamed -- built, and made the
 *                  default of both precisions, when every slice's rows use
 *                  at most 32 distinct offsets and the layout holds at most
 *                  1/4 more entries than the nonzeros (stencil-like
 *                  patterns: meshes of regular topology in natural node
 *                  order); the SELL-64 and CSR-stream layouts of such a
 *                  pattern are then built when first selected, tuned or
 *                  queried.  Node-interleaved 2- or 3-DOF patterns get
 *                  the same with one node offset and one mask bit per
 *                  D x D block (ehip_spmv_shared_layout's node_block).
 * Every format sums each row in CRS order; FP64 results are the same bits in
 * all of them. */
#define EHIP_SPMV_CSR 0
#define EHIP_SPMV_SELL 1
#define EHIP_SPMV_SELL_SHARED 2
/* SpMV configuration of one precision (0 = FP64 values, 1 = FP32): format,
 * workgroup size, nonzeros per chunk (CSR; 0 for SELL), number of chunks
 * (CSR) or slices (SELL), and the column index width the kernels read (16:
 * offsets from a per-segment or per-slice base when every span is < 65536
 * columns; 32: int32 indices; 0: none, EHIP_SPMV_SELL_SHARED).  Any pointer
 * may be NULL. */
int ehip_spmv_config(ehip_system *s, int precision, int *format, int *block, int *nnz_per_block, int *chunks,
                     int *index_bits);
/* Shape of the SELL-64 layout (EHIP_EUNSUP when the pattern has none):
 * sort_window = 0 when slices take rows in order, else the window (256,
 * 1024 or 4096 rows) inside which rows were sorted by length, longest
 * first, to keep the padding under the limit; node_block = D when the
 * pattern is made of whole D x D blocks of node-interleaved unknowns
 * (D = DOFs per node, 2 or 3) and one 16-bit node offset per block
 * replaces D column indices, else 0; lane_unroll = nonzeros (node blocks
 * with node_block) each lane loads before using them (8, or 9 when that
 * wastes fewer loads on the slice widths); pair_unroll = value pairs each
 * lane loads at once in the pair-packed FP64 layout (4 or 7, chosen the
 * same way; 0 when FP64 values are not pair-packed).  Any pointer may be
 * NULL.
 * No counterpart in the reference (CRS_MatrixVectorProd, CRSMatrix.src:
 * 3515-3640, reads CRS directly). */
int ehip_spmv_sell_layout(ehip_system *s, int *sort_window, int *node_block, int *lane_unroll,
                          int *pair_unroll);
/* Shape of the slice-shared offset layout (EHIP_EUNSUP when the pattern has
 * none): node_block = D when the pattern is made of whole D x D blocks of
 * node-interleaved unknowns (D = 2, 3) and the lists hold node offsets,
 * one mask bit per block, else 0; lane_unroll = entries (blocks, with
 * node_block) each lane loads before using them (8, or 9 when that wastes
 * fewer loads on the slice widths; 9 blocks with node_block); entries =
 * stored values, absent ones and the padding lanes of the last slice
 * included.  Any pointer may be NULL.  No counterpart in the reference. */
int ehip_spmv_shared_layout(ehip_system *s, int *node_block, int *lane_unroll, int64_t *entries);
/* Selects the format for one precision (EHIP_EUNSUP: the pattern has no
 * layout of that format). */
int ehip_spmv_set_format(ehip_system *s, int precision, int format);
/* Autotuner: times every compiled CSR schedule and the SELL formats on this
 * matrix, keeps the fastest.  Every choice sums each row in the same order. */
int ehip_spmv_tune(ehip_system *s, int precision);
/* Sets a schedule explicitly (EHIP_EUNSUP if it is not compiled in).  Every
 * schedule sums each row in the same order, so results do not change. */
int ehip_spmv_set_schedule(ehip_system *s, int precision, int block, int nnz_per_block);

/* ---- batching: many independent systems in one multi-system launch ----
 * No counterpart in the reference (a new form of the same solves).
 * Systems keep their own device memory.  The batch builds one device work
 * list of (system, row range) chunks over all of them; every launch of an
 * iteration walks that list, each system keeping its own scalars (alpha,
 * beta, omega, rho, convergence) and phase, and the iteration is replayed
 * from a recorded hipGraph.  Optionally (ehip_batch_set_lane) systems with
 * at least lane_min_dofs unknowns are instead solved whole by ehip_solve on
 * lane_workers host threads, concurrently with the multi-system launches;
 * their results are exactly their ehip_solve results.  Off by default. */
typedef struct ehip_batch ehip_batch;
int ehip_batch_create(int nsys, ehip_system **systems, ehip_batch **out);
int ehip_batch_free(ehip_batch *b);
/* workers = 0 (the default): every system joins the multi-system launches */
int ehip_batch_set_lane(ehip_batch *b, int64_t lane_min_dofs, int lane_workers);
/* reports[nsys]; solutions stay on the device (ehip_batch_get_solution).
 * t_ms: device time of the multi-system launches, or the wall time of the
 * whole batch when some systems took the concurrent lane.  In the
 * multi-system launches p->precision selects FP64 or mixed precision.  FP64
 * with EHIP_FAITHFUL runs HUTI's loop order (true residual every iteration,
 * HUTI's exits); with EHIP_FAST it runs the recurrence residual and, when
 * that meets tol, or on the S-norm, omega or rho breakdown, recomputes the
 * true residual from x and either ends the solve or restarts the Krylov
 * sequence from it (report.outer_iters = restarts).  Systems on the
 * concurrent lane run ehip_solve with p as given. */
int ehip_batch_solve(ehip_batch *b, const ehip_params *p, ehip_report *reports, double *t_ms);
/* Schedule search before the next ehip_batch_solve: mode 1 times the chunk
 * sizes (1024, 2048, 4096 nonzeros per work item) for every size class of
 * the batch that the process-wide schedule table does not hold yet, then
 * the workgroup count (1024, 2048, 4096), each over `replays` graph replays
 * (0 = 3) from the admitted state; mode 2 searches everything again; 0
 * (the default) uses the table, or chunk size 2048 and 2048 workgroups where
 * it holds nothing.  Size classes by nonzeros: < 2^15, < 2^17, < 2^19,
 * < 2^21, < 2^23, more.  A system's results depend on its class's chunk
 * size only (never on the other systems of the batch or the workgroup
 * count), so a system solved alone and in a batch under the same table
 * gives the same bits.  No counterpart in the reference. */
int ehip_batch_set_tuning(ehip_batch *b, int mode, int replays);
/* The table entry for solver form p (precision, method, mode) and a system
 * of nnz nonzeros: its chunk size, and the form's workgroup count. */
int ehip_batch_get_schedule(const ehip_params *p, int64_t nnz, int *chunk_nnz, int *grid);
/* Sets it: chunk_nnz 1024 | 2048 | 4096, 0 = default, -1 = unchanged;
 * grid a multiple of 8 up to 65536, 0 = default, -1 = unchanged.  Applies to
 * engines created afterwards (batches and servers). */
int ehip_batch_set_schedule(const ehip_params *p, int64_t nnz, int chunk_nnz, int grid);
int ehip_batch_get_solution(ehip_batch *b, int i, double *x);

/* ---- serving: continuous batching ----
 * A server owns the multi-system engine.  Systems (assembled, Dirichlet
 * rows applied) are admitted while a slot is free and the HBM budget
 * (mem_budget bytes, 0 = none) holds their solver workspace; each step
 * replays the iteration graph once for every live system and retires the
 * finished ones, whose solutions and reports become available.  Requests
 * that arrive while others run join at the next step. */
typedef struct ehip_batch ehip_server;
int ehip_server_create(const ehip_params *p, int max_slots, int64_t mem_budget, ehip_server **out);
/* EHIP_ECAP: no free slot or over budget -- keep the request queued.
 * Admission only enqueues work (the system's preparation on its own
 * stream, which the server's stream then waits for); a system whose
 * ehip_assemble without t_ms is still running is not waited for, and its
 * degenerate-element check happens at retirement: a failure there gives
 * the report info 3 (EHIP_HUTI_DIVERGENCE), with the count in
 * ehip_last_error. */
int ehip_server_submit(ehip_server *srv, ehip_system *s, int *slot);
int ehip_server_step(ehip_server *srv, int *ndone, ehip_system **done_systems, ehip_report *done_reports,
                     int max_done);
int ehip_server_status(const ehip_server *srv, int *live, int64_t *bytes);
/* Concurrent lane: systems with at least min_dofs unknowns are not batched
 * but solved whole (ehip_solve) on one of `workers` host threads, each on
 * the system's own stream, overlapping one another and the batched systems
 * on the device.  Admitted systems queue in the lane (up to 4 per worker)
 * so a worker starts its next solve as soon as it finishes one; beyond that
 * a submission waits (EHIP_ECAP).  workers = 0 turns the lane off (the
 * default: everything is batched). */
int ehip_server_set_lane(ehip_server *srv, int64_t min_dofs, int workers);
int ehip_server_free(ehip_server *srv);

#ifdef __cplusplus
}
#endif
#endif
