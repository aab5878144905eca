/*
 * femorax_amd — C ABI of the MI355X-native hex8 matrix-free operator path.
 *
 * The cut follows the reference's type-erased per-pair backend InternalForce::Batch
 * (amnp95/FemoraX src/exec/internal_force_batch.hpp:46-100) and the solver interface
 * (src/solvers/linear_solver.hpp:57-105): a C++ Batch / LinearOperator subclass on the
 * reference side forwards its virtual calls here, so src/exec/implicit_stage.cpp,
 * src/exec/explicit_analysis.cpp and the solvers' callers stay unchanged.
 *
 * Conventions (the reference's, src/exec/internal_force.hpp:80-244):
 *   - every entry point returns FX_OK (0) or a negative FX_E* code; fx_last_error(ctx)
 *     gives the message (the reference throws std::runtime_error / io::ModelError, or calls
 *     Kokkos::abort on misuse);
 *   - construction (fx_batch_create) uploads and allocates everything; the per-call entry
 *     points allocate nothing;
 *   - vectors x, y, b, u, v, f are borrowed DEVICE pointers in equation numbering, FP64
 *     unless the name says FP32; x and y never alias (linear_solver.hpp:62);
 *   - calls are blocking at this boundary unless the name says _async (the reference's
 *     dots and norms read back to the host: src/solvers/vector_ops.cpp:62);
 *   - a material failure is data (NaN stress), not an error (pressure_independ_multi_yield.hpp:515-518).
 */
#ifndef FEMORAX_AMD_H
#define FEMORAX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FX_OK 0
#define FX_EINVAL -1   /* bad argument or shape */
#define FX_EDEVICE -2  /* HIP runtime error */
#define FX_ENOMEM -3   /* over the HBM budget or out of device memory */
#define FX_EZERO -4    /* Jacobi: zero diagonal entries (jacobi.cpp:19-22) */
#define FX_EUNSUP -5   /* unsupported element/material/option */
#define FX_EFULL -6    /* a column pool has no room for the column now (retry after columns finish) */
#define FX_EIO -7      /* a recorder file could not be opened or written */

typedef struct fx_ctx fx_ctx;
typedef struct fx_mesh fx_mesh;
typedef struct fx_batch fx_batch;
typedef struct fx_recorder fx_recorder;

/* ---- context: one device, one stream, an HBM budget (exec/device_memory.cpp:26-38) ------- */
/* hbm_budget_bytes = 0: 90% of free memory at creation (explicit_analysis.cpp:446-454). */
int fx_ctx_create(int device, uint64_t hbm_budget_bytes, fx_ctx** out);
void fx_ctx_destroy(fx_ctx* ctx);
const char* fx_last_error(const fx_ctx* ctx);
int fx_sync(fx_ctx* ctx);
/* Context options (no counterpart in the reference, whose Kokkos launches are per kernel):
 * FX_OPT_GRAPHS (default 0): 1 replays the CG iterations of fx_cg / fx_cg_mixed (and the solves
 *   inside fx_newmark) from captured hipGraphs, 8 iterations each, cached per batch,
 *   coefficients and vectors (the same kernels and arguments, so the same bits). Capture needs
 *   no other thread to use the null stream meanwhile (serving loops with upload threads keep it
 *   off). */
#define FX_OPT_GRAPHS 1
/* FX_OPT_STREAM: the context's work goes to the caller's HIP stream (value: the hipStream_t, 0 to
 *   return to the context's own), e.g. the Kokkos HIP execution space's stream, so that the
 *   library's kernels and the caller's are ordered as the reference's Kokkos launches are. The
 *   work queued so far completes first. The caller keeps the stream alive and at the default
 *   priority.
 * FX_OPT_STREAM_ORDERED (default 0): 1 makes the drop-in operator calls (fx_operator_apply,
 *   fx_operator_apply_f32, fx_operator_diagonal, fx_run_operator) return once their kernels are
 *   queued, as Kokkos launches do; their results are ordered on the context's stream (fx_sync, or
 *   later work on that stream, sees them). Default: they return with the results in place. */
#define FX_OPT_STREAM 2
#define FX_OPT_STREAM_ORDERED 3
int fx_ctx_set_option(fx_ctx* ctx, int option, int64_t value);
/* Free and total device memory (src/native/device_memory_hip.cpp:6-11). */
int fx_device_memory(fx_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
uint64_t fx_budget_used(const fx_ctx* ctx);

/* Device buffers for callers without their own allocator (tests, the benchmark). */
int fx_alloc(fx_ctx* ctx, size_t bytes, void** dptr);
int fx_free(fx_ctx* ctx, void* dptr);
int fx_upload(fx_ctx* ctx, void* dst, const void* src, size_t bytes);
int fx_download(fx_ctx* ctx, void* dst, const void* src, size_t bytes);
int fx_memset0(fx_ctx* ctx, void* dst, size_t bytes);

/* ---- compile step (compile::compile_model for stdBrick batches, src/compile/batches.cpp:54-227,
 *      renumber.cpp:153-183, dof_numbering.cpp:269-463) ----------------------------------- */
typedef struct fx_model_desc {
  int32_t nnode;
  const double* xyz;        /* [nnode*3], row major, model order */
  int32_t nelem;
  const int32_t* conn;      /* [nelem*8] model node indices, OpenSees Brick node order */
  const int32_t* group;     /* [nelem] element group (one material each), NULL = all 0 */
  const int32_t* region;    /* [nelem] Rayleigh region or -1, NULL = none */
  int32_t nfix;
  const int32_t* fix_node;  /* model node index */
  const int32_t* fix_dof;   /* 0..2 */
  int32_t ntie;             /* equalDOF rows (retained, constrained, dof 0..2) */
  const int32_t* tie_retained;
  const int32_t* tie_constrained;
  const int32_t* tie_dof;
  const int32_t* tie_command; /* [ntie] equalDOF command of each row (EqualDof::command), NULL:
                                 one command per (retained, constrained) pair */
} fx_model_desc;
/* Node i of the model has tag i + 1 (the constraint rules process nodes in ascending tag order,
 * dof_numbering.cpp:117-129). fx_compile_model refuses, with FX_EINVAL and a message naming the
 * pattern, the models the reference refuses (dof_numbering.cpp:162-262, 303-339): a node tied
 * by two equalDOF commands, a fix on a tied DOF whose retained DOF is free, partial chains and
 * chains through a node of larger tag. */

typedef struct fx_mesh_view {
  int32_t nnode, nelem, padded, neq, nparam, ncolor;
  const int32_t* node_order;     /* [nnode] internal -> model node (Ordering::node_order) */
  const int32_t* node_rank;      /* [nnode] model -> internal */
  const int32_t* elem_order;     /* [nelem] batch position -> file element */
  const int32_t* node_eq;        /* [3*nnode] equation of internal node i, DOF d; -1 fixed */
  const double* coords;          /* [3*nnode] SoA x[N] y[N] z[N] (Compiled::coords) */
  const int32_t* node;           /* [8*padded] Batch::node, value[k*padded+e], -1 padding */
  const int32_t* eq;             /* [24*padded] Batch::eq */
  const int32_t* param;          /* [padded] Batch::param */
  const int32_t* param_group;    /* [nparam] ParamSet::group */
  const int32_t* param_region;   /* [nparam] ParamSet::region */
  const int32_t* color_start;    /* [ncolor+1] Batch::color_start (NULL without colouring) */
  const int32_t* color_elements; /* [nelem] */
} fx_mesh_view;

int fx_compile_model(const fx_model_desc* model, int renumber_rcm, int simd_width, int build_coloring,
                     fx_mesh** out);
int fx_mesh_get(const fx_mesh* mesh, fx_mesh_view* view);
void fx_mesh_destroy(fx_mesh* mesh);

/* ---- element parameters --------------------------------------------------------------- */
/* Rayleigh (src/physics/contracts.hpp:69-75). */
typedef struct fx_rayleigh { double alphaM, betaK, betaKinit, betaKcomm; } fx_rayleigh;

/* stdBrick<ElasticIsotropic>::Params (std_brick.hpp:181-185, elastic_isotropic.hpp:30-34)
 * with its ParamSet's damping (compile/batches.hpp:28-39). */
typedef struct fx_elastic_params {
  double E, nu, rho;
  double b[3];        /* body force per unit volume */
  int32_t mass_type;  /* 0 consistent, 1 lumped */
  fx_rayleigh damp;
} fx_elastic_params;

/* OperatorCoefficients (src/physics/operators/element_operator.hpp:33-40). */
typedef struct fx_coeffs { double stiffness, damping, mass; int32_t initial; } fx_coeffs;

/* ---- batch: InternalForce::Batch behind a C ABI (internal_force_batch.hpp:46-100) -------- */
typedef struct fx_batch_desc {
  int32_t count, padded;         /* real elements, padded count (Batch::count / ::padded) */
  const int32_t* node;           /* [8*padded] host, Batch::node */
  const int32_t* eq;             /* [24*padded] host, Batch::eq */
  const int32_t* param;          /* [padded] host, Batch::param */
  int32_t nparam;
  const fx_elastic_params* params; /* [nparam] host */
  int32_t nnode;                 /* nodes of the model */
  const double* coords;          /* [3*nnode] host SoA (Compiled::coords) */
  int32_t neq;                   /* equations (DofNumbering::num_equations) */
} fx_batch_desc;

/* Flags for fx_batch_create. */
#define FX_BATCH_FP32 1u    /* also build the FP32 operator (mixed precision solves); on
                               fx_pimy_batch_create, an FP32 copy of the stored tangents for it */
#define FX_BATCH_GENERAL 2u  /* do not use the structured-lattice fast path (testing) */
#define FX_BATCH_DYNAMIC 32u  /* allocate the time-integration workspace (12 * neq doubles) that
                                 fx_newmark needs (PIMY batches always have it) */
#define FX_BATCH_GMRES 64u    /* allocate the GMRES(30) workspace ((2 * 30 + 3) * neq doubles):
                                 fx_gmres, fx_krylov, and fx_newmark's CG -> GMRES fallback */
#define FX_BATCH_XCD_SLABS 8u /* lattice path: tiles placed in 8 slabs, one per XCD (default: lattice order,
                                which measured faster on MI355X; same bits either way) */
#define FX_BATCH_NO_TUNE 16u /* skip the kernel-variant autotune at creation (variant 0) */
#define FX_BATCH_TILE4 256u   /* lattice path: tiles of 8 x 8 x 4 cells (one pass of four cell planes per
                                 workgroup) */
#define FX_BATCH_TILE8 512u   /* lattice path: tiles of 8 x 8 x 8 cells (two passes). Neither flag: 8-deep
                                 tiles from 2^20 cells up (C2: 0.0625 ms against 0.0670 ms), 4-deep below
                                 (more workgroups for small latency-bound models) */
#define FX_BATCH_TILE16 1024u /* lattice path: tiles of 8 x 8 x 16 cells (four passes, streamed: x and words
                                 of five node planes in LDS at a time) */
#define FX_BATCH_SLOTS 2048u  /* keep the per-element operator forms of fx_run_operator ready: on elastic batches
                                 an element slot buffer [24 * padded] and its gather list (and, on the lattice
                                 path, the connectivity, coordinates and node equations on the device);
                                 PressureIndependMultiYield batches always have them */
#define FX_BATCH_SSPBRICK 4096u /* the element is SSPbrick<ElasticIsotropic> (ssp_brick.hpp: one mean-strain
                                 point plus the stabilisation Kstab = Mben^T interior Mben, lumped mass; the
                                 reference's batch of that (element, material) pair, batches.cpp:54-227)
                                 instead of stdBrick; mass_type is ignored. Lattice path: 4- or 8-deep tiles
                                 (TILE16 takes 8), no kernel variants; a lattice with equalDOF ties inside an
                                 element runs on the general path */
#define FX_BATCH_COMPACT 128u /* stdBrick<PressureIndependMultiYield> batches: the compacted return map
                                 (compacted_kernels.hpp:85-202): an elastic predictor per Gauss point, the
                                 plastic points listed, one full update per listed point, the element
                                 forces replayed from the stored results (same bits as the fused kernel) */

int fx_batch_create(fx_ctx* ctx, const fx_batch_desc* desc, uint32_t flags, fx_batch** out);
void fx_batch_destroy(fx_batch* b);

/* TangentSymmetry of the batch's material (src/physics/contracts.hpp:36-45). */
#define FX_TANGENT_SPD 0    /* SymmetricPositiveDefinite (ElasticIsotropic, elastic_isotropic.hpp:27) */
#define FX_TANGENT_SPSD 1   /* SymmetricPositiveSemidefinite (PressureIndependMultiYield,
                               pressure_independ_multi_yield.hpp:70) */
typedef struct fx_batch_info {
  int32_t count, neq, nnode, nblocks, ncolors, tied;
  int32_t lattice;          /* 1: the batch runs on the structured-lattice fast path */
  int32_t tiles;            /* lattice tiles (workgroups per apply) */
  int32_t variant;          /* tuned kernel variant */
  int32_t tangent_symmetry; /* FX_TANGENT_* (Batch::tangent_symmetry, internal_force_batch.hpp:84) */
  uint64_t device_bytes;
} fx_batch_info;
int fx_batch_info_get(const fx_batch* b, fx_batch_info* info);
/* Kernel-variant choice (the reference's force-settings autotune, src/exec/force_settings.cpp:
 * 135-205): variant >= 0 sets it, -1 times the variants on this batch and keeps the fastest.
 * Lattice batches have variants 0-15 (bits 0-1: the tile kernel's register budget, the compiler's
 * choice or at least 5, 7, 8 waves per SIMD; bit 2: persistent workgroups that prefetch their
 * next tile; bit 3: streamed 8-deep tiles, x and words of five node planes in LDS at a time;
 * 16-deep tiles are always streamed); general batches have one. The autotune times 0-7 (4-deep
 * tiles), 0-11 (8-deep) or 0-3 (16-deep), for the operator without mass terms and with them, and
 * keeps a variant for each (16 (m + 1) + v: v without mass terms, m with them; 0-15 sets both).
 * fx_batch_info.variant reports v. Every variant gives the same bits. */
int fx_batch_tune(fx_batch* b, int variant);

/* y = K_eff x, K_eff = cK K_t + cC C + cM M (InternalForce::apply_operator,
 * internal_force.cpp:545-573; element_operator_apply, element_operator.hpp:82-130).
 * Deterministic: the same bits for the same inputs, whatever else runs. */
int fx_operator_apply(fx_batch* b, const fx_coeffs* k, const double* x, double* y);
/* The same in FP32 (x, y FP32 device vectors; FP32 geometry and arithmetic). Needs FX_BATCH_FP32. */
int fx_operator_apply_f32(fx_batch* b, const fx_coeffs* k, const float* x, float* y);
/* d = diag(K_eff) (InternalForce::operator_diagonal, internal_force.cpp:575-599;
 * element_operator_diagonal, element_operator.hpp:138-209). */
int fx_operator_diagonal(fx_batch* b, const fx_coeffs* k, double* d);
/* The per-batch verb the reference's InternalForce calls once per batch per apply
 * (InternalForce::Batch::run_operator, internal_force_batch.hpp:80-82; launch_operator,
 * operator_kernels.hpp:75-141; emit, assembly.hpp:28-36), for a binding that keeps
 * InternalForce::apply_operator / operator_diagonal / mass_times_ground unchanged
 * (internal_force.cpp:545-626):
 *   out = FX_OUT_SLOTS: slots[slot_offset + a * padded + e] = value a (0..23) of element e of the
 *     batch (e < count), fixed DOFs included, as Out::Slots writes them; the caller's SlotGather
 *     then sums them (assembly.cpp:30-38). Elastic batches need FX_BATCH_SLOTS.
 *   out = FX_OUT_PLAIN or FX_OUT_ATOMIC: y[q] += this batch's sum at equation q (q >= 0); the
 *     caller zeroed y before the first batch, as apply_operator does (internal_force.cpp:565).
 *     The adds run in a fixed order here (the reference's Atomic order varies between runs).
 *   diagonal = 1: the element diagonals (element_operator_diagonal, element_operator.hpp:138-209);
 *     x unused.
 *   unit_dir = 0..2: x is replaced by the rigid unit translation in that direction at every
 *     element DOF, fixed ones included (the ground-load vector M r of mass_times_ground with
 *     k = {0, 0, 1}); -1: x is used. With PLAIN/ATOMIC, elastic batches need FX_BATCH_SLOTS.
 * The nodal masses of fx_batch_set_nodal_mass are not added (InternalForce adds them). x, y and
 * slots are device pointers in the reference's equation numbering. */
#define FX_OUT_PLAIN 0
#define FX_OUT_ATOMIC 1
#define FX_OUT_SLOTS 2
int fx_run_operator(fx_batch* b, int diagonal, const fx_coeffs* k, int out, const double* x, double* y,
                    double* slots, int64_t slot_offset, int unit_dir);
/* Nodal masses of the model (Compiled::nodal_mass / nodal_alpha_mass, batches.cpp:249-261): host
 * [neq] in the reference numbering, NULL for none. Like InternalForce, the batch then adds
 * cM m x + cC am x to fx_operator_apply and its diagonal (add_nodal_mass, internal_force.cpp:
 * 442-448, 571-572, 597-598; also inside fx_cg, fx_newmark and the other solvers), am v to
 * fx_force (:366-374), m and am to the lumped masses of fx_lumped_mass and fx_explicit_cd
 * (batches.cpp:252-258), and m to fx_mass_times_ground (:619-625). In a model of several batches
 * give them to one batch. Set-up call (uploads; replaces the previous nodal masses). */
int fx_batch_set_nodal_mass(fx_batch* b, const double* mass, const double* alpha_mass);
/* InternalForce::mass_times_ground (internal_force.cpp:601-626) for a model of this one batch:
 * y = M r for the rigid unit translation r in `direction` (0..2) at every element DOF, fixed ones
 * included, plus the nodal masses of that direction. y: device [neq], reference numbering. */
int fx_mass_times_ground(fx_batch* b, int direction, double* y);
/* Timing aid: the operator kernel alone, `reps` times back to back, device time in ms
 * (HIP events on the batch's stream). */
int fx_operator_time(fx_batch* b, const fx_coeffs* k, int precision, const void* x, void* y, int reps,
                     float* ms);

/* ---- solvers (src/solvers/linear_solver.hpp:79-105, cg.cpp:24-72, jacobi.cpp:10-26) ------- */
typedef struct fx_solve_options {
  double relative_tolerance; /* stop when ||r|| <= max(rtol ||b||, atol) (cg.cpp:34) */
  double absolute_tolerance;
  int32_t max_iterations;
} fx_solve_options;

typedef struct fx_solve_result {
  int32_t converged, iterations, indefinite;
  int32_t outer_iterations;  /* iterative refinement sweeps (mixed precision), else 0 */
  double rhs_norm, initial_residual, final_residual;
  double true_residual;      /* ||b - A x|| recomputed in FP64 at the end */
} fx_solve_result;

/* VectorOps::copy (vector_ops.cpp:79): dst = src, n doubles, as a plain device kernel. */
int fx_copy(fx_ctx* ctx, int64_t n, const double* src, double* dst);
/* VectorOps::dot (vector_ops.cpp:66-73): chunks of 2048, 64 lanes, fixed order. */
int fx_dot(fx_ctx* ctx, int64_t n, const double* x, const double* y, double* out);

/* The same dot into device memory, allocating nothing (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture, tight serving
 * loops). *out_dev (device) receives fx_dot's bits; the work is enqueued on the context's stream
 * and the call returns without waiting (fx_sync or a later blocking call waits). The reduction
 * scratch is a fixed 1.05 MB array of the library on each device, so n is limited to
 * FX_DOT_STATIC_MAX_N (FX_EINVAL above it), and calls that overlap on one device (several
 * contexts or threads on one GPU) race on it. */
#define FX_DOT_STATIC_MAX_N 268435456
int fx_dot_static(fx_ctx* ctx, int64_t n, const double* x, const double* y, double* out_dev);

/* Preconditioned CG with Jacobi, FP64 (ConjugateGradient::solve + Jacobi). x: initial guess in,
 * solution out. The plain drop-in path. */
int fx_cg(fx_batch* b, const fx_coeffs* k, const fx_solve_options* opt, const double* rhs, double* x,
          fx_solve_result* res);
/* FP32 inner CG (FP32 operator, FP32 vectors, FP64 dot products) inside FP64 iterative
 * refinement (FP64 residual b - A x with the FP64 operator, FP64 update); same stop rule on the
 * FP64 residual. inner_rtol: each inner solve's relative tolerance. */
/* Several CG + Jacobi solves advanced together, one per batch (no counterpart in the reference,
 * whose solvers take one system at a time; for serving loops that admit many small systems per
 * round): each poll round launches every unfinished solve's next iterations on the context's
 * stream (captured graphs with FX_OPT_GRAPHS) and reads every solve's state with one
 * synchronisation. Solve i gives the bits fx_cg(batches[i], k, opt, rhs[i], x[i], &res[i]) gives.
 * Elastic batches (fx_batch_create) of one context, each at most once. */
int fx_cg_many(int32_t n, fx_batch* const* batches, const fx_coeffs* k, const fx_solve_options* opt,
               const double* const* rhs, double* const* x, fx_solve_result* res);
/* FGMRES(30), CGS2, right Jacobi (Gmres::solve, gmres.cpp:32-123). Needs FX_BATCH_GMRES. */
int fx_gmres(fx_batch* b, const fx_coeffs* k, const fx_solve_options* opt, const double* rhs, double* x,
             fx_solve_result* res);
/* KrylovSolver::setup + solve (krylov.cpp:10-42) on the operator of k, whose symmetry is
 * operator_symmetry(batch tangent symmetry, k) (matrix_free_operator.cpp:5-16): a symmetric
 * positive definite operator runs CG and, if CG meets p^T A p <= 0, GMRES from the same initial
 * guess (iterations add up, indefinite set); any other operator (a PressureIndependMultiYield
 * batch without a mass term, negative coefficients) runs GMRES from the start. Needs
 * FX_BATCH_GMRES for the GMRES runs (FX_EINVAL when one is needed and the workspace is missing;
 * without it an SPD operator runs CG only). */
int fx_krylov(fx_batch* b, const fx_coeffs* k, const fx_solve_options* opt, const double* rhs, double* x,
              fx_solve_result* res);
int fx_cg_mixed(fx_batch* b, const fx_coeffs* k, const fx_solve_options* opt, double inner_rtol,
                const double* rhs, double* x, fx_solve_result* res);

/* ---- internal force and the explicit CentralDifference loop -------------------------------- */
/* Lumped mass m and lumped alphaM mass am per equation (batches.cpp:204-224); am may be NULL. */
int fx_lumped_mass(fx_batch* b, double* m, double* am);
/* F_body: the body force's nodal loads b * sum_gp N_j dvol (std_brick.hpp:519-542). */
int fx_body_load(fx_batch* b, double* fb);
/* f = internal force at (u, v) (InternalForce::compute, internal_force.cpp:354-377): for
 * ElasticIsotropic f = K u + C v - F_body. v may be NULL (no damping force). */
int fx_force(fx_batch* b, const double* u, const double* v, double* f);
/* Smallest element stable step (InternalForce::min_critical_dt; std_brick.hpp:658-694,
 * explicit_stability.hpp:15-32): scheme 0 CentralDifference, 1 ExplicitDifference. Host work;
 * lattice batches solve one eigenproblem per parameter set. */
int fx_critical_dt(fx_batch* b, int scheme, double* dt);

/* An excitation's time series (GroundMotions, src/exec/ground_motion.cpp:17-74): with time NULL,
 * PathSeries (uniform step dt from start, use_last holding the last value past the end; the
 * caller prepends OpenSees' leading zero if it wants one); with time set, PathTimeSeries
 * (linear between the given nondecreasing times, 0 outside them, use_last, dt and start unused),
 * whose interval walk is kept in *loc between calls as GroundMotions keeps it per pattern (loc
 * NULL: each fx_newmark / fx_explicit_cd call walks from interval 0). Either form is 0 before
 * time 0. */
typedef struct fx_time_series {
  const double* values;       /* host [count] */
  int32_t count;
  double dt, factor, start;
  int32_t use_last;
  const double* time;         /* host [count] or NULL */
  int32_t* loc;               /* host: the walk's interval, read at the start and written at the end of
                                 the call, or NULL */
} fx_time_series;
#define FX_MAX_POINT_LOADS 32
typedef struct fx_cd_desc {
  double dt;
  int32_t nsteps;
  const double* dash;        /* [neq] device: diagonal dampers outside the elements (e.g. Lysmer-
                                Kuhlemeyer boundaries), NULL = none: their force dash * v_e joins
                                the element forces at the garbage velocity and dash joins the
                                tangent's C (ExplicitDifference: the force dash * Utdot, no tangent
                                term) */
  int32_t nload;             /* point loads, <= FX_MAX_POINT_LOADS */
  const int32_t* load_eq;    /* [nload] host: equations */
  const double* load_hist;   /* [nsteps * nload] host: P(t_k) of load j at [k * nload + j] */
  int32_t scheme;            /* 0 CentralDifference, 1 ExplicitDifference */
  int32_t ground_dir;        /* UniformExcitation direction 0..2, or -1 */
  const double* ground_accel; /* host [nsteps]: a_g at each step's load time (t(n) for CentralDifference,
                                 t(n+1) for ExplicitDifference), used when ground_series is NULL */
  const fx_time_series* ground_series; /* the excitation's series, evaluated at those times, or NULL */
  double ground_factor;      /* UniformExcitation factor applied to ground_series */
  double time;               /* the stage's time before the first step */
  fx_recorder* recorder;     /* NULL, or records u, the recorded v and a after every step
                                (explicit_analysis.cpp:724-731); equations in the model numbering */
} fx_cd_desc;
/* nsteps explicit steps with the batch's lumped mass (the nodal masses included), the material
 * history committed after every step (Domain::commit), and the ground load -m a_g of a uniform
 * excitation at the equations of its direction (ground_load, explicit_analysis.cpp:48-58):
 *   scheme 0, CentralDifference (:87-117, 661-662, 712-722), with the lumped alphaM mass: from
 *     u = U(n), um = U(n-1), v, a (device, in/out) to u = U(n+nsteps), um = U(n+nsteps-1) and the
 *     recorded v, a of the last step;
 *   scheme 1, ExplicitDifference (:60-85, 699-710): u, a and um = the element velocity Utdot
 *     (the caller sets it to the committed velocity at a stage's start, :680) in/out, v the
 *     recorded velocity out; the elements' damping acts at Utdot; needs a positive lumped mass
 *     at every equation (check_masses). */
int fx_explicit_cd(fx_batch* b, const fx_cd_desc* d, double* u, double* um, double* v, double* a);

/* ---- PressureIndependMultiYield (J2 multi-surface soil) ------------------------------------ */
/* PressureIndependMultiYield<NS = 20>::Params from the constructor's arguments
 * (pressure_independ_multi_yield.hpp:56-61, 110-139) plus the element's body force and mass type. */
typedef struct fx_pimy_material {
  double rho, G, K, cohesion, gamma_peak, phi, pref, pdc; /* refShearModul, refBulkModul, cohesi,
                                                             peakShearStra, frictionAng, refPress,
                                                             pressDependCoe */
  int32_t nsurf;          /* yield surfaces, <= 20 (0: 10) */
  const double* gredu;    /* host [2 * nsurf] (gamma, G/Gmax) pairs, or NULL: automatic surfaces */
  double b[3];            /* body force per unit volume */
  int32_t mass_type;      /* 0 consistent, 1 lumped */
} fx_pimy_material;
#define FX_PIMY_NSTATE 190  /* doubles of history per material point (30 + 8 NS) */
/* The backbone surfaces (setUpSurfaces, :275-381), host only: size[21], modulus[21] (index 0
 * unused). FX_EINVAL with the reason for parameters the reference rejects. */
int fx_pimy_surfaces(const fx_pimy_material* m, double* size, double* modulus);
/* Material-point driver on device arrays, structure of arrays (field k of point i at
 * [k * npts + i]). mat[npts] (device) picks one of nmat <= 8 materials; load_stage[nmat] (host)
 * is each material's stage (updateMaterialStage: 0 elastic, 1 plastic).
 *   fx_pimy_init_state: history at rest (initial_state, :142-150), state[FX_PIMY_NSTATE * npts].
 *   fx_pimy_points: update at strain eps[6 * npts] (engineering shear, order 11 22 33 12 23 31)
 *     from the committed history into the trial history (update :228-240); sig[6 * npts],
 *     D[36 * npts] (row-major tangent, may be NULL). NaN stress marks a failed point. */
int fx_pimy_init_state(fx_ctx* ctx, int32_t nmat, const fx_pimy_material* mats, int64_t npts, const int32_t* mat,
                       double* state);
int fx_pimy_points(fx_ctx* ctx, int32_t nmat, const fx_pimy_material* mats, const int32_t* load_stage, int64_t npts,
                   const int32_t* mat, const double* committed, double* trial, const double* eps, double* sig,
                   double* D);

/* ---- stdBrick<PressureIndependMultiYield> batches --------------------------------------------- */
typedef struct fx_pimy_batch_desc {
  int32_t count, padded;
  const int32_t* node;             /* [8 * padded] as fx_batch_desc */
  const int32_t* eq;               /* [24 * padded] */
  const int32_t* param;            /* [padded] */
  int32_t nparam;
  const fx_pimy_material* params;  /* [nparam] */
  int32_t nnode;
  const double* coords;            /* [3 * nnode] SoA */
  int32_t neq;
  const fx_rayleigh* damp;         /* [nparam] Rayleigh damping per parameter set, or NULL */
} fx_pimy_batch_desc;
/* A batch whose material points (8 per element, point e * 8 + gp) carry PIMY history (committed
 * and trial, at rest at creation). fx_force runs the return map at every point (the trial
 * history; InternalForce::compute), with the Rayleigh damping force at v when v is given
 * (std_brick.hpp:409-464: betaK on the updated tangent, betaKinit on the initial, betaKcomm on the
 * committed tangent, alphaM on the mass); fx_operator_apply / _diagonal / fx_cg use the tangents
 * kept by the last fx_store_tangents (store_tangents, internal_force.cpp:532-543; D_eff = cK D_t +
 * c0 D_0 + cc D_c with the factors of element_operator.hpp:45-56), fx_explicit_cd commits after
 * every step. No FP32 operator. */
/* With FX_BATCH_SSPBRICK: SSPbrick<PressureIndependMultiYield> (ssp_brick.hpp:1056-1156), one material
 * point per element (the element's point 0 of the eight-point history layout), Kstab from the
 * material's initial tangent (element_constants, :881), the lumped mass, the initial stiffness
 * taken as the tangent in the operator's factors; no compacted return map, FP32 operator or
 * stable step (FX_EUNSUP). */
int fx_pimy_batch_create(fx_ctx* ctx, const fx_pimy_batch_desc* desc, uint32_t flags, fx_batch** out);
/* updateMaterialStage for every parameter set (0 elastic, 1 plastic). No-op on elastic batches. */
int fx_set_material_stage(fx_batch* b, int32_t stage);
/* The return map at u (trial history written), the tangents kept for the operator; f (may be
 * NULL) receives the internal force. No-op on elastic batches. */
int fx_store_tangents(fx_batch* b, const double* u, double* f);
/* commit_state: the trial history becomes the committed one (Domain::commit). */
int fx_commit_state(fx_batch* b);
/* Device copy of the history, [FX_PIMY_NSTATE * npts] structure of arrays (committed or trial). */
int fx_pimy_state(fx_batch* b, int committed, double* dst);

/* ---- implicit Newmark stage with Newton (implicit_stage.cpp) ---------------------------------- */
typedef struct fx_newmark_desc {
  double gamma, beta, dt;     /* Newmark constants and step */
  int32_t nsteps;
  int32_t test;               /* 0 NormUnbalance, 1 NormDispIncr, 2 EnergyIncr (0.5 |dU . B|) */
  double tol;
  int32_t max_iter;           /* Newton iterations per step */
  double cg_rtol;             /* the solver's relative tolerance (the reference's default 1e-12) */
  int32_t cg_max_it;
  int32_t ground_dir;         /* UniformExcitation direction 0..2, or -1 */
  const double* ground_accel; /* host [nsteps]: the ground acceleration at the end of each step (used
                                 when ground_series is NULL) */
  /* recovery of a failed step (implicit_stage.cpp:262-267, 385-402): max_retries more attempts at
   * the same size, then, while the level is below num_sublevels, num_substeps steps of 1/num_substeps
   * the size, each recovered the same way one level down. 0, 0: none. */
  int32_t max_retries, num_sublevels, num_substeps;
  const fx_time_series* ground_series; /* the excitation's acceleration series, evaluated at the end time
                                 of every attempt (GroundMotions::evaluate), or NULL; needed to substep
                                 under an excitation */
  double ground_factor;       /* UniformExcitation factor applied to ground_series (OpenSees -fact) */
  double time;                /* the stage's start time (the domain time before the first step) */
  int32_t mixed;              /* 1: each Newton system by mixed-precision CG (FP32 operator and inner
                                 CG to inner_rtol, FP64 residuals and refinement to cg_rtol; batches
                                 created with FX_BATCH_FP32), in place of FP64 CG (and its GMRES
                                 fallback) */
  double inner_rtol;
  /* the stage's analysis and algorithm (ImplicitStage::run, implicit_stage.cpp:118-402) */
  int32_t analysis;           /* 0 transient (Newmark form D), 1 static (LoadControl: the elements see
                                 v = 0, no inertia, u += dU; gamma, beta and dt unused; v and a kept) */
  int32_t algorithm;          /* 0 Newton; 1 Linear (one solve per step, which must converge; refused
                                 for materials with history, :130-140); 2 ModifiedNewton (the tangent once
                                 per attempt, after the first unbalance, :324-328) */
  int32_t modified_initial;   /* ModifiedNewton -initial: the initial stiffness K_0 in the stiffness term
                                 (OperatorCoefficients::initial, :146-150) */
  int32_t factor_once;        /* ModifiedNewton -factoronce: the tangent kept across steps until a test
                                 fails or the step size changes (:166-170, 324-328, 343) */
  double load_increment;      /* LoadControl dlambda (static): the time advances by it per step */
  double load_min_increment, load_max_increment; /* LoadControl dlambdaMin, dlambdaMax; both 0: not
                                 given (the increment stays fixed) */
  int32_t load_num_iter;      /* LoadControl numIter: with min and max given, dlambda *= numIter / (the last
                                 attempt's iterations), clamped to [min, max] (:262-269, 277-283) */
  fx_recorder* recorder;      /* NULL, or records u, v, a after every commit, substeps included
                                 (implicit_stage.cpp:371-381); equations in the model numbering */
} fx_newmark_desc;
typedef struct fx_newmark_result {
  int32_t steps;              /* converged steps of size dt (each may have been substepped) */
  int32_t newton_iterations, max_newton_iterations, krylov_iterations;
  int32_t failed;             /* 1: step `steps` could not be recovered; its start state is restored */
  double last_test;
  int32_t commits;            /* converged attempts (substeps included) */
  int32_t failed_attempts, retries, substepped;
  double time;                /* the domain time after the last converged attempt */
} fx_newmark_result;
/* nsteps steps of an implicit stage (ImplicitStage::run): transient Newmark steps, each solved by
 * the algorithm on K + gamma/(beta dt) C + 1/(beta dt^2) M, or static LoadControl steps on K
 * (stored tangents for history materials), committing the material history after each converged
 * attempt. The solver is the reference's KrylovSolver choice (krylov.cpp:10-22): PCG + Jacobi for
 * a symmetric positive definite operator (with the GMRES fallback when the batch has
 * FX_BATCH_GMRES), GMRES(30) + Jacobi from the start otherwise (a PressureIndependMultiYield
 * batch in a static stage: FX_BATCH_GMRES needed). u, v, a, du: device vectors in/out (du: the
 * Krylov initial guess, the last increment out). Needs FX_BATCH_DYNAMIC (or a PIMY batch). A step
 * that fails is retried and (transient only) substepped as desc asks (the reference's recover,
 * :385-402); one that cannot be recovered returns FX_OK with result.failed = 1 (the reference
 * stops the stage with a ModelError). */
int fx_newmark(fx_batch* b, const fx_newmark_desc* d, double* u, double* v, double* a, double* du,
               fx_newmark_result* result);

/* ---- batched soil columns ------------------------------------------------------------------
 * Many independent 1 x 1 x N stdBrick<PressureIndependMultiYield> columns (the layered soil
 * column of synthetic_models.hpp:14-91: base fixed, the side nodes of each level tied by
 * equalDOF in DOFs 1 2 3), each running its own implicit Newmark stage with Newton and PCG +
 * Jacobi (implicit_stage.cpp:150-372, cg.cpp:24-72), advanced together on one GPU. No
 * counterpart entry point in the reference (it runs one model per analysis); each column's
 * result is that analysis' result within the stated tolerance, and does not depend on which
 * other columns share the pool.
 *
 * The four nodes of a level share their three equations, so on the column's box elements the
 * strain is uniform: eps33, gamma23, gamma31 from the level displacement differences, the rest
 * zero. All eight Gauss points of an element see the same strain history and hold the same
 * state, so the pool keeps one material point per element (one history, one return map) and
 * integrates it with weight 8. Columns advance asynchronously: each has its own step, Newton
 * iteration and CG scalars, so columns can be added while others run. */
#define FX_COLUMN_MAX_ELEMS 512
#define FX_COLUMN_MIXED 1u    /* inner CG in FP32 on FP32 blocks, inside FP64 iterative refinement */
#define FX_COLUMN_NO_GRAPH 2u /* launch the iteration kernels one by one instead of replaying a hipGraph */
#define FX_COLUMN_WAVES(n) ((uint32_t)(n) << 8) /* register budget of the return-map kernels: at least n
                                 waves per SIMD (3 or 4; 1: the compiler's choice); default 3 */
#define FX_COLUMN_ONE_STREAM 8u /* the four solve sizes one after another on the pool's stream (default:
                                 forked onto four side streams so they overlap; one stream leaves the
                                 other hardware queues to co-running work) */
#define FX_COLUMN_TUNE 16u    /* the return-map kernels' register budget (1, 3 or 4 waves per SIMD) by autotune:
                                 each timed on the first running columns (the same results: the budgets
                                 differ only in register allocation), the fastest kept per device and pool
                                 in the autotune cache; FX_COLUMN_WAVES(n) fixes it instead */
#define FX_COLUMN_COMPACT 4u  /* compacted return map (compacted_kernels.hpp:85-202): an elastic predictor
                                 per point, the plastic points listed, one full update per listed point
                                 (same bits as the fused kernel) */
typedef struct fx_column_pool fx_column_pool;
typedef struct fx_column_pool_desc {
  int64_t max_elements;       /* element slots of all resident columns together */
  int32_t max_columns;        /* columns resident at once */
  int32_t max_materials;      /* parameter sets resident at once */
  int64_t max_record;         /* doubles for excitation records and per-step top-level histories */
  double gamma, beta, dt;     /* Newmark constants and step (shared by the pool) */
  int32_t test;               /* 0 NormUnbalance, 1 NormDispIncr */
  double tol;
  int32_t max_iter;           /* Newton iterations per step */
  double cg_rtol;             /* PCG relative tolerance (the reference's default 1e-12) */
  int32_t cg_max_it;
  uint32_t flags;             /* FX_COLUMN_* */
  double inner_rtol;          /* FX_COLUMN_MIXED: relative reduction of each FP32 inner solve */
  /* recovery of a failed step, per column, as fx_newmark (implicit_stage.cpp:385-402): max_retries
   * more attempts at the same size, then num_substeps steps of 1/num_substeps the size down to
   * num_sublevels levels (at most 6); a substep's ground acceleration is the record read as a path
   * series with a leading zero (linear within a step from the previous step's value). 0, 0: a
   * failed step fails the column. */
  int32_t max_retries, num_sublevels, num_substeps;
} fx_column_pool_desc;
typedef struct fx_column_desc {
  int32_t nelem;              /* elements, bottom to top (1..FX_COLUMN_MAX_ELEMS) */
  double hx, hy, hz;          /* element size */
  int32_t nmat;               /* parameter sets of this column */
  const fx_pimy_material* mats;
  const int32_t* mat_of;      /* host [nelem]: parameter set of each element, bottom to top */
  int32_t stage;              /* material stage: 0 elastic, 1 elastoplastic */
  int32_t nsteps;
  int32_t ground_dir;         /* UniformExcitation direction 0..2 */
  const double* ground_accel; /* host [nsteps]: the ground acceleration at the end of each step */
  int32_t record;             /* 1: keep u, v, a of the top level after each step */
} fx_column_desc;
typedef struct fx_column_result {
  int32_t status;             /* 0 running, 1 finished, 2 failed (the failed step's start is kept) */
  int32_t steps;              /* converged (committed) steps */
  int32_t newton_iterations, max_newton_iterations, krylov_iterations;
  double last_test;
  int32_t commits;            /* converged attempts (substeps included) */
  int32_t failed_attempts, retries, substepped;
} fx_column_result;
/* Allocates all device memory of the pool (nothing is allocated per column); FX_ENOMEM over the
 * context's budget. */
int fx_column_pool_create(fx_ctx* ctx, const fx_column_pool_desc* d, fx_column_pool** out);
void fx_column_pool_destroy(fx_column_pool* p);
/* Admits a column at rest (u = v = a = 0, histories at their initial state) and starts its
 * first step; *id names it until fx_column_remove. FX_EFULL when the pool lacks element slots,
 * materials, record space or a column slot now. */
int fx_column_add(fx_column_pool* p, const fx_column_desc* d, int32_t* id);
/* Runs `iterations` Newton iterations of every running column (asynchronous columns: a column
 * that converges commits, records and predicts its next step within the same iteration);
 * *running: columns still running afterwards. Blocking. */
int fx_column_run(fx_column_pool* p, int32_t iterations, int32_t* running);
int fx_column_status(fx_column_pool* p, int32_t id, fx_column_result* r);
/* Host only (no device call): the columns that had finished or failed at the end of the last
 * fx_column_run and are not removed yet; up to cap ids, *count = how many there are. */
int fx_column_finished(fx_column_pool* p, int32_t* ids, int32_t cap, int32_t* count);
/* Host copies of a column: u, v, a [3 nelem] (level 1..N, x y z; null to skip) and the
 * recorded history [9 nsteps] (u, v, a of the top level per committed step; null to skip). */
int fx_column_fetch(fx_column_pool* p, int32_t id, double* u, double* v, double* a, double* history);
/* The column's committed histories [FX_PIMY_NSTATE * nelem], field-major (host). */
int fx_column_state(fx_column_pool* p, int32_t id, double* state);
/* Frees the column's slots (a running column is stopped). */
int fx_column_remove(fx_column_pool* p, int32_t id);

/* Node recorders (exec/recorders.hpp:5-20; recorders.cpp:193-240, 272-298): OpenSees'
 * `recorder Node -file f [-time] [-dT dT] [-precision p] -node ... -dof ... disp|vel|accel`, fed
 * from device vectors. A recorder is due after a commit at `time` when dT == 0 or
 * time - next >= -dT * 1e-5 (then next = time + dT); a record is one line: the time (echo_time),
 * then each column's value, every number printed as %.{precision}g, separated by one space. A
 * column is an equation of the model numbering, or -1 for a fixed DOF (printed 0). At a record
 * one gather launch copies the due recorders' columns into a device staging array, one
 * asynchronous copy on the context's stream takes them to pinned host memory (a ring of 2 to 32
 * buffers within 64 MiB: the host waits for a copy only when every buffer holds a record in
 * flight), and a writer thread formats and writes the files. */
int fx_recorder_create(fx_ctx* ctx, fx_recorder** out);
/* Adds a Node recorder (response 0 disp, 1 vel, 2 accel; ncolumns equations, node-major with the
 * DOFs of each node); *index names it. Every recorder is added before the first activation. */
int fx_recorder_add_node(fx_recorder* r, const char* file, int32_t response, int32_t echo_time, double dT,
                         int32_t precision, int64_t ncolumns, const int32_t* eq, int32_t* index);
/* Starts recording (the reference's `recorder` process step, recorders.cpp:181-193): its file is
 * created or truncated by the writer. */
int fx_recorder_activate(fx_recorder* r, int32_t index);
/* After a committed step at `time`: records every active recorder that is due, from device
 * vectors of n equations in the model numbering (a vector nobody due records may be NULL).
 * Asynchronous: the host waits only when every pinned buffer holds a record in flight. */
int fx_recorder_record(fx_recorder* r, double time, int64_t n, const double* u, const double* v, const double* a);
/* Writes every pending record and closes the files; FX_EIO names a file that failed. Records
 * after it are refused. */
int fx_recorder_finish(fx_recorder* r);
int64_t fx_recorder_records(const fx_recorder* r);  /* lines written so far */
void fx_recorder_destroy(fx_recorder* r);           /* finishes first */

#ifdef __cplusplus
}
#endif
#endif
