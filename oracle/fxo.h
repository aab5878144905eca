/*
 * fxo — the CPU oracle for the hex8 matrix-free operator path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's algorithms
 * (amnp95/FemoraX) for the hot path: the compile-step data the kernels read (RCM node order,
 * element order, DOF numbering, SoA maps, greedy colouring), the stdBrick element with
 * ElasticIsotropic and PressureIndependMultiYield, the matrix-free element operator and its
 * diagonal, the deterministic gather assembly, the fixed-order dot product, CG + Jacobi,
 * Newmark and CentralDifference. Only the tests, the smoke check and the CPU-baseline leg of
 * the benchmark load it, and only as the checker. The product (femorax_amd) never links it.
 *
 * Every function names the reference file:line it restates. Where the restatement keeps the
 * reference's arithmetic order (so that the goldens the reference holds match bit for bit on
 * a compiler that does not contract multiply-adds), that is said at the function.
 *
 * Pinning (how each part is checked, in tests/):
 *   brick geometry/force/tangent/mass/damping  reference goldens tests/reference/data/element
 *                                              (Brick_*), and oracle/_ref (reference headers)
 *   SSPbrick force/tangent/mass/damping        reference goldens element/SSPbrick_* (16 files)
 *   ElasticIsotropic                           goldens nd/ElasticIsotropic_*, and oracle/_ref
 *   PIMY return map                            goldens nd/PressureIndependMultiYield_* (34)
 *   element operator / diagonal                K·v and diag(K) from the Brick goldens, oracle/_ref
 *   RCM, element order, colouring              closed forms on small graphs (path, grid, star),
 *                                              invariants (bandwidth, BFS level structure)
 *   DOF numbering                              goldens constraints/... (OpenSees equivalence classes)
 *   gather assembly / apply                    brute-force assembled K·x on small meshes
 *   dot                                        exact rational sums; lane order vs naive sum
 *   CG + Jacobi                                dense direct solve on small meshes
 * Nothing here is "parity unpinned" except where a header comment says so.
 */
#ifndef FXO_H
#define FXO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- LCG (tests/reference/drivers/driver_common.hpp:21-32) ------------------------------ */
double fxo_lcg_symmetric(uint64_t* state);

/* ---- Model (the subset of io::Model the path uses, src/io/model.hpp:74-186) -------------- */
typedef struct fxo_model {
  int32_t nnode;
  const double* xyz;       /* [nnode*3], row major, model (file) order */
  int32_t nelem;
  const int32_t* conn;     /* [nelem*8] model node indices, OpenSees Brick node order */
  const int32_t* group;    /* [nelem] element group (file order; one material per group) */
  const int32_t* region;   /* [nelem] Rayleigh region or -1; NULL = none */
  int32_t nfix;
  const int32_t* fix_node; /* model node index */
  const int32_t* fix_dof;  /* 0..2 */
  int32_t ntie;            /* equalDOF rows */
  const int32_t* tie_retained;
  const int32_t* tie_constrained;
  const int32_t* tie_dof;  /* 0..2 */
  const int32_t* tie_command; /* equalDOF command id per row, NULL: per (retained, constrained) */
} fxo_model;

typedef struct fxo_compiled {
  int32_t nnode, nelem, padded, neq, nparam, ncolor;
  int32_t* node_order;   /* [nnode] internal -> model index */
  int32_t* node_rank;    /* [nnode] model -> internal index */
  int32_t* elem_order;   /* [nelem] batch position -> global element index */
  int32_t* node_eq;      /* [3*nnode] by internal node: equation or -1 */
  double* coords;        /* [3*nnode] SoA by internal node */
  int32_t* node;         /* [8*padded] internal node, -1 padding */
  int32_t* eq;           /* [24*padded] equation, -1 fixed/padding */
  int32_t* param;        /* [padded] param set, 0 padding */
  int32_t* param_group;  /* [nparam] */
  int32_t* param_region; /* [nparam] */
  int32_t* color_start;  /* [ncolor+1] */
  int32_t* color_elements; /* [nelem] */
  /* the deterministic gather (internal_force.cpp:121-134, assembly.cpp:64-83): each
   * equation's slots k*padded+e in (element, DOF) order, CSR by equation */
  int64_t* gather_start; /* [neq+1] */
  int64_t* gather_list;  /* [nnz] */
  double* slots;         /* [24*padded] work */
} fxo_compiled;

/* src/compile/renumber.cpp:30-183, dof_numbering.cpp:21-40,269-485, batches.cpp:16-227.
 * Returns 0, or the pattern of a model the reference refuses (dof_numbering.cpp:162-339);
 * node i has tag i + 1. */
#define FXO_REFUSE_COMMANDS 2      /* a node tied by two equalDOF commands */
#define FXO_REFUSE_FIX_LOST 3      /* a fix on a tied DOF whose retained DOF is free */
#define FXO_REFUSE_PARTIAL_CHAIN 4 /* C -> R -> M with different DOF lists */
#define FXO_REFUSE_LARGER_TAG 5    /* C -> R -> M with R's tag larger than C's */
#define FXO_REFUSE_OTHER 6
int fxo_compile(const fxo_model* m, int renumber_rcm, int simd_width, int build_coloring,
                fxo_compiled* out);
void fxo_compiled_free(fxo_compiled* c);
/* Largest |rank(a)-rank(b)| over node pairs sharing an element (renumber.cpp:185-199). */
int32_t fxo_node_bandwidth(const fxo_model* m, const int32_t* node_rank);

/* ---- Element parameters ------------------------------------------------------------------ */
typedef struct fxo_rayleigh { double alphaM, betaK, betaKinit, betaKcomm; } fxo_rayleigh;

/* stdBrick<ElasticIsotropic>::Params (std_brick.hpp:181-185, elastic_isotropic.hpp:30-34), or
 * SSPbrick<ElasticIsotropic>::Params (ssp_brick.hpp:860-863; always lumped) when element = 1:
 * the element functions below and every mesh-level function dispatch on it */
typedef struct fxo_elastic {
  double E, nu, rho;
  double b[3];
  int32_t mass_type; /* 0 consistent, 1 lumped */
  fxo_rayleigh damp;
  int32_t element;   /* 0 stdBrick, 1 SSPbrick */
} fxo_elastic;

typedef struct fxo_coeffs { double stiffness, damping, mass; int32_t initial; } fxo_coeffs;

/* ---- stdBrick (std_brick.hpp) -------------------------------------------------------------- */
/* shp3d at Gauss point gp (std_brick.hpp:45-152, 203-211): shp[4][8] (dN/dx, dN/dy, dN/dz, N) */
void fxo_brick_shape(const double xl[3][8], int gp, double shp[4][8], double* dvol);
/* Elastic D (elastic_isotropic.hpp:66-94). */
void fxo_elastic_tangent(double E, double nu, double D[6][6]);
/* Brick::compute<WantK> with ElasticIsotropic (std_brick.hpp:481-543): fe and optionally Ke. */
void fxo_brick_force_elastic(const fxo_elastic* p, const double xe[8][3], const double* ue,
                             const double* ve, double* fe, double* Ke /* 24*24 or NULL */);
/* Brick::mass (std_brick.hpp:582-605) and lumped_mass (:609-623). */
void fxo_brick_mass(const fxo_elastic* p, const double xe[8][3], double* Me);
void fxo_brick_lumped_mass(double rho, const double xe[8][3], double m[8]);
/* element_operator_apply / _diagonal with a constant D (element_operator.hpp:82-209).
 * D: the tangent, D0: the initial tangent (may equal D), Dc: committed tangent (NULL: = D). */
void fxo_brick_operator_apply(const fxo_elastic* p, const fxo_coeffs* k, const double xe[8][3],
                              const double* D, const double* ve, double* ye);
void fxo_brick_operator_diagonal(const fxo_elastic* p, const fxo_coeffs* k, const double xe[8][3],
                                 const double* D, const int32_t* eq, double* de);

/* ---- SSPbrick (ssp_brick.hpp, fxo_ssp.c) ---------------------------------------------------- */
/* geom[33]: dNmod (node-major), mVol, the nodes' volume weights (opensees_mass: OpenSees' polyJac
 * formula, else the exact integral of N_i); consts[110]: gst gut gus gstu, then interior's upper
 * triangle (element_constants, ssp_brick.hpp:898-914) */
void fxo_ssp_element(const fxo_elastic* p, const double xe[8][3], int opensees_mass, double* geom, double* consts);
/* compute<WantK> (ssp_brick.hpp:1142-1193) with the set's Rayleigh damping */
void fxo_ssp_force(const fxo_elastic* p, const double xe[8][3], int opensees_mass, const double* ue, const double* ve,
                   double* fe, double* Ke);
void fxo_ssp_mass(const fxo_elastic* p, const double xe[8][3], int opensees_mass, double* Me);
void fxo_ssp_operator_apply(const fxo_elastic* p, const fxo_coeffs* k, const double xe[8][3], const double* ve,
                            double* ye);
void fxo_ssp_operator_diagonal(const fxo_elastic* p, const fxo_coeffs* k, const double xe[8][3], const int32_t* eq,
                               double* de);
double fxo_ssp_max_omega_sq(const fxo_elastic* p, const double xe[8][3]);

/* ---- Global operator (gather assembly, internal_force.cpp:121-134, 545-599) ---------------- */
/* y = K_eff x; params[nparam]. x, y in equation space. Deterministic slot-gather order. */
void fxo_apply(const fxo_compiled* c, const fxo_elastic* params, const fxo_coeffs* k,
               const double* x, double* y, int nthreads);
void fxo_diagonal(const fxo_compiled* c, const fxo_elastic* params, const fxo_coeffs* k, double* d,
                  int nthreads);
/* Lumped mass per equation (batches.cpp:204-224). */
void fxo_lumped_mass(const fxo_compiled* c, const fxo_elastic* params, double* m, double* am);
/* Body-force load vector: -f(u=0) restricted to the body term (std_brick.hpp:329-333). */
void fxo_body_load(const fxo_compiled* c, const fxo_elastic* params, double* b);

/* ---- VectorOps (solvers/vector_ops.cpp:19-124) -------------------------------------------- */
double fxo_dot(int64_t n, const double* x, const double* y, int lanes);

/* ---- CG + Jacobi (solvers/cg.cpp:24-72, jacobi.cpp:10-26) ---------------------------------- */
typedef struct fxo_solve_result {
  int32_t converged, iterations, indefinite;
  double rhs_norm, initial_residual, final_residual;
} fxo_solve_result;
int fxo_cg(const fxo_compiled* c, const fxo_elastic* params, const fxo_coeffs* k, const double* b,
           double* x, double rtol, double atol, int32_t max_it, int lanes, int nthreads,
           fxo_solve_result* res);

/* ---- PressureIndependMultiYield<NS> (pressure_independ_multi_yield.hpp) -------------------- */
#define FXO_PIMY_NS 20
typedef struct fxo_pimy_params {
  int32_t load_stage, nsurf;
  double rho, G, K, phi, gamma_peak, pref, cohesion, pdc, residual_press;
  double size[FXO_PIMY_NS + 1], modulus[FXO_PIMY_NS + 1];
} fxo_pimy_params;
/* in: rho, refShearModul, refBulkModul, cohesi, peakShearStra, frictionAng, refPress, pressDependCoe
 * (the constructor's order, :56-61); gredu: NULL or nsurf (gamma, G/Gmax) pairs. 0 = ok. */
int fxo_pimy_setup(const double* in, int nsurf, const double* gredu, fxo_pimy_params* p);
int fxo_pimy_nstate(void);
void fxo_pimy_initial_state(const fxo_pimy_params* p, double* state);
/* update (:228-240): sig, D at the trial strain eps from the committed buffer; trial written. */
void fxo_pimy_update(const fxo_pimy_params* p, const double* committed, double* trial, const double eps[6],
                     double sig[6], double D[36]);

/* ---- stdBrick<PressureIndependMultiYield> (fxo_brick.c, fxo_ops.c) ------------------------ */
/* committed_tangent (:162-168): getTangent from the committed history (stress, active surface,
 * surfaces, moduli) */
void fxo_pimy_committed_tangent(const fxo_pimy_params* p, const double* committed, double D[36]);
typedef struct fxo_pimy_elem {
  fxo_pimy_params mat;
  double b[3];       /* body force per unit volume */
  int32_t mass_type; /* 0 consistent, 1 lumped (SSPbrick: always lumped) */
  fxo_rayleigh damp; /* the element's Rayleigh damping */
  int32_t element;   /* 0 stdBrick (8 points), 1 SSPbrick (one point: the element's point 0) */
} fxo_pimy_elem;
/* the elastic tangent of the material's reference moduli (initial_tangent, :156-158) */
void fxo_pimy_initial_tangent(const fxo_pimy_params* p, double D[36]);
/* SSPbrick<PressureIndependMultiYield> (fxo_ssp.c): compute_with (ssp_brick.hpp:1107-1156) with the
 * point's update (point_update_shape :1056-1083, point_node_contribution :1087-1098) on history
 * C0 -> T0, Kstab from the material's initial tangent (element_constants :881); D: the point's
 * tangent out, or NULL. The operator and its diagonal on the stored tangent (element_operator.hpp
 * with initial_stiffness_is_tangent; the committed tangent from C0 when betaKcomm enters). */
void fxo_ssp_force_pimy(const fxo_pimy_elem* p, const double xe[8][3], const double* ue, const double* ve,
                        const double* C0, double* T0, double* fe, double* D);
void fxo_ssp_operator_apply_pimy(const fxo_pimy_elem* p, const fxo_coeffs* k, const double xe[8][3], const double* Dt,
                                 const double* C0, const double* ve, double* ye);
void fxo_ssp_operator_diagonal_pimy(const fxo_pimy_elem* p, const fxo_coeffs* k, const double xe[8][3],
                                    const double* Dt, const double* C0, const int32_t* eq, double* de);
/* ve: the element velocities (the Rayleigh damping force, std_brick.hpp:409-464), or NULL */
void fxo_brick_force_pimy(const fxo_pimy_elem* p, const double xe[8][3], const double* ue, const double* ve,
                          const double* const* C, double* const* T, double* fe, double* Dgp);
/* C: the points' committed histories (the committed tangent of betaKcomm), or NULL */
void fxo_brick_operator_apply_gp(const fxo_pimy_elem* p, const fxo_coeffs* k, const double xe[8][3],
                                 const double* Dgp, const double* const* C, const double* ve, double* ye);
void fxo_brick_operator_diagonal_gp(const fxo_pimy_elem* p, const fxo_coeffs* k, const double xe[8][3],
                                    const double* Dgp, const double* const* C, const int32_t* eq, double* de);
/* global: material points e * 8 + gp (batch order), history point-major (FXO_PIMY_NSTATE each),
 * tangents D[point][36]; v may be NULL (no damping velocity) */
void fxo_pimy_force(const fxo_compiled* c, const fxo_pimy_elem* params, const double* C, double* T,
                    const double* u, const double* v, double* f, double* D, int nthreads);
void fxo_apply_gp(const fxo_compiled* c, const fxo_pimy_elem* params, const fxo_coeffs* k, const double* D,
                  const double* C, const double* x, double* y, int nthreads);
void fxo_diagonal_gp(const fxo_compiled* c, const fxo_pimy_elem* params, const fxo_coeffs* k, const double* D,
                     const double* C, double* d, int nthreads);

/* ---- implicit Newmark + Newton (fxo_newmark.c) -------------------------------------------- */
typedef struct fxo_system {
  const fxo_compiled* c;
  const fxo_elastic* elastic; /* elastic parameter sets, or NULL */
  const fxo_pimy_elem* pimy;  /* PIMY parameter sets, or NULL */
  double* C;                  /* PIMY history (point-major), committed and trial */
  double* T;
  double* D;                  /* PIMY stored tangents [points * 36] */
  int nthreads;
} fxo_system;
/* PathSeries with a uniform step (ground_motion.cpp:17-33), or, with time set, PathTimeSeries
 * (:43-74: linear between the given times, its interval walk kept in *loc between calls; a NULL
 * loc starts every call from interval 0) */
typedef struct fxo_series {
  const double* values;
  int32_t count;
  double dt, factor, start;
  int32_t use_last;
  const double* time;
  int32_t* loc;
} fxo_series;
double fxo_series_value(const fxo_series* s, double fact, double t);
typedef struct fxo_newmark_opts {
  double gamma, beta, dt;
  int32_t nsteps;
  int32_t test;          /* 0 NormUnbalance, 1 NormDispIncr, 2 EnergyIncr */
  double tol;
  int32_t max_iter;
  double cg_rtol;
  int32_t cg_max_it;
  int32_t lanes;         /* dot order: 64 as the device */
  const double* ground_accel; /* [nsteps] a_g at the end of each step, or NULL */
  int32_t max_retries, num_sublevels, num_substeps; /* recover (implicit_stage.cpp:385-402) */
  const fxo_series* ground_series; /* a_g = ground_factor * series(t) at each attempt's end time, or NULL */
  double ground_factor;
  double time;           /* the stage's start time */
  /* ImplicitStage::run (implicit_stage.cpp:118-402) */
  int32_t analysis;      /* 0 transient, 1 static (LoadControl) */
  int32_t algorithm;     /* 0 Newton, 1 Linear, 2 ModifiedNewton */
  int32_t modified_initial, factor_once;
  double load_increment, load_min_increment, load_max_increment; /* min = max = 0: not given */
  int32_t load_num_iter;
} fxo_newmark_opts;
typedef struct fxo_newmark_stats {
  int32_t steps, newton_iterations, max_newton_iterations, krylov_iterations;
  double last_test;
  int32_t commits, failed_attempts, retries, substepped;
  double time;
} fxo_newmark_stats;
/* PCG + Jacobi on a system (cg.cpp:24-72) */
int fxo_cg_system(fxo_system* s, const fxo_coeffs* k, const double* b, double* x, double rtol, double atol,
                  int32_t max_it, int lanes, fxo_solve_result* res);
/* FGMRES(m), CGS2, right Jacobi (gmres.cpp:32-123, block_ops.cpp:59-90) on a system */
int fxo_gmres_system(fxo_system* s, const fxo_coeffs* k, const double* b, double* x, double rtol, double atol,
                     int32_t max_it, int32_t m, int lanes, fxo_solve_result* res);
/* y = M r, r the rigid unit translation in direction dir at every element DOF (fixed ones
 * included): InternalForce::mass_times_ground */
void fxo_mass_times_ground(const fxo_compiled* c, const fxo_elastic* elastic, const fxo_pimy_elem* pimy, int dir,
                           double* y, int nthreads);
/* nsteps Newmark steps with Newton (implicit_stage.cpp); ground_direction: the excitation's
 * direction 0..2 (its load -a_g M r from fxo_mass_times_ground), or -1. Returns 0, or 1 at the
 * first failed step (its state restored). du: the Krylov initial guess in, the last increment out. */
int fxo_newmark(fxo_system* s, const fxo_newmark_opts* o, int ground_direction, double* u, double* v, double* a,
                double* du, fxo_newmark_stats* st);

/* ---- Internal force, stable step, CentralDifference (fxo_time.c) ------------------------- */
/* f = gather of the elements' stdBrick forces at u (damping force from v; body force), the
 * flat force path (internal_force.cpp:354-377, element_kernels.hpp:198-235). */
void fxo_force(const fxo_compiled* c, const fxo_elastic* params, const double* u, const double* v, double* f,
               int nthreads);
/* sym_eigen.hpp:11-49 (overwrites A) */
double fxo_max_eigenvalue_symmetric24(double A[24][24]);
/* largest omega^2 of M_l^-1/2 K M_l^-1/2 (std_brick.hpp:658-687) */
double fxo_brick_max_omega_sq(const fxo_elastic* p, const double xe[8][3]);
/* scheme 0 CentralDifference, 1 ExplicitDifference (explicit_stability.hpp:15-32) */
double fxo_stable_dt(int scheme, const fxo_rayleigh* damp, double omega_sq);
double fxo_critical_dt(const fxo_compiled* c, const fxo_elastic* params, int scheme);
/* nsteps CentralDifference steps (explicit_analysis.cpp:87-117, 712-722) from (u = U_t,
 * um = U_t-1, v, a); m, cm: the lumped mass and the damping diagonal of the tangent
 * c3 M + c2 C (the reference's lumped alphaM mass; NULL = 0); dash: a diagonal damper outside the
 * elements (NULL = none), whose force dash * v_elem joins the element forces at the garbage
 * velocity v_elem = -c2 um, as element damping does; point loads load_hist[k*nload+j] on
 * equation load_eq[j] at t_k join the residual. */
void fxo_cd_run(const fxo_compiled* c, const fxo_elastic* params, const double* m, const double* cm,
                const double* dash, int32_t nload, const int32_t* load_eq, const double* load_hist, double dt, int32_t nsteps,
                double* u, double* um, double* v, double* a, int nthreads);
/* The explicit stage loop (explicit_analysis.cpp:48-117, 687-722): scheme 0 CentralDifference (um =
 * U(n-1)), 1 ExplicitDifference (um = the element velocity Utdot); ground_dir >= 0: the uniform
 * excitation's load -m a_g at the equations of that direction, ground_accel[k] at step k's load
 * time (t(n) CentralDifference, t(n+1) ExplicitDifference). */
void fxo_explicit_run(const fxo_compiled* c, const fxo_elastic* params, const double* m, const double* cm,
                      const double* dash, int32_t nload, const int32_t* load_eq, const double* load_hist, double dt,
                      int32_t nsteps, int scheme, int ground_dir, const double* ground_accel, double* u, double* um,
                      double* v, double* a, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
