// Kernel argument blocks and launcher declarations shared by the HIP translation units.
#pragma once

#include "fx_internal.hpp"

namespace fx {

template <typename T>
struct KernelArgs {
  fx_coeffs k;
  int32_t padded, nnode;
  const int32_t* node;
  const int32_t* param;
  const int32_t* node_eq;
  const double* coords64;
  const float* coords32;
  const DevParam* params;
  const int32_t* elems;
  const int32_t* ln_start;
  const uint32_t* lnode;
  const uint32_t* ln_slots;
  const uint16_t* slot_list;
  const int32_t* ln_part;
  const uint16_t* lcorner;  // block-local node of each element corner (block order)
  int64_t ln_total;
  const int32_t* ln_eq;     // local nodes' equations and coordinates, [d * ln_total + ln]
  const double* ln_xyz64;
  const float* ln_xyz32;
  const T* x;
  T* y;
  T* ynode;
  int accumulate;
  const int32_t* done;  // solver stop flag: the launch is a no-op once it is set
  const double* ssp;    // SSPbrick: the elements' constants, value[k * ssp_stride + at], k < kSspConsts
  int64_t ssp_stride;   // padded (element order: the lattice path's per-element forms) or nblocks * kBlockElems
  const int32_t* ssp_pos;  // block order: each element's index at (null: at = e)
};

// Effective factors of element_operator.hpp:45-56 for an elastic material (D_0 = D_c = D):
// D_eff = (cK + c0 + cc) D, mass factor cM' = cM + cC alphaM.
__host__ __device__ inline void operator_factors(const DevParam& p, const fx_coeffs& k, double& fD, double& cM) {
  const double cK = k.initial ? k.damping * p.betaK : k.stiffness + k.damping * p.betaK;
  const double c0 = k.initial ? k.stiffness + k.damping * p.betaKinit : k.damping * p.betaKinit;
  const double cc = k.damping * p.betaKcomm;
  fD = cK + c0 + cc;
  cM = k.mass + k.damping * p.alphaM;
}

// ref = 1: x and y in the reference numbering (the batch's reference-numbering words)
template <typename T, bool DIAG>
int launch_box(fx_batch* b, const fx_coeffs& k, const T* x, T* y, int accumulate, const int32_t* done, int ref = 0);

int operator_apply64(fx_batch* b, const fx_coeffs& k, const double* x, double* y, const int32_t* done = nullptr);
int operator_apply32(fx_batch* b, const fx_coeffs& k, const float* x, float* y, const int32_t* done = nullptr);
int operator_diagonal64(fx_batch* b, const fx_coeffs& k, double* d);
// The same without the nodal masses (one batch's elements only).
int elements_apply64(fx_batch* b, const fx_coeffs& k, const double* x, double* y, const int32_t* done = nullptr);
int elements_diagonal64(fx_batch* b, const fx_coeffs& k, double* d);
// fx_run_operator's pieces: the per-element values into slots[a * padded + e] (elastic,
// fx_operator.hip; PressureIndependMultiYield, fx_pimy.hip), the batch's gather of such slots into
// y (accumulate: y += the sums), and y += the fast-path operator or diagonal (reference numbering).
int element_slots64(fx_batch* b, bool diagonal, const fx_coeffs& k, const double* x, int unit_dir, double* slots);
int pimy_slots64(fx_batch* b, bool diagonal, const fx_coeffs& k, const double* x, int unit_dir, double* slots);
int slots_gather64(fx_batch* b, const double* slots, double* y, int accumulate);
int add_elements64(fx_batch* b, bool diagonal, const fx_coeffs& k, const double* x, double* y);
// f += am v with the nodal masses' alphaM part (solver numbering).
int nodal_damping_force64(fx_batch* b, const double* v, double* f);

// The C-ABI operator entry points (reference equation numbering; lattice batches permute to
// and from their internal numbering around the internal apply), fx_operator.hip.
int external_apply64(fx_batch* b, const fx_coeffs& k, const double* x, double* y);
int external_apply32(fx_batch* b, const fx_coeffs& k, const float* x, float* y);
int external_diagonal64(fx_batch* b, const fx_coeffs& k, double* d);
// Permutations between the reference and the internal numbering (identity for general batches).
template <typename T>
int to_internal(fx_batch* b, const T* ext, T* in);
template <typename T>
int to_external(fx_batch* b, const T* in, T* ext);

// Solver-side launchers (fx_solver.hip).
int copy64(fx_ctx* ctx, int64_t n, const double* a, double* b);
int dot64(fx_ctx* ctx, int64_t n, const double* x, const double* y, double* out);
// out[c] = V(:, c) . w for c < ncol (V column-major, stride n), each in dot64's order, one host
// synchronisation; scratch: ncol * (chunks + level buffers + 1) doubles (gmres_workspace)
int multidot64(fx_ctx* ctx, int64_t n, const double* V, int ncol, const double* w, double* scratch, double* out);
int cg64(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* rhs, double* x,
         fx_solve_result& r);
int jacobi_inverse_public(fx_batch* b, const fx_coeffs& k, double* d, double* inv);
// GMRES(30) + Jacobi, and CG with the GMRES fallback, in the solver numbering (fx_gmres.hip);
// the C-ABI wrapper (method 1 GMRES, 2 CG then GMRES).
int gmres64_solver(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* bi, double* xi,
                   fx_solve_result& r);
int krylov64_solver(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* bi, double* xi,
                    fx_solve_result& r);
int krylov64(fx_batch* b, int method, const fx_coeffs& k, const fx_solve_options& o, const double* rhs, double* x,
             fx_solve_result& r);
// CG + Jacobi in the solver numbering; uses the batch's work vectors 0-5.
int cg64_many(int32_t nb, fx_batch* const* bs, const fx_coeffs& k, const fx_solve_options& o,
              const double* const* rhs, double* const* x, fx_solve_result* res);
size_t cg_red_doubles(int64_t n);
int cg64_solver(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* bi, double* xi,
                fx_solve_result& r);
int cg_mixed(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, double inner_rtol, const double* rhs,
             double* x, fx_solve_result& r);
// the same on internal-numbering vectors (bi, xi outside the batch's work arrays)
int cg_mixed_solver(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, double inner_rtol, const double* bi,
                    double* xi, fx_solve_result& r);

// stdBrick<PressureIndependMultiYield> batches (fx_pimy.hip).
int pimy_batch_init(fx_batch* b);
int pimy_force64(fx_batch* b, const double* u, const double* v, double* f, bool store_tangents);
int pimy_apply64(fx_batch* b, const fx_coeffs& k, const double* x, double* y, const int32_t* done = nullptr);
int pimy_apply32(fx_batch* b, const fx_coeffs& k, const float* x, float* y, const int32_t* done = nullptr);
int pimy_diagonal64(fx_batch* b, const fx_coeffs& k, double* d);

// Internal force and the CentralDifference loop (fx_explicit.hip), solver numbering.
int force64(fx_batch* b, const double* u, const double* v, double* f, double* scratch);
int cd_run(fx_batch* b, const fx_cd_desc& d, double* u, double* um, double* v, double* a);
// Node recorders (fx_recorder.hip): whether r records anything; the device map from the model
// numbering to the batch's solver numbering (null when they are the same); a record of the
// solver-numbered vectors u, v, a of n equations at `time` (nothing when r is null).
bool recorder_active(const fx_recorder* r);
bool recorder_reads_rates(const fx_recorder* r);  // an active recorder of velocities or accelerations
int batch_internal_map(fx_batch* b, const int32_t** map);
int recorder_record(fx_recorder* r, double time, int64_t n, const double* u, const double* v, const double* a,
                    const int32_t* map);

// z = x * y elementwise (fx_implicit.hip).
int multiply64(fx_ctx* ctx, int32_t n, const double* x, const double* y, double* z);

// An excitation's value at time t (PathSeries / PathTimeSeries times the pattern factor,
// ground_motion.cpp:17-74, 121-145); loc: the time array's interval walk (fx_implicit.hip).
double series_value(const fx_time_series& s, double fact, double t, int32_t& loc);

// The implicit Newmark stage (fx_implicit.hip).
int newmark_run(fx_batch* b, const fx_newmark_desc& d, double* u, double* v, double* a, double* du,
                fx_newmark_result& out);

}  // namespace fx
