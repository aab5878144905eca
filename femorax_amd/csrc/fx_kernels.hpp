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
  const T* x;
  T* y;
  T* ynode;
  int accumulate;
  const int32_t* done;  // solver stop flag: the launch is a no-op once it is set
};

int operator_apply64(fx_batch* b, const fx_coeffs& k, const double* x, double* y, const int32_t* done = nullptr);
int operator_apply32(fx_batch* b, const fx_coeffs& k, const float* x, float* y, const int32_t* done = nullptr);
int operator_diagonal64(fx_batch* b, const fx_coeffs& k, double* d);

// Solver-side launchers (fx_solver.hip).
int dot64(fx_ctx* ctx, int64_t n, const double* x, const double* y, double* out);
int cg64(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* rhs, double* x,
         fx_solve_result& r);
int cg_mixed(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, double inner_rtol, const double* rhs,
             double* x, fx_solve_result& r);

}  // namespace fx
