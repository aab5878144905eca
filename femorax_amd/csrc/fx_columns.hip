// Batched soil columns: many independent 1 x 1 x N stdBrick<PressureIndependMultiYield> columns
// (the layered column of synthetic_models.hpp:14-91: base fixed, side nodes of each level tied
// in DOFs 1 2 3), each running its own implicit Newmark stage with Newton and PCG + Jacobi
// (implicit_stage.cpp:29-68, 150-372; cg.cpp:24-72; jacobi.cpp:10-26), advanced together.
//
// Tied column algebra. Level L (0 = the fixed base, 1..N free) has three equations shared by
// its four nodes. On an axis-aligned box element (size hx, hy, hz) the sums over a face's four
// shape functions have gradient (0, 0, -1/hz) (lower face) and (0, 0, 1/hz) (upper face), so
// the strain is uniform over the element:
//   eps33 = du_z / hz, gamma23 = du_y / hz, gamma31 = du_x / hz (du: upper - lower level),
// the other three components zero. Every Gauss point of an element sees the same strain
// history and so holds the same PIMY state: the pool keeps one material point per element and
// integrates it with weight 8 (sum of the eight dvol = V). The element's upper-level force is
// F = A (sigma31, sigma23, sigma33) (A = hx hy), the lower-level force -F; its tangent in the
// tied space is [[S, -S], [-S, S]] with S = (A / hz) D restricted to rows and columns (31, 23,
// 33). Mass per element and direction: lumped rho V / 2 at each level; consistent rho V [[1/3,
// 1/6], [1/6, 1/3]] (the exact integral of the face sums (1 -+ zeta) / 2, which the 2-point
// Gauss rule reproduces). Body load per element: V b / 2 at each level.
//
// Kernels, one pair per Newton iteration of every running column (FX_COLUMN_COMPACT splits the
// first into the reference's predict / compact / correct passes, below):
//  - column_force_kernel: one thread per element slot of the pool: strain from the level
//    displacements, the return map (committed history of the column's current buffer, trial
//    history into the other), F and S written per element;
//  - column_solve_kernel<WAVES, MIXED>: one workgroup of 64 WAVES threads per column of at most
//    64 WAVES elements, one level per thread. The column's own Newmark/Newton state machine:
//    unbalance B = -f - M a - a_g M r (r the rigid unit translation, base included); the test (NormUnbalance ||B||, or NormDispIncr ||dU||
//    of the last solve); on convergence the commit (buffer swap), the recorded top-level values
//    and the next step's predictor; otherwise PCG + Jacobi on K_eff = K_t + c3 M (block
//    tridiagonal, 3 x 3 blocks held in registers; neighbour values through LDS), warm-started
//    from the previous increment, and the update. MIXED: the same solve as FP32 PCG (FP32
//    blocks, FP32 sums) inside FP64 iterative refinement, stopped on the FP64 true residual.
//    Workgroup sums are wave butterflies then a fixed-order sum over waves: a column's bits
//    do not depend on the other columns.
// The four solve sizes run on side streams forked from the pool's stream (they touch disjoint
// columns); the iteration is captured once as a hipGraph (branches for the forks) and replayed.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "fx_internal.hpp"
#include "fx_pimy.hpp"
#include "fx_pimy_point.hpp"

namespace fx {
namespace {

enum : int32_t { kIdle = 0, kRunning = 1, kFinished = 2, kFailed = 3 };
constexpr int kMaxSublevels = 6;  // recovery depth of a failed step (num_sublevels)

struct ColumnDev {
  int32_t status, step, nsteps, iter;
  int32_t cur, nelem, ground_dir, record;
  int32_t newton, max_newton, krylov, pad;
  int64_t e0;    // first element slot; level j (1-based j + 1) has equations 3 (e0 + j) + d
  int64_t rec;   // ground record offset in the record arena
  int64_t hist;  // history offset (9 doubles per step), or -1
  double hx, hy, hz;
  double dnorm;      // ||dU|| of the last solve
  double last_test;  // the last test value
  // recovery of a failed step (implicit_stage.cpp:385-402): the attempt's level (its size is
  // dt / num_substeps^level), the retries it has used, the substep reached at each level, and
  // the counters of fx_newmark's result
  int32_t level, tries;
  int32_t commits, failed_attempts, retries, substepped;
  int32_t sub[kMaxSublevels + 1];
};

struct PoolArgs {
  int64_t cap;                  // element slots
  int32_t max_columns;
  int32_t test, max_iter, cg_max_it;
  double tol, cg_rtol, inner_rtol;
  // per recovery level (size h = dt / num_substeps^level): predictor constants a1..a4, update
  // constants c2 = gamma / (beta h), c3 = 1 / (beta h^2)
  double nk[kMaxSublevels + 1][6];
  int32_t max_retries, num_sublevels, num_substeps, pad;
  ColumnDev* cols;
  const int32_t* elem_col;      // [cap] column slot or -1
  const int32_t* elem_param;    // [cap] parameter set
  const PimyParams* params;
  double* state;                // 2 buffers of [kPimyNState * cap]
  int32_t* touched;             // [cap] the highest surface whose centre an update has written
  double* eout;                 // [9 * cap]: F (3), S upper triangle (6)
  double* u;                    // [3 cap] each
  double* v;
  double* a;
  double* ut;
  double* vt;
  double* at;
  double* du;
  double* du0;                  // du at the attempt's start (the solver's initial guess, restored)
  const double* record;         // record arena
  double* rec_out;              // same arena, written (histories)
  // compacted return map (FX_COLUMN_COMPACT): the plastic points' strains and list
  double* ceps;                 // [3 * cap]: eps33, gamma23, gamma31
  int32_t* plist;               // [cap] element slots
  int32_t* pcount;              // [1]
};

__device__ inline void column_outputs(const PoolArgs& a, int64_t i, double A, double hz, const double (&sig)[6],
                                      const double (&D)[36]);

// ---- the return map, one thread per element slot ------------------------------------------------

template <int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void column_force_kernel(PoolArgs a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.cap) return;
  const int32_t c = a.elem_col[i];
  if (c < 0) return;
  const ColumnDev& col = a.cols[c];
  if (col.status != kRunning) return;
  const int64_t j = i - col.e0;
  const double hz = col.hz, A = col.hx * col.hy;
  double lo[3] = {0.0, 0.0, 0.0}, up[3];
  for (int d = 0; d < 3; ++d) {
    up[d] = a.u[3 * i + d];
    if (j > 0) lo[d] = a.u[3 * (i - 1) + d];
  }
  double eps[6] = {0.0, 0.0, (up[2] - lo[2]) / hz, 0.0, (up[1] - lo[1]) / hz, (up[0] - lo[0]) / hz};
  const size_t NS = static_cast<size_t>(kPimyNState) * static_cast<size_t>(a.cap);
  const double* C = a.state + (col.cur ? NS : 0) + i;
  double* T = a.state + (col.cur ? 0 : NS) + i;
  double sig[6], D[36];
  pimy_update(a.params[a.elem_param[i]], C, T, a.cap, eps, sig, D, a.touched + i);
  column_outputs(a, i, A, hz, sig, D);
}

// ---- the compacted return map (compacted_kernels.hpp:85-202) -------------------------------------
// 1. predict: one thread per element slot; a point that needs no stress correction is finished
//    (F, S written); a plastic one keeps its strain and joins the list (one atomic per wave);
// 2. correct: a fixed grid walks the list, one full update per plastic point (the same
//    function as the fused kernel, so the same bits), F and S written;
// 3. the list count is cleared for the next iteration (a kernel, not a memset: graph replays).

__device__ inline void column_outputs(const PoolArgs& a, int64_t i, double A, double hz, const double (&sig)[6],
                                      const double (&D)[36]) {
  const int64_t P = a.cap;
  a.eout[0 * P + i] = A * sig[5];
  a.eout[1 * P + i] = A * sig[4];
  a.eout[2 * P + i] = A * sig[2];
  const double s = A / hz;
  constexpr int idx[3] = {5, 4, 2};  // x, y, z <- gamma31, gamma23, eps33
  int w = 3;
  for (int r = 0; r < 3; ++r)
    for (int q = r; q < 3; ++q, ++w) a.eout[w * P + i] = s * D[6 * idx[r] + idx[q]];
}

template <int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void column_predict_kernel(PoolArgs a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool plastic = false;
  if (i < a.cap) {
    const int32_t c = a.elem_col[i];
    if (c >= 0 && a.cols[c].status == kRunning) {
      const ColumnDev& col = a.cols[c];
      const int64_t j = i - col.e0;
      const double hz = col.hz, A = col.hx * col.hy;
      double lo[3] = {0.0, 0.0, 0.0}, up[3];
      for (int d = 0; d < 3; ++d) {
        up[d] = a.u[3 * i + d];
        if (j > 0) lo[d] = a.u[3 * (i - 1) + d];
      }
      double eps[6] = {0.0, 0.0, (up[2] - lo[2]) / hz, 0.0, (up[1] - lo[1]) / hz, (up[0] - lo[0]) / hz};
      const size_t NS = static_cast<size_t>(kPimyNState) * static_cast<size_t>(a.cap);
      double sig[6], D[36];
      plastic = pimy_predict(a.params[a.elem_param[i]], a.state + (col.cur ? NS : 0) + i,
                             a.state + (col.cur ? 0 : NS) + i, a.cap, eps, sig, D, a.touched + i);
      if (plastic) {
        a.ceps[i] = eps[2];
        a.ceps[a.cap + i] = eps[4];
        a.ceps[2 * a.cap + i] = eps[5];
      } else {
        column_outputs(a, i, A, hz, sig, D);
      }
    }
  }
  const uint64_t mask = __ballot(plastic);
  if (mask == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(a.pcount, __popcll(mask));
  base = __shfl(base, leader, 64);
  if (plastic) a.plist[base + __popcll(mask & ((uint64_t(1) << lane) - 1))] = static_cast<int32_t>(i);
}

template <int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void column_correct_kernel(PoolArgs a) {
  const int32_t n = *a.pcount;
  const size_t NS = static_cast<size_t>(kPimyNState) * static_cast<size_t>(a.cap);
  for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int64_t i = a.plist[k];
    const ColumnDev& col = a.cols[a.elem_col[i]];
    double eps[6] = {0.0, 0.0, a.ceps[i], 0.0, a.ceps[a.cap + i], a.ceps[2 * a.cap + i]};
    double sig[6], D[36];
    pimy_update(a.params[a.elem_param[i]], a.state + (col.cur ? NS : 0) + i, a.state + (col.cur ? 0 : NS) + i, a.cap,
                eps, sig, D, a.touched + i);
    column_outputs(a, i, col.hx * col.hy, col.hz, sig, D);
  }
}

__global__ void column_clear_count_kernel(int32_t* count) {
  if (threadIdx.x == 0) count[0] = 0;
}

// ---- per-column solve ---------------------------------------------------------------------------

template <typename T>
struct Sym3 {  // xx xy xz yy yz zz
  T m[6];
  __device__ void mul_add(const T (&x)[3], T (&y)[3]) const {
    y[0] += m[0] * x[0] + m[1] * x[1] + m[2] * x[2];
    y[1] += m[1] * x[0] + m[3] * x[1] + m[4] * x[2];
    y[2] += m[2] * x[0] + m[4] * x[1] + m[5] * x[2];
  }
};

// DPP lane moves within rows of 16 lanes (quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror,
// row_mirror): every source lane is in the same row, so all lanes are written.
template <int CTRL>
__device__ inline float dpp_move(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ inline double dpp_move(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ inline float lane_value(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}
__device__ inline double lane_value(double x, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                          __builtin_amdgcn_readlane(__double2loint(x), lane));
}

// Sum over a full wave in a fixed order: pairs, quads, halves and rows of 16 by DPP (each step
// adds two equal-shape partial sums, so every lane of a row ends with the same bits), then the
// four row sums ((r0 + r1) + r2) + r3. All 64 lanes must be active.
template <typename T>
__device__ inline T wave_sum(T x) {
  x += dpp_move<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_move<0x4E>(x);   // quad_perm [2,3,0,1]
  x += dpp_move<0x141>(x);  // row_half_mirror
  x += dpp_move<0x140>(x);  // row_mirror
  T s = lane_value(x, 0);
  s += lane_value(x, 16);
  s += lane_value(x, 32);
  s += lane_value(x, 48);
  return s;
}

// Workgroup sums of two values: butterflies within each wave, then the wave sums in wave order.
template <int WAVES, typename T>
struct Reducer {
  T* red;  // LDS [2 * WAVES]
  // one value: the same sums, in the same order, as the first of sum2
  __device__ void sum1(T& x) const {
    x = wave_sum(x);
    if constexpr (WAVES > 1) {
      const int w = threadIdx.x / 64;
      if ((threadIdx.x & 63) == 0) red[2 * w] = x;
      __syncthreads();
      x = red[0];
      for (int k = 1; k < WAVES; ++k) x += red[2 * k];
    }
  }
  __device__ void sum2(T& x, T& y) const {
    x = wave_sum(x);
    y = wave_sum(y);
    if constexpr (WAVES > 1) {
      const int w = threadIdx.x / 64;
      if ((threadIdx.x & 63) == 0) {
        red[2 * w] = x;
        red[2 * w + 1] = y;
      }
      __syncthreads();
      x = red[0];
      y = red[1];
      for (int k = 1; k < WAVES; ++k) {
        x += red[2 * k];
        y += red[2 * k + 1];
      }
    }
  }
};

// Neighbouring levels' values of a level vector (one level per thread) through LDS. Two
// buffers in turn: exchange k + 2 writes the buffer exchange k read, and every thread has left
// exchange k's reads before it passes exchange k + 1's barrier, so one barrier per exchange.
template <typename T, int NT>
struct Exchange {
  T (*xs)[3][NT];
  int t, N;
  bool live;
  int phase = 0;
  __device__ void operator()(const T (&x)[3], T (&xm)[3], T (&xp)[3]) {
    T (*b)[NT] = xs[phase];
    phase ^= 1;
    for (int d = 0; d < 3; ++d) b[d][t] = x[d];
    __syncthreads();
    for (int d = 0; d < 3; ++d) {
      xm[d] = (live && t > 0) ? b[d][t - 1] : T(0);
      xp[d] = (live && t + 1 < N) ? b[d][t + 1] : T(0);
    }
  }
};

// y = K x on the column's block-tridiagonal operator (blocks of this thread's level).
template <typename T, int NT>
__device__ inline void column_apply(Exchange<T, NT>& ex, const Sym3<T>& Klo, const Sym3<T>& Kd,
                                    const Sym3<T>& Khi, const T (&x)[3], T (&y)[3]) {
  T xm[3], xp[3];
  ex(x, xm, xp);
  for (int d = 0; d < 3; ++d) y[d] = T(0);
  if (!ex.live) return;
  Klo.mul_add(xm, y);
  Kd.mul_add(x, y);
  Khi.mul_add(xp, y);
}

constexpr int kMaxRefinements = 50;

template <int WAVES, bool MIXED>
__global__ __launch_bounds__(64 * WAVES) void column_solve_kernel(PoolArgs a) {
  constexpr int NT = 64 * WAVES;
  __shared__ double xs[2][3][NT];
  __shared__ float xs32[2][3][NT];
  __shared__ double redA[2 * WAVES], redB[2 * WAVES], redC[2 * WAVES];
  __shared__ float redD[2 * WAVES], redE[2 * WAVES];
  const int32_t c = blockIdx.x;
  ColumnDev& col = a.cols[c];
  const int32_t status = col.status, N = col.nelem;
  if (status != kRunning || N > NT || (WAVES > 1 && N <= NT / 2)) return;
  const int t = threadIdx.x;
  const bool live = t < N;
  const int64_t e = col.e0 + t;  // the element below this level, whose top the level is
  const int64_t P = a.cap;
  const double V = col.hx * col.hy * col.hz;
  const Reducer<WAVES, double> RA{redA}, RB{redB}, RC{redC};
  Exchange<double, NT> neighbours{xs, t, N, live};
  // mass of this level: diagonal and couplings (lumped: no coupling); mr: the row of M times the
  // rigid unit translation, the base included (mass_times_ground), for the ground load
  double m = 0.0, mlo = 0.0, mhi = 0.0, mr = 0.0, body[3] = {0, 0, 0};
  const bool above = live && t + 1 < N;
  if (live) {
    const PimyParams& pe = a.params[a.elem_param[e]];
    const bool lumped = pe.mass_type != 0;
    m += lumped ? pe.rho * V / 2.0 : pe.rho * V / 3.0;
    if (!lumped && t > 0) mlo = pe.rho * V / 6.0;
    if (!lumped) mr = pe.rho * V / 6.0;  // the coupling to the level below, the fixed base too
    for (int d = 0; d < 3; ++d) body[d] += V * pe.b[d] / 2.0;
    if (above) {
      const PimyParams& pa = a.params[a.elem_param[e + 1]];
      const bool lumped_a = pa.mass_type != 0;
      m += lumped_a ? pa.rho * V / 2.0 : pa.rho * V / 3.0;
      if (!lumped_a) mhi = pa.rho * V / 6.0;
      for (int d = 0; d < 3; ++d) body[d] += V * pa.b[d] / 2.0;
    }
  }
  // the unbalance B = -f - M a - a_g M r at the current u (forces from column_force_kernel)
  double B[3] = {0, 0, 0}, av[3] = {0, 0, 0};
  if (live)
    for (int d = 0; d < 3; ++d) av[d] = a.a[3 * e + d];
  double am[3], ap[3];
  neighbours(av, am, ap);
  const int step = col.step;
  const int level = col.level;
  const double* K0 = a.nk[level];  // a1 a2 a3 a4 c2 c3 of this attempt's size
  double ag = a.record[col.rec + step];
  if (level > 0) {
    // a substep: the record read as a path series with a leading zero (PathSeries,
    // ground_motion.cpp:17-33), linear within the step from the previous step's value to this
    // step's, at the attempt's end (fraction pos / num_substeps^level of the step)
    int64_t pos = 1, scale = 1;
    for (int l = level; l >= 1; --l) {
      pos += col.sub[l] * scale;
      scale *= a.num_substeps;
    }
    const double prev = step > 0 ? a.record[col.rec + step - 1] : 0.0;
    ag = prev + (ag - prev) * (static_cast<double>(pos) / static_cast<double>(scale));
  }
  // a failed attempt (implicit_stage.cpp:385-402): its start restored with the solver's initial
  // guess; then a retry at the same size, else substeps one level down, else the column fails.
  // Every thread derives the same choice from the state read above; thread 0 stores it.
  auto fail_attempt = [&](int newton_its, int krylov_its) {
    const int tries = col.tries;
    const int outcome = tries < a.max_retries ? 0 : level < a.num_sublevels ? 1 : 2;
    const int ln = outcome == 1 ? level + 1 : level;
    const double* K1 = a.nk[ln];
    if (live)
      for (int d = 0; d < 3; ++d) {
        const int64_t q = 3 * e + d;
        const double u0 = a.ut[q], v0 = a.vt[q], a0 = a.at[q];
        a.u[q] = u0;
        a.v[q] = outcome == 2 ? v0 : v0 * K1[0] + a0 * K1[1];
        a.a[q] = outcome == 2 ? a0 : a0 * K1[3] + v0 * K1[2];
        a.du[q] = a.du0[q];
      }
    __syncthreads();
    if (t == 0) {
      col.krylov += krylov_its;
      col.newton += newton_its;
      col.max_newton = max(col.max_newton, newton_its);
      ++col.failed_attempts;
      col.iter = 0;
      if (outcome == 0) {
        ++col.tries;
        ++col.retries;
      } else if (outcome == 1) {
        ++col.substepped;
        col.level = ln;
        col.sub[ln] = 0;
        col.tries = 0;
      } else {
        col.status = kFailed;
      }
    }
  };
  if (live)
    for (int d = 0; d < 3; ++d) {
      double f = a.eout[d * P + e];
      if (above) f -= a.eout[d * P + e + 1];
      f -= body[d];
      double ma = m * av[d];
      ma += mlo * am[d];
      ma += mhi * ap[d];
      double r = -f;
      r -= ma;
      if (d == col.ground_dir) r -= ag * (m + mr + mhi);
      B[d] = r;
    }
  double bb = 0.0;
  for (int d = 0; d < 3; ++d) bb += B[d] * B[d];
  RA.sum1(bb);
  const double bnorm = sqrt(bb);
  const int iter = col.iter;
  if (iter > 0) {
    const double value = a.test == 0 ? bnorm : col.dnorm;
    bool done = value <= a.tol;
    if (done) {  // commit; the next attempt's predictor; the record once the step is whole
      // a converged substep moves on within its level or, as the last of its level, completes the
      // attempt one level up (the recursion of recover, implicit_stage.cpp:385-402)
      int sub[kMaxSublevels + 1];
      for (int l = 0; l <= kMaxSublevels; ++l) sub[l] = col.sub[l];
      int L = level;
      while (L > 0 && ++sub[L] >= a.num_substeps) --L;
      const bool whole = L == 0;
      const int next = whole ? step + 1 : step;
      const double* K1 = a.nk[L];
      if (whole && live && t == N - 1 && col.hist >= 0) {
        double* h = a.rec_out + col.hist + 9 * static_cast<int64_t>(step);
        for (int d = 0; d < 3; ++d) {
          h[d] = a.u[3 * e + d];
          h[3 + d] = a.v[3 * e + d];
          h[6 + d] = a.a[3 * e + d];
        }
      }
      if (live && !(whole && next >= col.nsteps))
        for (int d = 0; d < 3; ++d) {
          const int64_t q = 3 * e + d;
          const double u0 = a.u[q], v0 = a.v[q], a0 = a.a[q];
          a.ut[q] = u0;
          a.vt[q] = v0;
          a.at[q] = a0;
          a.du0[q] = a.du[q];
          a.v[q] = v0 * K1[0] + a0 * K1[1];
          a.a[q] = a0 * K1[3] + v0 * K1[2];
        }
      __syncthreads();
      if (t == 0) {
        col.last_test = value;
        col.newton += iter;
        col.max_newton = max(col.max_newton, iter);
        ++col.commits;
        col.cur ^= 1;
        col.step = next;
        col.iter = 0;
        col.level = L;
        col.tries = 0;
        for (int l = 1; l <= kMaxSublevels; ++l) col.sub[l] = sub[l];
        if (whole && next >= col.nsteps) col.status = kFinished;
      }
      return;
    }
    if (iter >= a.max_iter) {  // the attempt failed
      if (t == 0) col.last_test = value;
      fail_attempt(iter, 0);
      return;
    }
  }
  // K_eff blocks: diagonal Kd = S_e + S_{e+1} + c3 m I; couplings Klo (to the level below)
  // = -S_e + c3 mlo I, Khi (to the level above) = -S_{e+1} + c3 mhi I
  Sym3<double> Kd{}, Klo{}, Khi{};
  double inv[3] = {0, 0, 0};
  if (live) {
    for (int w = 0; w < 6; ++w) {
      const double se = a.eout[(3 + w) * P + e];
      const double sa = above ? a.eout[(3 + w) * P + e + 1] : 0.0;
      Kd.m[w] = se + sa;
      Klo.m[w] = -se;
      Khi.m[w] = -sa;
    }
    constexpr int diag[3] = {0, 3, 5};
    for (int d = 0; d < 3; ++d) {
      Kd.m[diag[d]] += K0[5] * m;
      Klo.m[diag[d]] += K0[5] * mlo;
      Khi.m[diag[d]] += K0[5] * mhi;
    }
    for (int d = 0; d < 3; ++d) inv[d] = 1.0 / Kd.m[diag[d]];
  }
  auto apply = [&](const double (&x)[3], double (&y)[3]) { column_apply(neighbours, Klo, Kd, Khi, x, y); };
  double x[3] = {0, 0, 0}, r[3], Ap[3];
  if (live)
    for (int d = 0; d < 3; ++d) x[d] = a.du[3 * e + d];
  const double tol = a.cg_rtol * bnorm;
  double rnorm = 0.0;
  int its = 0;
  bool converged = false;
  if constexpr (!MIXED) {
    // PCG (cg.cpp:24-72): warm start from the previous increment, stop on the recursive residual
    double z[3], p[3];
    apply(x, Ap);
    double rr = 0.0, rz = 0.0;
    for (int d = 0; d < 3; ++d) {
      r[d] = B[d] - Ap[d];
      z[d] = inv[d] * r[d];
      p[d] = z[d];
      rr += r[d] * r[d];
      rz += r[d] * z[d];
    }
    RB.sum2(rr, rz);
    rnorm = sqrt(rr);
    converged = rnorm <= tol;
    while (!converged && its < a.cg_max_it) {
      apply(p, Ap);
      double pAp = 0.0;
      for (int d = 0; d < 3; ++d) pAp += p[d] * Ap[d];
      RC.sum1(pAp);
      if (!(pAp > 0.0)) break;  // indefinite
      const double alpha = rz / pAp;
      double rr2 = 0.0, rz2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        x[d] += alpha * p[d];
        r[d] -= alpha * Ap[d];
        z[d] = inv[d] * r[d];
        rr2 += r[d] * r[d];
        rz2 += r[d] * z[d];
      }
      RB.sum2(rr2, rz2);
      ++its;
      rnorm = sqrt(rr2);
      if (rnorm <= tol) {
        converged = true;
        break;
      }
      const double beta = rz2 / rz;
      rz = rz2;
      for (int d = 0; d < 3; ++d) p[d] = z[d] + beta * p[d];
    }
  } else {
    // FP64 iterative refinement around FP32 PCG: r = B - K x in FP64 (the true residual, the
    // stop test), the correction K e = r solved on FP32 blocks with FP32 sums to inner_rtol
    const Reducer<WAVES, float> RD{redD}, RE{redE};
    Exchange<float, NT> neighbours32{xs32, t, N, live};
    Sym3<float> Klo32, Kd32, Khi32;
    for (int w = 0; w < 6; ++w) {
      Klo32.m[w] = static_cast<float>(Klo.m[w]);
      Kd32.m[w] = static_cast<float>(Kd.m[w]);
      Khi32.m[w] = static_cast<float>(Khi.m[w]);
    }
    float inv32[3];
    for (int d = 0; d < 3; ++d) inv32[d] = live ? 1.0f / Kd32.m[d == 0 ? 0 : d == 1 ? 3 : 5] : 0.0f;
    const float inner_rtol = static_cast<float>(a.inner_rtol);
    for (int cycle = 0;; ++cycle) {
      apply(x, Ap);
      double rr = 0.0;
      for (int d = 0; d < 3; ++d) {
        r[d] = B[d] - Ap[d];
        rr += r[d] * r[d];
      }
      RB.sum1(rr);
      rnorm = sqrt(rr);
      converged = rnorm <= tol;
      if (converged || its >= a.cg_max_it || cycle >= kMaxRefinements) break;
      // the correction solves K e = r / ||r|| (unit right-hand side: no FP32 underflow of the
      // inner dots as the Newton residual shrinks) and is scaled back in FP64
      const double scale = rnorm > 0.0 ? 1.0 / rnorm : 0.0;
      float rf[3], zf[3], pf[3], Apf[3], ef[3] = {0.f, 0.f, 0.f};
      float rrf = 0.f, rzf = 0.f;
      for (int d = 0; d < 3; ++d) {
        rf[d] = static_cast<float>(r[d] * scale);
        zf[d] = inv32[d] * rf[d];
        pf[d] = zf[d];
        rrf += rf[d] * rf[d];
        rzf += rf[d] * zf[d];
      }
      RD.sum2(rrf, rzf);
      const float itol = inner_rtol * sqrtf(rrf);
      while (its < a.cg_max_it) {
        column_apply(neighbours32, Klo32, Kd32, Khi32, pf, Apf);
        float pAp = 0.f;
        for (int d = 0; d < 3; ++d) pAp += pf[d] * Apf[d];
        RE.sum1(pAp);
        if (!(pAp > 0.f)) break;
        const float alpha = rzf / pAp;
        float rr2 = 0.f, rz2 = 0.f;
        for (int d = 0; d < 3; ++d) {
          ef[d] += alpha * pf[d];
          rf[d] -= alpha * Apf[d];
          zf[d] = inv32[d] * rf[d];
          rr2 += rf[d] * rf[d];
          rz2 += rf[d] * zf[d];
        }
        RD.sum2(rr2, rz2);
        ++its;
        if (sqrtf(rr2) <= itol) break;
        const float beta = rz2 / rzf;
        rzf = rz2;
        for (int d = 0; d < 3; ++d) pf[d] = zf[d] + beta * pf[d];
      }
      for (int d = 0; d < 3; ++d) x[d] += rnorm * static_cast<double>(ef[d]);
    }
  }
  const bool accept = converged || rnorm <= 1e-6 * bnorm;
  if (!accept) {  // the solve failed: the attempt fails (implicit_stage.cpp:226-238)
    fail_attempt(iter + 1, its);
    return;
  }
  double dd = 0.0;
  if (live)
    for (int d = 0; d < 3; ++d) {
      const int64_t q = 3 * e + d;
      a.du[q] = x[d];
      a.u[q] += x[d];
      a.v[q] += x[d] * K0[4];
      a.a[q] += x[d] * K0[5];
      dd += x[d] * x[d];
    }
  RA.sum1(dd);
  __syncthreads();
  if (t == 0) {
    col.dnorm = sqrt(dd);
    col.krylov += its;
    col.iter = iter + 1;
  }
}

__global__ void column_init_kernel(int64_t cap, int64_t e0, int32_t n, const PimyParams* __restrict__ params,
                                   const int32_t* __restrict__ elem_param, double* state, int32_t* touched) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t i = e0 + k;
  touched[i] = 0;  // every centre at its initial value in both buffers
  const PimyParams& p = params[elem_param[i]];
  const size_t NS = static_cast<size_t>(kPimyNState) * static_cast<size_t>(cap);
  for (double* S : {state, state + NS}) {
    for (int f = 0; f < kPimyNState; ++f) S[f * cap + i] = 0.0;
    S[kPG * cap + i] = p.G;
    S[kPK * cap + i] = p.K;
    for (int s = 1; s <= kPimyNS; ++s) {
      S[(kPSizes + s - 1) * cap + i] = p.size[s];
      S[(kPModuli + s - 1) * cap + i] = p.modulus[s];
    }
  }
}

__global__ void column_count_kernel(const ColumnDev* cols, int32_t n, int32_t* out, int8_t* status) {
  __shared__ int32_t part[256];
  int32_t k = 0;
  for (int32_t c = threadIdx.x; c < n; c += blockDim.x) {
    const int32_t s = cols[c].status;
    status[c] = static_cast<int8_t>(s);
    k += s == kRunning ? 1 : 0;
  }
  part[threadIdx.x] = k;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

// First-fit allocator of [offset, offset + length) ranges with coalescing.
class Ranges {
 public:
  void reset(int64_t cap) {
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  int64_t take(int64_t n) {
    if (n <= 0) return 0;
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        const int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return off;
      }
    return -1;
  }
  void give(int64_t off, int64_t n) {
    if (n <= 0) return;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<int64_t, int64_t> free_;
};

struct HostColumn {
  bool used = false;
  int64_t e0 = 0, rec = 0, hist = -1, rec_len = 0;
  int32_t nelem = 0, m0 = 0, nmat = 0, nsteps = 0;
};

}  // namespace
}  // namespace fx

struct fx_column_pool {
  fx_ctx* ctx = nullptr;
  fx_column_pool_desc d{};
  fx::PoolArgs args{};
  std::vector<void*> allocs;
  uint64_t bytes = 0;
  fx::Ranges elems, mats, recs;
  std::vector<fx::HostColumn> host;
  int32_t* count = nullptr;       // device: running columns, then the status byte of each column
  int32_t* count_host = nullptr;  // pinned copy of the same
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int wpe = 3;                    // the return-map kernels' register budget (waves per SIMD)
  bool tuned = true;              // false until FX_COLUMN_TUNE has chosen wpe
  bool return_map_only = false;   // launch_iteration stops after the return maps (the tuner's samples)
  hipStream_t side[4] = {nullptr, nullptr, nullptr, nullptr};  // default priority
  hipEvent_t fork = nullptr, join[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace fx {
namespace {

int fail(fx_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  else set_thread_error(msg);
  return code;
}

template <typename T>
int pool_alloc(fx_column_pool* p, T*& dst, size_t n) {
  dst = nullptr;
  const size_t sz = std::max<size_t>(n, 1) * sizeof(T);
  fx_ctx* ctx = p->ctx;
  if (ctx->used + sz > ctx->budget) return fail(ctx, FX_ENOMEM, "column pool: HBM budget exceeded");
  FX_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&dst), sz));
  p->allocs.push_back(dst);
  ctx->used += sz;
  p->bytes += sz;
  FX_HIP(ctx, hipMemset(dst, 0, sz));
  return FX_OK;
}

// One Newton iteration of every running column: the return maps, then the solves in four
// workgroup sizes (1, 2, 4, 8 waves: columns of up to 64, 128, 256, 512 elements).
int launch_iteration(fx_column_pool* p, hipStream_t st) {
  const PoolArgs& a = p->args;
  const unsigned slots = static_cast<unsigned>((a.cap + 255) / 256);
  // register budget of the return-map kernels: waves per SIMD the compiler must allow
  const int wpe = p->wpe;
  auto return_map = [&](auto w) {
    constexpr int W = decltype(w)::value;
    if (p->d.flags & FX_COLUMN_COMPACT) {
      hipLaunchKernelGGL(column_predict_kernel<W>, dim3(slots), dim3(256), 0, st, a);
      hipLaunchKernelGGL(column_correct_kernel<W>, dim3(std::min(slots, 2048u)), dim3(256), 0, st, a);
      hipLaunchKernelGGL(column_clear_count_kernel, dim3(1), dim3(64), 0, st, a.pcount);
    } else {
      hipLaunchKernelGGL(column_force_kernel<W>, dim3(slots), dim3(256), 0, st, a);
    }
  };
  if (wpe >= 4) return_map(std::integral_constant<int, 4>{});
  else if (wpe == 3) return_map(std::integral_constant<int, 3>{});
  else return_map(std::integral_constant<int, 1>{});
  if (p->return_map_only) return FX_OK;
  // the four solve sizes touch disjoint columns: forked onto side streams so they overlap
  // (graph branches when captured), joined back before the next iteration
  const unsigned nc = static_cast<unsigned>(a.max_columns);
  const bool mixed = (p->d.flags & FX_COLUMN_MIXED) != 0;
  const bool fork = !(p->d.flags & FX_COLUMN_ONE_STREAM);
  if (fork) FX_HIP(p->ctx, hipEventRecord(p->fork, st));
  for (int k = 0; k < 4; ++k) {
    hipStream_t s = fork ? p->side[k] : st;
    if (fork) FX_HIP(p->ctx, hipStreamWaitEvent(s, p->fork, 0));
    const dim3 g(nc), b(64u << k);
    switch (k + (mixed ? 4 : 0)) {
      case 0: hipLaunchKernelGGL((column_solve_kernel<1, false>), g, b, 0, s, a); break;
      case 1: hipLaunchKernelGGL((column_solve_kernel<2, false>), g, b, 0, s, a); break;
      case 2: hipLaunchKernelGGL((column_solve_kernel<4, false>), g, b, 0, s, a); break;
      case 3: hipLaunchKernelGGL((column_solve_kernel<8, false>), g, b, 0, s, a); break;
      case 4: hipLaunchKernelGGL((column_solve_kernel<1, true>), g, b, 0, s, a); break;
      case 5: hipLaunchKernelGGL((column_solve_kernel<2, true>), g, b, 0, s, a); break;
      case 6: hipLaunchKernelGGL((column_solve_kernel<4, true>), g, b, 0, s, a); break;
      default: hipLaunchKernelGGL((column_solve_kernel<8, true>), g, b, 0, s, a); break;
    }
    if (fork) {
      FX_HIP(p->ctx, hipEventRecord(p->join[k], s));
      FX_HIP(p->ctx, hipStreamWaitEvent(st, p->join[k], 0));
    }
  }
  FX_HIP(p->ctx, hipGetLastError());
  return FX_OK;
}

}  // namespace
}  // namespace fx

namespace fx {
namespace {

// The iteration captured once as a hipGraph (the forks as branches), unless FX_COLUMN_NO_GRAPH.
int capture_iteration(fx_column_pool* p) {
  if (p->d.flags & FX_COLUMN_NO_GRAPH) return FX_OK;
  fx_ctx* ctx = p->ctx;
  hipStream_t st = ctx->stream;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(ctx, FX_EDEVICE, "hipStreamBeginCapture failed");
  const int lrc = launch_iteration(p, st);
  const hipError_t ec = hipStreamEndCapture(st, &p->graph);
  if (lrc || ec != hipSuccess || hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
    return fail(ctx, FX_EDEVICE, "column pool: graph capture failed");
  return FX_OK;
}

// FX_COLUMN_TUNE: the return maps alone (a pure function of the columns' displacements and
// committed histories: a repeat writes the same trial histories, element results and copy
// bounds, so the iteration that follows is unchanged) timed for each register budget on the
// running columns, the fastest kept, per device and pool in the autotune cache (the reference's
// choose_force_settings, force_settings.cpp:179-205).
int tune_return_map(fx_column_pool* p) {
  fx_ctx* ctx = p->ctx;
  hipStream_t st = ctx->stream;
  const int64_t cap = p->args.cap;
  const int log2cap = cap > 0 ? static_cast<int>(std::floor(std::log2(static_cast<double>(cap)))) : 0;
  const std::string key = tune_machine_key(ctx->device) + "|column_pool,slots=2^" + std::to_string(log2cap) +
                          ",compact=" + ((p->d.flags & FX_COLUMN_COMPACT) ? "1" : "0");
  int depth = 0, v = -1;
  if (tune_cache_get(key, depth, v) && (v == 1 || v == 3 || v == 4)) {
    p->wpe = v;
    return FX_OK;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  FX_HIP(ctx, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(ctx, FX_EDEVICE, "hipEventCreate failed");
  }
  const int cand[3] = {3, 4, 1};
  float best_ms = 1e30f;
  int best = 3, rc = FX_OK;
  p->return_map_only = true;
  for (int round = 0; round < 2 && rc == FX_OK; ++round)
    for (int w : cand) {
      p->wpe = w;
      if ((rc = launch_iteration(p, st))) break;  // warm-up
      if (hipEventRecord(e0, st) != hipSuccess) rc = FX_EDEVICE;
      for (int r = 0; r < 3 && rc == FX_OK; ++r) rc = launch_iteration(p, st);
      float ms = 0.0f;
      if (rc || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = rc ? rc : fail(ctx, FX_EDEVICE, "column pool: autotune timing failed");
        break;
      }
      if (ms < best_ms) {
        best_ms = ms;
        best = w;
      }
    }
  p->return_map_only = false;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  p->wpe = best;
  tune_cache_put(key, 8, best);
  return FX_OK;
}

}  // namespace
}  // namespace fx

extern "C" {

void fx_column_pool_destroy(fx_column_pool* p) {
  if (!p) return;
  fx_ctx* ctx = p->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (int k = 0; k < 4; ++k) {
    if (p->side[k]) (void)hipStreamDestroy(p->side[k]);
    if (p->join[k]) (void)hipEventDestroy(p->join[k]);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  for (void* q : p->allocs) (void)hipFree(q);
  if (p->count_host) (void)hipHostFree(p->count_host);
  ctx->used -= std::min<uint64_t>(ctx->used, p->bytes);
  delete p;
}

int fx_column_pool_create(fx_ctx* ctx, const fx_column_pool_desc* d, fx_column_pool** out) {
  using namespace fx;
  if (!d || !out) return fail(ctx, FX_EINVAL, "fx_column_pool_create: null argument");
  *out = nullptr;
  // the descriptor is checked before the context, so a bad one is refused without a device
  if (d->max_elements <= 0 || d->max_columns <= 0 || d->max_materials <= 0 || d->max_record <= 0 ||
      !(d->dt > 0.0) || !(d->beta > 0.0) || d->max_iter <= 0 || d->cg_max_it <= 0 || (d->test != 0 && d->test != 1))
    return fail(ctx, FX_EINVAL, "fx_column_pool_create: bad descriptor");
  if (d->max_retries < 0 || d->num_sublevels < 0 || d->num_sublevels > kMaxSublevels ||
      (d->num_sublevels > 0 && d->num_substeps < 1))
    return fail(ctx, FX_EINVAL, "fx_column_pool_create: max_retries >= 0, num_sublevels in [0, 6], num_substeps >= 1");
  {
    // a substep's position within its step is counted in units of num_substeps^level (int64 on
    // the device); the deepest level's count must stay within INT32_MAX
    int64_t units = 1;
    for (int l = 0; l < d->num_sublevels; ++l) {
      units *= d->num_substeps;
      if (units > INT32_MAX)
        return fail(ctx, FX_EINVAL, "fx_column_pool_create: num_substeps^num_sublevels exceeds 2^31 - 1");
    }
  }
  if (!ctx) return fail(ctx, FX_EINVAL, "fx_column_pool_create: null context");
  FX_HIP(ctx, hipSetDevice(ctx->device));
  fx_column_pool* p = new (std::nothrow) fx_column_pool;
  if (!p) return fail(ctx, FX_ENOMEM, "host allocation failed");
  p->ctx = ctx;
  p->d = *d;
  PoolArgs& a = p->args;
  a.cap = d->max_elements;
  a.max_columns = d->max_columns;
  a.test = d->test;
  a.max_iter = d->max_iter;
  a.cg_max_it = d->cg_max_it;
  a.tol = d->tol;
  a.cg_rtol = d->cg_rtol;
  a.inner_rtol = d->inner_rtol > 0.0 ? d->inner_rtol : 1e-5;
  const double gamma = d->gamma, beta = d->beta, h = d->dt;
  // the constants of each recovery level's size, h / num_substeps per level down (as fx_newmark
  // divides it)
  a.max_retries = d->max_retries;
  a.num_sublevels = d->num_sublevels;
  a.num_substeps = d->num_sublevels > 0 ? d->num_substeps : 1;
  double hl = h;
  for (int l = 0; l <= kMaxSublevels; ++l) {
    double* K = a.nk[l];
    K[0] = 1.0 - gamma / beta;
    K[1] = hl * (1.0 - 0.5 * gamma / beta);
    K[2] = -1.0 / (beta * hl);
    K[3] = 1.0 - 0.5 / beta;
    K[4] = gamma / (beta * hl);
    K[5] = 1.0 / (beta * hl * hl);
    hl = hl / a.num_substeps;
  }
  const size_t cap = static_cast<size_t>(a.cap);
  int rc = FX_OK;
  ColumnDev* cols = nullptr;
  int32_t *elem_col = nullptr, *elem_param = nullptr;
  PimyParams* params = nullptr;
  double* rec = nullptr;
  if ((rc = pool_alloc(p, cols, d->max_columns)) || (rc = pool_alloc(p, elem_col, cap)) ||
      (rc = pool_alloc(p, elem_param, cap)) || (rc = pool_alloc(p, params, d->max_materials)) ||
      (rc = pool_alloc(p, a.state, 2 * kPimyNState * cap)) || (rc = pool_alloc(p, a.touched, cap)) ||
      (rc = pool_alloc(p, a.eout, 9 * cap)) ||
      (rc = pool_alloc(p, a.u, 3 * cap)) || (rc = pool_alloc(p, a.v, 3 * cap)) || (rc = pool_alloc(p, a.a, 3 * cap)) ||
      (rc = pool_alloc(p, a.ut, 3 * cap)) || (rc = pool_alloc(p, a.vt, 3 * cap)) ||
      (rc = pool_alloc(p, a.at, 3 * cap)) || (rc = pool_alloc(p, a.du, 3 * cap)) || (rc = pool_alloc(p, a.du0, 3 * cap)) ||
      (rc = pool_alloc(p, rec, static_cast<size_t>(d->max_record))) ||
      ((d->flags & FX_COLUMN_COMPACT) &&
       ((rc = pool_alloc(p, a.ceps, 3 * cap)) || (rc = pool_alloc(p, a.plist, cap)) ||
        (rc = pool_alloc(p, a.pcount, 1)))) ||
      (rc = pool_alloc(p, p->count, 1 + (static_cast<size_t>(d->max_columns) + 3) / 4))) {
    fx_column_pool_destroy(p);
    return rc;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&p->count_host),
                    sizeof(int32_t) * (1 + (static_cast<size_t>(d->max_columns) + 3) / 4)) != hipSuccess) {
    fx_column_pool_destroy(p);
    return fail(ctx, FX_EDEVICE, "hipHostMalloc failed");
  }
  a.cols = cols;
  a.elem_col = elem_col;
  a.elem_param = elem_param;
  a.params = params;
  a.record = rec;
  a.rec_out = rec;
  // no element slot belongs to a column yet
  std::vector<int32_t> none(cap, -1);
  if (hipMemcpy(elem_col, none.data(), cap * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
    fx_column_pool_destroy(p);
    return fail(ctx, FX_EDEVICE, "hipMemcpy failed");
  }
  bool streams_ok = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; k < 4 && streams_ok; ++k)
    streams_ok = hipStreamCreateWithFlags(&p->side[k], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&p->join[k], hipEventDisableTiming) == hipSuccess;
  if (!streams_ok) {
    fx_column_pool_destroy(p);
    return fail(ctx, FX_EDEVICE, "column pool: stream or event creation failed");
  }
  p->elems.reset(a.cap);
  p->mats.reset(d->max_materials);
  p->recs.reset(d->max_record);
  p->host.assign(d->max_columns, HostColumn{});
  const uint32_t bits = (d->flags >> 8) & 0xfu;
  p->wpe = bits ? static_cast<int>(bits) : 3;  // 3: measured best on C4
  p->tuned = bits != 0 || !(d->flags & FX_COLUMN_TUNE);
  if (p->tuned) {
    if (int rc = capture_iteration(p)) {
      fx_column_pool_destroy(p);
      return rc;
    }
  }
  *out = p;
  return FX_OK;
}

int fx_column_add(fx_column_pool* p, const fx_column_desc* d, int32_t* id) {
  using namespace fx;
  if (!p || !d || !id) return fail(p ? p->ctx : nullptr, FX_EINVAL, "fx_column_add: null argument");
  fx_ctx* ctx = p->ctx;
  *id = -1;
  const int32_t N = d->nelem;
  if (N < 1 || N > FX_COLUMN_MAX_ELEMS || d->nmat < 1 || !d->mats || !d->mat_of || d->nsteps < 1 ||
      !d->ground_accel || d->ground_dir < 0 || d->ground_dir > 2 || !(d->hx > 0) || !(d->hy > 0) || !(d->hz > 0) ||
      (d->stage != 0 && d->stage != 1))
    return fail(ctx, FX_EINVAL, "fx_column_add: bad descriptor");
  for (int32_t k = 0; k < N; ++k)
    if (d->mat_of[k] < 0 || d->mat_of[k] >= d->nmat) return fail(ctx, FX_EINVAL, "fx_column_add: mat_of out of range");
  std::vector<PimyParams> pp(d->nmat);
  for (int32_t i = 0; i < d->nmat; ++i) {
    std::string err;
    if (int rc = pimy_setup(d->mats[i], pp[i], err)) return fail(ctx, rc, err + " (set " + std::to_string(i) + ")");
    pp[i].load_stage = d->stage;
  }
  int32_t c = -1;
  for (int32_t k = 0; k < p->d.max_columns; ++k)
    if (!p->host[k].used) {
      c = k;
      break;
    }
  const int64_t rec_len = d->nsteps + (d->record ? 9 * static_cast<int64_t>(d->nsteps) : 0);
  if (c < 0) return fail(ctx, FX_EFULL, "fx_column_add: no free column slot");
  const int64_t e0 = p->elems.take(N);
  if (e0 < 0) return fail(ctx, FX_EFULL, "fx_column_add: no room for the elements");
  const int64_t m0 = p->mats.take(d->nmat);
  if (m0 < 0) {
    p->elems.give(e0, N);
    return fail(ctx, FX_EFULL, "fx_column_add: no room for the parameter sets");
  }
  const int64_t r0 = p->recs.take(rec_len);
  if (r0 < 0) {
    p->elems.give(e0, N);
    p->mats.give(m0, d->nmat);
    return fail(ctx, FX_EFULL, "fx_column_add: no room for the record");
  }
  HostColumn& hc = p->host[c];
  hc.used = true;
  hc.e0 = e0;
  hc.nelem = N;
  hc.m0 = static_cast<int32_t>(m0);
  hc.nmat = d->nmat;
  hc.rec = r0;
  hc.rec_len = rec_len;
  hc.hist = d->record ? r0 + d->nsteps : -1;
  hc.nsteps = d->nsteps;
  PoolArgs& a = p->args;
  hipStream_t st = ctx->stream;
  std::vector<int32_t> ecol(N, c), epar(N);
  for (int32_t k = 0; k < N; ++k) epar[k] = static_cast<int32_t>(m0) + d->mat_of[k];
  ColumnDev cd{};
  cd.status = kRunning;
  cd.nsteps = d->nsteps;
  cd.nelem = N;
  cd.ground_dir = d->ground_dir;
  cd.record = d->record;
  cd.e0 = e0;
  cd.rec = r0;
  cd.hist = hc.hist;
  cd.hx = d->hx;
  cd.hy = d->hy;
  cd.hz = d->hz;
  const size_t n3 = 3 * static_cast<size_t>(N) * sizeof(double);
  FX_HIP(ctx, hipMemcpyAsync(const_cast<PimyParams*>(a.params) + m0, pp.data(), sizeof(PimyParams) * d->nmat,
                             hipMemcpyHostToDevice, st));
  FX_HIP(ctx, hipMemcpyAsync(const_cast<int32_t*>(a.elem_param) + e0, epar.data(), sizeof(int32_t) * N,
                             hipMemcpyHostToDevice, st));
  FX_HIP(ctx, hipMemcpyAsync(const_cast<double*>(a.record) + r0, d->ground_accel, sizeof(double) * d->nsteps,
                             hipMemcpyHostToDevice, st));
  for (double* q : {a.u, a.v, a.a, a.ut, a.vt, a.at, a.du, a.du0}) FX_HIP(ctx, hipMemsetAsync(q + 3 * e0, 0, n3, st));
  hipLaunchKernelGGL(column_init_kernel, dim3(static_cast<unsigned>((N + 255) / 256)), dim3(256), 0, st, a.cap, e0, N,
                     a.params, a.elem_param, a.state, a.touched);
  FX_HIP(ctx, hipGetLastError());
  // the column joins the running set last: slot ownership, then its status
  FX_HIP(ctx, hipMemcpyAsync(const_cast<int32_t*>(a.elem_col) + e0, ecol.data(), sizeof(int32_t) * N,
                             hipMemcpyHostToDevice, st));
  FX_HIP(ctx, hipMemcpyAsync(a.cols + c, &cd, sizeof cd, hipMemcpyHostToDevice, st));
  FX_HIP(ctx, hipStreamSynchronize(st));
  *id = c;
  return FX_OK;
}

int fx_column_run(fx_column_pool* p, int32_t iterations, int32_t* running) {
  using namespace fx;
  if (!p || iterations < 0) return fail(p ? p->ctx : nullptr, FX_EINVAL, "fx_column_run: bad argument");
  fx_ctx* ctx = p->ctx;
  hipStream_t st = ctx->stream;
  if (!p->tuned && iterations > 0) {
    bool any = false;
    for (const auto& h : p->host) any |= h.used;
    if (any) {  // the first running columns: choose the register budget, then capture
      if (int rc = tune_return_map(p)) return rc;
      p->tuned = true;
      if (int rc = capture_iteration(p)) return rc;
    }
  }
  for (int32_t k = 0; k < iterations; ++k) {
    if (p->exec) FX_HIP(ctx, hipGraphLaunch(p->exec, st));
    else if (int rc = launch_iteration(p, st)) return rc;
  }
  hipLaunchKernelGGL(column_count_kernel, dim3(1), dim3(256), 0, st, p->args.cols, p->args.max_columns, p->count,
                     reinterpret_cast<int8_t*>(p->count + 1));
  FX_HIP(ctx, hipGetLastError());
  FX_HIP(ctx, hipMemcpyAsync(p->count_host, p->count, sizeof(int32_t) + static_cast<size_t>(p->args.max_columns),
                             hipMemcpyDeviceToHost, st));
  FX_HIP(ctx, hipStreamSynchronize(st));
  if (running) *running = *p->count_host;
  return FX_OK;
}

int fx_column_finished(fx_column_pool* p, int32_t* ids, int32_t cap, int32_t* count) {
  using namespace fx;
  if (!p || !count || (cap > 0 && !ids)) return fail(p ? p->ctx : nullptr, FX_EINVAL, "fx_column_finished: bad argument");
  const int8_t* s = reinterpret_cast<const int8_t*>(p->count_host + 1);
  int32_t k = 0;
  for (int32_t c = 0; c < p->d.max_columns; ++c)
    if (p->host[c].used && (s[c] == kFinished || s[c] == kFailed)) {
      if (k < cap) ids[k] = c;
      ++k;
    }
  *count = k;
  return FX_OK;
}

int fx_column_status(fx_column_pool* p, int32_t id, fx_column_result* r) {
  using namespace fx;
  if (!p || !r || id < 0 || id >= p->d.max_columns || !p->host[id].used)
    return fail(p ? p->ctx : nullptr, FX_EINVAL, "fx_column_status: bad column");
  fx_ctx* ctx = p->ctx;
  ColumnDev cd{};
  FX_HIP(ctx, hipMemcpyAsync(&cd, p->args.cols + id, sizeof cd, hipMemcpyDeviceToHost, ctx->stream));
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  r->status = cd.status == kRunning ? 0 : cd.status == kFinished ? 1 : 2;
  r->steps = cd.step;
  r->newton_iterations = cd.newton;
  r->max_newton_iterations = cd.max_newton;
  r->krylov_iterations = cd.krylov;
  r->last_test = cd.last_test;
  r->commits = cd.commits;
  r->failed_attempts = cd.failed_attempts;
  r->retries = cd.retries;
  r->substepped = cd.substepped;
  return FX_OK;
}

int fx_column_fetch(fx_column_pool* p, int32_t id, double* u, double* v, double* a, double* history) {
  using namespace fx;
  if (!p || id < 0 || id >= p->d.max_columns || !p->host[id].used)
    return fail(p ? p->ctx : nullptr, FX_EINVAL, "fx_column_fetch: bad column");
  const HostColumn& hc = p->host[id];
  if (history && hc.hist < 0) return fail(p->ctx, FX_EINVAL, "fx_column_fetch: the column keeps no history");
  fx_ctx* ctx = p->ctx;
  hipStream_t st = ctx->stream;
  const size_t n3 = 3 * static_cast<size_t>(hc.nelem) * sizeof(double);
  const double* src[3] = {p->args.u, p->args.v, p->args.a};
  double* dst[3] = {u, v, a};
  for (int k = 0; k < 3; ++k)
    if (dst[k]) FX_HIP(ctx, hipMemcpyAsync(dst[k], src[k] + 3 * hc.e0, n3, hipMemcpyDeviceToHost, st));
  if (history)
    FX_HIP(ctx, hipMemcpyAsync(history, p->args.record + hc.hist, sizeof(double) * 9 * hc.nsteps,
                               hipMemcpyDeviceToHost, st));
  FX_HIP(ctx, hipStreamSynchronize(st));
  return FX_OK;
}

int fx_column_state(fx_column_pool* p, int32_t id, double* state) {
  using namespace fx;
  if (!p || !state || id < 0 || id >= p->d.max_columns || !p->host[id].used)
    return fail(p ? p->ctx : nullptr, FX_EINVAL, "fx_column_state: bad column");
  const HostColumn& hc = p->host[id];
  fx_ctx* ctx = p->ctx;
  hipStream_t st = ctx->stream;
  ColumnDev cd{};
  FX_HIP(ctx, hipMemcpyAsync(&cd, p->args.cols + id, sizeof cd, hipMemcpyDeviceToHost, st));
  FX_HIP(ctx, hipStreamSynchronize(st));
  const size_t cap = static_cast<size_t>(p->args.cap);
  const double* S = p->args.state + (cd.cur ? static_cast<size_t>(kPimyNState) * cap : 0) + hc.e0;
  FX_HIP(ctx, hipMemcpy2DAsync(state, sizeof(double) * hc.nelem, S, sizeof(double) * cap, sizeof(double) * hc.nelem,
                               kPimyNState, hipMemcpyDeviceToHost, st));
  FX_HIP(ctx, hipStreamSynchronize(st));
  return FX_OK;
}

int fx_column_remove(fx_column_pool* p, int32_t id) {
  using namespace fx;
  if (!p || id < 0 || id >= p->d.max_columns || !p->host[id].used)
    return fail(p ? p->ctx : nullptr, FX_EINVAL, "fx_column_remove: bad column");
  fx_ctx* ctx = p->ctx;
  hipStream_t st = ctx->stream;
  HostColumn& hc = p->host[id];
  ColumnDev cd{};
  cd.status = kIdle;
  std::vector<int32_t> none(hc.nelem, -1);
  FX_HIP(ctx, hipMemcpyAsync(p->args.cols + id, &cd, sizeof cd, hipMemcpyHostToDevice, st));
  FX_HIP(ctx, hipMemcpyAsync(const_cast<int32_t*>(p->args.elem_col) + hc.e0, none.data(), sizeof(int32_t) * hc.nelem,
                             hipMemcpyHostToDevice, st));
  FX_HIP(ctx, hipStreamSynchronize(st));
  p->elems.give(hc.e0, hc.nelem);
  p->mats.give(hc.m0, hc.nmat);
  p->recs.give(hc.rec, hc.rec_len);
  hc = HostColumn{};
  return FX_OK;
}

}  // extern "C"
