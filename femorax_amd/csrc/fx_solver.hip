// Device-resident Krylov solvers for the hex8 operator.
//
// ConjugateGradient::solve with Jacobi (src/solvers/cg.cpp:24-72, jacobi.cpp:10-26) with its
// scalars kept on the device: each kernel reads a `done` flag and becomes a no-op once the
// stop rule ||r|| <= max(rtol ||b||, atol) (cg.cpp:34, 60) or p.Ap <= 0 (cg.cpp:49-53) has
// fired, so the host only polls every few iterations and the iterate is exactly the one the
// reference stops at. Dot products follow VectorOps' order (vector_ops.cpp:19-64): chunks of
// 2048 entries, 64 lanes, lane l summing entries l, l+64, ...; lane sums in lane order; chunk
// sums reduced level by level. They accumulate in FP64 whatever the vector type.
//
// The mixed-precision solve wraps an FP32 CG (FP32 operator and vectors, FP64 dots) in FP64
// iterative refinement: r = b - A x with the FP64 operator, d = A32^-1 r approximately,
// x += d in FP64, until the FP64 residual meets the same stop rule.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "fx_internal.hpp"
#include "fx_kernels.hpp"

namespace fx {

constexpr int kChunk = 2048;
constexpr int kLanes = 64;

struct CgState {
  double rz, pap, alpha, beta, rr, tol, rhs_norm, init_res, final_res, scale;
  int32_t it, done, converged, indefinite, max_it, pad;
};

__device__ __forceinline__ double as_double(double v) { return v; }
__device__ __forceinline__ double as_double(float v) { return static_cast<double>(v); }

// Partial sums per chunk of x.y (and, when y2 != null, of x2.y2) in VectorOps' lane order.
template <typename T>
__global__ __launch_bounds__(kLanes) void chunk_dot_kernel(int64_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                           const T* __restrict__ x2, const T* __restrict__ y2,
                                                           double* __restrict__ partial, const int32_t* done) {
  if (done && *done) return;
  __shared__ double lane_sum[2][kLanes];
  const int64_t c = blockIdx.x;
  const int64_t begin = c * kChunk;
  const int64_t len = (n - begin < kChunk) ? n - begin : kChunk;
  const int l = threadIdx.x;
  double s = 0.0, s2 = 0.0;
  for (int64_t k = l; k < len; k += kLanes) {
    s += as_double(x[begin + k]) * as_double(y[begin + k]);
    if (y2) s2 += as_double(x2[begin + k]) * as_double(y2[begin + k]);
  }
  lane_sum[0][l] = s;
  lane_sum[1][l] = s2;
  __syncthreads();
  if (l == 0) {
    double t = 0.0, t2 = 0.0;
    for (int j = 0; j < kLanes; ++j) t += lane_sum[0][j];
    for (int j = 0; j < kLanes; ++j) t2 += lane_sum[1][j];
    if (y2) {
      partial[2 * c] = t;
      partial[2 * c + 1] = t2;
    } else {
      partial[c] = t;
    }
  }
}

enum ReduceOp : int { kStore = 0, kRhsNorm, kInitRes, kRz0, kPap, kIter };

// One chunk of a level, summed by one wave in VectorOps' order: lane l adds entries l, l + 64,
// ... in turn, then lane 0 adds the 64 lane sums in lane order (through this wave's lane_sum
// row in LDS: a wave's LDS accesses complete in program order). A full chunk issues each lane's
// 32 loads together; a partial one (the last of a level, or a small vector) walks its entries.
// Returns the total in lane 0.
__device__ __forceinline__ double wave_chunk_sum(const double* src, int64_t stride, int64_t begin, int64_t len, int l,
                                                 double* lane_sum) {
  double s = 0.0;
  if (len == kChunk) {
    double v[kChunk / kLanes];
#pragma unroll
    for (int j = 0; j < kChunk / kLanes; ++j) v[j] = src[(begin + l + static_cast<int64_t>(j) * kLanes) * stride];
#pragma unroll
    for (int j = 0; j < kChunk / kLanes; ++j) s += v[j];
  } else {
    for (int64_t k = l; k < len; k += kLanes) s += src[(begin + k) * stride];
  }
  lane_sum[l] = s;
  __builtin_amdgcn_wave_barrier();
  double t = 0.0;
  if (l == 0)
    for (int j = 0; j < kLanes; ++j) t += lane_sum[j];
  __builtin_amdgcn_wave_barrier();
  return t;
}

// Reduce `nv` (1 or 2) interleaved sequences of chunk partials level by level in one workgroup,
// in VectorOps' order, then apply the solver step `op` to the totals. The first level's chunks
// go to the workgroup's waves (red_threads: one wave per chunk, at most kRedWaves; their totals
// in LDS); wave 0 sums those (one chunk: fewer than 2^31 entries have at most kRedTot first-level
// chunks). The same bits as one wave walking the levels in turn.
constexpr int kRedWaves = 8, kRedThreads = kRedWaves * kLanes, kRedTot = 512;
static_assert(static_cast<int64_t>(kRedTot) * kChunk * kChunk >= (int64_t(1) << 31), "first-level totals of 2^31 entries");
__host__ __device__ inline int64_t red_chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }
// the reduction's workgroup for n entries: a wave per first-level chunk of the n / 2048 partials
inline dim3 red_threads(int64_t n) {
  const int64_t c = red_chunks(red_chunks(n));
  return dim3(static_cast<unsigned>(kLanes * (c < 1 ? 1 : c > kRedWaves ? kRedWaves : c)));
}
__global__ __launch_bounds__(kRedThreads) void reduce_kernel(int64_t npart, int nv, double* __restrict__ partial,
                                                             double* __restrict__ tmp, CgState* st, int op,
                                                             double rtol, double atol, double* out) {
  if (op == kPap || op == kIter)
    if (st->done) return;
  (void)tmp;
  __shared__ double tot[2][kRedTot];
  __shared__ double lane_sum[kRedWaves][kLanes];
  const int l = threadIdx.x % kLanes, w = threadIdx.x / kLanes;
  const int64_t nch = (npart + kChunk - 1) / kChunk;
  double total[2] = {0.0, 0.0};
  if (npart > 1) {
    for (int v = 0; v < nv; ++v)
      for (int64_t c = w; c < nch; c += static_cast<int64_t>(blockDim.x / kLanes)) {
        const int64_t begin = c * kChunk;
        const double t =
            wave_chunk_sum(partial + v, nv, begin, npart - begin < kChunk ? npart - begin : kChunk, l, lane_sum[w]);
        if (l == 0) tot[v][c] = t;
      }
    __syncthreads();
    if (w == 0)
      for (int v = 0; v < nv; ++v) total[v] = nch > 1 ? wave_chunk_sum(tot[v], 1, 0, nch, l, lane_sum[0]) : tot[v][0];
  } else {
    for (int v = 0; v < nv; ++v) total[v] = partial[v];
  }
  if (threadIdx.x != 0) return;
  switch (op) {
    case kStore:
      out[0] = total[0];
      break;
    case kRhsNorm:
      st->rhs_norm = sqrt(total[0]);
      st->tol = fmax(rtol * st->rhs_norm, atol);
      break;
    case kInitRes: {
      const double r = sqrt(total[0]);
      st->init_res = st->final_res = r;
      st->it = 0;
      st->indefinite = 0;
      st->converged = r <= st->tol;
      st->done = st->converged || st->max_it <= 0;
      break;
    }
    case kRz0:
      st->rz = total[0];
      break;
    case kPap:
      st->pap = total[0];
      if (!(total[0] > 0.0)) {
        st->indefinite = 1;
        st->done = 1;
      } else {
        st->alpha = st->rz / total[0];
      }
      break;
    case kIter: {
      const double r = sqrt(total[0]);
      st->it += 1;
      st->final_res = r;
      if (r <= st->tol) {
        st->converged = 1;
        st->done = 1;
      } else {
        st->beta = total[1] / st->rz;
        st->rz = total[1];
        if (st->it >= st->max_it) st->done = 1;
      }
      break;
    }
  }
}

template <typename T>
__global__ void residual_kernel(int64_t n, const T* __restrict__ b, const T* __restrict__ ap, T* __restrict__ r) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] + T(-1) * ap[i];
}

// z = inv r, p = z (first direction)
template <typename T>
__global__ void precond_first_kernel(int64_t n, const T* __restrict__ inv, const T* __restrict__ r, T* __restrict__ z,
                                     T* __restrict__ p) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    const T zi = inv[i] * r[i];
    z[i] = zi;
    p[i] = zi;
  }
}

// x += alpha p; r -= alpha ap; z = inv r   (cg.cpp:55-56, 64), skipped once done
// p = z + beta p (cg.cpp:68)
template <typename T>
__global__ void direction_kernel(int64_t n, const CgState* st, const T* __restrict__ z, T* __restrict__ p) {
  if (st->done) return;
  const T beta = static_cast<T>(st->beta);
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + beta * p[i];
}

// update_kernel followed by chunk_dot_kernel(r, r, r, z) in one pass: workgroup c updates x, r, z
// on chunk c (lane l on entries l, l + 64, ...) and forms the chunk's lane sums of r.r and r.z
// from the updated values in VectorOps' order, so x, r, z and the partials have the bits of the
// two kernels.
template <typename T>
__global__ __launch_bounds__(kLanes) void update_dot_kernel(int64_t n, const CgState* st, T* __restrict__ x,
                                                            T* __restrict__ r, T* __restrict__ z, const T* __restrict__ p,
                                                            const T* __restrict__ ap, const T* __restrict__ inv,
                                                            double* __restrict__ partial) {
  if (st->done) return;
  __shared__ double lane_sum[2][kLanes];
  const T alpha = static_cast<T>(st->alpha);
  const int64_t c = blockIdx.x;
  const int64_t begin = c * kChunk;
  const int64_t len = (n - begin < kChunk) ? n - begin : kChunk;
  const int l = threadIdx.x;
  double s = 0.0, s2 = 0.0;
  for (int64_t k = l; k < len; k += kLanes) {
    const int64_t i = begin + k;
    x[i] += alpha * p[i];
    const T ri = r[i] + (-alpha) * ap[i];
    r[i] = ri;
    const T zi = inv[i] * ri;
    z[i] = zi;
    s += as_double(ri) * as_double(ri);
    s2 += as_double(ri) * as_double(zi);
  }
  lane_sum[0][l] = s;
  lane_sum[1][l] = s2;
  __syncthreads();
  if (l == 0) {
    double t = 0.0, t2 = 0.0;
    for (int j = 0; j < kLanes; ++j) t += lane_sum[0][j];
    for (int j = 0; j < kLanes; ++j) t2 += lane_sum[1][j];
    partial[2 * c] = t;
    partial[2 * c + 1] = t2;
  }
}


template <typename T>
__global__ void reciprocal_kernel(int64_t n, const double* __restrict__ d, T* __restrict__ inv, int32_t* zeros) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    if (d[i] == 0.0) atomicAdd(zeros, 1);
    inv[i] = static_cast<T>(1.0 / d[i]);
  }
}

// x64 += scale * d32; r32 = scale_in * r64 (conversions of the refinement)
__global__ void refine_update_kernel(int64_t n, double* __restrict__ x, const float* __restrict__ d, double s) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] += s * static_cast<double>(d[i]);
}
__global__ void to_f32_kernel(int64_t n, const double* __restrict__ a, float* __restrict__ b, double s) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = static_cast<float>(a[i] * s);
}
template <typename T>
__global__ void zero_kernel(int64_t n, T* __restrict__ a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = T(0);
}

// Plain vector copy (also the calibration kernel of the HBM traffic counters: it moves exactly
// n * 16 bytes).
__global__ void copy_kernel(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

int copy64(fx_ctx* ctx, int64_t n, const double* a, double* b) {
  if (n <= 0) return FX_OK;
  hipLaunchKernelGGL(copy_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, ctx->stream, n, a, b);
  FX_HIP(ctx, hipGetLastError());
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

namespace {

inline dim3 grid_of(int64_t n, int bs = 256) { return dim3(static_cast<unsigned>((n + bs - 1) / bs)); }
inline int64_t chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }

// Reduction scratch: [2 * chunks] partials + a level buffer.
int ensure_red(fx_ctx* ctx, int64_t n) {
  const size_t need = static_cast<size_t>(2 * chunks(n) + 2 * (chunks(chunks(n)) + 2) + 16);
  if (ctx->red_cap >= need) return FX_OK;
  release_cg_graphs(ctx, nullptr);  // captured graphs read the old buffer
  if (ctx->red) (void)hipFree(ctx->red);
  ctx->red = nullptr;
  FX_HIP(ctx, hipMalloc(&ctx->red, need * sizeof(double)));
  ctx->red_cap = need;
  return FX_OK;
}

}  // namespace

// chunk partials of the columns' dots: block (chunk, column), the lane order of chunk_dot_kernel
__global__ __launch_bounds__(kLanes) void multi_chunk_dot_kernel(int64_t n, const double* __restrict__ V,
                                                                 const double* __restrict__ w,
                                                                 double* __restrict__ partial, int64_t nch) {
  __shared__ double lane_sum[kLanes];
  const int64_t c = blockIdx.x, col = blockIdx.y;
  const double* x = V + col * n;
  const int64_t begin = c * kChunk;
  const int64_t len = (n - begin < kChunk) ? n - begin : kChunk;
  const int l = threadIdx.x;
  double s = 0.0;
  for (int64_t k = l; k < len; k += kLanes) s += x[begin + k] * w[begin + k];
  lane_sum[l] = s;
  __syncthreads();
  if (l == 0) {
    double t = 0.0;
    for (int j = 0; j < kLanes; ++j) t += lane_sum[j];
    partial[col * nch + c] = t;
  }
}

// one workgroup per column: its chunk partials reduced level by level as reduce_kernel does
__global__ __launch_bounds__(kLanes) void multi_reduce_kernel(int64_t npart, const double* __restrict__ partial,
                                                              double* __restrict__ levels, int64_t lvl,
                                                              double* __restrict__ out) {
  __shared__ double lane_sum[kLanes];
  const int col = blockIdx.x, l = threadIdx.x;
  const double* src = partial + col * npart;
  double* bufs[2] = {levels + col * lvl, levels + col * lvl + lvl / 2};
  int which = 0;
  int64_t m = npart;
  while (m > 1) {
    const int64_t nch = (m + kChunk - 1) / kChunk;
    double* dst = bufs[which];
    for (int64_t c = 0; c < nch; ++c) {
      const int64_t begin = c * kChunk;
      const int64_t len = (m - begin < kChunk) ? m - begin : kChunk;
      double s = 0.0;
      for (int64_t k = l; k < len; k += kLanes) s += src[begin + k];
      lane_sum[l] = s;
      __syncthreads();
      if (l == 0) {
        double t = 0.0;
        for (int j = 0; j < kLanes; ++j) t += lane_sum[j];
        dst[c] = t;
      }
      __syncthreads();
    }
    m = nch;
    src = dst;
    which ^= 1;
  }
  if (l == 0) out[col] = src[0];
}

int multidot64(fx_ctx* ctx, int64_t n, const double* V, int ncol, const double* w, double* scratch, double* out) {
  if (ncol <= 0) return FX_OK;
  if (n <= 0) {
    for (int c = 0; c < ncol; ++c) out[c] = 0.0;
    return FX_OK;
  }
  const int64_t nch = chunks(n), lvl = 2 * (chunks(nch) + 2);
  double* partial = scratch;
  double* levels = partial + static_cast<size_t>(ncol) * nch;
  double* dres = levels + static_cast<size_t>(ncol) * lvl;
  hipLaunchKernelGGL(multi_chunk_dot_kernel, dim3(static_cast<unsigned>(nch), static_cast<unsigned>(ncol)), dim3(kLanes),
                     0, ctx->stream, n, V, w, partial, nch);
  hipLaunchKernelGGL(multi_reduce_kernel, dim3(static_cast<unsigned>(ncol)), dim3(kLanes), 0, ctx->stream, nch,
                     partial, levels, lvl, dres);
  FX_HIP(ctx, hipGetLastError());
  FX_HIP(ctx, hipMemcpyAsync(out, dres, sizeof(double) * ncol, hipMemcpyDeviceToHost, ctx->stream));
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

int dot64(fx_ctx* ctx, int64_t n, const double* x, const double* y, double* out) {
  if (n <= 0) {
    *out = 0.0;
    return FX_OK;
  }
  if (int rc = ensure_red(ctx, n)) return rc;
  double* part = ctx->red;
  double* tmp = ctx->red + 2 * chunks(n);
  hipLaunchKernelGGL(chunk_dot_kernel<double>, dim3(static_cast<unsigned>(chunks(n))), dim3(kLanes), 0, ctx->stream,
                     n, x, y, static_cast<const double*>(nullptr), static_cast<const double*>(nullptr), part,
                     static_cast<const int32_t*>(nullptr));
  double* dres = tmp + chunks(chunks(n)) + 8;
  hipLaunchKernelGGL(reduce_kernel, dim3(1), red_threads(n), 0, ctx->stream, chunks(n), 1, part, tmp,
                     static_cast<CgState*>(nullptr), static_cast<int>(kStore), 0.0, 0.0, dres);
  FX_HIP(ctx, hipGetLastError());
  FX_HIP(ctx, hipMemcpyAsync(out, dres, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

void release_cg_graphs(fx_ctx* ctx, const void* owner) {
  auto& g = ctx->cg_graphs;
  for (auto it = g.begin(); it != g.end();) {
    if (owner && it->owner != owner) {
      ++it;
      continue;
    }
    if (it->exec) (void)hipGraphExecDestroy(it->exec);
    if (it->graph) (void)hipGraphDestroy(it->graph);
    it = g.erase(it);
  }
}

namespace {

constexpr int kGraphIters = 8;     // CG iterations per captured graph
constexpr size_t kMaxGraphs = 64;  // per context (a serving worker keeps a few dozen block sizes)

struct CgGraphKey {
  const void* owner;
  int64_t n;
  int32_t tsize, initial;
  double ks, kd, km, rtol, atol;
  const void* ptr[10];
  // the batch arrays the operator kernels read whose pointers change between calls: the
  // committed and trial histories swap on every commit (the committed tangent is formed from
  // the committed history), so each parity of the swap has its own graph
  const void* state[2];
};

// The captured graph of kGraphIters iterations for this key, capturing it on first use.
template <typename Body>
hipGraphExec_t cg_graph(fx_ctx* ctx, const CgGraphKey& key, Body body) {
  std::vector<unsigned char> k(sizeof key);
  std::memcpy(k.data(), &key, sizeof key);
  for (auto& g : ctx->cg_graphs)
    if (g.key == k) return g.exec;
  if (ctx->cg_graphs.size() >= kMaxGraphs) {  // drop the oldest
    fx_cg_graph& old = ctx->cg_graphs.front();
    if (old.exec) (void)hipGraphExecDestroy(old.exec);
    if (old.graph) (void)hipGraphDestroy(old.graph);
    ctx->cg_graphs.erase(ctx->cg_graphs.begin());
  }
  hipStream_t s = ctx->stream;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  int rc = FX_OK;
  for (int i = 0; i < kGraphIters && rc == FX_OK; ++i) rc = body();
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  hipGraphExec_t exec = nullptr;
  if (rc != FX_OK || ec != hipSuccess || !graph || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return nullptr;  // the caller launches the iterations itself
  }
  ctx->cg_graphs.push_back(fx_cg_graph{std::move(k), key.owner, graph, exec});
  return exec;
}

// The CG setup on the device for vector type T: ||b||, r = b - A x, z = inv r, p = z, r.z; the
// state block st and the reduction scratch red (2 chunks(n) + levels) are the solve's own.
template <typename T, typename Apply>
int cg_setup(fx_ctx* ctx, int64_t n, Apply apply, const T* inv, const T* b, T* x, T* r, T* z, T* p, T* ap,
             CgState* st, double* red, double rtol, double atol, int max_it) {
  hipStream_t s = ctx->stream;
  double* part = red;
  double* tmp = red + 2 * chunks(n);
  const dim3 g = grid_of(n), cg_chunks(static_cast<unsigned>(chunks(n)));
  CgState init{};
  init.max_it = max_it;
  FX_HIP(ctx, hipMemcpyAsync(st, &init, sizeof(CgState), hipMemcpyHostToDevice, s));
  const T* nt = nullptr;
  // ||b||
  hipLaunchKernelGGL(chunk_dot_kernel<T>, cg_chunks, dim3(kLanes), 0, s, n, b, b, nt, nt, part,
                     static_cast<const int32_t*>(nullptr));
  hipLaunchKernelGGL(reduce_kernel, dim3(1), red_threads(n), 0, s, chunks(n), 1, part, tmp, st,
                     static_cast<int>(kRhsNorm), rtol, atol, static_cast<double*>(nullptr));
  // r = b - A x
  if (int rc = apply(x, ap, static_cast<const int32_t*>(nullptr))) return rc;
  hipLaunchKernelGGL(residual_kernel<T>, g, dim3(256), 0, s, n, b, ap, r);
  hipLaunchKernelGGL(chunk_dot_kernel<T>, cg_chunks, dim3(kLanes), 0, s, n, r, r, nt, nt, part,
                     static_cast<const int32_t*>(nullptr));
  hipLaunchKernelGGL(reduce_kernel, dim3(1), red_threads(n), 0, s, chunks(n), 1, part, tmp, st,
                     static_cast<int>(kInitRes), rtol, atol, static_cast<double*>(nullptr));
  // z = inv r, p = z, rz = r.z
  hipLaunchKernelGGL(precond_first_kernel<T>, g, dim3(256), 0, s, n, inv, r, z, p);
  hipLaunchKernelGGL(chunk_dot_kernel<T>, cg_chunks, dim3(kLanes), 0, s, n, r, z, nt, nt, part,
                     static_cast<const int32_t*>(nullptr));
  hipLaunchKernelGGL(reduce_kernel, dim3(1), red_threads(n), 0, s, chunks(n), 1, part, tmp, st,
                     static_cast<int>(kRz0), rtol, atol, static_cast<double*>(nullptr));
  FX_HIP(ctx, hipGetLastError());
  return FX_OK;
}

// One CG iteration (cg.cpp:44-69): A p, p.Ap, the fused x, r, z update with r.r and r.z, the
// direction; a no-op once the state's stop flag is set.
template <typename T, typename Apply>
int cg_iteration(fx_ctx* ctx, int64_t n, Apply apply, const T* inv, T* x, T* r, T* z, T* p, T* ap, CgState* st,
                 double* red, double rtol, double atol) {
  hipStream_t s = ctx->stream;
  double* part = red;
  double* tmp = red + 2 * chunks(n);
  const dim3 g = grid_of(n), cg_chunks(static_cast<unsigned>(chunks(n)));
  const T* nt = nullptr;
  const int32_t* done = &st->done;
  if (int rc = apply(p, ap, done)) return rc;
  hipLaunchKernelGGL(chunk_dot_kernel<T>, cg_chunks, dim3(kLanes), 0, s, n, p, ap, nt, nt, part, done);
  hipLaunchKernelGGL(reduce_kernel, dim3(1), red_threads(n), 0, s, chunks(n), 1, part, tmp, st,
                     static_cast<int>(kPap), rtol, atol, static_cast<double*>(nullptr));
  hipLaunchKernelGGL(update_dot_kernel<T>, cg_chunks, dim3(kLanes), 0, s, n, st, x, r, z, p, ap, inv, part);
  hipLaunchKernelGGL(reduce_kernel, dim3(1), red_threads(n), 0, s, chunks(n), 2, part, tmp, st,
                     static_cast<int>(kIter), rtol, atol, static_cast<double*>(nullptr));
  hipLaunchKernelGGL(direction_kernel<T>, g, dim3(256), 0, s, n, st, z, p);
  return FX_OK;
}

// The captured graph of a solve's iterations (null: graphs off, or the capture failed).
template <typename T, typename Body>
hipGraphExec_t cg_solve_graph(fx_ctx* ctx, int64_t n, const T* inv, const T* b, T* x, T* r, T* z, T* p, T* ap,
                              CgState* st, double* red, double rtol, double atol, int max_it, const fx_batch* owner,
                              const fx_coeffs* kc, Body body) {
  if (!ctx->use_graphs || !owner || !kc || max_it <= 0) return nullptr;
  CgGraphKey key{};
  key.owner = owner;
  key.n = n;
  key.tsize = static_cast<int32_t>(sizeof(T));
  key.initial = kc->initial;
  key.ks = kc->stiffness;
  key.kd = kc->damping;
  key.km = kc->mass;
  key.rtol = rtol;
  key.atol = atol;
  const void* ptrs[10] = {inv, b, x, r, z, p, ap, st, red, red + 2 * chunks(n)};
  for (int i = 0; i < 10; ++i) key.ptr[i] = ptrs[i];
  key.state[0] = owner->state_c;
  key.state[1] = owner->state_t;
  return cg_graph(ctx, key, body);
}

// The CG iteration on the device for vector type T. apply(x, y) computes y = A x for the
// operator of `owner` with coefficients k. The iterations run in groups between host polls of
// the device stop flag (4, 8, ..., 64 iterations); iterations after the stop are no-ops. With
// graphs on, each group replays a captured graph of kGraphIters iterations (the same kernels
// and arguments, so the same bits as launching them).
template <typename T, typename Apply>
int cg_device(fx_ctx* ctx, int64_t n, Apply apply, const T* inv, const T* b, T* x, T* r, T* z, T* p, T* ap,
              CgState* st, double rtol, double atol, int max_it, CgState& host, const fx_batch* owner = nullptr,
              const fx_coeffs* kc = nullptr) {
  if (int rc = ensure_red(ctx, n)) return rc;
  double* red = ctx->red;
  hipStream_t s = ctx->stream;
  if (int rc = cg_setup<T>(ctx, n, apply, inv, b, x, r, z, p, ap, st, red, rtol, atol, max_it)) return rc;
  auto body = [&]() -> int { return cg_iteration<T>(ctx, n, apply, inv, x, r, z, p, ap, st, red, rtol, atol); };
  hipGraphExec_t exec = cg_solve_graph<T>(ctx, n, inv, b, x, r, z, p, ap, st, red, rtol, atol, max_it, owner, kc, body);
  int launched = 0, poll = 4;
  for (;;) {
    FX_HIP(ctx, hipMemcpyAsync(&host, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    FX_HIP(ctx, hipStreamSynchronize(s));
    if (host.done || launched >= max_it) break;
    const int batch = std::min(poll, max_it - launched);
    if (exec) {
      const int graphs = (batch + kGraphIters - 1) / kGraphIters;
      for (int i = 0; i < graphs; ++i) FX_HIP(ctx, hipGraphLaunch(exec, s));
      launched += graphs * kGraphIters;
    } else {
      for (int i = 0; i < batch; ++i)
        if (int rc = body()) return rc;
      launched += batch;
    }
    FX_HIP(ctx, hipGetLastError());
    poll = std::min(poll * 2, 64);
  }
  return FX_OK;
}

int jacobi_inverse(fx_batch* b, const fx_coeffs& k, double* d, void* inv, bool f32) {
  fx_ctx* ctx = b->ctx;
  if (int rc = operator_diagonal64(b, k, d)) return rc;
  int32_t* zeros = reinterpret_cast<int32_t*>(ctx->scalars) + 64;
  FX_HIP(ctx, hipMemsetAsync(zeros, 0, sizeof(int32_t), ctx->stream));
  if (f32)
    hipLaunchKernelGGL(reciprocal_kernel<float>, grid_of(b->neq), dim3(256), 0, ctx->stream, b->neq, d,
                       static_cast<float*>(inv), zeros);
  else
    hipLaunchKernelGGL(reciprocal_kernel<double>, grid_of(b->neq), dim3(256), 0, ctx->stream, b->neq, d,
                       static_cast<double*>(inv), zeros);
  FX_HIP(ctx, hipGetLastError());
  int32_t hz = 0;
  FX_HIP(ctx, hipMemcpyAsync(&hz, zeros, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (hz > 0) {
    ctx->error = "Jacobi: " + std::to_string(hz) + " zero diagonal entries";
    return FX_EZERO;
  }
  return FX_OK;
}

}  // namespace

int jacobi_inverse_public(fx_batch* b, const fx_coeffs& k, double* d, double* inv) {
  return jacobi_inverse(b, k, d, inv, false);
}

// CG + Jacobi on vectors already in the solver numbering (no permutation, no true residual).
int cg64_solver(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* bi, double* xi,
                fx_solve_result& res) {
  fx_ctx* ctx = b->ctx;
  const int64_t n = b->neq;
  double *r = b->work, *z = r + n, *p = z + n, *ap = p + n, *inv = ap + n, *d = inv + n;
  if (int rc = jacobi_inverse(b, k, d, inv, false)) return rc;
  CgState* st = static_cast<CgState*>(ctx->scalars);
  CgState h{};
  auto apply = [&](const double* xx, double* yy, const int32_t* dn) { return operator_apply64(b, k, xx, yy, dn); };
  if (int rc = cg_device<double>(ctx, n, apply, inv, bi, xi, r, z, p, ap, st, o.relative_tolerance,
                                 o.absolute_tolerance, o.max_iterations, h, b, &k))
    return rc;
  res = fx_solve_result{};
  res.converged = h.converged;
  res.iterations = h.it;
  res.indefinite = h.indefinite;
  res.rhs_norm = h.rhs_norm;
  res.initial_residual = h.init_res;
  res.final_residual = h.final_res;
  return FX_OK;
}

namespace {
int cg64_once(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* rhs, double* x,
              fx_solve_result& res) {
  fx_ctx* ctx = b->ctx;
  const int64_t n = b->neq;
  double *r = b->work, *ap = r + 3 * n, *bi = r + 6 * n, *xi = r + 7 * n;
  if (int rc = to_internal<double>(b, rhs, bi)) return rc;
  if (int rc = to_internal<double>(b, x, xi)) return rc;
  if (int rc = cg64_solver(b, k, o, bi, xi, res)) return rc;
  // true residual ||b - A x||
  if (int rc = operator_apply64(b, k, xi, ap)) return rc;
  hipLaunchKernelGGL(residual_kernel<double>, grid_of(n), dim3(256), 0, ctx->stream, n, bi, ap, r);
  double rr = 0;
  if (int rc = dot64(ctx, n, r, r, &rr)) return rc;
  res.true_residual = std::sqrt(rr);
  return to_external<double>(b, xi, x);
}
}  // namespace

int cg64(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* rhs, double* x,
         fx_solve_result& res) {
  return cg64_once(b, k, o, rhs, x, res);
}

size_t cg_red_doubles(int64_t n) { return static_cast<size_t>(2 * chunks(n) + 2 * (chunks(chunks(n)) + 2) + 16); }

// Several FP64 CG solves advanced together on one stream (the serving loop's admission round): each
// poll round launches every unfinished solve's next group of iterations (a captured graph per
// solve with graphs on) and then reads every solve's state with one synchronisation. A solve
// keeps its own state block and reduction scratch (the batch's), and its kernels, arguments and
// order are those of cg64 on its own batch, so its iterates, result and x are the same bits.
int cg64_many(int32_t nb, fx_batch* const* bs, const fx_coeffs& k, const fx_solve_options& o,
              const double* const* rhs, double* const* x, fx_solve_result* res) {
  if (nb <= 0) return FX_OK;
  fx_ctx* ctx = bs[0]->ctx;
  hipStream_t s = ctx->stream;
  struct Solve {
    fx_batch* b;
    double *r, *z, *p, *ap, *inv, *d, *bi, *xi;
    CgState* st;
    double* red;
    hipGraphExec_t exec;
    int launched;
    bool active;
  };
  std::vector<Solve> sv(static_cast<size_t>(nb));
  std::vector<CgState> host(static_cast<size_t>(nb));
  for (int32_t i = 0; i < nb; ++i) {
    fx_batch* b = bs[i];
    const int64_t n = b->neq;
    Solve& v = sv[i];
    v.b = b;
    v.r = b->work;
    v.z = v.r + n;
    v.p = v.z + n;
    v.ap = v.p + n;
    v.inv = v.ap + n;
    v.d = v.inv + n;
    v.bi = v.r + 6 * n;
    v.xi = v.r + 7 * n;
    v.st = reinterpret_cast<CgState*>(b->cg_state);
    v.red = b->cg_red;
    v.launched = 0;
    v.active = true;
    if (int rc = to_internal<double>(b, rhs[i], v.bi)) return rc;
    if (int rc = to_internal<double>(b, x[i], v.xi)) return rc;
    if (int rc = jacobi_inverse(b, k, v.d, v.inv, false)) return rc;
    auto apply = [b, &k](const double* xx, double* yy, const int32_t* dn) { return operator_apply64(b, k, xx, yy, dn); };
    if (int rc = cg_setup<double>(ctx, n, apply, v.inv, v.bi, v.xi, v.r, v.z, v.p, v.ap, v.st, v.red,
                                  o.relative_tolerance, o.absolute_tolerance, o.max_iterations))
      return rc;
    auto body = [&, b, n]() -> int {
      return cg_iteration<double>(ctx, n, apply, v.inv, v.xi, v.r, v.z, v.p, v.ap, v.st, v.red, o.relative_tolerance,
                                  o.absolute_tolerance);
    };
    v.exec = cg_solve_graph<double>(ctx, n, v.inv, v.bi, v.xi, v.r, v.z, v.p, v.ap, v.st, v.red, o.relative_tolerance,
                                    o.absolute_tolerance, o.max_iterations, b, &k, body);
  }
  const int max_it = o.max_iterations;
  for (int poll = 4, left = nb; left > 0; poll = std::min(poll * 2, 64)) {
    for (int32_t i = 0; i < nb; ++i)
      if (sv[i].active)
        FX_HIP(ctx, hipMemcpyAsync(&host[i], sv[i].st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    FX_HIP(ctx, hipStreamSynchronize(s));
    for (int32_t i = 0; i < nb; ++i) {
      Solve& v = sv[i];
      if (!v.active) continue;
      const CgState& h = host[i];
      if (h.done || v.launched >= max_it) {
        // finished: cg64_solver's result, then cg64_once's true residual and permutation
        fx_batch* b = v.b;
        const int64_t n = b->neq;
        fx_solve_result& rs = res[i];
        rs = fx_solve_result{};
        rs.converged = h.converged;
        rs.iterations = h.it;
        rs.indefinite = h.indefinite;
        rs.rhs_norm = h.rhs_norm;
        rs.initial_residual = h.init_res;
        rs.final_residual = h.final_res;
        if (int rc = operator_apply64(b, k, v.xi, v.ap)) return rc;
        hipLaunchKernelGGL(residual_kernel<double>, grid_of(n), dim3(256), 0, s, n, v.bi, v.ap, v.r);
        double rr = 0;
        if (int rc = dot64(ctx, n, v.r, v.r, &rr)) return rc;
        rs.true_residual = std::sqrt(rr);
        if (int rc = to_external<double>(b, v.xi, x[i])) return rc;
        v.active = false;
        --left;
        continue;
      }
      const int batch = std::min(poll, max_it - v.launched);
      if (v.exec) {
        const int graphs = (batch + kGraphIters - 1) / kGraphIters;
        for (int g = 0; g < graphs; ++g) FX_HIP(ctx, hipGraphLaunch(v.exec, s));
        v.launched += graphs * kGraphIters;
      } else {
        const int64_t n = v.b->neq;
        fx_batch* b = v.b;
        auto apply = [b, &k](const double* xx, double* yy, const int32_t* dn) {
          return operator_apply64(b, k, xx, yy, dn);
        };
        for (int it = 0; it < batch; ++it)
          if (int rc = cg_iteration<double>(ctx, n, apply, v.inv, v.xi, v.r, v.z, v.p, v.ap, v.st, v.red,
                                            o.relative_tolerance, o.absolute_tolerance))
            return rc;
        v.launched += batch;
      }
      FX_HIP(ctx, hipGetLastError());
    }
  }
  return FX_OK;
}

// GMRES (method 1) or CG with the GMRES fallback (method 2) at the C ABI: permutations around
// the solver-numbering solve, the true residual reported.
int krylov64(fx_batch* b, int method, const fx_coeffs& k, const fx_solve_options& o, const double* rhs, double* x,
             fx_solve_result& res) {
  fx_ctx* ctx = b->ctx;
  const int64_t n = b->neq;
  double *r = b->work, *ap = r + 3 * n, *bi = r + 6 * n, *xi = r + 7 * n;
  if (int rc = to_internal<double>(b, rhs, bi)) return rc;
  if (int rc = to_internal<double>(b, x, xi)) return rc;
  if (int rc = method == 1 ? gmres64_solver(b, k, o, bi, xi, res) : krylov64_solver(b, k, o, bi, xi, res)) return rc;
  if (int rc = operator_apply64(b, k, xi, ap)) return rc;
  hipLaunchKernelGGL(residual_kernel<double>, grid_of(n), dim3(256), 0, ctx->stream, n, bi, ap, r);
  double rr = 0;
  if (int rc = dot64(ctx, n, r, r, &rr)) return rc;
  res.true_residual = std::sqrt(rr);
  return to_external<double>(b, xi, x);
}

int cg_mixed(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, double inner_rtol, const double* rhs,
             double* x, fx_solve_result& res) {
  const int64_t n = b->neq;
  double *bi = b->work + 6 * n, *xi = b->work + 7 * n;
  if (int rc = to_internal<double>(b, rhs, bi)) return rc;
  if (int rc = to_internal<double>(b, x, xi)) return rc;
  if (int rc = cg_mixed_solver(b, k, o, inner_rtol, bi, xi, res)) return rc;
  return to_external<double>(b, xi, x);
}

int cg_mixed_solver(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, double inner_rtol, const double* bi,
                    double* xi, fx_solve_result& res) {
  fx_ctx* ctx = b->ctx;
  if (!b->has_f32) {
    ctx->error = "fx_cg_mixed: the batch was created without FX_BATCH_FP32";
    return FX_EUNSUP;
  }
  const int64_t n = b->neq;
  double *r64 = b->work, *ap64 = r64 + 3 * n, *d64 = r64 + 5 * n;
  float *inv = b->work32, *r = inv + n, *z = r + n, *p = z + n, *ap = p + n, *dx = ap + n, *rhs32 = dx + n;
  if (int rc = jacobi_inverse(b, k, d64, inv, true)) return rc;
  CgState* st = static_cast<CgState*>(ctx->scalars);
  hipStream_t s = ctx->stream;
  double bb = 0;
  if (int rc = dot64(ctx, n, bi, bi, &bb)) return rc;
  const double bnorm = std::sqrt(bb);
  const double tol = std::max(o.relative_tolerance * bnorm, o.absolute_tolerance);
  res = fx_solve_result{};
  res.rhs_norm = bnorm;
  int total_inner = 0;
  auto apply32 = [&](const float* xx, float* yy, const int32_t* dn) { return operator_apply32(b, k, xx, yy, dn); };
  for (int outer = 0;; ++outer) {
    // r = b - A x in FP64
    if (int rc = operator_apply64(b, k, xi, ap64)) return rc;
    hipLaunchKernelGGL(residual_kernel<double>, grid_of(n), dim3(256), 0, s, n, bi, ap64, r64);
    double rr = 0;
    if (int rc = dot64(ctx, n, r64, r64, &rr)) return rc;
    const double rn = std::sqrt(rr);
    if (outer == 0) res.initial_residual = rn;
    res.final_residual = res.true_residual = rn;
    res.outer_iterations = outer;
    if (rn <= tol) {
      res.converged = 1;
      break;
    }
    if (total_inner >= o.max_iterations || outer >= 100) break;
    // inner FP32 solve of A d = r / ||r||
    hipLaunchKernelGGL(to_f32_kernel, grid_of(n), dim3(256), 0, s, n, r64, rhs32, 1.0 / rn);
    hipLaunchKernelGGL(zero_kernel<float>, grid_of(n), dim3(256), 0, s, n, dx);
    CgState h{};
    if (int rc = cg_device<float>(ctx, n, apply32, inv, rhs32, dx, r, z, p, ap, st, inner_rtol, 0.0,
                                  o.max_iterations - total_inner, h, b, &k))
      return rc;
    total_inner += h.it;
    if (h.indefinite) res.indefinite = 1;
    hipLaunchKernelGGL(refine_update_kernel, grid_of(n), dim3(256), 0, s, n, xi, dx, rn);
    FX_HIP(ctx, hipGetLastError());
    if (h.it == 0) break;  // no progress possible
  }
  res.iterations = total_inner;
  return FX_OK;
}

}  // namespace fx

// ---- VectorOps::dot without allocation (fx_dot_static) ------------------------------------
// The reduction scratch of fx_dot_static: a fixed array of the library on every device it runs
// on, sized for FX_DOT_STATIC_MAX_N entries (one chunk partial per 2048 entries, then the level
// buffers reduce_kernel walks: 131,220 doubles, 1.05 MB). In host code `scratch` names a host
// stand-in, never the device memory: host code uses only the address hipGetSymbolAddress gives
// for the current device, kept per device in scratch_ptr.
namespace fx {
namespace {

constexpr int64_t kStaticDotChunks = FX_DOT_STATIC_MAX_N / kChunk;
constexpr size_t kStaticScratch = kStaticDotChunks + 2 * (kStaticDotChunks / kChunk + 2) + 16;
static_assert(FX_DOT_STATIC_MAX_N % kChunk == 0, "whole chunks");

static __device__ __align__(256) double scratch[kStaticScratch];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_ptr[kMaxDevices] = {};

__global__ void store_zero_kernel(double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = 0.0;
}

}  // namespace

// Once per device, with that device current (fx_ctx_create): the device address of the array.
int static_scratch_init(fx_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) {
    ctx->error = "fx_ctx_create: device index beyond the static scratch table";
    return FX_EINVAL;
  }
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (scratch_ptr[ctx->device]) return FX_OK;
  void* p = nullptr;
  FX_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) {
    ctx->error = "hipGetSymbolAddress(scratch): null address";
    return FX_EDEVICE;
  }
  scratch_ptr[ctx->device] = static_cast<double*>(p);
  return FX_OK;
}

// The kernels and arguments of dot64 on the device's fixed scratch, the result left in device
// memory: no allocation, no host synchronisation.
int dot64_static(fx_ctx* ctx, int64_t n, const double* x, const double* y, double* out_dev) {
  double* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mutex);
    base = scratch_ptr[ctx->device];
  }
  if (!base) {
    ctx->error = "fx_dot_static: no static scratch on this device (context set-up failed)";
    return FX_EDEVICE;
  }
  hipStream_t s = ctx->stream;
  if (n == 0) {
    hipLaunchKernelGGL(store_zero_kernel, dim3(1), dim3(64), 0, s, out_dev);
    FX_HIP(ctx, hipGetLastError());
    return FX_OK;
  }
  double* part = base;
  double* tmp = base + chunks(n);
  hipLaunchKernelGGL(chunk_dot_kernel<double>, dim3(static_cast<unsigned>(chunks(n))), dim3(kLanes), 0, s, n, x, y,
                     static_cast<const double*>(nullptr), static_cast<const double*>(nullptr), part,
                     static_cast<const int32_t*>(nullptr));
  hipLaunchKernelGGL(reduce_kernel, dim3(1), red_threads(n), 0, s, chunks(n), 1, part, tmp,
                     static_cast<CgState*>(nullptr), static_cast<int>(kStore), 0.0, 0.0, out_dev);
  FX_HIP(ctx, hipGetLastError());
  return FX_OK;
}

}  // namespace fx

extern "C" int fx_dot_static(fx_ctx* ctx, int64_t n, const double* x, const double* y, double* out_dev) {
  if (!ctx || !out_dev || (n > 0 && (!x || !y))) {
    if (ctx) ctx->error = "fx_dot_static: null argument";
    else fx::set_thread_error("fx_dot_static: null argument");
    return FX_EINVAL;
  }
  if (n < 0 || n > FX_DOT_STATIC_MAX_N) {
    ctx->error = "fx_dot_static: n = " + std::to_string(n) + " outside [0, FX_DOT_STATIC_MAX_N = " +
                 std::to_string(static_cast<int64_t>(FX_DOT_STATIC_MAX_N)) + "]";
    return FX_EINVAL;
  }
  return fx::dot64_static(ctx, n, x, y, out_dev);
}
