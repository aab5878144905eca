// Node recorders (OpenSees `recorder Node -file f [-time] [-dT dT] [-precision p] -node ... -dof
// ... disp|vel|accel`) fed from device vectors.
//
// The reference's Recorders (exec/recorders.hpp:5-20, recorders.cpp:193-240, 272-298):
//   - a recorder is due after a committed step at `time` when dT == 0 or
//     time - next >= -dT * 1e-5, and then next = time + dT (NodeRecorder.cpp's rule);
//   - a record is one text line: the time (with -time), then each column's value, every number
//     printed as printf's %.{precision}g and separated by one space (DataFileStream's
//     setprecision in the default float format);
//   - a column is a node DOF: its equation's value, or 0 for a fixed DOF.
// Here: at a record, one gather kernel per due recorder copies its columns into a device
// staging array and only those values go to pinned host memory by an asynchronous copy on the
// context's stream, into one of a ring of buffers (the host waits for a copy only when every
// buffer holds a record in flight). A writer thread formats and writes the files, so the stream
// never waits for the disk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fx_internal.hpp"
#include "fx_kernels.hpp"

namespace {

int rec_fail(fx_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  return code;
}

// The due recorders of one record (up to kGatherSegs per launch): recorder s copies its columns
// into stage[off[s] ...]: src[s][map ? map[eq[j]] : eq[j]] for a free DOF, 0 for a fixed one.
constexpr int kGatherSegs = 8;
struct GatherSegs {
  const int32_t* eq[kGatherSegs];
  const double* src[kGatherSegs];
  int64_t off[kGatherSegs];
  int64_t cols[kGatherSegs];
};

__global__ __launch_bounds__(256) void recorder_gather_kernel(GatherSegs g, const int32_t* __restrict__ map,
                                                              double* __restrict__ stage) {
  const int s = blockIdx.y;
  const int32_t* __restrict__ eq = g.eq[s];
  const double* __restrict__ src = g.src[s];
  double* __restrict__ dst = stage + g.off[s];
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < g.cols[s];
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int32_t q = eq[j];
    if (q >= 0 && map) q = map[q];
    dst[j] = q >= 0 ? src[q] : 0.0;
  }
}

// Records in flight: one pinned buffer and event each, as many as fit 64 MiB (2 to 32).
constexpr int kMaxSlots = 32;
constexpr size_t kPinnedBudget = size_t(64) << 20;

}  // namespace

struct fx_recorder {
  struct Node {
    std::string file;
    int32_t response = 0;  // 0 disp, 1 vel, 2 accel
    bool echo_time = false;
    double dT = 0.0, next = 0.0;
    int32_t precision = 6;
    bool active = false;
    int64_t columns = 0, offset = 0;
    int32_t max_eq = -1;
    std::vector<int32_t> eq;
    int32_t* eq_dev = nullptr;
  };
  struct Job {
    enum Type { Open, Frame } type = Frame;
    int rec = 0;
    double time = 0.0;
    std::vector<double> values;
  };
  struct Pending {
    int slot = 0;
    std::vector<std::pair<int, double>> due;
  };

  fx_ctx* ctx = nullptr;
  std::vector<Node> recs;
  bool started = false, finished = false, any_active = false;
  int64_t total = 0;
  double* stage = nullptr;
  int nslots = 0;
  double* pinned[kMaxSlots] = {};
  hipEvent_t copied[kMaxSlots] = {};
  int next_slot = 0;
  std::deque<Pending> pending;
  std::atomic<int64_t> records{0};

  // the writer thread's files and queue
  std::vector<FILE*> files;
  std::thread writer;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> jobs;
  bool stop = false;
  std::string error;

  void enqueue(std::vector<Job>&& js) {
    {
      std::lock_guard<std::mutex> lock(mu);
      for (auto& j : js) jobs.push_back(std::move(j));
    }
    cv.notify_one();
  }

  void set_error(const std::string& e) {
    std::lock_guard<std::mutex> lock(mu);
    if (error.empty()) error = e;
  }

  void write(Job& job) {
    const Node& r = recs[static_cast<size_t>(job.rec)];
    FILE*& f = files[static_cast<size_t>(job.rec)];
    if (job.type == Job::Open) {
      if (f) std::fclose(f);
      f = std::fopen(r.file.c_str(), "w");
      if (!f) set_error("cannot open recorder file " + r.file);
      return;
    }
    if (!f) return;
    std::string s;
    s.reserve(static_cast<size_t>(r.columns + 1) * 26);
    char buf[64];
    if (r.echo_time) {
      std::snprintf(buf, sizeof buf, "%.*g", r.precision, job.time);
      s += buf;
    }
    for (double v : job.values) {
      std::snprintf(buf, sizeof buf, "%.*g", r.precision, v);
      if (!s.empty()) s += ' ';
      s += buf;
    }
    s += '\n';
    if (std::fwrite(s.data(), 1, s.size(), f) != s.size()) set_error("cannot write recorder file " + r.file);
    records.fetch_add(1);
  }

  void writer_loop() {
    for (;;) {
      Job job;
      {
        std::unique_lock<std::mutex> lock(mu);
        cv.wait(lock, [&] { return stop || !jobs.empty(); });
        if (jobs.empty()) break;  // stopped and drained
        job = std::move(jobs.front());
        jobs.pop_front();
      }
      write(job);
    }
    for (FILE*& f : files)
      if (f) {
        if (std::fclose(f) != 0) set_error("cannot close a recorder file");
        f = nullptr;
      }
  }

  // hands the oldest pending record to the writer once its copy has landed
  int flush_oldest() {
    Pending p = std::move(pending.front());
    pending.pop_front();
    FX_HIP(ctx, hipEventSynchronize(copied[p.slot]));
    std::vector<Job> js;
    for (const auto& [i, t] : p.due) {
      const Node& r = recs[static_cast<size_t>(i)];
      Job j;
      j.rec = i;
      j.time = t;
      j.values.assign(pinned[p.slot] + r.offset, pinned[p.slot] + r.offset + r.columns);
      js.push_back(std::move(j));
    }
    enqueue(std::move(js));
    return FX_OK;
  }

  int start() {
    total = 0;
    for (Node& r : recs) {
      r.offset = total;
      total += r.columns;
    }
    const size_t sz = sizeof(double) * static_cast<size_t>(std::max<int64_t>(total, 1));
    nslots = static_cast<int>(std::min<size_t>(kMaxSlots, std::max<size_t>(2, kPinnedBudget / sz)));
    FX_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&stage), sz));
    for (int s = 0; s < nslots; ++s) {
      FX_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&pinned[s]), sz, hipHostMallocDefault));
      FX_HIP(ctx, hipEventCreateWithFlags(&copied[s], hipEventDisableTiming));
    }
    for (Node& r : recs) {
      if (r.columns == 0) continue;
      FX_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&r.eq_dev), sizeof(int32_t) * r.eq.size()));
      FX_HIP(ctx, hipMemcpy(r.eq_dev, r.eq.data(), sizeof(int32_t) * r.eq.size(), hipMemcpyHostToDevice));
    }
    started = true;
    return FX_OK;
  }

  int record(double time, int64_t n, const double* u, const double* v, const double* a, const int32_t* map) {
    if (finished) return rec_fail(ctx, FX_EINVAL, "fx_recorder_record: the recorder has finished");
    std::vector<std::pair<int, double>> due;
    for (size_t i = 0; i < recs.size(); ++i) {
      const Node& r = recs[i];
      if (r.active && (r.dT == 0.0 || time - r.next >= -r.dT * 1e-5)) due.emplace_back(static_cast<int>(i), time);
    }
    if (due.empty()) return FX_OK;
    for (const auto& [i, t] : due) {
      const Node& r = recs[static_cast<size_t>(i)];
      if (r.max_eq >= n) return rec_fail(ctx, FX_EINVAL, "fx_recorder_record: a recorded equation is past the vectors");
      const double* src = r.response == 0 ? u : r.response == 1 ? v : a;
      if (!src && r.columns > 0) return rec_fail(ctx, FX_EINVAL, "fx_recorder_record: null vector for a due recorder");
    }
    for (const auto& [i, t] : due) {
      Node& r = recs[static_cast<size_t>(i)];
      if (r.dT != 0.0) r.next = time + r.dT;
    }
    if (!started)
      if (int rc = start()) return rc;
    // hand the records whose copies have landed to the writer; wait only when every slot is in use
    while (!pending.empty()) {
      const hipError_t q = hipEventQuery(copied[pending.front().slot]);
      if (q == hipErrorNotReady) break;
      FX_HIP(ctx, q);
      if (int rc = flush_oldest()) return rc;
    }
    if (static_cast<int>(pending.size()) >= nslots)
      if (int rc = flush_oldest()) return rc;
    const int slot = next_slot;
    next_slot = (next_slot + 1) % nslots;
    hipStream_t s = ctx->stream;
    // one gather launch per kGatherSegs due recorders, one copy of the span they fill
    int64_t lo = total, hi = 0, maxcols = 0;
    GatherSegs g{};
    int nseg = 0;
    auto launch = [&]() -> int {
      if (nseg == 0) return FX_OK;
      const int64_t blocks = std::min<int64_t>((maxcols + 255) / 256, 1024);
      hipLaunchKernelGGL(recorder_gather_kernel, dim3(static_cast<unsigned>(blocks), static_cast<unsigned>(nseg)),
                         dim3(256), 0, s, g, map, stage);
      FX_HIP(ctx, hipGetLastError());
      nseg = 0;
      maxcols = 0;
      return FX_OK;
    };
    for (const auto& [i, t] : due) {
      const Node& r = recs[static_cast<size_t>(i)];
      if (r.columns == 0) continue;
      g.eq[nseg] = r.eq_dev;
      g.src[nseg] = r.response == 0 ? u : r.response == 1 ? v : a;
      g.off[nseg] = r.offset;
      g.cols[nseg] = r.columns;
      ++nseg;
      maxcols = std::max(maxcols, r.columns);
      lo = std::min(lo, r.offset);
      hi = std::max(hi, r.offset + r.columns);
      if (nseg == kGatherSegs)
        if (int rc = launch()) return rc;
    }
    if (int rc = launch()) return rc;
    if (hi > lo)
      FX_HIP(ctx, hipMemcpyAsync(pinned[slot] + lo, stage + lo, sizeof(double) * static_cast<size_t>(hi - lo),
                                 hipMemcpyDeviceToHost, s));
    FX_HIP(ctx, hipEventRecord(copied[slot], s));
    Pending p;
    p.slot = slot;
    p.due = std::move(due);
    pending.push_back(std::move(p));
    return FX_OK;
  }

  int finish() {
    if (finished) return FX_OK;
    finished = true;
    int rc = FX_OK;
    while (!pending.empty() && rc == FX_OK) rc = flush_oldest();
    {
      std::lock_guard<std::mutex> lock(mu);
      stop = true;
    }
    cv.notify_one();
    if (writer.joinable()) writer.join();
    if (rc) return rc;
    if (!error.empty()) return rec_fail(ctx, FX_EIO, "fx_recorder: " + error);
    return FX_OK;
  }
};

namespace fx {
bool recorder_active(const fx_recorder* r) { return r && r->any_active && !r->finished; }

bool recorder_reads_rates(const fx_recorder* r) {
  if (!recorder_active(r)) return false;
  for (const auto& nd : r->recs)
    if (nd.active && nd.response != 0) return true;
  return false;
}

int batch_internal_map(fx_batch* b, const int32_t** map) {
  *map = nullptr;
  if (!b->internal) return FX_OK;
  if (!b->internal_of_dev) {
    const size_t sz = sizeof(int32_t) * b->internal_of.size();
    if (b->ctx->used + sz > b->ctx->budget) return rec_fail(b->ctx, FX_ENOMEM, "HBM budget exceeded");
    FX_HIP(b->ctx, hipMalloc(reinterpret_cast<void**>(&b->internal_of_dev), sz));
    b->ctx->used += sz;
    b->device_bytes += sz;
    FX_HIP(b->ctx, hipMemcpy(b->internal_of_dev, b->internal_of.data(), sz, hipMemcpyHostToDevice));
  }
  *map = b->internal_of_dev;
  return FX_OK;
}

int recorder_record(fx_recorder* r, double time, int64_t n, const double* u, const double* v, const double* a,
                    const int32_t* map) {
  if (!r) return FX_OK;
  return r->record(time, n, u, v, a, map);
}
}  // namespace fx

extern "C" {

int fx_recorder_create(fx_ctx* ctx, fx_recorder** out) {
  if (!ctx || !out) return rec_fail(ctx, FX_EINVAL, "fx_recorder_create: null argument");
  auto* r = new fx_recorder();
  r->ctx = ctx;
  r->writer = std::thread([r] { r->writer_loop(); });
  *out = r;
  return FX_OK;
}

int fx_recorder_add_node(fx_recorder* r, const char* file, int32_t response, int32_t echo_time, double dT,
                         int32_t precision, int64_t ncolumns, const int32_t* eq, int32_t* index) {
  if (!r || !file || (ncolumns > 0 && !eq) || ncolumns < 0)
    return rec_fail(r ? r->ctx : nullptr, FX_EINVAL, "fx_recorder_add_node: null argument");
  // the writer thread reads the recorder list once a recorder is active
  if (r->any_active || r->started || r->finished)
    return rec_fail(r->ctx, FX_EINVAL, "fx_recorder_add_node: add every recorder before the first activation");
  if (response < 0 || response > 2 || !(dT >= 0.0) || !std::isfinite(dT) || precision < 0 || precision > 40)
    return rec_fail(r->ctx, FX_EINVAL, "fx_recorder_add_node: response 0..2, dT >= 0, precision 0..40");
  fx_recorder::Node nd;
  nd.file = file;
  nd.response = response;
  nd.echo_time = echo_time != 0;
  nd.dT = dT;
  nd.precision = precision;
  nd.columns = ncolumns;
  nd.eq.assign(eq, eq + ncolumns);
  for (int32_t q : nd.eq) {
    if (q < -1) return rec_fail(r->ctx, FX_EINVAL, "fx_recorder_add_node: equations are >= 0, or -1 for a fixed DOF");
    nd.max_eq = std::max(nd.max_eq, q);
  }
  r->recs.push_back(std::move(nd));
  r->files.push_back(nullptr);
  if (index) *index = static_cast<int32_t>(r->recs.size() - 1);
  return FX_OK;
}

int fx_recorder_activate(fx_recorder* r, int32_t index) {
  if (!r) return FX_EINVAL;
  if (index < 0 || index >= static_cast<int32_t>(r->recs.size()) || r->finished)
    return rec_fail(r->ctx, FX_EINVAL, "fx_recorder_activate: no such recorder, or finished");
  fx_recorder::Node& nd = r->recs[static_cast<size_t>(index)];
  if (nd.active) return FX_OK;
  nd.active = true;
  r->any_active = true;
  std::vector<fx_recorder::Job> js(1);
  js[0].type = fx_recorder::Job::Open;
  js[0].rec = index;
  r->enqueue(std::move(js));
  return FX_OK;
}

int fx_recorder_record(fx_recorder* r, double time, int64_t n, const double* u, const double* v, const double* a) {
  if (!r || n < 0) return rec_fail(r ? r->ctx : nullptr, FX_EINVAL, "fx_recorder_record: bad argument");
  FX_HIP(r->ctx, hipSetDevice(r->ctx->device));
  return r->record(time, n, u, v, a, nullptr);
}

int fx_recorder_finish(fx_recorder* r) {
  if (!r) return FX_EINVAL;
  FX_HIP(r->ctx, hipSetDevice(r->ctx->device));
  return r->finish();
}

int64_t fx_recorder_records(const fx_recorder* r) { return r ? r->records.load() : 0; }

void fx_recorder_destroy(fx_recorder* r) {
  if (!r) return;
  (void)hipSetDevice(r->ctx->device);
  (void)r->finish();
  (void)hipStreamSynchronize(r->ctx->stream);
  for (auto& nd : r->recs)
    if (nd.eq_dev) (void)hipFree(nd.eq_dev);
  if (r->stage) (void)hipFree(r->stage);
  for (int s = 0; s < kMaxSlots; ++s) {
    if (r->pinned[s]) (void)hipHostFree(r->pinned[s]);
    if (r->copied[s]) (void)hipEventDestroy(r->copied[s]);
  }
  delete r;
}

}  // extern "C"
