// Replays the reference's InternalForce::apply_operator / operator_diagonal / mass_times_ground
// (amnp95/FemoraX src/exec/internal_force.cpp:442-448, 545-626) over the C ABI's per-batch verb,
// as the reference-side binding of INTEGRATION.md does: a Batch interface with the reference's
// run_operator signature (internal_force_batch.hpp:80-82), one FxBatch per element batch
// forwarding to fx_run_operator, and the InternalForce steps around it (y zeroed unless the
// scatter is Gather, every batch's run_operator, the SlotGather over (batch, element, DOF) pairs,
// the nodal masses). The SlotGather and nodal-mass loops, which are the reference's own kernels,
// run on the host here.
//
// Test driver (tests/test_gpu_binding_replay.py writes the input, compares the output with the
// oracle): replay_apply_operator IN OUT.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "femorax_amd.h"

namespace {

enum class Out : int { Plain = 0, Atomic = 1, Slots = 2 };  // assembly.hpp:24
enum class Scatter { Plain, Atomic, Gather };               // the OperatorSettings choice

struct Reader {
  FILE* f;
  template <typename T>
  std::vector<T> read(size_t n) {
    std::vector<T> v(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short input");
    return v;
  }
  template <typename T>
  T one() { return read<T>(1)[0]; }
};

void check(fx_ctx* ctx, int rc, const char* what) {
  if (rc != FX_OK) throw std::runtime_error(std::string(what) + ": " + fx_last_error(ctx));
}

// A device vector owned by the replay (the reference's Kokkos::View<double*>).
struct Vec {
  fx_ctx* ctx = nullptr;
  double* p = nullptr;
  size_t n = 0;
  Vec(fx_ctx* c, size_t count) : ctx(c), n(count) {
    void* d = nullptr;
    check(ctx, fx_alloc(ctx, sizeof(double) * (n ? n : 1), &d), "fx_alloc");
    p = static_cast<double*>(d);
  }
  ~Vec() { fx_free(ctx, p); }
  Vec(const Vec&) = delete;
  Vec& operator=(const Vec&) = delete;
  void upload(const std::vector<double>& h) { check(ctx, fx_upload(ctx, p, h.data(), sizeof(double) * n), "upload"); }
  std::vector<double> get() const {
    std::vector<double> h(n);
    check(ctx, fx_download(ctx, h.data(), p, sizeof(double) * n), "download");
    return h;
  }
  void zero() { check(ctx, fx_memset0(ctx, p, sizeof(double) * n), "memset"); }  // Kokkos::deep_copy(y, 0.0)
};

// InternalForce::Batch's operator verb (internal_force_batch.hpp:80-82), the settings, coords
// and u arguments omitted: the device batch holds them.
struct Batch {
  virtual ~Batch() = default;
  virtual void run_operator(bool diagonal, const fx_coeffs& k, Out out, const double* x, double* y, double* slots,
                            int unit_dir) = 0;
  virtual int64_t padded() const = 0;
  int64_t slot_offset = 0;  // Batch::slot_offset (operator_kernels.hpp:91)
};

// The binding: one FxBatch per compiled batch, forwarding to the C ABI.
struct FxBatch final : Batch {
  fx_ctx* ctx;
  fx_batch* b = nullptr;
  int32_t P;
  FxBatch(fx_ctx* c, fx_batch* batch, int32_t pad) : ctx(c), b(batch), P(pad) {}
  ~FxBatch() override { fx_batch_destroy(b); }
  void run_operator(bool diagonal, const fx_coeffs& k, Out out, const double* x, double* y, double* slots,
                    int unit_dir) override {
    check(ctx, fx_run_operator(b, diagonal ? 1 : 0, &k, static_cast<int>(out), x, y, slots, slot_offset, unit_dir),
          "fx_run_operator");
  }
  int64_t padded() const override { return P; }
};

// The steps of InternalForce that surround the per-batch verb.
struct InternalForceReplay {
  fx_ctx* ctx;
  int32_t neq;
  std::vector<std::unique_ptr<Batch>> batches;
  std::vector<std::vector<int32_t>> eqs;       // each batch's eq[a * P + e] (host)
  std::vector<int32_t> counts;
  std::unique_ptr<Vec> slots;                  // SlotGather::slots()
  std::vector<std::pair<int32_t, int64_t>> pairs;  // (target, slot) in summation order
  std::vector<double> nodal_m, nodal_am;
  std::vector<int8_t> eq_dof;                  // eq_dof_: the DOF direction of each equation

  void finish() {  // the gather pairs (internal_force.cpp:121-134): batches, elements, DOFs in order
    int64_t total = 0;
    for (size_t i = 0; i < batches.size(); ++i) {
      batches[i]->slot_offset = total;
      const int64_t P = batches[i]->padded();
      for (int32_t e = 0; e < counts[i]; ++e)
        for (int a = 0; a < 24; ++a) {
          const int32_t q = eqs[i][static_cast<size_t>(a) * P + e];
          if (q >= 0) pairs.emplace_back(q, total + a * P + e);
        }
      total += 24 * P;
    }
    slots = std::make_unique<Vec>(ctx, static_cast<size_t>(total));
  }
  void gather(Vec& y) const {  // SlotGather::gather (assembly.cpp:30-38): overwrites every target
    const std::vector<double> s = slots->get();
    std::vector<double> t(neq, 0.0);
    for (const auto& pr : pairs) t[pr.first] += s[pr.second];
    y.upload(t);
  }
  void add_nodal_mass(double c, const std::vector<double>& m, const Vec* x, Vec& y, bool diagonal) const {
    if (c == 0.0) return;  // internal_force.cpp:442-448
    std::vector<double> h = y.get();
    const std::vector<double> xv = diagonal ? std::vector<double>() : x->get();
    for (int32_t q = 0; q < neq; ++q) h[q] += c * m[q] * (diagonal ? 1.0 : xv[q]);
    y.upload(h);
  }
  static Out out_of(Scatter s) { return s == Scatter::Gather ? Out::Slots : s == Scatter::Atomic ? Out::Atomic : Out::Plain; }

  void apply_operator(const fx_coeffs& k, Scatter s, const Vec& x, Vec& y) {  // :545-573
    const Out out = out_of(s);
    if (out != Out::Slots) y.zero();
    for (auto& b : batches) b->run_operator(false, k, out, x.p, y.p, slots->p, -1);
    if (out == Out::Slots) gather(y);
    add_nodal_mass(k.mass, nodal_m, &x, y, false);
    add_nodal_mass(k.damping, nodal_am, &x, y, false);
  }
  void operator_diagonal(const fx_coeffs& k, Scatter s, Vec& d) {  // :575-599
    const Out out = out_of(s);
    if (out != Out::Slots) d.zero();
    for (auto& b : batches) b->run_operator(true, k, out, d.p, d.p, slots->p, -1);
    if (out == Out::Slots) gather(d);
    add_nodal_mass(k.mass, nodal_m, nullptr, d, true);
    add_nodal_mass(k.damping, nodal_am, nullptr, d, true);
  }
  void mass_times_ground(int direction, Scatter s, Vec& y) {  // :601-626
    const Out out = out_of(s);
    if (out != Out::Slots) y.zero();
    const fx_coeffs m{0.0, 0.0, 1.0, 0};
    for (auto& b : batches) b->run_operator(false, m, out, y.p, y.p, slots->p, direction);
    if (out == Out::Slots) gather(y);
    std::vector<double> h = y.get();
    for (int32_t q = 0; q < neq; ++q)
      if (eq_dof[q] == direction) h[q] += nodal_m[q];
    y.upload(h);
  }
};

int run(const char* in_path, const char* out_path) {
  FILE* f = fopen(in_path, "rb");
  if (!f) throw std::runtime_error("cannot open the input");
  Reader r{f};
  if (r.one<int32_t>() != 0x46584250) throw std::runtime_error("bad input");
  const int32_t neq = r.one<int32_t>(), nnode = r.one<int32_t>();
  const std::vector<double> coords = r.read<double>(3 * static_cast<size_t>(nnode));
  fx_ctx* ctx = nullptr;
  if (fx_ctx_create(0, 0, &ctx) != FX_OK) throw std::runtime_error("fx_ctx_create failed");
  InternalForceReplay force;
  force.ctx = ctx;
  force.neq = neq;
  force.eq_dof.assign(neq, -1);
  const int32_t nbatch = r.one<int32_t>();
  std::vector<double> u;
  std::vector<std::pair<fx_batch*, int32_t>> pimy;  // batches whose tangents are stored at u, with their stage
  for (int32_t i = 0; i < nbatch; ++i) {
    const int32_t kind = r.one<int32_t>(), count = r.one<int32_t>(), P = r.one<int32_t>(), nparam = r.one<int32_t>();
    const auto node = r.read<int32_t>(8 * static_cast<size_t>(P));
    const auto eq = r.read<int32_t>(24 * static_cast<size_t>(P));
    const auto param = r.read<int32_t>(static_cast<size_t>(P));
    fx_batch* b = nullptr;
    if (kind == 0) {
      const auto ps = r.read<fx_elastic_params>(nparam);
      fx_batch_desc d{count, P, node.data(), eq.data(), param.data(), nparam, ps.data(), nnode, coords.data(), neq};
      check(ctx, fx_batch_create(ctx, &d, FX_BATCH_SLOTS | FX_BATCH_GENERAL | FX_BATCH_NO_TUNE, &b), "fx_batch_create");
    } else {
      auto ps = r.read<fx_pimy_material>(nparam);
      for (auto& m : ps) m.gredu = nullptr;
      const int32_t stage = r.one<int32_t>();
      fx_pimy_batch_desc d{count, P, node.data(), eq.data(), param.data(), nparam, ps.data(), nnode, coords.data(),
                           neq, nullptr};
      check(ctx, fx_pimy_batch_create(ctx, &d, 0, &b), "fx_pimy_batch_create");
      check(ctx, fx_set_material_stage(b, stage), "fx_set_material_stage");
      pimy.emplace_back(b, stage);
    }
    for (int32_t e = 0; e < count; ++e)
      for (int a = 0; a < 24; ++a) {
        const int32_t q = eq[static_cast<size_t>(a) * P + e];
        if (q >= 0) force.eq_dof[q] = static_cast<int8_t>(a % 3);
      }
    force.batches.push_back(std::make_unique<FxBatch>(ctx, b, P));
    force.eqs.push_back(eq);
    force.counts.push_back(count);
  }
  const auto xh = r.read<double>(neq);
  u = r.read<double>(neq);
  force.nodal_m = r.read<double>(neq);
  force.nodal_am = r.read<double>(neq);
  const int32_t ncoef = r.one<int32_t>();
  const auto coefs = r.read<fx_coeffs>(ncoef);
  fclose(f);
  force.finish();
  Vec x(ctx, neq), y(ctx, neq), uu(ctx, neq);
  x.upload(xh);
  uu.upload(u);
  for (auto& pb : pimy) check(ctx, fx_store_tangents(pb.first, uu.p, nullptr), "fx_store_tangents");
  FILE* o = fopen(out_path, "wb");
  if (!o) throw std::runtime_error("cannot open the output");
  auto put = [&](const Vec& v) {
    const auto h = v.get();
    fwrite(h.data(), sizeof(double), h.size(), o);
  };
  for (Scatter s : {Scatter::Gather, Scatter::Plain, Scatter::Atomic}) {
    for (const fx_coeffs& k : coefs) {
      force.apply_operator(k, s, x, y);
      put(y);
      force.operator_diagonal(k, s, y);
      put(y);
    }
    for (int dir = 0; dir < 3; ++dir) {
      force.mass_times_ground(dir, s, y);
      put(y);
    }
  }
  fclose(o);
  force.batches.clear();
  force.slots.reset();
  fx_ctx_destroy(ctx);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 2;
  }
  try {
    return run(argv[1], argv[2]);
  } catch (const std::exception& e) {
    fprintf(stderr, "replay_apply_operator: %s\n", e.what());
    return 1;
  }
}
