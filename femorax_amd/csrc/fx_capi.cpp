// The extern "C" entry points of include/femorax_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <cmath>
#include <cstring>
#include <new>
#include <set>
#include <thread>
#include <vector>

#include "fx_internal.hpp"
#include "fx_kernels.hpp"
#include "fx_pimy.hpp"

struct fx_mesh {
  fx::Mesh m;
};

namespace {

thread_local std::string g_error;  // errors before a context exists

int fail(fx_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  else g_error = msg;
  return code;
}

template <typename T>
int dev_upload(fx_ctx* ctx, T*& dst, const T* src, size_t n, uint64_t& bytes) {
  dst = nullptr;
  if (n == 0) return FX_OK;
  const size_t sz = n * sizeof(T);
  if (ctx->used + sz > ctx->budget) return fail(ctx, FX_ENOMEM, "HBM budget exceeded");
  FX_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&dst), sz));
  ctx->used += sz;
  bytes += sz;
  if (src) FX_HIP(ctx, hipMemcpy(dst, src, sz, hipMemcpyHostToDevice));
  return FX_OK;
}

template <typename T>
int dev_zeros(fx_ctx* ctx, T*& dst, size_t n, uint64_t& bytes) {
  if (int rc = dev_upload<T>(ctx, dst, nullptr, n, bytes)) return rc;
  if (n) FX_HIP(ctx, hipMemset(dst, 0, n * sizeof(T)));
  return FX_OK;
}

void dev_free(fx_ctx* ctx, void* p, uint64_t& bytes, size_t sz) {
  if (!p) return;
  (void)hipFree(p);
  ctx->used -= std::min<uint64_t>(ctx->used, sz);
  bytes -= std::min<uint64_t>(bytes, sz);
}

}  // namespace

void fx::set_thread_error(const std::string& msg) { g_error = msg; }

// The gather of a batch's element slots k * padded + e: each equation's slots in (element, DOF)
// order (SlotGather's pairs, internal_force.cpp:121-134), and the equations with more than 64 of
// them (gathered one wave each).
static void build_gather(int32_t count, int32_t P, const int32_t* eq, int32_t neq, std::vector<int32_t>& start,
                  std::vector<int32_t>& list, std::vector<int32_t>& long_eqs) {
  start.assign(static_cast<size_t>(neq) + 1, 0);
  for (int32_t e = 0; e < count; ++e)
    for (int k = 0; k < 24; ++k) {
      const int32_t q = eq[static_cast<size_t>(k) * P + e];
      if (q >= 0) ++start[q + 1];
    }
  for (int32_t q = 0; q < neq; ++q) start[q + 1] += start[q];
  long_eqs.clear();
  for (int32_t q = 0; q < neq; ++q)
    if (start[q + 1] - start[q] > 64) long_eqs.push_back(q);
  list.assign(start[neq], 0);
  std::vector<int32_t> fill(start.begin(), start.end() - 1);
  for (int32_t e = 0; e < count; ++e)
    for (int k = 0; k < 24; ++k) {
      const int32_t q = eq[static_cast<size_t>(k) * P + e];
      if (q >= 0) list[fill[q]++] = k * P + e;
    }
}

extern "C" {

int fx_ctx_create(int device, uint64_t hbm_budget_bytes, fx_ctx** out) {
  if (!out) return fail(nullptr, FX_EINVAL, "fx_ctx_create: null out");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(nullptr, FX_EDEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(nullptr, FX_EINVAL, "fx_ctx_create: no such device");
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, FX_EDEVICE, "hipSetDevice failed");
  auto* c = new (std::nothrow) fx_ctx();
  if (!c) return fail(nullptr, FX_ENOMEM, "host allocation failed");
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(nullptr, FX_EDEVICE, "hipStreamCreate failed");
  }
  c->own_stream = c->stream;

  size_t fr = 0, tot = 0;
  (void)hipMemGetInfo(&fr, &tot);
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : static_cast<uint64_t>(fr / 10 * 9);
  if (hipMalloc(&c->scalars, 4096) != hipSuccess ||
      hipMalloc(&c->pimy_params, sizeof(fx::PimyParams) * fx::kMaxPointMaterials) != hipSuccess) {
    if (c->scalars) (void)hipFree(c->scalars);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(nullptr, FX_EDEVICE, "hipMalloc failed");
  }
  if (int rc = fx::static_scratch_init(c)) {
    const std::string msg = c->error;
    (void)hipFree(c->scalars);
    (void)hipFree(c->pimy_params);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(nullptr, rc, msg);
  }
  *out = c;
  return FX_OK;
}

void fx_ctx_destroy(fx_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  fx::release_cg_graphs(ctx, nullptr);
  if (ctx->red) (void)hipFree(ctx->red);
  if (ctx->scalars) (void)hipFree(ctx->scalars);
  if (ctx->pimy_params) (void)hipFree(ctx->pimy_params);
  (void)hipStreamSynchronize(ctx->own_stream);
  (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
}

int fx_ctx_set_option(fx_ctx* ctx, int option, int64_t value) {
  if (!ctx) return fail(nullptr, FX_EINVAL, "fx_ctx_set_option: null context");
  switch (option) {
    case FX_OPT_GRAPHS:
      ctx->use_graphs = value != 0;
      if (!ctx->use_graphs) fx::release_cg_graphs(ctx, nullptr);
      return FX_OK;
    case FX_OPT_STREAM: {
      // the work queued so far completes first; graphs captured on the old stream are dropped
      FX_HIP(ctx, hipSetDevice(ctx->device));
      FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
      fx::release_cg_graphs(ctx, nullptr);
      ctx->stream = value ? reinterpret_cast<hipStream_t>(static_cast<intptr_t>(value)) : ctx->own_stream;
      return FX_OK;
    }
    case FX_OPT_STREAM_ORDERED:
      ctx->stream_ordered = value != 0;
      return FX_OK;
    default:
      return fail(ctx, FX_EINVAL, "fx_ctx_set_option: unknown option");
  }
}

const char* fx_last_error(const fx_ctx* ctx) { return ctx ? ctx->error.c_str() : g_error.c_str(); }

int fx_sync(fx_ctx* ctx) {
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

int fx_device_memory(fx_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
  size_t f = 0, t = 0;
  FX_HIP(ctx, hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return FX_OK;
}

uint64_t fx_budget_used(const fx_ctx* ctx) { return ctx ? ctx->used : 0; }

int fx_alloc(fx_ctx* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return fail(ctx, FX_EINVAL, "fx_alloc: null argument");
  *dptr = nullptr;
  if (bytes == 0) bytes = 8;
  FX_HIP(ctx, hipMalloc(dptr, bytes));
  return FX_OK;
}
int fx_free(fx_ctx* ctx, void* dptr) {
  if (dptr) FX_HIP(ctx, hipFree(dptr));
  return FX_OK;
}
int fx_upload(fx_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes) FX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}
int fx_download(fx_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes) FX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}
int fx_memset0(fx_ctx* ctx, void* dst, size_t bytes) {
  if (bytes) FX_HIP(ctx, hipMemsetAsync(dst, 0, bytes, ctx->stream));
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

// ---- compile step -------------------------------------------------------------------------
int fx_compile_model(const fx_model_desc* model, int renumber_rcm, int simd_width, int build_coloring,
                     fx_mesh** out) {
  if (!model || !out) return fail(nullptr, FX_EINVAL, "fx_compile_model: null argument");
  *out = nullptr;
  if (model->nnode < 0 || model->nelem < 0 || (model->nelem > 0 && (!model->conn || !model->xyz)))
    return fail(nullptr, FX_EINVAL, "fx_compile_model: bad sizes");
  auto* m = new (std::nothrow) fx_mesh();
  if (!m) return fail(nullptr, FX_ENOMEM, "host allocation failed");
  std::string err;
  const int rc = fx::compile_model(*model, renumber_rcm, simd_width, build_coloring, m->m, err);
  if (rc != FX_OK) {
    delete m;
    return fail(nullptr, rc, err);
  }
  *out = m;
  return FX_OK;
}

int fx_mesh_get(const fx_mesh* mesh, fx_mesh_view* v) {
  if (!mesh || !v) return fail(nullptr, FX_EINVAL, "fx_mesh_get: null argument");
  const fx::Mesh& m = mesh->m;
  v->nnode = m.nnode;
  v->nelem = m.nelem;
  v->padded = m.padded;
  v->neq = m.neq;
  v->nparam = m.nparam;
  v->ncolor = m.ncolor;
  v->node_order = m.node_order.data();
  v->node_rank = m.node_rank.data();
  v->elem_order = m.elem_order.data();
  v->node_eq = m.node_eq.data();
  v->coords = m.coords.data();
  v->node = m.node.data();
  v->eq = m.eq.data();
  v->param = m.param.data();
  v->param_group = m.param_group.data();
  v->param_region = m.param_region.data();
  v->color_start = m.color_start.empty() ? nullptr : m.color_start.data();
  v->color_elements = m.color_elements.empty() ? nullptr : m.color_elements.data();
  return FX_OK;
}

void fx_mesh_destroy(fx_mesh* mesh) { delete mesh; }

// ---- batch ------------------------------------------------------------------------------------
void fx_batch_destroy(fx_batch* b) {
  if (!b) return;
  fx_ctx* ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  fx::release_cg_graphs(ctx, b);  // they read the batch's arrays
  void* ptrs[] = {b->node, b->param, b->node_eq, b->coords, b->coords32, b->params, b->blk.elems,
                  b->blk.ln_start, b->blk.lnode, b->blk.ln_part, b->blk.ln_slots, b->blk.slot_list,
                  b->blk.sh_node, b->blk.sh_start, b->blk.pos, b->blk.lcorner, b->blk.ln_eq, b->blk.ln_xyz,
                  b->blk.ln_xyz32, b->ynode, b->ynode32, b->con_start, b->con_list, b->work, b->work32,
                  b->tiles, b->words, b->tile_slots, b->cell_param, b->orphans, b->shared_word, b->shared_start,
                  b->shared_meta, b->tied_word, b->tied_start, b->tie_idx, b->tie_val, b->partial64, b->partial32, b->perm, b->lumped_m, b->lumped_am, b->body, b->pimy_params,
                  b->state_c, b->state_t, b->tangent, b->tangent32, b->slots, b->gather_start, b->gather_list, b->long_eqs, b->pimy_touched, b->dyn, b->gm,
                  b->cres, b->clist, b->ccount, b->nodal_m, b->nodal_am, b->nodal_m32, b->nodal_am32, b->eq_dir_dev,
                  b->words_ref, b->shared_word_ref, b->tied_word_ref, b->orphans_ref, b->ssp_elem, b->ssp_box,
                  b->kx, b->ky, b->kz, b->cg_state, b->cg_red, b->internal_of_dev};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  ctx->used -= std::min<uint64_t>(ctx->used, b->device_bytes);
  delete b;
}

namespace {
constexpr uint32_t kDepthFlags = FX_BATCH_TILE4 | FX_BATCH_TILE8 | FX_BATCH_TILE16;
constexpr int kTuneMinElements = 32768;

// The model half of the autotune key (tune_key, force_settings.cpp:157-177): lattice size and the
// operator forms' terms.
std::string lattice_model_key(const fx_batch_desc* d, int forced_depth) {
  bool mass = false, damping = false;
  for (int32_t i = 0; i < d->nparam; ++i) {
    mass |= d->params[i].rho != 0.0;
    const fx_rayleigh& r = d->params[i].damp;
    damping |= r.alphaM != 0.0 || r.betaK != 0.0 || r.betaKinit != 0.0 || r.betaKcomm != 0.0;
  }
  return fx::tune_model_key(d->count, mass, damping, forced_depth);
}

uint32_t depth_flag(int depth) {
  return depth == 4 ? FX_BATCH_TILE4 : depth == 16 ? FX_BATCH_TILE16 : FX_BATCH_TILE8;
}
}  // namespace

int fx_batch_create(fx_ctx* ctx, const fx_batch_desc* d, uint32_t flags, fx_batch** out) {
  if (!ctx || !d || !out) return fail(ctx, FX_EINVAL, "fx_batch_create: null argument");
  *out = nullptr;
  // Tile depth by autotune (choose_force_settings, force_settings.cpp:179-205): a large lattice
  // without a depth flag tries two depths (8 and 16 cells from 2^20 cells up, 4 and 8 below), each
  // with its variant autotune, and keeps the faster; the choice is kept in the autotune cache per
  // device and model, so a known model is built once, with the kept depth and variant.
  if (!(flags & (kDepthFlags | FX_BATCH_NO_TUNE | FX_BATCH_GENERAL | FX_BATCH_SSPBRICK)) &&
      d->count >= kTuneMinElements && d->params && d->nparam > 0) {
    const std::string key = fx::tune_machine_key(ctx->device) + "|" + lattice_model_key(d, 0);
    int depth = 0, variant = -1;
    if (fx::tune_cache_get(key, depth, variant)) {
      fx_batch* b = nullptr;
      if (int rc = fx_batch_create(ctx, d, flags | depth_flag(depth) | FX_BATCH_NO_TUNE, &b)) return rc;
      if (b->lattice) fx_batch_tune(b, variant);
      *out = b;
      return FX_OK;
    }
    const int cand[2] = {d->count >= (1 << 20) ? 8 : 4, d->count >= (1 << 20) ? 16 : 8};
    fx_batch* best = nullptr;
    int best_depth = 0;
    for (int c : cand) {
      fx_batch* b = nullptr;
      if (int rc = fx_batch_create(ctx, d, flags | depth_flag(c), &b)) {
        fx_batch_destroy(best);
        return rc;
      }
      if (!b->lattice) {  // not a lattice: no depth to choose
        fx_batch_destroy(best);
        *out = b;
        return FX_OK;
      }
      if (!best || b->tuned_ms < best->tuned_ms) {
        fx_batch_destroy(best);
        best = b;
        best_depth = c;
      } else {
        fx_batch_destroy(b);
      }
    }
    fx::tune_cache_put(key, best_depth, best->box_variant + 16 * (best->box_variant_mass + 1));
    *out = best;
    return FX_OK;
  }
  if (d->count < 0 || d->padded < d->count || d->nparam <= 0 || d->nnode <= 0 || d->neq < 0 || !d->node ||
      !d->param || !d->params || !d->coords || (!d->eq))
    return fail(ctx, FX_EINVAL, "fx_batch_create: bad descriptor");
  FX_HIP(ctx, hipSetDevice(ctx->device));
  const int32_t P = d->padded, E = d->count, N = d->nnode;
  // validate maps and derive the node -> equation map from Batch::eq (3 DOFs per node)
  std::vector<int32_t> node_eq(static_cast<size_t>(3) * N, -1);
  std::vector<char> seen(static_cast<size_t>(3) * N, 0);
  for (int32_t e = 0; e < E; ++e) {
    const int32_t pi = d->param[e];
    if (pi < 0 || pi >= d->nparam) return fail(ctx, FX_EINVAL, "fx_batch_create: param index out of range");
    for (int a = 0; a < 8; ++a) {
      const int32_t n = d->node[static_cast<size_t>(a) * P + e];
      if (n < 0 || n >= N) return fail(ctx, FX_EINVAL, "fx_batch_create: node index out of range");
      for (int k = 0; k < 3; ++k) {
        const int32_t q = d->eq[static_cast<size_t>(3 * a + k) * P + e];
        if (q < -1 || q >= d->neq) return fail(ctx, FX_EINVAL, "fx_batch_create: equation out of range");
        const size_t at = static_cast<size_t>(3) * n + k;
        if (seen[at] && node_eq[at] != q)
          return fail(ctx, FX_EINVAL, "fx_batch_create: a node DOF maps to two equations");
        seen[at] = 1;
        node_eq[at] = q;
      }
    }
  }
  for (int32_t i = 0; i < d->nparam; ++i) {
    const auto& p = d->params[i];
    if (!(p.E > 0) || !(p.nu > -1.0 && p.nu < 0.5) || p.rho < 0 || (p.mass_type != 0 && p.mass_type != 1))
      return fail(ctx, FX_EINVAL, "fx_batch_create: invalid elastic parameters in set " + std::to_string(i));
  }
  auto* b = new (std::nothrow) fx_batch();
  if (!b) return fail(ctx, FX_ENOMEM, "host allocation failed");
  b->ctx = ctx;
  b->count = E;
  b->padded = P;
  b->nnode = N;
  b->neq = d->neq;
  b->nparam = d->nparam;
  b->has_f32 = (flags & FX_BATCH_FP32) != 0;
  b->ssp = (flags & FX_BATCH_SSPBRICK) != 0;
  for (int32_t i = 0; i < d->nparam; ++i) {
    const auto& p = d->params[i];
    fx::DevParam q{};
    // ElasticIsotropic (elastic_isotropic.hpp:45-48): mu2 = E/(1+nu), lam = nu mu2/(1-2nu), mu = mu2/2
    const double mu2 = p.E / (1.0 + p.nu);
    q.lam = p.nu * mu2 / (1.0 - 2.0 * p.nu);
    q.mu = 0.5 * mu2;
    q.rho = p.rho;
    for (int k = 0; k < 3; ++k) q.b[k] = p.b[k];
    q.mass_type = b->ssp ? 1 : p.mass_type;  // SSPbrick's mass is always lumped
    q.alphaM = p.damp.alphaM;
    q.betaK = p.damp.betaK;
    q.betaKinit = p.damp.betaKinit;
    q.betaKcomm = p.damp.betaKcomm;
    b->host_params.push_back(q);
  }
  // equations no element touches (zeroed after each apply, as the reference's gather does)
  std::vector<char> covered(static_cast<size_t>(d->neq) + 1, 0);
  for (int32_t e = 0; e < E; ++e)
    for (int k = 0; k < 24; ++k) {
      const int32_t q = d->eq[static_cast<size_t>(k) * P + e];
      if (q >= 0) covered[q] = 1;
    }
  std::vector<int32_t> orphans;
  for (int32_t q = 0; q < d->neq; ++q)
    if (!covered[q]) orphans.push_back(q);
  fx::Lattice lat;
  const int depth = (flags & FX_BATCH_TILE4)    ? 4
                    : (flags & FX_BATCH_TILE8)  ? 8
                    : (flags & FX_BATCH_TILE16) ? (b->ssp ? 8 : 16)
                    : E >= (1 << 20)           ? 8
                                               : 4;
  // SSPbrick's lattice constants assume one cell size: rectilinear SSPbrick lattices take the
  // general path; rectilinear lattices run 4- or 8-deep tiles
  bool use_lattice = !(flags & FX_BATCH_GENERAL) &&
                     fx::build_lattice(E, P, d->node, d->param, N, d->coords, node_eq.data(), d->neq,
                                       (flags & FX_BATCH_XCD_SLABS) != 0, depth, !b->ssp, lat);
  if (use_lattice && lat.rect && depth == 16)
    use_lattice = fx::build_lattice(E, P, d->node, d->param, N, d->coords, node_eq.data(), d->neq,
                                    (flags & FX_BATCH_XCD_SLABS) != 0, 8, true, lat);
  // an SSPbrick lattice with DOFs one element ties together: the general path (the diagonal's
  // cross terms come from each element's constants there)
  if (use_lattice && b->ssp && !lat.tied_word.empty()) {
    use_lattice = false;
    lat = fx::Lattice{};
  }
  if (use_lattice) {  // orphans in the internal numbering
    std::vector<int32_t> internal_orphans;
    for (int32_t i = 0; i < d->neq; ++i)
      if (!covered[lat.perm[i]]) internal_orphans.push_back(i);
    orphans.swap(internal_orphans);
  }
  fx::Blocks blk;
  fx::Contraction con;
  if (!use_lattice) {
    fx::build_blocks(E, P, d->node, N, d->coords, blk);
    fx::build_contraction(N, d->neq, node_eq.data(), con);
  }
  b->tied = con.tied;
  b->lattice = use_lattice;
  uint64_t& by = b->device_bytes;
  int rc = FX_OK;
  auto ok = [&](int r) {
    if (r != FX_OK && rc == FX_OK) rc = r;
    return rc == FX_OK;
  };
  ok(dev_upload(ctx, b->params, b->host_params.data(), b->host_params.size(), by)) &&
      ok(dev_zeros(ctx, b->work, static_cast<size_t>(8) * std::max(1, d->neq), by)) &&
      ok(dev_zeros(ctx, b->cg_state, 32, by)) &&
      ok(dev_zeros(ctx, b->cg_red, fx::cg_red_doubles(std::max(1, d->neq)), by)) &&
      ok(b->has_f32 ? dev_zeros(ctx, b->work32, static_cast<size_t>(9) * std::max(1, d->neq), by) : FX_OK) &&
      ok(orphans.empty() ? FX_OK : dev_upload(ctx, b->orphans, orphans.data(), orphans.size(), by));
  b->norphans = static_cast<int32_t>(orphans.size());
  if (use_lattice) {
    b->tile_z = lat.tz;
    b->hx = lat.hx;
    b->hy = lat.hy;
    b->hz = lat.hz;
    b->rect = lat.rect;
    if (lat.rect) {
      // each cell interval's (kappa = 1 / (4 h), h) along each axis
      const std::vector<double>* planes[3] = {&lat.px, &lat.py, &lat.pz};
      double** dst[3] = {&b->kx, &b->ky, &b->kz};
      for (int ax = 0; ax < 3; ++ax) {
        const auto& pl = *planes[ax];
        std::vector<double> k(2 * (pl.size() - 1));
        for (size_t i = 0; i + 1 < pl.size(); ++i) {
          const double h = pl[i + 1] - pl[i];
          k[2 * i] = 0.25 / h;
          k[2 * i + 1] = h;
        }
        ok(dev_upload(ctx, *dst[ax], k.data(), k.size(), by));
      }
    }
    b->ntiles = static_cast<int32_t>(lat.tiles.size());
    b->internal = true;
    b->nshared = static_cast<int32_t>(lat.shared_word.size());
    b->ntied = static_cast<int32_t>(lat.tied_word.size());
    ok(dev_upload(ctx, b->tiles, lat.tiles.data(), lat.tiles.size(), by)) &&
        ok(dev_upload(ctx, b->words, lat.words.data(), lat.words.size(), by)) &&
        ok(dev_upload(ctx, b->tile_slots, lat.slots.data(), lat.slots.size(), by)) &&
        ok(dev_upload(ctx, b->cell_param, lat.cell_param.data(), lat.cell_param.size(), by)) &&
        ok(lat.shared_word.empty() ? FX_OK : dev_upload(ctx, b->shared_word, lat.shared_word.data(), lat.shared_word.size(), by)) &&
        ok(lat.shared_start.empty() ? FX_OK : dev_upload(ctx, b->shared_start, lat.shared_start.data(), lat.shared_start.size(), by)) &&
        ok(lat.shared_meta.empty() ? FX_OK : dev_upload(ctx, b->shared_meta, lat.shared_meta.data(), lat.shared_meta.size(), by)) &&
        ok(lat.tied_word.empty() ? FX_OK : dev_upload(ctx, b->tied_word, lat.tied_word.data(), lat.tied_word.size(), by)) &&
        ok(lat.tied_word.empty() ? FX_OK : dev_upload(ctx, b->tied_start, lat.tied_start.data(), lat.tied_start.size(), by)) &&
        ok(dev_zeros(ctx, b->partial64, static_cast<size_t>(std::max<int64_t>(lat.nslot, 1)) * 3, by)) &&
        ok(b->has_f32 ? dev_zeros(ctx, b->partial32, static_cast<size_t>(std::max<int64_t>(lat.nslot, 1)) * 3, by) : FX_OK) &&
        ok(dev_upload(ctx, b->perm, lat.perm.data(), lat.perm.size(), by));
    // the node words in the reference numbering, where each node's free DOFs have consecutive
    // reference equations (so a word's first equation and mask still name them)
    bool consecutive = true;
    auto to_ref = [&](uint32_t w) {
      const uint32_t mask = (w >> 28) & 7u;
      if (mask == 0u) return w;
      const int32_t i0 = static_cast<int32_t>(w & 0x0fffffffu);
      const int cnt = ((mask >> 0) & 1u) + ((mask >> 1) & 1u) + ((mask >> 2) & 1u);
      if (i0 + cnt > d->neq) {
        consecutive = false;
        return w;
      }
      const int32_t q0 = lat.perm[i0];
      for (int j = 1; j < cnt; ++j) consecutive &= lat.perm[i0 + j] == q0 + j;
      return (w & 0xf0000000u) | static_cast<uint32_t>(q0);
    };
    std::vector<uint32_t> wr(lat.words.size()), sr(lat.shared_word.size()), tr(lat.tied_word.size());
    for (size_t i = 0; i < wr.size(); ++i) wr[i] = to_ref(lat.words[i]);
    for (size_t i = 0; i < sr.size(); ++i) sr[i] = to_ref(lat.shared_word[i]);
    for (size_t i = 0; i < tr.size(); ++i) tr[i] = to_ref(lat.tied_word[i]);
    if (consecutive && rc == FX_OK) {
      std::vector<int32_t> orr(orphans.size());
      for (size_t i = 0; i < orr.size(); ++i) orr[i] = lat.perm[orphans[i]];
      ok(dev_upload(ctx, b->words_ref, wr.data(), wr.size(), by)) &&
          ok(sr.empty() ? FX_OK : dev_upload(ctx, b->shared_word_ref, sr.data(), sr.size(), by)) &&
          ok(tr.empty() ? FX_OK : dev_upload(ctx, b->tied_word_ref, tr.data(), tr.size(), by)) &&
          ok(orr.empty() ? FX_OK : dev_upload(ctx, b->orphans_ref, orr.data(), orr.size(), by));
    }
  } else {
    std::vector<float> c32(d->coords, d->coords + static_cast<size_t>(3) * N);
    ok(dev_upload(ctx, b->node, d->node, static_cast<size_t>(8) * P, by)) &&
        ok(dev_upload(ctx, b->param, d->param, static_cast<size_t>(P), by)) &&
        ok(dev_upload(ctx, b->node_eq, node_eq.data(), node_eq.size(), by)) &&
        ok(dev_upload(ctx, b->coords, d->coords, static_cast<size_t>(3) * N, by)) &&
        ok(b->has_f32 ? dev_upload(ctx, b->coords32, c32.data(), c32.size(), by) : FX_OK) &&
        ok(dev_upload(ctx, b->blk.elems, blk.elems.data(), blk.elems.size(), by)) &&
        ok(dev_upload(ctx, b->blk.ln_start, blk.ln_start.data(), blk.ln_start.size(), by)) &&
        ok(dev_upload(ctx, b->blk.lnode, blk.lnode.data(), blk.lnode.size(), by)) &&
        ok(dev_upload(ctx, b->blk.ln_part, blk.ln_part.data(), blk.ln_part.size(), by)) &&
        ok(dev_upload(ctx, b->blk.ln_slots, blk.ln_slots.data(), blk.ln_slots.size(), by)) &&
        ok(dev_upload(ctx, b->blk.slot_list, blk.slot_list.data(), blk.slot_list.size(), by)) &&
        ok(blk.sh_node.empty() ? FX_OK : dev_upload(ctx, b->blk.sh_node, blk.sh_node.data(), blk.sh_node.size(), by)) &&
        ok(dev_upload(ctx, b->blk.sh_start, blk.sh_start.data(), blk.sh_start.size(), by)) &&
        ok(b->ssp ? dev_upload(ctx, b->blk.pos, blk.pos.data(), blk.pos.size(), by) : FX_OK);
    // the blocks' local nodes in block order: their equations and coordinates (loaded once per
    // block and node, coalesced), and each element corner's local node
    {
      const int64_t tot = static_cast<int64_t>(blk.lnode.size());
      std::vector<int32_t> leq(static_cast<size_t>(3 * tot));
      std::vector<double> lxyz(static_cast<size_t>(3 * tot));
      for (int64_t ln = 0; ln < tot; ++ln)
        for (int dd = 0; dd < 3; ++dd) {
          leq[dd * tot + ln] = node_eq[3 * static_cast<size_t>(blk.lnode[ln]) + dd];
          lxyz[dd * tot + ln] = d->coords[static_cast<size_t>(dd) * N + blk.lnode[ln]];
        }
      std::vector<float> lxyz32(b->has_f32 ? lxyz.size() : 0);
      for (size_t i = 0; i < lxyz32.size(); ++i) lxyz32[i] = static_cast<float>(lxyz[i]);
      b->blk.ln_total = tot;
      ok(dev_upload(ctx, b->blk.lcorner, blk.lcorner.data(), blk.lcorner.size(), by)) &&
          ok(dev_upload(ctx, b->blk.ln_eq, leq.data(), leq.size(), by)) &&
          ok(dev_upload(ctx, b->blk.ln_xyz, lxyz.data(), lxyz.size(), by)) &&
          ok(b->has_f32 ? dev_upload(ctx, b->blk.ln_xyz32, lxyz32.data(), lxyz32.size(), by) : FX_OK);
    }
    // the block-boundary nodes' partial sums (three values per (node, block))
    const size_t npart = static_cast<size_t>(std::max(1, blk.sh_start.back())) * 3;
    ok(dev_zeros(ctx, b->partial64, npart, by)) && ok(b->has_f32 ? dev_zeros(ctx, b->partial32, npart, by) : FX_OK);
    b->blk.nshared = static_cast<int32_t>(blk.sh_node.size());
    if (rc == FX_OK && b->tied) {
      ok(dev_zeros(ctx, b->ynode, static_cast<size_t>(3) * N, by)) &&
          ok(b->has_f32 ? dev_zeros(ctx, b->ynode32, static_cast<size_t>(3) * N, by) : FX_OK) &&
          ok(dev_upload(ctx, b->con_start, con.start.data(), con.start.size(), by)) &&
          ok(dev_upload(ctx, b->con_list, con.list.data(), con.list.size(), by));
    }
  }
  // SSPbrick constants: per parameter set on a lattice (the modal stabilisation of its box and
  // diag(K_e)); per element on the general path and for the per-element forms of fx_run_operator
  // (SSPbrick::point_store_geometry / element_constants, ssp_brick.hpp:890-914, kept as the
  // reference's geometry cache and constants arrays, element_kernels.hpp:47)
  if (rc == FX_OK && b->ssp) {
    if (use_lattice) {
      std::vector<double> box(static_cast<size_t>(d->nparam) * fx::kSspBox);
      for (int32_t i = 0; i < d->nparam; ++i)
        fx::ssp_box_constants(lat.hx, lat.hy, lat.hz, b->host_params[i].lam, b->host_params[i].mu,
                              box.data() + static_cast<size_t>(i) * fx::kSspBox);
      ok(dev_upload(ctx, b->ssp_box, box.data(), box.size(), by));
    }
    if (!use_lattice || (flags & FX_BATCH_SLOTS)) {
      // the general path reads them in block order (position b * kBlockElems + i), the lattice
      // path's per-element forms in element order
      const int64_t stride = use_lattice ? P : static_cast<int64_t>(blk.nblocks) * fx::kBlockElems;
      std::vector<double> cst(static_cast<size_t>(fx::kSspConsts) * std::max<int64_t>(1, stride), 0.0);
      const int nth = static_cast<int>(std::min<int64_t>(16, std::max<int64_t>(1, E / 4096)));
      std::vector<std::thread> pool;
      for (int t = 0; t < nth; ++t)
        pool.emplace_back([&, t] {
          for (int32_t e = t; e < E; e += nth) {
            double xe[8][3];
            for (int a = 0; a < 8; ++a) {
              const int32_t nd = d->node[static_cast<size_t>(a) * P + e];
              for (int k = 0; k < 3; ++k) xe[a][k] = d->coords[static_cast<size_t>(k) * N + nd];
            }
            const fx::DevParam& q = b->host_params[d->param[e]];
            fx::ssp_constants(xe, q.lam, q.mu, cst.data() + (use_lattice ? e : blk.pos[e]), stride);
          }
        });
      for (auto& th : pool) th.join();
      ok(dev_upload(ctx, b->ssp_elem, cst.data(), cst.size(), by));
    }
  }
  // lumped mass, lumped alphaM mass and body load (batches.cpp:204-224; std_brick.hpp:519-542,
  // 609-623): rho, alphaM rho and b times the node weights sum_gp N_j dvol, in element order.
  // The congruent boxes of a lattice batch share one set of weights.
  if (rc == FX_OK) {
    const int32_t n = d->neq;
    std::vector<double> lm(n, 0.0), la(n, 0.0), fb(n, 0.0);
    bool any_am = false, any_body = false;
    double w[8];
    auto element_xe = [&](int32_t e, double (&xe)[8][3]) {
      for (int a = 0; a < 8; ++a) {
        const int32_t nd = d->node[static_cast<size_t>(a) * P + e];
        for (int k = 0; k < 3; ++k) xe[a][k] = d->coords[static_cast<size_t>(k) * N + nd];
      }
    };
    // a tied lattice: the unit stiffness parts and consistent mass of its (congruent) box, for
    // the diagonal's cross terms of DOFs that one element ties together
    double Kl[24][24], Km[24][24];
    double Mn[8][8] = {};
    const bool tied_lattice = use_lattice && !lat.tied_word.empty();
    std::map<std::pair<int32_t, int32_t>, std::array<double, 3>> cross;  // (equation, set) -> sums
    if (use_lattice) {
      const double box[8][3] = {{0, 0, 0}, {lat.hx, 0, 0}, {lat.hx, lat.hy, 0}, {0, lat.hy, 0},
                                {0, 0, lat.hz}, {lat.hx, 0, lat.hz}, {lat.hx, lat.hy, lat.hz}, {0, lat.hy, lat.hz}};
      fx::brick_node_weights(box, w);
      if (tied_lattice) {
        fx::brick_stiffness(box, 1.0, 0.0, Kl);
        fx::brick_stiffness(box, 0.0, 1.0, Km);
        fx::brick_mass_unit(box, Mn);
      }
    }
    std::set<std::array<double, 4>> rect_boxes;  // (hx, hy, hz, set) of a rectilinear lattice's cells
    for (int32_t e = 0; e < E; ++e) {
      if (!use_lattice || lat.rect) {
        double xe[8][3];
        element_xe(e, xe);
        fx::brick_node_weights(xe, w);
        if (lat.rect) {
          const double hx = xe[1][0] - xe[0][0], hy = xe[3][1] - xe[0][1], hz = xe[4][2] - xe[0][2];
          rect_boxes.insert({hx, hy, hz, static_cast<double>(d->param[e])});
          if (tied_lattice) {  // this cell's own unit stiffness parts and mass (when it ties DOFs)
            bool ties = false;
            for (int a = 0; a < 24 && !ties; ++a) {
              const int32_t q = d->eq[static_cast<size_t>(a) * P + e];
              for (int c = a + 1; c < 24 && q >= 0; ++c) ties |= d->eq[static_cast<size_t>(c) * P + e] == q;
            }
            if (ties) {
              const double box[8][3] = {{0, 0, 0}, {hx, 0, 0}, {hx, hy, 0}, {0, hy, 0},
                                        {0, 0, hz}, {hx, 0, hz}, {hx, hy, hz}, {0, hy, hz}};
              fx::brick_stiffness(box, 1.0, 0.0, Kl);
              fx::brick_stiffness(box, 0.0, 1.0, Km);
              fx::brick_mass_unit(box, Mn);
            }
          }
        }
      }
      const fx::DevParam& p = b->host_params[d->param[e]];
      if (tied_lattice)
        for (int a = 0; a < 24; ++a) {
          const int32_t q = d->eq[static_cast<size_t>(a) * P + e];
          if (q < 0) continue;
          for (int c = 0; c < 24; ++c) {
            if (c == a || d->eq[static_cast<size_t>(c) * P + e] != q) continue;
            auto& s = cross[{q, d->param[e]}];
            s[0] += Kl[a][c];
            s[1] += Km[a][c];
            if (a % 3 == c % 3) s[2] += Mn[a / 3][c / 3];
          }
        }
      for (int k = 0; k < 24; ++k) {
        const int32_t q = d->eq[static_cast<size_t>(k) * P + e];
        if (q < 0) continue;
        const double mj = p.rho * w[k / 3];
        lm[q] += mj;
        la[q] += p.alphaM * mj;
        fb[q] += p.b[k % 3] * w[k / 3];
        any_am |= p.alphaM != 0.0;
        any_body |= p.b[k % 3] != 0.0;
      }
    }
    for (const auto& bx : rect_boxes) b->rect_boxes.push_back(bx);
    if (use_lattice) {
      b->internal_of.assign(n, -1);
      for (int32_t i = 0; i < n; ++i) b->internal_of[lat.perm[i]] = i;
      auto to_int = [&](std::vector<double>& v) {
        std::vector<double> t(n);
        for (int32_t i = 0; i < n; ++i) t[i] = v[lat.perm[i]];
        v.swap(t);
      };
      to_int(lm);
      to_int(la);
      to_int(fb);
      for (const auto& kv : cross) {
        b->tie_q.push_back(b->internal_of[kv.first.first]);
        b->tie_p.push_back(kv.first.second);
        b->tie_kl.push_back(kv.second[0]);
        b->tie_km.push_back(kv.second[1]);
        b->tie_mc.push_back(kv.second[2]);
        if (b->tie_eqs.empty() || b->tie_eqs.back() != b->tie_q.back()) b->tie_eqs.push_back(b->tie_q.back());
      }
      if (!b->tie_eqs.empty()) {
        b->tie_host.assign(b->tie_eqs.size(), 0.0);
        ok(dev_upload(ctx, b->tie_idx, b->tie_eqs.data(), b->tie_eqs.size(), by)) &&
            ok(dev_zeros(ctx, b->tie_val, b->tie_eqs.size(), by));
      }
    } else {
      b->host_node.assign(d->node, d->node + static_cast<size_t>(8) * P);
      b->host_param.assign(d->param, d->param + P);
      b->host_coords.assign(d->coords, d->coords + static_cast<size_t>(3) * N);
    }
    b->host_lm = lm;
    b->host_la = la;
    ok(n > 0 ? dev_upload(ctx, b->lumped_m, lm.data(), lm.size(), by) : FX_OK) &&
        ok(any_am ? dev_upload(ctx, b->lumped_am, la.data(), la.size(), by) : FX_OK) &&
        ok(any_body ? dev_upload(ctx, b->body, fb.data(), fb.size(), by) : FX_OK);
  }
  if (rc == FX_OK) {
    b->eq_dir.assign(static_cast<size_t>(d->neq), -1);
    for (int32_t e = 0; e < E; ++e)
      for (int k = 0; k < 24; ++k) {
        const int32_t q = d->eq[static_cast<size_t>(k) * P + e];
        if (q >= 0) b->eq_dir[q] = static_cast<int8_t>(k % 3);
      }
    {
      std::vector<int8_t> dir(static_cast<size_t>(std::max(1, d->neq)), -1);
      for (int32_t q = 0; q < d->neq; ++q) dir[b->internal ? b->internal_of[q] : q] = b->eq_dir[q];
      ok(dev_upload(ctx, b->eq_dir_dev, dir.data(), dir.size(), by));
    }
    if (flags & FX_BATCH_DYNAMIC) ok(dev_zeros(ctx, b->dyn, static_cast<size_t>(13) * std::max(1, d->neq), by));
    if (flags & FX_BATCH_GMRES)
      ok(dev_zeros(ctx, b->gm,
                   fx::gmres_workspace(d->neq), by));
  }
  if (rc == FX_OK && (flags & FX_BATCH_SLOTS)) {
    // fx_run_operator's per-element forms: the connectivity the lattice path does not keep, an
    // element slot buffer and its gather (reference numbering)
    if (use_lattice)
      ok(dev_upload(ctx, b->node, d->node, static_cast<size_t>(8) * P, by)) &&
          ok(dev_upload(ctx, b->param, d->param, static_cast<size_t>(P), by)) &&
          ok(dev_upload(ctx, b->node_eq, node_eq.data(), node_eq.size(), by)) &&
          ok(dev_upload(ctx, b->coords, d->coords, static_cast<size_t>(3) * N, by));
    std::vector<int32_t> start, list, long_eqs;
    build_gather(E, P, d->eq, d->neq, start, list, long_eqs);
    ok(dev_zeros(ctx, b->slots, static_cast<size_t>(24) * std::max(1, P), by)) &&
        ok(dev_upload(ctx, b->gather_start, start.data(), start.size(), by)) &&
        ok(list.empty() ? FX_OK : dev_upload(ctx, b->gather_list, list.data(), list.size(), by)) &&
        ok(long_eqs.empty() ? FX_OK : dev_upload(ctx, b->long_eqs, long_eqs.data(), long_eqs.size(), by));
    b->nlong = static_cast<int32_t>(long_eqs.size());
  }
  if (rc != FX_OK) {
    fx_batch_destroy(b);
    return rc;
  }
  b->blk.nblocks = blk.nblocks;
  b->blk.max_ln = blk.max_ln;
  b->info.count = E;
  b->info.neq = d->neq;
  b->info.nnode = N;
  b->info.nblocks = blk.nblocks;
  b->info.ncolors = blk.nblocks > 0 ? 1 : 0;  // one launch of every block
  b->info.tied = b->tied ? 1 : 0;
  b->info.lattice = b->lattice ? 1 : 0;
  if (b->lattice && !b->ssp && !b->rect && !(flags & FX_BATCH_NO_TUNE) && E >= kTuneMinElements) {
    // the variant autotune, kept per device and model (with this tile depth) in the cache
    const std::string key = fx::tune_machine_key(ctx->device) + "|" + lattice_model_key(d, b->tile_z);
    int depth = 0, variant = -1;
    if (fx::tune_cache_get(key, depth, variant) && depth == b->tile_z) {
      fx_batch_tune(b, variant);
    } else {
      const int rc2 = fx_batch_tune(b, -1);
      if (rc2 != FX_OK) {
        fx_batch_destroy(b);
        return rc2;
      }
      fx::tune_cache_put(key, b->tile_z, b->box_variant + 16 * (b->box_variant_mass + 1));
    }
  }
  b->info.variant = b->box_variant;
  b->info.tiles = static_cast<int32_t>(lat.tiles.size());
  b->info.device_bytes = by;
  *out = b;
  return FX_OK;
}

int fx_batch_tune(fx_batch* b, int variant) {
  if (!b) return fail(nullptr, FX_EINVAL, "fx_batch_tune: null batch");
  if (variant >= 0) {
    if (variant > 16 * 16 + 15) return fail(b->ctx, FX_EINVAL, "fx_batch_tune: variant out of range");
    b->box_variant = variant & 15;
    b->box_variant_mass = variant > 15 ? (variant >> 4) - 1 : variant;
    b->info.variant = b->box_variant;
    return FX_OK;
  }
  if (!b->lattice || b->ssp || b->rect) return FX_OK;  // one variant only
  // time each variant on scratch vectors (the reference's force-settings autotune,
  // src/exec/force_settings.cpp:135-155): three rounds over the variants, each sample a warm-up
  // apply then ten back-to-back applies between two events (one apply is a few tens of
  // microseconds: single-apply samples picked a different variant from run to run), and the
  // best sample per variant kept. Both operator forms are timed, the stiffness alone and with
  // damping and mass terms (another kernel), and each keeps its fastest variant.
  fx_ctx* ctx = b->ctx;
  const fx_coeffs forms[2] = {{1.0, 0.0, 0.0, 0}, {1.0, 1.0, 1.0, 0}};
  double* x = b->work;
  double* y = b->work + b->neq;
  {
    // a nonzero input: the apply draws more power on real data than on zeros (about 5% slower on
    // the C2 box), and the variants are ranked as they will run
    std::vector<double> pattern(static_cast<size_t>(std::max(1, b->neq)));
    for (size_t i = 0; i < pattern.size(); ++i) pattern[i] = std::sin(0.001 * static_cast<double>(i));
    FX_HIP(ctx, hipMemcpyAsync(x, pattern.data(), sizeof(double) * pattern.size(), hipMemcpyHostToDevice, ctx->stream));
    FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  struct Events {  // destroyed on every return
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } ev;
  FX_HIP(ctx, hipEventCreate(&ev.e0));
  FX_HIP(ctx, hipEventCreate(&ev.e1));
  const hipEvent_t e0 = ev.e0, e1 = ev.e1;
  // 16-deep tiles: the four register budgets of the streamed kernel; 8-deep: held and streamed
  // tiles; 4-deep: held tiles only
  const int nvar = b->tile_z == 16 ? 4 : b->tile_z == 8 ? 12 : 8;
  constexpr int kSampleApplies = 10;
  {
    // the GPU's clocks rise over the first tens of milliseconds of work: applies until 30 ms have
    // run (at most 2000) before any sample, so that the first variants are not timed slow
    float warm = 0.0f;
    for (int n = 0; n < 2000 && warm < 30.0f;) {
      FX_HIP(ctx, hipEventRecord(e0, ctx->stream));
      for (int r = 0; r < 20; ++r, ++n)
        if (int rc = fx::operator_apply64(b, forms[0], x, y)) return rc;
      FX_HIP(ctx, hipEventRecord(e1, ctx->stream));
      FX_HIP(ctx, hipEventSynchronize(e1));
      float ms = 0;
      FX_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
      warm += ms;
    }
  }
  std::vector<float> vbest(static_cast<size_t>(nvar), 1e30f);
  std::vector<float> mbest(static_cast<size_t>(nvar), 1e30f);
  for (int round = 0; round < 3; ++round)
    for (int v = 0; v < nvar; ++v)
      for (int f = 0; f < 2; ++f) {
        const fx_coeffs& k = forms[f];
        b->box_variant = v;
        if (int rc = fx::operator_apply64(b, k, x, y)) return rc;
        FX_HIP(ctx, hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < kSampleApplies; ++r)
          if (int rc = fx::operator_apply64(b, k, x, y)) return rc;
        FX_HIP(ctx, hipEventRecord(e1, ctx->stream));
        FX_HIP(ctx, hipEventSynchronize(e1));
        float ms = 0;
        FX_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        float& best = f == 0 ? vbest[v] : mbest[v];
        best = std::min(best, ms);
      }
  int best_v = 0, best_m = 0;
  for (int v = 1; v < nvar; ++v) {
    if (vbest[v] < vbest[best_v]) best_v = v;
    if (mbest[v] < mbest[best_m]) best_m = v;
  }
  b->box_variant = best_v;
  b->box_variant_mass = best_m;
  b->info.variant = best_v;
  b->tuned_ms = vbest[best_v] + mbest[best_m];
  return FX_OK;
}

int fx_batch_info_get(const fx_batch* b, fx_batch_info* info) {
  if (!b || !info) return fail(nullptr, FX_EINVAL, "fx_batch_info_get: null argument");
  *info = b->info;
  info->device_bytes = b->device_bytes;
  info->tangent_symmetry = b->tangent_symmetry;
  return FX_OK;
}

int fx_operator_apply(fx_batch* b, const fx_coeffs* k, const double* x, double* y) {
  if (!b || !k || !x || !y) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_operator_apply: null argument");
  if (x == y) return fail(b->ctx, FX_EINVAL, "fx_operator_apply: x and y alias");
  return fx::external_apply64(b, *k, x, y);
}

int fx_operator_apply_f32(fx_batch* b, const fx_coeffs* k, const float* x, float* y) {
  if (!b || !k || !x || !y) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_operator_apply_f32: null argument");
  if (!b->has_f32) return fail(b->ctx, FX_EUNSUP, "fx_operator_apply_f32: batch created without FX_BATCH_FP32");
  if (static_cast<const void*>(x) == static_cast<const void*>(y))
    return fail(b->ctx, FX_EINVAL, "fx_operator_apply_f32: x and y alias");
  return fx::external_apply32(b, *k, x, y);
}

int fx_operator_diagonal(fx_batch* b, const fx_coeffs* k, double* d) {
  if (!b || !k || !d) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_operator_diagonal: null argument");
  return fx::external_diagonal64(b, *k, d);
}

int fx_run_operator(fx_batch* b, int diagonal, const fx_coeffs* k, int out, const double* x, double* y, double* slots,
                    int64_t slot_offset, int unit_dir) {
  if (!b || !k) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_run_operator: null argument");
  fx_ctx* ctx = b->ctx;
  if (out != FX_OUT_PLAIN && out != FX_OUT_ATOMIC && out != FX_OUT_SLOTS)
    return fail(ctx, FX_EINVAL, "fx_run_operator: out must be FX_OUT_PLAIN, FX_OUT_ATOMIC or FX_OUT_SLOTS");
  if (unit_dir < -1 || unit_dir > 2) return fail(ctx, FX_EINVAL, "fx_run_operator: unit_dir outside -1..2");
  if (diagonal != 0 && diagonal != 1) return fail(ctx, FX_EINVAL, "fx_run_operator: diagonal must be 0 or 1");
  if (out == FX_OUT_SLOTS && (!slots || slot_offset < 0))
    return fail(ctx, FX_EINVAL, "fx_run_operator: FX_OUT_SLOTS needs slots and slot_offset >= 0");
  if (out != FX_OUT_SLOTS && !y) return fail(ctx, FX_EINVAL, "fx_run_operator: null y");
  if (!diagonal && unit_dir < 0 && !x) return fail(ctx, FX_EINVAL, "fx_run_operator: null x");
  if (!diagonal && unit_dir < 0 && out != FX_OUT_SLOTS && x == y)
    return fail(ctx, FX_EINVAL, "fx_run_operator: x and y alias");
  FX_HIP(ctx, hipSetDevice(ctx->device));
  const bool per_element = out == FX_OUT_SLOTS || unit_dir >= 0 || b->pimy;
  if (per_element && !b->slots)
    return fail(ctx, FX_EINVAL,
                "fx_run_operator: FX_OUT_SLOTS and unit_dir >= 0 need the per-element forms (create the batch with "
                "FX_BATCH_SLOTS)");
  int rc = FX_OK;
  if (per_element) {
    double* dst = out == FX_OUT_SLOTS ? slots + slot_offset : b->slots;
    rc = b->pimy ? fx::pimy_slots64(b, diagonal != 0, *k, x, unit_dir, dst)
                 : fx::element_slots64(b, diagonal != 0, *k, x, unit_dir, dst);
    if (!rc && out != FX_OUT_SLOTS) rc = fx::slots_gather64(b, b->slots, y, 1);
  } else {
    rc = fx::add_elements64(b, diagonal != 0, *k, x, y);
  }
  if (rc) return rc;
  if (!ctx->stream_ordered) FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

int fx_batch_set_nodal_mass(fx_batch* b, const double* mass, const double* alpha_mass) {
  if (!b) return fail(nullptr, FX_EINVAL, "fx_batch_set_nodal_mass: null batch");
  fx_ctx* ctx = b->ctx;
  FX_HIP(ctx, hipSetDevice(ctx->device));
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int32_t n = b->neq;
  const size_t N = static_cast<size_t>(n);
  uint64_t& by = b->device_bytes;
  for (double** p : {&b->nodal_m, &b->nodal_am}) dev_free(ctx, *p, by, sizeof(double) * N), *p = nullptr;
  for (float** p : {&b->nodal_m32, &b->nodal_am32}) dev_free(ctx, *p, by, sizeof(float) * N), *p = nullptr;
  // the solver numbering of the batch (internal_of maps a reference equation to it)
  auto to_solver = [&](const double* src, std::vector<double>& dst) {
    dst.assign(N, 0.0);
    for (int32_t q = 0; q < n; ++q) {
      if (!std::isfinite(src[q]) || src[q] < 0.0) return false;
      dst[b->internal ? b->internal_of[q] : q] = src[q];
    }
    return true;
  };
  std::vector<double> m, am;
  if (mass && !to_solver(mass, m)) return fail(ctx, FX_EINVAL, "fx_batch_set_nodal_mass: negative or non-finite mass");
  if (alpha_mass && !to_solver(alpha_mass, am))
    return fail(ctx, FX_EINVAL, "fx_batch_set_nodal_mass: negative or non-finite alphaM mass");
  int rc = FX_OK;
  auto ok = [&](int r) {
    if (r != FX_OK && rc == FX_OK) rc = r;
    return rc == FX_OK;
  };
  auto any = [](const std::vector<double>& v) {
    for (double x : v)
      if (x != 0.0) return true;
    return false;
  };
  const bool has_m = any(m), has_am = any(am);
  if (has_m) {
    ok(dev_upload(ctx, b->nodal_m, m.data(), N, by));
    if (b->has_f32) {
      std::vector<float> m32(m.begin(), m.end());
      ok(dev_upload(ctx, b->nodal_m32, m32.data(), N, by));
    }
  }
  if (has_am) {
    ok(dev_upload(ctx, b->nodal_am, am.data(), N, by));
    if (b->has_f32) {
      std::vector<float> a32(am.begin(), am.end());
      ok(dev_upload(ctx, b->nodal_am32, a32.data(), N, by));
    }
  }
  // the lumped masses: the elements' sums, then the nodal masses (batches.cpp:249-261)
  if (rc == FX_OK && n > 0) {
    std::vector<double> lm = b->host_lm, la = b->host_la;
    bool any_am = false;
    for (size_t q = 0; q < N; ++q) {
      if (has_m) lm[q] += m[q];
      if (has_am) la[q] += am[q];
      any_am |= la[q] != 0.0;
    }
    FX_HIP(ctx, hipMemcpy(b->lumped_m, lm.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    if (any_am && !b->lumped_am) ok(dev_zeros(ctx, b->lumped_am, N, by));
    if (rc == FX_OK && b->lumped_am)
      FX_HIP(ctx, hipMemcpy(b->lumped_am, la.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  }
  fx::release_cg_graphs(ctx, b);  // captured applies would miss the nodal terms
  b->info.device_bytes = by;
  return rc;
}

int fx_mass_times_ground(fx_batch* b, int direction, double* y) {
  if (!b || !y || direction < 0 || direction > 2)
    return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_mass_times_ground: bad argument");
  fx_ctx* ctx = b->ctx;
  FX_HIP(ctx, hipSetDevice(ctx->device));
  // The element mass times a rigid translation is its row sums rho sum_gp N_j dvol (sum_k N_k = 1),
  // so the elements' M r is their lumped mass at the direction's equations; the nodal masses of
  // the direction follow (internal_force.cpp:619-625), as the lumped mass holds them.
  const size_t N = static_cast<size_t>(b->neq);
  std::vector<double> r(N, 0.0);
  for (int32_t q = 0; q < b->neq; ++q)
    if (b->eq_dir[q] == direction) r[b->internal ? b->internal_of[q] : q] = 1.0;
  if (N == 0) return FX_OK;
  double* scr = b->work + 6 * N;
  double* mr = b->work + 7 * N;
  FX_HIP(ctx, hipMemcpyAsync(scr, r.data(), sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
  if (int rc = fx::multiply64(ctx, b->neq, b->lumped_m, scr, mr)) return rc;
  if (int rc = fx::to_external<double>(b, mr, y)) return rc;
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

int fx_operator_time(fx_batch* b, const fx_coeffs* k, int precision, const void* x, void* y, int reps, float* ms) {
  if (!b || !k || !x || !y || !ms || reps <= 0) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_operator_time: bad argument");
  if (precision == 32 && !b->has_f32) return fail(b->ctx, FX_EUNSUP, "fx_operator_time: no FP32 operator");
  fx_ctx* ctx = b->ctx;
  hipEvent_t e0, e1;
  FX_HIP(ctx, hipEventCreate(&e0));
  FX_HIP(ctx, hipEventCreate(&e1));
  FX_HIP(ctx, hipEventRecord(e0, ctx->stream));
  for (int i = 0; i < reps; ++i) {
    // the operator as the solvers run it (internal numbering of lattice batches)
    const int rc = precision == 32
                       ? fx::operator_apply32(b, *k, static_cast<const float*>(x), static_cast<float*>(y))
                       : fx::operator_apply64(b, *k, static_cast<const double*>(x), static_cast<double*>(y));
    if (rc) return rc;
  }
  FX_HIP(ctx, hipEventRecord(e1, ctx->stream));
  FX_HIP(ctx, hipEventSynchronize(e1));
  FX_HIP(ctx, hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return FX_OK;
}

int fx_copy(fx_ctx* ctx, int64_t n, const double* src, double* dst) {
  if (!ctx || (n > 0 && (!src || !dst))) return fail(ctx, FX_EINVAL, "fx_copy: null argument");
  return fx::copy64(ctx, n, src, dst);
}

int fx_dot(fx_ctx* ctx, int64_t n, const double* x, const double* y, double* out) {
  if (!ctx || !out || (n > 0 && (!x || !y))) return fail(ctx, FX_EINVAL, "fx_dot: null argument");
  return fx::dot64(ctx, n, x, y, out);
}

int fx_cg(fx_batch* b, const fx_coeffs* k, const fx_solve_options* opt, const double* rhs, double* x,
          fx_solve_result* res) {
  if (!b || !k || !opt || !rhs || !x || !res) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_cg: null argument");
  return fx::cg64(b, *k, *opt, rhs, x, *res);
}

int fx_cg_many(int32_t n, fx_batch* const* batches, const fx_coeffs* k, const fx_solve_options* opt,
               const double* const* rhs, double* const* x, fx_solve_result* res) {
  if (n < 0 || (n > 0 && (!batches || !k || !opt || !rhs || !x || !res)))
    return fail(nullptr, FX_EINVAL, "fx_cg_many: null argument");
  if (n == 0) return FX_OK;
  fx_ctx* ctx = batches[0] ? batches[0]->ctx : nullptr;
  for (int32_t i = 0; i < n; ++i) {
    fx_batch* b = batches[i];
    if (!b || !rhs[i] || !x[i]) return fail(ctx, FX_EINVAL, "fx_cg_many: null batch or vector");
    if (b->ctx != ctx) return fail(ctx, FX_EINVAL, "fx_cg_many: the batches belong to different contexts");
    if (!b->cg_state || !b->cg_red)
      return fail(ctx, FX_EUNSUP, "fx_cg_many: elastic batches (fx_batch_create) only");
    for (int32_t j = 0; j < i; ++j)
      if (batches[j] == b) return fail(ctx, FX_EINVAL, "fx_cg_many: a batch appears twice (one solve per batch)");
  }
  FX_HIP(ctx, hipSetDevice(ctx->device));
  return fx::cg64_many(n, batches, *k, *opt, rhs, x, res);
}

int fx_cg_mixed(fx_batch* b, const fx_coeffs* k, const fx_solve_options* opt, double inner_rtol, const double* rhs,
                double* x, fx_solve_result* res) {
  if (!b || !k || !opt || !rhs || !x || !res)
    return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_cg_mixed: null argument");
  if (!(inner_rtol > 0 && inner_rtol < 1)) return fail(b->ctx, FX_EINVAL, "fx_cg_mixed: inner_rtol outside (0, 1)");
  return fx::cg_mixed(b, *k, *opt, inner_rtol, rhs, x, *res);
}

}  // extern "C"

int fx_lumped_mass(fx_batch* b, double* m, double* am) {
  if (!b || !m) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_lumped_mass: null argument");
  FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
  if (int rc = fx::to_external<double>(b, b->lumped_m, m)) return rc;
  if (am) {
    if (b->lumped_am) {
      if (int rc = fx::to_external<double>(b, b->lumped_am, am)) return rc;
    } else {
      FX_HIP(b->ctx, hipMemsetAsync(am, 0, sizeof(double) * b->neq, b->ctx->stream));
    }
  }
  FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return FX_OK;
}

int fx_body_load(fx_batch* b, double* fb) {
  if (!b || !fb) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_body_load: null argument");
  FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
  if (b->body) {
    if (int rc = fx::to_external<double>(b, b->body, fb)) return rc;
  } else {
    FX_HIP(b->ctx, hipMemsetAsync(fb, 0, sizeof(double) * b->neq, b->ctx->stream));
  }
  FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return FX_OK;
}

int fx_force(fx_batch* b, const double* u, const double* v, double* f) {
  if (!b || !u || !f) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_force: null argument");
  FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
  const size_t n = static_cast<size_t>(b->neq);
  double* W = b->work;
  if (b->internal) {
    if (int rc = fx::to_internal<double>(b, u, W)) return rc;
    if (v)
      if (int rc = fx::to_internal<double>(b, v, W + n)) return rc;
    if (int rc = fx::force64(b, W, v ? W + n : nullptr, W + 2 * n, W + 3 * n)) return rc;
    if (int rc = fx::to_external<double>(b, W + 2 * n, f)) return rc;
  } else {
    if (int rc = fx::force64(b, u, v, f, W)) return rc;
  }
  FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return FX_OK;
}

int fx_critical_dt(fx_batch* b, int scheme, double* dt) {
  if (!b || !dt || (scheme != 0 && scheme != 1)) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_critical_dt: bad argument");
  if (b->pimy && b->ssp)
    return fail(b->ctx, FX_EUNSUP, "fx_critical_dt: no stable step for SSPbrick<PressureIndependMultiYield> batches");
  double best = INFINITY;
  if (b->lattice && b->rect) {
    // a rectilinear lattice: each distinct (cell size, parameter set) once
    for (const auto& bx : b->rect_boxes) {
      const double hx = bx[0], hy = bx[1], hz = bx[2];
      const double box[8][3] = {{0, 0, 0}, {hx, 0, 0}, {hx, hy, 0}, {0, hy, 0}, {0, 0, hz}, {hx, 0, hz}, {hx, hy, hz}, {0, hy, hz}};
      const fx::DevParam& p = b->host_params[static_cast<size_t>(bx[3])];
      best = std::min(best, fx::stable_dt(scheme, p, fx::brick_max_omega_sq(box, p.lam, p.mu, p.rho)));
    }
  } else if (b->lattice) {
    const double hx = b->hx, hy = b->hy, hz = b->hz;
    const double box[8][3] = {{0, 0, 0}, {hx, 0, 0}, {hx, hy, 0}, {0, hy, 0}, {0, 0, hz}, {hx, 0, hz}, {hx, hy, hz}, {0, hy, hz}};
    for (const fx::DevParam& p : b->host_params)
      best = std::min(best, fx::stable_dt(scheme, p,
                                          b->ssp ? fx::ssp_max_omega_sq(box, p.lam, p.mu, p.rho)
                                                 : fx::brick_max_omega_sq(box, p.lam, p.mu, p.rho)));
  } else {
    // PIMY points: the stability tangent of the committed state (stability_moduli,
    // pressure_independ_multi_yield.hpp:174-203): the point's moduli, scaled as paramScaling will
    // scale them while the switch to the plastic stage is pending
    std::vector<double> vol, e2p, G, K;
    if (b->pimy && b->npts > 0) {
      const size_t np = static_cast<size_t>(b->npts);
      vol.resize(np);
      e2p.resize(np);
      G.resize(np);
      K.resize(np);
      const std::pair<double*, int> fields[4] = {{vol.data(), fx::kPStress + 12}, {e2p.data(), fx::kPE2P},
                                                 {G.data(), fx::kPG}, {K.data(), fx::kPK}};
      FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
      for (const auto& fd : fields)
        FX_HIP(b->ctx, hipMemcpyAsync(fd.first, b->state_c + static_cast<size_t>(fd.second) * np, sizeof(double) * np,
                                      hipMemcpyDeviceToHost, b->ctx->stream));
      FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    }
    const int32_t P = b->padded, N = b->nnode;
    for (int32_t e = 0; e < b->count; ++e) {
      double xe[8][3];
      for (int a = 0; a < 8; ++a) {
        const int32_t nd = b->host_node[static_cast<size_t>(a) * P + e];
        for (int k = 0; k < 3; ++k) xe[a][k] = b->host_coords[static_cast<size_t>(k) * N + nd];
      }
      const fx::DevParam& p = b->host_params[b->host_param[e]];
      if (!b->pimy) {
        best = std::min(best, fx::stable_dt(scheme, p,
                                            b->ssp ? fx::ssp_max_omega_sq(xe, p.lam, p.mu, p.rho)
                                                   : fx::brick_max_omega_sq(xe, p.lam, p.mu, p.rho)));
        continue;
      }
      const fx::PimyParams& q = b->host_pimy[b->host_param[e]];
      double lam[8], mu[8];
      for (int gp = 0; gp < 8; ++gp) {
        const size_t i = static_cast<size_t>(e) * 8 + gp;
        double g = G[i], k = K[i];
        if (q.load_stage == 1 && e2p[i] == 0.0 && q.phi != 0.) {
          double v = vol[i];
          if (v > 0.) v = 0.;  // elast2Plast drops the tension first
          const double con = -(v - q.residual_press);
          const double scale = std::pow(-con / (q.pref - q.residual_press), q.pdc);
          g *= scale;
          k *= scale;
        }
        lam[gp] = k - 2. * g / 3.;
        mu[gp] = g;
      }
      best = std::min(best, fx::stable_dt(scheme, p, fx::brick_max_omega_sq_points(xe, lam, mu, q.rho)));
    }
  }
  *dt = best;
  return FX_OK;
}

int fx_explicit_cd(fx_batch* b, const fx_cd_desc* d, double* u, double* um, double* v, double* a) {
  if (!b || !d || !u || !um || !v || !a) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_explicit_cd: null argument");
  if (!(d->dt > 0) || d->nsteps < 0 || d->nload < 0 || d->nload > FX_MAX_POINT_LOADS ||
      (d->nload > 0 && (!d->load_eq || !d->load_hist)) || (d->scheme != 0 && d->scheme != 1) || d->ground_dir < -1 ||
      d->ground_dir > 2 || (d->ground_dir >= 0 && !d->ground_accel && !d->ground_series))
    return fail(b->ctx, FX_EINVAL, "fx_explicit_cd: bad descriptor");
  if (const fx_time_series* gs = d->ground_series) {
    if (!gs->values || gs->count < 1 || (!gs->time && !(gs->dt > 0)))
      return fail(b->ctx, FX_EINVAL, "fx_explicit_cd: bad ground_series");
    if (gs->time) {
      for (int32_t i = 0; i < gs->count; ++i)
        if (!std::isfinite(gs->time[i]) || (i > 0 && gs->time[i] < gs->time[i - 1]))
          return fail(b->ctx, FX_EINVAL, "fx_explicit_cd: ground_series times must be finite and nondecreasing");
      if (gs->loc && (*gs->loc < 0 || *gs->loc >= gs->count))
        return fail(b->ctx, FX_EINVAL, "fx_explicit_cd: ground_series interval (*loc) out of range");
    }
  }
  if (d->scheme == 1 && b->neq > 0) {  // check_masses: a = R / m needs a mass at every equation
    std::vector<double> m(static_cast<size_t>(b->neq));
    FX_HIP(b->ctx, hipMemcpy(m.data(), b->lumped_m, sizeof(double) * m.size(), hipMemcpyDeviceToHost));
    for (double x : m)
      if (!(x > 0.0))
        return fail(b->ctx, FX_EINVAL, "fx_explicit_cd: ExplicitDifference needs a positive lumped mass at every equation");
  }
  for (int j = 0; j < d->nload; ++j)
    if (d->load_eq[j] < 0 || d->load_eq[j] >= b->neq) return fail(b->ctx, FX_EINVAL, "fx_explicit_cd: load equation out of range");
  FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
  return fx::cd_run(b, *d, u, um, v, a);
}

int fx_pimy_surfaces(const fx_pimy_material* m, double* size, double* modulus) {
  if (!m || !size || !modulus) return fail(nullptr, FX_EINVAL, "fx_pimy_surfaces: null argument");
  fx::PimyParams p;
  std::string err;
  if (int rc = fx::pimy_setup(*m, p, err)) return fail(nullptr, rc, err);
  for (int i = 0; i <= fx::kPimyNS; ++i) {
    size[i] = p.size[i];
    modulus[i] = p.modulus[i];
  }
  return FX_OK;
}

namespace {
int upload_point_materials(fx_ctx* ctx, int32_t nmat, const fx_pimy_material* mats, const int32_t* load_stage) {
  if (nmat <= 0 || nmat > fx::kMaxPointMaterials || !mats)
    return fail(ctx, FX_EINVAL, "fx_pimy: 1 to 8 materials");
  fx::PimyParams p[fx::kMaxPointMaterials];
  for (int32_t i = 0; i < nmat; ++i) {
    std::string err;
    if (int rc = fx::pimy_setup(mats[i], p[i], err)) return fail(ctx, rc, err);
    p[i].load_stage = load_stage ? load_stage[i] : 0;
  }
  FX_HIP(ctx, hipMemcpyAsync(ctx->pimy_params, p, sizeof(fx::PimyParams) * nmat, hipMemcpyHostToDevice, ctx->stream));
  return FX_OK;
}
}  // namespace

int fx_pimy_init_state(fx_ctx* ctx, int32_t nmat, const fx_pimy_material* mats, int64_t npts, const int32_t* mat,
                       double* state) {
  if (!ctx || (npts > 0 && (!mat || !state))) return fail(ctx, FX_EINVAL, "fx_pimy_init_state: null argument");
  FX_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = upload_point_materials(ctx, nmat, mats, nullptr)) return rc;
  if (int rc = fx::pimy_init(ctx, npts, static_cast<const fx::PimyParams*>(ctx->pimy_params), mat, state)) return rc;
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

int fx_pimy_points(fx_ctx* ctx, int32_t nmat, const fx_pimy_material* mats, const int32_t* load_stage, int64_t npts,
                   const int32_t* mat, const double* committed, double* trial, const double* eps, double* sig,
                   double* D) {
  if (!ctx || (npts > 0 && (!mat || !committed || !trial || !eps || !sig)))
    return fail(ctx, FX_EINVAL, "fx_pimy_points: null argument");
  if (committed == trial) return fail(ctx, FX_EINVAL, "fx_pimy_points: committed and trial history alias");
  FX_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = upload_point_materials(ctx, nmat, mats, load_stage)) return rc;
  if (int rc = fx::pimy_points(ctx, npts, static_cast<const fx::PimyParams*>(ctx->pimy_params), mat, committed, trial,
                               eps, sig, D))
    return rc;
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

int fx_pimy_batch_create(fx_ctx* ctx, const fx_pimy_batch_desc* d, uint32_t flags, fx_batch** out) {
  if (!ctx || !d || !out || !d->params || d->nparam <= 0) return fail(ctx, FX_EINVAL, "fx_pimy_batch_create: null argument");
  *out = nullptr;
  if ((flags & FX_BATCH_SSPBRICK) && (flags & (FX_BATCH_COMPACT | FX_BATCH_FP32)))
    return fail(ctx, FX_EUNSUP,
                "fx_pimy_batch_create: SSPbrick<PressureIndependMultiYield> has no compacted return map or FP32 operator");
  std::vector<fx::PimyParams> pp(d->nparam);
  std::vector<fx_elastic_params> ep(d->nparam);
  for (int32_t i = 0; i < d->nparam; ++i) {
    std::string err;
    if (int rc = fx::pimy_setup(d->params[i], pp[i], err)) return fail(ctx, rc, err + " (set " + std::to_string(i) + ")");
    // the elastic moduli of the reference G, K carry the lumped mass, body load and the
    // general-path geometry the batch shares with elastic batches
    const double G = d->params[i].G, K = d->params[i].K;
    fx_elastic_params& e = ep[i];
    std::memset(&e, 0, sizeof e);
    e.E = 9.0 * K * G / (3.0 * K + G);
    e.nu = (3.0 * K - 2.0 * G) / (2.0 * (3.0 * K + G));
    e.rho = pp[i].rho;
    for (int k = 0; k < 3; ++k) e.b[k] = d->params[i].b[k];
    e.mass_type = d->params[i].mass_type;
    if (d->damp) {  // the lumped alphaM mass and the stable step see the damping too
      e.damp = d->damp[i];
      pp[i].alphaM = d->damp[i].alphaM;
      pp[i].betaK = d->damp[i].betaK;
      pp[i].betaKinit = d->damp[i].betaKinit;
      pp[i].betaKcomm = d->damp[i].betaKcomm;
    }
  }
  fx_batch_desc ed{};
  ed.count = d->count;
  ed.padded = d->padded;
  ed.node = d->node;
  ed.eq = d->eq;
  ed.param = d->param;
  ed.nparam = d->nparam;
  ed.params = ep.data();
  ed.nnode = d->nnode;
  ed.coords = d->coords;
  ed.neq = d->neq;
  fx_batch* b = nullptr;
  if (int rc = fx_batch_create(ctx, &ed,
                               (flags & ~(FX_BATCH_FP32 | FX_BATCH_SLOTS)) | FX_BATCH_GENERAL | FX_BATCH_NO_TUNE |
                                   FX_BATCH_DYNAMIC,
                               &b))
    return rc;
  b->pimy = true;
  b->tangent_symmetry = FX_TANGENT_SPSD;
  b->host_pimy = pp;
  for (const auto& q : pp)
    b->pimy_damped |= q.alphaM != 0.0 || q.betaK != 0.0 || q.betaKinit != 0.0 || q.betaKcomm != 0.0;
  b->npts = static_cast<int64_t>(d->count) * 8;
  // the gather: each equation's slots k * padded + e in (element, DOF) order
  const int32_t P = d->padded;
  std::vector<int32_t> start, list, long_eqs;  // long equations: gathered one wave each
  build_gather(d->count, P, d->eq, d->neq, start, list, long_eqs);
  uint64_t& by = b->device_bytes;
  int rc = FX_OK;
  auto ok = [&](int r) {
    if (r != FX_OK && rc == FX_OK) rc = r;
    return rc == FX_OK;
  };
  const size_t np = static_cast<size_t>(std::max<int64_t>(b->npts, 1));
  ok(dev_upload(ctx, b->pimy_params, pp.data(), pp.size(), by)) &&
      ok(dev_zeros(ctx, b->state_c, np * fx::kPimyNState, by)) && ok(dev_zeros(ctx, b->state_t, np * fx::kPimyNState, by)) &&
      ok(dev_zeros(ctx, b->pimy_touched, np, by)) &&
      ok(dev_zeros(ctx, b->tangent, np * 21, by)) &&
      // FX_BATCH_FP32: the FP32 operator (an FP32 copy of the tangents) and the mixed CG's FP32 vectors
      ok((flags & FX_BATCH_FP32) ? dev_zeros(ctx, b->tangent32, np * 21, by) : FX_OK) &&
      ok((flags & FX_BATCH_FP32) ? dev_zeros(ctx, b->work32, static_cast<size_t>(9) * std::max(1, d->neq), by) : FX_OK) &&
      ok((b->has_f32 = (flags & FX_BATCH_FP32) != 0, FX_OK)) &&
      ok(dev_zeros(ctx, b->slots, static_cast<size_t>(24) * std::max(1, P), by)) &&
      ok(dev_upload(ctx, b->gather_start, start.data(), start.size(), by)) &&
      ok(list.empty() ? FX_OK : dev_upload(ctx, b->gather_list, list.data(), list.size(), by)) &&
      ok(long_eqs.empty() ? FX_OK : dev_upload(ctx, b->long_eqs, long_eqs.data(), long_eqs.size(), by)) &&
      ok((b->nlong = static_cast<int32_t>(long_eqs.size()), FX_OK)) &&
      (!(flags & FX_BATCH_COMPACT) ||
       (ok(dev_zeros(ctx, b->cres, np * 33, by)) && ok(dev_zeros(ctx, b->clist, np, by)) &&
        ok(dev_zeros(ctx, b->ccount, static_cast<size_t>(1), by)))) &&
      ok(fx::pimy_batch_init(b));
  if (rc == FX_OK) {
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, FX_EDEVICE, "fx_pimy_batch_create: init failed");
  }
  if (rc != FX_OK) {
    fx_batch_destroy(b);
    return rc;
  }
  b->info.device_bytes = by;
  *out = b;
  return FX_OK;
}

int fx_set_material_stage(fx_batch* b, int32_t stage) {
  if (!b || (stage != 0 && stage != 1)) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_set_material_stage: bad argument");
  if (!b->pimy) return FX_OK;  // elastic materials have one stage
  for (auto& p : b->host_pimy) p.load_stage = stage;
  FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
  FX_HIP(b->ctx, hipMemcpyAsync(b->pimy_params, b->host_pimy.data(), sizeof(fx::PimyParams) * b->host_pimy.size(),
                                hipMemcpyHostToDevice, b->ctx->stream));
  FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return FX_OK;
}

int fx_store_tangents(fx_batch* b, const double* u, double* f) {
  if (!b || !u) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_store_tangents: null argument");
  if (!b->pimy) return FX_OK;  // elastic tangents are constant
  FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
  if (int rc = fx::pimy_force64(b, u, nullptr, f ? f : b->work, true)) return rc;
  FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return FX_OK;
}

int fx_commit_state(fx_batch* b) {
  if (!b) return fail(nullptr, FX_EINVAL, "fx_commit_state: null batch");
  if (b->pimy) std::swap(b->state_c, b->state_t);
  return FX_OK;
}

int fx_pimy_state(fx_batch* b, int committed, double* dst) {
  if (!b || !dst || !b->pimy) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_pimy_state: not a PIMY batch");
  FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
  FX_HIP(b->ctx, hipMemcpyAsync(dst, committed ? b->state_c : b->state_t, sizeof(double) * fx::kPimyNState * b->npts,
                                hipMemcpyDeviceToDevice, b->ctx->stream));
  FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return FX_OK;
}

int fx_newmark(fx_batch* b, const fx_newmark_desc* d, double* u, double* v, double* a, double* du,
               fx_newmark_result* result) {
  if (!b || !d || !u || !v || !a || !du || !result) return fail(b ? b->ctx : nullptr, FX_EINVAL, "fx_newmark: null argument");
  if (!b->dyn) return fail(b->ctx, FX_EINVAL, "fx_newmark: the batch was created without FX_BATCH_DYNAMIC");
  const bool is_static = d->analysis == 1;
  if ((d->analysis != 0 && d->analysis != 1) || (!is_static && (!(d->dt > 0) || !(d->beta > 0))) || d->nsteps < 0 ||
      d->max_iter <= 0 || d->ground_dir > 2 || d->test < 0 || d->test > 2 ||
      (d->ground_dir >= 0 && !d->ground_accel && !d->ground_series))
    return fail(b->ctx, FX_EINVAL, "fx_newmark: bad descriptor");
  if (d->algorithm < 0 || d->algorithm > 2)
    return fail(b->ctx, FX_EINVAL, "fx_newmark: algorithm must be 0 (Newton), 1 (Linear) or 2 (ModifiedNewton)");
  if (d->algorithm == 1 && b->pimy)  // implicit_stage.cpp:130-140
    return fail(b->ctx, FX_EUNSUP,
                "fx_newmark: the Linear algorithm with a material that has history state (OpenSees would commit a "
                "stress inconsistent with the strain); use Newton or ModifiedNewton");
  if (is_static) {
    const bool adapt = d->load_min_increment != 0.0 && d->load_max_increment != 0.0;
    if (!std::isfinite(d->load_increment) || (adapt && (!(d->load_min_increment <= d->load_max_increment) ||
                                                        d->load_num_iter <= 0)))
      return fail(b->ctx, FX_EINVAL, "fx_newmark: bad LoadControl increment, bounds or numIter");
  }
  {
    const double h = is_static ? 1.0 : d->dt;
    const fx_coeffs k = is_static ? fx_coeffs{1.0, 0.0, 0.0, 0}
                                  : fx_coeffs{1.0, d->gamma / (d->beta * h), 1.0 / (d->beta * h * h), 0};
    if (!fx::operator_spd(b, k)) {
      if (d->mixed)
        return fail(b->ctx, FX_EINVAL, "fx_newmark: mixed-precision CG needs a symmetric positive definite operator");
      if (!b->gm)
        return fail(b->ctx, FX_EINVAL,
                    "fx_newmark: the stage's operator is not symmetric positive definite (a history material in a "
                    "static stage), so its solver is GMRES: create the batch with FX_BATCH_GMRES");
    }
  }
  if (d->max_retries < 0 || d->num_sublevels < 0 || (d->num_sublevels > 0 && d->num_substeps < 1))
    return fail(b->ctx, FX_EINVAL, "fx_newmark: max_retries and num_sublevels must be non-negative, num_substeps positive");
  if (d->ground_dir >= 0 && d->num_sublevels > 0 && !d->ground_series)
    return fail(b->ctx, FX_EINVAL, "fx_newmark: substeps under an excitation need its time series (ground_series)");
  if (d->mixed && (!b->has_f32 || !(d->inner_rtol > 0 && d->inner_rtol < 1)))
    return fail(b->ctx, FX_EINVAL, "fx_newmark: mixed needs a batch created with FX_BATCH_FP32 and inner_rtol in (0, 1)");
  if (const fx_time_series* gs = d->ground_series) {
    if (!gs->values || gs->count < 1 || (!gs->time && !(gs->dt > 0)))
      return fail(b->ctx, FX_EINVAL, "fx_newmark: bad ground_series");
    if (gs->time) {
      for (int32_t i = 0; i < gs->count; ++i)
        if (!std::isfinite(gs->time[i]) || (i > 0 && gs->time[i] < gs->time[i - 1]))
          return fail(b->ctx, FX_EINVAL, "fx_newmark: ground_series times must be finite and nondecreasing");
      if (gs->loc && (*gs->loc < 0 || *gs->loc >= gs->count))
        return fail(b->ctx, FX_EINVAL, "fx_newmark: ground_series interval (*loc) out of range");
    }
  }
  FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
  return fx::newmark_run(b, *d, u, v, a, du, *result);
}

namespace {
int krylov_entry(fx_batch* b, int method, const fx_coeffs* k, const fx_solve_options* opt, const double* rhs,
                 double* x, fx_solve_result* res, const char* name) {
  if (!b || !k || !opt || !rhs || !x || !res) return fail(b ? b->ctx : nullptr, FX_EINVAL, std::string(name) + ": null argument");
  if (method == 1 && !b->gm) return fail(b->ctx, FX_EINVAL, std::string(name) + ": the batch was created without FX_BATCH_GMRES");
  FX_HIP(b->ctx, hipSetDevice(b->ctx->device));
  const int rc = fx::krylov64(b, method, *k, *opt, rhs, x, *res);
  if (rc == FX_OK) FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return rc;
}
}  // namespace

int fx_gmres(fx_batch* b, const fx_coeffs* k, const fx_solve_options* opt, const double* rhs, double* x,
             fx_solve_result* res) {
  return krylov_entry(b, 1, k, opt, rhs, x, res, "fx_gmres");
}

int fx_krylov(fx_batch* b, const fx_coeffs* k, const fx_solve_options* opt, const double* rhs, double* x,
              fx_solve_result* res) {
  return krylov_entry(b, 2, k, opt, rhs, x, res, "fx_krylov");
}
