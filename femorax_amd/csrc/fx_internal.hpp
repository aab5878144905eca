// Internal types of the femorax_amd library (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <array>
#include <vector>

#include "../../include/femorax_amd.h"
#include "fx_pimy.hpp"

namespace fx {

// Compile-step output (compile::Compiled for one stdBrick batch).
struct Mesh {
  int32_t nnode = 0, nelem = 0, padded = 0, neq = 0, nparam = 0, ncolor = 0;
  std::vector<int32_t> node_order, node_rank, elem_order, node_eq;
  std::vector<double> coords;
  std::vector<int32_t> node, eq, param, param_group, param_region, color_start, color_elements;
};

// Constraint refusals of the compile step (fx_constraints.cpp).
int check_equal_dof_commands(const fx_model_desc& m, std::string& err);
int check_transformation_handler(const fx_model_desc& m, const std::vector<int32_t>& node_rank,
                                 const std::vector<int32_t>& node_eq, std::string& err);
int compile_model(const fx_model_desc& m, int renumber_rcm, int simd_width, int build_coloring, Mesh& out,
                  std::string& err);

// Elements per block: one workgroup of this many threads, one element per thread.
constexpr int kBlockElems = 256;

// Host view of the block decomposition the kernels walk (built once per batch).
struct Blocks {
  int32_t nblocks = 0;
  std::vector<int32_t> elems;        // [nblocks * kBlockElems] batch position or -1
  std::vector<int32_t> pos;          // [count] block position b * kBlockElems + i of each element
  std::vector<int32_t> ln_start;     // [nblocks + 1] into lnode
  std::vector<uint32_t> lnode;       // global node id of each block-local node
  std::vector<int32_t> ln_part;      // per local node: its partial-sum slot (a node several blocks touch), or -1
  std::vector<uint32_t> ln_slots;    // per local node: start (low 16) | count (high 16)
  std::vector<uint16_t> slot_list;   // [nblocks * 8 * kBlockElems] elem_local * 8 + corner
  std::vector<int32_t> sh_node;      // nodes several blocks touch
  std::vector<int32_t> sh_start;     // [nshared + 1]: node i's partial slots, in block order
  std::vector<uint16_t> lcorner;     // [nblocks * 8 * kBlockElems] at b * 8 * kBlockElems + corner * kBlockElems + i:
                                     // the block-local node of element i's corner
  int32_t max_ln = 0;
};

// A block whose local nodes number at most kStageNodes stages them (coordinates, equations, x)
// in the LDS its element slots later use: 6 values per node in the 24 * kBlockElems slots.
constexpr int kStageNodes = 4 * kBlockElems;

// Partition the batch's elements into compact blocks (recursive coordinate bisection of the
// element centroids). A node only one block touches is written by that block; the others get a
// partial-sum slot per (node, block), summed in block order by a second kernel.
void build_blocks(int32_t count, int32_t padded, const int32_t* node, int32_t nnode, const double* coords,
                  Blocks& out);

// Structured fast path: tiles of kTileX x kTileY x (4 or 8) lattice cells (fx_lattice.cpp); a
// workgroup of kTileCells threads takes kTileZ = 4 cell planes per pass (one pass for 4-deep
// tiles, two for 8-deep). kTileNodes: the nodes of a 4-deep tile.
constexpr int kTileX = 8, kTileY = 8, kTileZ = 4;
constexpr int kTileCells = kTileX * kTileY * kTileZ;
constexpr int kTileNodes = (kTileX + 1) * (kTileY + 1) * (kTileZ + 1);
static_assert(kTileCells == kBlockElems, "one cell per thread");

struct TileHeader {
  int32_t nx, ny, nz;       // cells of this tile (<= kTileX, kTileY, tile depth)
  int32_t word_offset;      // into the node words: tile * 81 (tz + 1), node (x, y, z) at x + 9 (y + 9 z)
  int32_t cell_offset;      // into the cell parameters: tile * 64 tz (full-tile order, x fastest)
  int32_t interior_base;    // a full tile whose strictly-interior nodes are its own, with three free DOFs
                            // each, numbered in interior order (x fastest): their first equation; else -1
  int32_t ox, oy, oz;       // the lattice cell of the tile's first cell (rectilinear lattices: its spacings)
  int32_t pad;
};

struct Lattice {
  double hx = 0, hy = 0, hz = 0;      // the cell size (a uniform lattice)
  // a rectilinear lattice (the plane spacings differ along some axis): the plane coordinates along
  // each axis and each element's lattice cell (i, j, k)
  bool rect = false;
  std::vector<double> px, py, pz;
  std::vector<int32_t> ecell;         // [3 * count]
  int tz = 4;                         // cells per tile in z
  std::vector<TileHeader> tiles;      // in launch order (XCD-aware slabs or lattice order)
  std::vector<uint32_t> words;        // per tile node: first equation (28 bits) | free-DOF mask << 28 | shared << 31
  std::vector<int32_t> slots;         // per tile node: the tile's partial-sum slot of a shared node, else -1
  std::vector<uint16_t> cell_param;   // per tile cell: parameter set, 0xffff = no element
  std::vector<int32_t> cell_elem;     // per tile cell: batch element, -1 = none (host)
  std::vector<int32_t> perm;          // internal equation -> reference equation [neq]
  // nodes touched by several tiles, in group order: word (first equation | mask), first slot,
  // and (slot stride | tile count << 24); tile j's partial sum sits at slot first + j * stride,
  // tiles in parity-colour order
  // tied equation sets (nodes holding the same equations): word (first equation | mask) and
  // [tied_start[u], tied_start[u + 1]) the slots of every (node, tile) of set u
  std::vector<uint32_t> tied_word;
  std::vector<int32_t> tied_start;
  std::vector<uint32_t> shared_word;
  std::vector<int32_t> shared_start;
  std::vector<uint32_t> shared_meta;
  int64_t nslot = 0;
};

// tile_z: cells per tile in z (4 or 8; tiles are 8 x 8 x tile_z cells)
// allow_rect: accept rectilinear lattices (else only uniform ones)
bool build_lattice(int32_t count, int32_t padded, const int32_t* node, const int32_t* param, int32_t nnode,
                   const double* coords, const int32_t* node_eq, int32_t neq, bool xcd_order, int tile_z,
                   bool allow_rect, Lattice& out);

// Contraction of node-space values into equations when equations are shared by several
// nodes (equalDOF ties): eq -> list of node DOFs (3 n + d), in ascending order.
struct Contraction {
  bool tied = false;
  std::vector<int32_t> start, list;
};
void build_contraction(int32_t nnode, int32_t neq, const int32_t* node_eq, Contraction& out);

}  // namespace fx

// A captured run of CG iterations (cg_device): the key holds everything the kernels' arguments
// depend on, owner the batch whose arrays they read (its graphs go when the batch goes).
struct fx_cg_graph {
  std::vector<unsigned char> key;
  const void* owner = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// Context and batch (opaque at the C ABI).
struct fx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0, used = 0;
  std::string error;
  // device scratch for reductions and solver scalars
  double* red = nullptr;  // partial sums
  size_t red_cap = 0;
  void* scalars = nullptr;  // solver state block (device)
  void* host_pinned = nullptr;
  void* pimy_params = nullptr;  // material-point driver: up to kMaxPointMaterials parameter sets
  bool use_graphs = false;      // CG iterations replayed from captured graphs (FX_OPT_GRAPHS)
  hipStream_t own_stream = nullptr;  // the stream the context created (stream is the caller's with FX_OPT_STREAM)
  bool stream_ordered = false;  // the drop-in operator calls return without synchronising (FX_OPT_STREAM_ORDERED)
  std::vector<fx_cg_graph> cg_graphs;
};

namespace fx {
// The message fx_last_error(NULL) returns (errors raised before a context exists).
void set_thread_error(const std::string& msg);
// The device address of fx_dot_static's fixed scratch array, looked up once per device (at
// context creation, with the device current).
int static_scratch_init(fx_ctx* ctx);
// Drop the captured CG graphs of one owner (nullptr: all).
void release_cg_graphs(fx_ctx* ctx, const void* owner);
// The persistent autotune cache (fx_tune.cpp): keys for the device and the lattice model, and the
// tile depth and kernel variant kept for a key.
std::string tune_machine_key(int device);
std::string tune_model_key(int64_t cells, bool mass, bool damping, int forced_depth);
bool tune_cache_get(const std::string& key, int& depth, int& variant);
void tune_cache_put(const std::string& key, int depth, int variant);
}  // namespace fx

namespace fx {
struct DevBlocks {
  int32_t nblocks = 0, nshared = 0;
  int32_t max_ln = 0;  // the most local nodes of a block
  int32_t* elems = nullptr;
  int32_t* pos = nullptr;  // SSPbrick batches: each element's block position (its constants' index)
  int32_t* ln_start = nullptr;
  uint32_t* lnode = nullptr;
  int32_t* ln_part = nullptr;
  uint32_t* ln_slots = nullptr;
  uint16_t* slot_list = nullptr;
  int32_t* sh_node = nullptr;
  int32_t* sh_start = nullptr;
  uint16_t* lcorner = nullptr;
  int64_t ln_total = 0;      // local nodes over all blocks
  int32_t* ln_eq = nullptr;  // [3 * ln_total] at d * ln_total + ln: the local node's equations
  double* ln_xyz = nullptr;  // [3 * ln_total] at d * ln_total + ln: its coordinates
  float* ln_xyz32 = nullptr;
};

struct DevParam {  // per parameter set, device
  double lam, mu, rho;
  double b[3];
  int32_t mass_type, pad;
  double alphaM, betaK, betaKinit, betaKcomm;
};
}  // namespace fx

namespace fx {
constexpr int kMaxPointLoads = 32;
constexpr int kMaxPointMaterials = 8;
constexpr int kGmresRestart = 30;  // GMRES(30) (gmres.hpp default restart)
// doubles of the GMRES workspace (FX_BATCH_GMRES): V (m + 1 columns), Z (m), the initial guess,
// the combine coefficients, and the block dots' chunk partials and level buffers
inline size_t gmres_workspace(int64_t n) {
  const int64_t ch = (n + 2047) / 2048, lvl = 2 * ((ch + 2047) / 2048 + 2);
  return static_cast<size_t>(2 * kGmresRestart + 3) * static_cast<size_t>(n > 0 ? n : 1) + kGmresRestart + 1 +
         static_cast<size_t>(kGmresRestart + 1) * static_cast<size_t>(ch + lvl + 1);
}
struct PimyParams;
// PressureIndependMultiYield (fx_pimy_host.cpp, fx_pimy.hip)
int pimy_setup(const fx_pimy_material& m, PimyParams& p, std::string& err);
int pimy_points(fx_ctx* ctx, int64_t npts, const PimyParams* params, const int32_t* param, const double* C, double* T,
                const double* eps, double* sig, double* D);
int pimy_init(fx_ctx* ctx, int64_t npts, const PimyParams* params, const int32_t* param, double* S);
// Host stdBrick helpers (fx_brick_host.cpp).
void brick_node_weights(const double (&xe)[8][3], double (&w)[8]);
void brick_stiffness(const double (&xe)[8][3], double lam, double mu, double (&K)[24][24]);
double max_eigenvalue_symmetric(double (&A)[24][24]);
double brick_max_omega_sq(const double (&xe)[8][3], double lam, double mu, double rho);
void brick_mass_unit(const double (&xe)[8][3], double (&M)[8][8]);
void brick_stiffness_points(const double (&xe)[8][3], const double (&lam)[8], const double (&mu)[8],
                            double (&K)[24][24]);
double brick_max_omega_sq_points(const double (&xe)[8][3], const double (&lam)[8], const double (&mu)[8], double rho);
// Host SSPbrick helpers (fx_ssp_host.cpp). An element's constants, kSspConsts values at c[k * stride]:
// dNmod (node-major, 24), mVol, the node volume weights (8), the hourglass vectors gst gut gus gstu
// (32), interior's upper triangle (78).
constexpr int kSspConsts = 143;
constexpr int kSspGamma = 33;
constexpr int kSspInterior = 65;
// A lattice cell's: the modal stabilisation (interior / 64, upper triangle, 78) and diag(K_e) of unit
// operator factor by lattice corner (component-major, 24).
constexpr int kSspBox = 102;
void ssp_constants(const double (&xe)[8][3], double lam, double mu, double* c, int64_t stride);
void ssp_stiffness(const double* c, double lam, double mu, double (&K)[24][24]);
double ssp_max_omega_sq(const double (&xe)[8][3], double lam, double mu, double rho);
void ssp_box_constants(double hx, double hy, double hz, double lam, double mu, double* out);
double stable_dt(int scheme, const DevParam& p, double omega_sq);
}  // namespace fx

struct fx_batch {
  fx_ctx* ctx = nullptr;
  int32_t count = 0, padded = 0, nnode = 0, neq = 0, nparam = 0;
  bool has_f32 = false;
  // device arrays
  int32_t* node = nullptr;      // [8 * padded]
  int32_t* param = nullptr;     // [padded]
  int32_t* node_eq = nullptr;   // [3 * nnode]
  double* coords = nullptr;     // [3 * nnode]
  float* coords32 = nullptr;    // [3 * nnode]
  fx::DevParam* params = nullptr;
  // SSPbrick batches (FX_BATCH_SSPBRICK): the element constants ([kSspConsts][padded], the general
  // path and the per-element forms) or the lattice's per-set modal constants ([nparam][kSspBox])
  // a CG solve's own state block and reduction scratch (fx_cg_many: several solves in flight)
  double* cg_state = nullptr;  // a CgState (fx_solver.hip)
  double* cg_red = nullptr;
  bool ssp = false;
  double* ssp_elem = nullptr;
  double* ssp_box = nullptr;
  // rectilinear lattices: (kappa = 1 / (4 h), h) of every cell interval along x, y, z, and the
  // distinct (hx, hy, hz, parameter set) of the cells (the stable step)
  bool rect = false;
  double* kx = nullptr;
  double* ky = nullptr;
  double* kz = nullptr;
  std::vector<std::array<double, 4>> rect_boxes;
  fx::DevBlocks blk;
  // tied meshes: node-space output and its contraction
  bool tied = false;
  double* ynode = nullptr;      // [3 * nnode]
  float* ynode32 = nullptr;
  int32_t* con_start = nullptr; // [neq + 1]
  int32_t* con_list = nullptr;
  // structured fast path (when the batch is an axis-aligned uniform lattice)
  bool lattice = false;
  double hx = 0, hy = 0, hz = 0;
  fx::TileHeader* tiles = nullptr;
  uint32_t* words = nullptr;
  int32_t* tile_slots = nullptr;
  uint16_t* cell_param = nullptr;
  int32_t ntiles = 0;
  int32_t nshared = 0;
  uint32_t* tied_word = nullptr;  // lattice tied equation sets (laminar ties)
  int32_t* tied_start = nullptr;
  int32_t ntied = 0;
  // the diagonal's cross terms of tied DOF pairs inside an element (element_operator.hpp:
  // 138-209): per (tied equation, parameter set) the sums of the unit-coefficient stiffness
  // (lambda part, mu part) and consistent-mass entries K_ab, M_ab over ordered pairs a != b
  std::vector<int32_t> tie_q;       // internal equation of each entry
  std::vector<int32_t> tie_p;       // parameter set of each entry
  std::vector<double> tie_kl, tie_km, tie_mc;
  std::vector<int32_t> tie_eqs;     // the distinct tied equations (internal numbering)
  std::vector<double> tie_host;     // their corrections for the current coefficients
  int32_t* tie_idx = nullptr;
  double* tie_val = nullptr;
  uint32_t* shared_word = nullptr;
  int32_t* shared_start = nullptr;
  uint32_t* shared_meta = nullptr;
  double* partial64 = nullptr;        // [nslot * 3] per-tile sums at shared nodes
  float* partial32 = nullptr;
  int box_variant = 0;                // lattice kernel variant of the operator without mass terms (tuned)
  int box_variant_mass = 0;           // the same with mass terms, and for the diagonal
  float tuned_ms = 0.0f;              // the tuned variants' sample times (both operator forms), 0 untuned
  int tile_z = 8;                     // lattice tiles of 8 x 8 x tile_z cells
  // internal (tile-major) equation numbering of lattice batches: perm[internal] = reference
  bool internal = false;
  int32_t* perm = nullptr;
  // the same node words in the reference numbering (where every node's free DOFs have consecutive
  // reference equations): the drop-in apply reads x and writes y in place, with no permutation
  uint32_t* words_ref = nullptr;
  uint32_t* shared_word_ref = nullptr;
  uint32_t* tied_word_ref = nullptr;
  int32_t* orphans_ref = nullptr;
  // free equations no element touches: zeroed after each apply
  int32_t* orphans = nullptr;
  int32_t norphans = 0;
  // lumped mass, lumped alphaM mass and body load in the solver numbering (null when all zero)
  double* lumped_m = nullptr;
  double* lumped_am = nullptr;
  double* body = nullptr;
  std::vector<int32_t> internal_of;   // host: reference equation -> internal (lattice batches)
  int32_t* internal_of_dev = nullptr; // its device copy, made for the first recorded stage
  // host copies for the stable time step of general batches (lattice batches need none)
  std::vector<int32_t> host_node, host_param;
  std::vector<double> host_coords;
  // stdBrick<PressureIndependMultiYield> batches: history, stored tangents, slots + gather
  bool pimy = false;
  int64_t npts = 0;                   // count * 8 Gauss points
  fx::PimyParams* pimy_params = nullptr;
  std::vector<fx::PimyParams> host_pimy;
  double* state_c = nullptr;          // committed history [kPimyNState * npts]
  double* state_t = nullptr;          // trial history
  double* tangent = nullptr;          // [21 * npts] upper triangles
  float* tangent32 = nullptr;         // their FP32 copy (FX_BATCH_FP32 PIMY batches)
  bool tangents_stored = false;
  double* slots = nullptr;            // [24 * padded]
  int32_t* gather_start = nullptr;    // [neq + 1]
  int32_t* gather_list = nullptr;     // slots in (element, DOF) order
  // compacted return map (FX_BATCH_COMPACT): per point stress (6), tangent upper triangle (21)
  // and strain (6); the plastic points' list and its length
  double* cres = nullptr;             // [33 * npts]
  int32_t* clist = nullptr;           // [npts]
  int32_t* ccount = nullptr;          // [1]
  // time-integration workspace (FX_BATCH_DYNAMIC): 12 * neq doubles; the DOF direction of each
  // equation (reference numbering, host)
  double* dyn = nullptr;
  bool pimy_damped = false;  // some PIMY parameter set has Rayleigh damping
  int32_t* long_eqs = nullptr;  // PIMY gather: the equations with more than 64 slots (tied levels)
  int32_t* pimy_touched = nullptr;  // per point: the highest surface whose centre an update has written
  int32_t nlong = 0;
  std::vector<int8_t> eq_dir;
  int8_t* eq_dir_dev = nullptr;  // the same on the device, solver numbering (explicit ground loads)
  // GMRES workspace (FX_BATCH_GMRES): V [(m + 1) n], Z [m n], coefficients [m + 1], CG start [n]
  double* gm = nullptr;
  // nodal masses of the model (fx_batch_set_nodal_mass), solver numbering; the elements' own
  // lumped masses (host, solver numbering) that the lumped arrays add them to
  double* nodal_m = nullptr;
  double* nodal_am = nullptr;
  float* nodal_m32 = nullptr;
  float* nodal_am32 = nullptr;
  std::vector<double> host_lm, host_la;
  int32_t tangent_symmetry = FX_TANGENT_SPD;
  // solver work (allocated at create: compute allocates nothing)
  double* work = nullptr;       // 8 * neq doubles: r z p ap inv d, permuted b and x
  float* work32 = nullptr;      // 5 * neq floats
  uint64_t device_bytes = 0;
  std::vector<fx::DevParam> host_params;
  fx_batch_info info{};
};

#define FX_HIP(ctx, call)                                                          \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      (ctx)->error = std::string(#call) + ": " + hipGetErrorString(e_);           \
      return FX_EDEVICE;                                                           \
    }                                                                              \
  } while (0)

namespace fx {
// operator_symmetry (matrix_free_operator.cpp:5-16): whether K_eff of k is symmetric positive
// definite, from the batch's tangent symmetry (elastic SPD, PressureIndependMultiYield SPSD: made
// definite by a mass term) and the coefficients' signs.
inline bool operator_spd(const fx_batch* b, const fx_coeffs& k) {
  const bool positive =
      k.stiffness >= 0.0 && k.damping >= 0.0 && k.mass >= 0.0 && (k.stiffness > 0.0 || k.mass > 0.0);
  if (b->tangent_symmetry == FX_TANGENT_SPD) return positive;
  return positive && k.mass > 0.0;
}
}  // namespace fx
