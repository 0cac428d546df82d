// hg_ctx.hpp -- the context behind an hg_ctx* handle (host-side layout of
// every device buffer) and the kernel argument blocks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "hgswe.h"

namespace hg {

// Arguments of the fused RHS / Euler-step kernel.  Cell data are in the
// library's (possibly renumbered) order; faces in ELL form, slot k of cell i
// at k * N + i, so every face load of a wavefront is coalesced.
struct StepArgs {
  int32_t N;  // cells per field (the mesh); the launch covers K fields of N cells
  int32_t maxdeg;
  int32_t K;                // fields sharing the mesh (1 unless the context holds K fields)
  int32_t nin_m, nex_m;     // inlet / exit boundaries per field
  const double* xi;   // state in
  const double* qx;
  const double* qy;
  double* xo;         // state out (step) or dQdt (rhs)
  double* qxo;
  double* qyo;
  const double* hstill;
  const double* zb;
  const double* n;
  const double* S0x;
  const double* S0y;
  const double* area;
  const int32_t* deg;
  const int32_t* nbr;  // >= 0 cell, < 0 ~ghost
  const double* fnx;
  const double* fny;
  const double* flen;
  const int32_t* g_kind;
  const int32_t* g_bc;
  const double* g_hstill;
  const double* g_zb;
  const double* g_L23;   // L^(2/3) of inlet ghosts
  const double* inletQ;
  const double* exitWSE;
  const double* totalA;  // per inlet, this step
  uint8_t* clamp;        // optional: 1 where the step's dry clamp fired (adjoint recompute)
  double g, half_g, kn2, hsmall, dt;
  const double* dtc;  // per-cell time step (batches with per-instance steps), or NULL: dt
  const double* fmag; // friction magnitude per cell (FlowResistance closure), or NULL: Manning friction
  int32_t* flag;      // [1] set to 1 when a step writes a NaN or infinite value
};

struct InletArgs {
  int32_t n_inlet;  // inlets over all fields (one work-group each)
  int32_t nin_m;    // inlets per field (n_inlet / K)
  int32_t Nm;       // cells per field
  const int32_t* off;
  const int32_t* cell;
  const double* L53;
  const double* xi;
  const double* hstill;
  const double* n;
  double hsmall;
  double* totalA;
  int32_t* flag;  // set to 1 when a conveyance is <= 1e-10
};

// One inlet-Q face as the inlet adjoint reads it (32 bytes): its ghost and
// cell, whether it is the cell's first face in the inlet and the cell's next
// face there (or -1), L^(2/3) and L^(5/3).
struct InletFaceRec {
  int32_t ghost, cell, first, next;
  double L23, L53;
};

// One work item of the tiled kernels, as they use it.
struct ItemRec {
  int32_t code;
  int32_t ref;
  double nx, ny, len;
};

// Face geometry of one work item.
struct ItemGeo {
  double nx, ny, len;
};

// The work items of one tiling in device memory.  Tile t's items are
// off[t] .. off[t+1]-1, its cross-tile and boundary items first; those, up to
// x0[t], also carry a ref (the global neighbour or ghost id), so the
// intra-tile items, most of them, are read as 28 bytes (code + geometry).
struct ItemTable {
  int32_t* off;
  int32_t* x0;
  int32_t* code;
  int32_t* ref;
  ItemGeo* geo;
};

// Arguments of the tiled (face-once) step kernel.  Cells are split into
// tiles of `T` consecutive (Morton-ordered) cells, one work-group per tile.
// Work items of tile t are items.off[t] .. items.off[t+1]-1:
//   code bits  0..9  a  (local cell of the face's first side)
//              10..13 ka (local face index of that side)
//              14..23 b  (local cell of the second side, intra-tile items)
//              24..27 kb
//              28..29 type: 0 intra-tile (both sides), 1 cross-tile
//                     (first side only; ref = global neighbour), 2 ghost
//                     (ref = ghost id)
struct TileArgs {
  int32_t N;  // cells per field (the mesh); the launch covers K fields of N cells
  int32_t K;
  int32_t nin_m, nex_m;  // inlet / exit boundaries per field
  int32_t T;
  int32_t maxdeg;
  int32_t n_inlet;  // inlets whose conveyance the kernel forms (0: done by a separate kernel)
  const double* xi;
  const double* qx;
  const double* qy;
  double* xo;
  double* qxo;
  double* qyo;
  const double* hstill;
  const double* zb;
  const double* n;
  const double* S0x;
  const double* S0y;
  const double* area;
  const int32_t* deg;  // faces per cell, or NULL: every cell has MD faces
  struct ItemTable items;
  const int32_t* tile_in_off;   // [ntiles+1] into tile_in_list
  const int32_t* tile_in_list;  // inlets with ghosts in the tile; a ghost item's local slot indexes this list
  const int32_t* g_kind;
  const int32_t* g_bc;
  const double* g_hstill;
  const double* g_zb;
  const double* g_L23;
  const double* inletQ;
  const double* exitWSE;
  const int32_t* in_off;
  const int32_t* in_cell;
  const double* in_L53;
  const double* totalA;  // used when n_inlet == 0 (conveyance from the separate kernel)
  double* totalA_out;    // optional record of the fused conveyances (per inlet)
  int32_t* flag;
  uint8_t* clamp;  // optional: 1 where the step's dry clamp fired (adjoint recompute)
  double g, half_g, kn2, hsmall, dt;
  const double* dtc;  // per-cell time step (batches with per-instance steps), or NULL: dt
  int32_t xcd_chunk;  // block -> tile mapping: 0 identity, G > 0 runs of G tiles per XCD in windows of 8G
  const double* fmag; // friction magnitude per cell (FlowResistance closure), or NULL: Manning friction
};

// Arguments of the adjoint (VJP) kernel of one RHS evaluation / one step.
struct AdjArgs {
  int32_t N;  // cells per field (the mesh); the launch covers K fields of N cells
  int32_t maxdeg;
  int32_t K;
  int32_t nin_m, nex_m, G_m;  // inlet / exit boundaries and ghosts per field
  const double* xi;  // state the RHS was evaluated at
  const double* qx;
  const double* qy;
  const double* lx;  // cotangent of the step output (STEP) or of the RHS
  const double* lqx;
  const double* lqy;
  const uint8_t* mask;  // STEP: dry-clamp flags of the step output (nullable)
  double* ox;           // cotangent of the state
  double* oqx;
  double* oqy;
  double* nbar;   // += d/dn per cell
  double* zbbar;  // += d/dzb per cell (exit-H depth)
  double* S0bar;  // [2N] d/dS0 of this evaluation
  double* vbar;   // [G] d/d(inlet ghost velocity) of this evaluation
  const double* hstill;
  const double* zb;
  const double* n;
  const double* S0x;
  const double* S0y;
  const double* area;
  const int32_t* deg;  // faces per cell, or NULL (tiled reverse step): every cell has MD faces
  const int32_t* nbr;
  const double* fnx;
  const double* fny;
  const double* flen;
  const int32_t* g_kind;
  const int32_t* g_bc;
  const double* g_hstill;
  const double* g_zb;
  const double* g_L23;
  const double* inletQ;
  const double* exitWSE;
  const double* totalA;
  double g, half_g, kn2, hsmall, dt;
  const double* dtc;  // per-cell time step (batches with per-instance steps), or NULL: dt
  const double* fmag; // FlowResistance closure: friction magnitude per cell (else NULL) ...
  double* fbar;       // ... and its cotangent
  int32_t nbar_set;   // 1: nbar holds this evaluation's n-bar alone (set, not added to)
  int32_t fbar_set;   // 1: fbar likewise
  // tiled reverse step (k_adj_tile): the step kernel's work items and block -> tile mapping
  struct ItemTable items;
  int32_t xcd_chunk;
};

// Per-field views of a launch over K fields that share one mesh: field f's
// state-like arrays start f N cells into the arrays over all fields (its
// inlet / exit / ghost arrays f nin_m / nex_m / G_m entries in, its S0
// cotangent f 2N), while the mesh geometry and per-cell constants are
// indexed by the cell within the mesh.  Field 0 is the arrays themselves.
template <class T>
__host__ __device__ __forceinline__ T* shift(T* p, int64_t o) { return p ? p + o : p; }

__host__ __device__ __forceinline__ StepArgs field_view(StepArgs a, int f) {
  const int64_t o = int64_t(f) * a.N;
  a.xi = shift(a.xi, o); a.qx = shift(a.qx, o); a.qy = shift(a.qy, o);
  a.xo = shift(a.xo, o); a.qxo = shift(a.qxo, o); a.qyo = shift(a.qyo, o);
  a.n = shift(a.n, o); a.clamp = shift(a.clamp, o); a.dtc = shift(a.dtc, o); a.fmag = shift(a.fmag, o);
  a.inletQ = shift(a.inletQ, int64_t(f) * a.nin_m); a.totalA = shift(a.totalA, int64_t(f) * a.nin_m);
  a.exitWSE = shift(a.exitWSE, int64_t(f) * a.nex_m);
  return a;
}

__host__ __device__ __forceinline__ TileArgs field_view(TileArgs a, int f) {
  const int64_t o = int64_t(f) * a.N;
  a.xi = shift(a.xi, o); a.qx = shift(a.qx, o); a.qy = shift(a.qy, o);
  a.xo = shift(a.xo, o); a.qxo = shift(a.qxo, o); a.qyo = shift(a.qyo, o);
  a.n = shift(a.n, o); a.clamp = shift(a.clamp, o); a.dtc = shift(a.dtc, o); a.fmag = shift(a.fmag, o);
  a.inletQ = shift(a.inletQ, int64_t(f) * a.nin_m); a.totalA = shift(a.totalA, int64_t(f) * a.nin_m);
  a.totalA_out = shift(a.totalA_out, int64_t(f) * a.nin_m);
  a.exitWSE = shift(a.exitWSE, int64_t(f) * a.nex_m);
  return a;
}

__host__ __device__ __forceinline__ AdjArgs field_view(AdjArgs a, int f) {
  const int64_t o = int64_t(f) * a.N;
  a.xi = shift(a.xi, o); a.qx = shift(a.qx, o); a.qy = shift(a.qy, o);
  a.lx = shift(a.lx, o); a.lqx = shift(a.lqx, o); a.lqy = shift(a.lqy, o);
  a.mask = shift(a.mask, o);
  a.ox = shift(a.ox, o); a.oqx = shift(a.oqx, o); a.oqy = shift(a.oqy, o);
  a.nbar = shift(a.nbar, o); a.zbbar = shift(a.zbbar, o); a.S0bar = shift(a.S0bar, 2 * o);
  a.vbar = shift(a.vbar, int64_t(f) * a.G_m);
  a.n = shift(a.n, o); a.dtc = shift(a.dtc, o); a.fmag = shift(a.fmag, o); a.fbar = shift(a.fbar, o);
  a.inletQ = shift(a.inletQ, int64_t(f) * a.nin_m); a.totalA = shift(a.totalA, int64_t(f) * a.nin_m);
  a.exitWSE = shift(a.exitWSE, int64_t(f) * a.nex_m);
  return a;
}

// A variable Manning-n law (kind HG_NLAW_*; 0: off) and its parameters.
enum { NLAW_OFF = 0, NLAW_POWER = 1, NLAW_SIGMOID = 2, NLAW_INVERSE = 3, NLAW_H_UMAG_KS = 4 };
struct NLaw {
  int32_t kind;
  double lo, hi, k, hmid;  // n_lower, n_upper, k, h_mid
};

// The reference's UDE network (create_NN_model, process_UDE.jl:2-65):
// per hidden layer Dense(in, w, act) -> LayerNorm(w), then Dense(w, 1) and
// the output bounds lo + (hi - lo) sigmoid(.).  Offsets index the flat
// parameter vector (see include/hgswe.h for its order).
constexpr int kLuxMaxLayers = 4, kLuxMaxW = 16, kLuxMaxP = 2048;
enum { LUX_OFF = 0, LUX_MANNING_H = 1, LUX_MANNING_H_UMAG_KS = 2, LUX_FLOW_RESISTANCE = 3 };
enum { ACT_RELU = 0, ACT_LEAKYRELU = 1, ACT_SIGMOID = 2, ACT_TANH = 3, ACT_SOFTPLUS = 4 };
struct LuxNet {
  int32_t choice, n_in, n_hidden, P;
  int32_t width[kLuxMaxLayers], act[kLuxMaxLayers];
  int32_t off_W[kLuxMaxLayers], off_b[kLuxMaxLayers], off_g[kLuxMaxLayers], off_be[kLuxMaxLayers];
  int32_t off_Wo, off_bo;
  double lo, hi;             // output bounds
  double in_lo[3], in_hi[3]; // input bounds: h, then |U| or |q|, then ks
  double eps;                // LayerNorm epsilon
};

template <class T>
struct DBuf {  // owning device buffer
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace hg

// Key of a cached hipGraph of the adjoint's per-segment work: kind 0 the
// recompute of ns steps into segment buffer b, kind 1 the reversal of ns
// steps in buffer b starting from cotangent buffer lam0; for time step dt.
struct SegGraphKey {
  int kind, b, lam0;
  int64_t ns;
  double dt;
  int want;  // the parameter cotangents the reversal accumulates (bit 0: n, bit 1: zb)
  bool operator<(const SegGraphKey& o) const {
    if (kind != o.kind) return kind < o.kind;
    if (b != o.b) return b < o.b;
    if (lam0 != o.lam0) return lam0 < o.lam0;
    if (want != o.want) return want < o.want;
    if (ns != o.ns) return ns < o.ns;
    return dt < o.dt;
  }
};

struct hg_ctx {
  std::string err;
  int64_t N = 0, G = 0;
  // K fields on one mesh (hg_fields_create; K = 1 otherwise): N, G, n_inlet,
  // n_exit and n_inlet_faces count all fields; the geometry and the per-cell
  // constants (h_still, zb, S0) are stored once for the mesh of Nm cells
  int32_t K = 1;
  int64_t Nm = 0, Gm = 0;
  int32_t nin_m = 0, nex_m = 0;
  int32_t maxdeg = 0;
  int32_t n_inlet = 0, n_exit = 0, n_inlet_faces = 0;
  hg_consts c{};
  hipStream_t stream = nullptr;
  int device = 0;
  bool fields_set = false;

  // renumbering: lib index -> caller index, and inverse
  std::vector<int32_t> perm, iperm;
  // host copies needed by parameter application
  std::vector<int32_t> h_zone;  // lib order
  int32_t n_zones = 0;
  int32_t active = HG_PARAM_NONE;

  // device buffers (lib order)
  double* st[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  int cur = 0;
  double *hstill = nullptr, *zb = nullptr, *n = nullptr, *S0x = nullptr, *S0y = nullptr, *area = nullptr;
  double *zb0 = nullptr, *n0 = nullptr, *S0x0 = nullptr, *S0y0 = nullptr;  // as set by hg_set_fields
  int32_t *deg = nullptr, *nbr = nullptr;
  double *fnx = nullptr, *fny = nullptr, *flen = nullptr;
  int32_t *g_kind = nullptr, *g_bc = nullptr, *g_cell = nullptr;
  double *g_hstill = nullptr, *g_zb = nullptr, *g_zb0 = nullptr, *g_L23 = nullptr;
  int32_t *in_off = nullptr, *in_cell = nullptr;
  double* in_L53 = nullptr;
  // per inlet face: ghost, cell, first face of its cell in the inlet, the
  // cell's next face there (or -1), L^(2/3), L^(5/3) (the inlet adjoint)
  hg::InletFaceRec* in_rec = nullptr;
  bool inlets_disjoint = true;  // no cell sits on two inlets: the inlet adjoint runs one work-group per inlet
  double *inletQ = nullptr, *inletQ0 = nullptr, *exitWSE = nullptr, *totalA = nullptr;
  int32_t* flags = nullptr;  // device address of flags_host: [0] conveyance error, [1] non-finite state
  volatile int32_t* flags_host = nullptr;  // pinned, coherent host memory: cleared and read by the host
  int32_t* zone = nullptr;
  // tiled kernel data
  bool tiled = false;
  int32_t tileT = 256, ntiles = 0;
  bool fuse_inlet = false;
  hg::ItemTable items{};  // the step kernel's work items
  int32_t tileMD = 3;
  bool deg_full = false;  // every cell has tileMD faces: the tile kernels need not read deg
  int32_t xcd_chunk = 0;  // identity block -> tile mapping by default (the autotuner re-checks)
  int32_t ncu = 256;      // compute units of the device
  // the reverse step's tiling (its own tile size; the step kernel's items
  // when those have it, else its own copy built on first use)
  int32_t adj_T_req = 0;  // hg_set_adjoint_tile (0: the library's choice)
  int32_t adj_T = 0, adj_ntiles = 0;
  hg::ItemTable adj_items{};  // the reverse step's: items or adj_own
  int32_t adj_own_T = 0, adj_own_ntiles = 0;
  hg::ItemTable adj_own{};
  int32_t *tile_in_off = nullptr, *tile_in_list = nullptr;
  std::vector<void*> tile_allocs;  // device buffers of the current tiling (freed on re-tiling)
  size_t lds_max = 65536;          // dynamic LDS a work-group may use on this device
  // host copies of the ELL face arrays and ghost tables (lib order), kept for re-tiling
  std::vector<int32_t> h_deg, h_nbr, h_gkind, h_gbc;
  std::vector<double> h_nx, h_ny, h_len;
  double* scratch = nullptr;  // 3N doubles for host transfers / params
  // host <-> device transfers in caller order: a 3N staging buffer and the
  // permutation (library cell i is caller cell perm[i]) on the device
  double* stage = nullptr;
  int32_t* d_perm = nullptr;
  // probe-loss buffers of hg_grad_probes (capacity probe_cap probes)
  int32_t* probe_i = nullptr;
  double* probe_d = nullptr;
  int64_t probe_cap = 0;
  size_t scratch_n = 0;
  double* red = nullptr;      // reduction partials
  double* noalloc_scratch = nullptr;  // device address of the static scratch of hg_cfl_dt_noalloc on this device

  // parameter cotangents the current adjoint call accumulates: those of the
  // per-cell n and of zb (through S0 and the exit-H depth) are skipped when
  // the caller does not ask for them
  bool adj_want_n = true, adj_want_zb = true;
  int64_t last_kck = 0;  // checkpoint interval of the last hg_adjoint / hg_grad_probes call
  // a forward sweep recorded by hg_adjoint_forward, awaiting hg_adjoint_reverse
  // (cleared by any other gradient call and by the calls that set fields or
  // parameters)
  bool adj_pending = false;
  double pend_dt = 0.0;
  int64_t pend_nsteps = 0, pend_kck = 0;
  // adjoint work buffers (allocated on first use)
  double* adj_lam[2] = {nullptr, nullptr};  // 3N each
  double *adj_nbar = nullptr, *adj_zbbar = nullptr, *adj_S0bar = nullptr, *adj_vbar = nullptr, *adj_Qbar = nullptr;
  double* adj_seg[2] = {nullptr, nullptr};    // (kck + 1) x 3N states of a checkpoint segment (two buffers)
  uint8_t* adj_mask[2] = {nullptr, nullptr};  // kck x N clamp flags of each
  double* adj_ckpt = nullptr;   // nseg x 3N checkpoints
  double* adj_totalA[2] = {nullptr, nullptr};  // kck x n_inlet conveyances recorded by each segment's recompute
  hipStream_t stream2 = nullptr;  // the adjoint's segment recompute, overlapping the reverse sweep
  hipEvent_t ev_seg[2] = {nullptr, nullptr}, ev_rev[2] = {nullptr, nullptr}, ev_fwd = nullptr;
  size_t adj_seg_cap[2] = {0, 0}, adj_ckpt_cap = 0;
  double* adj_nseg[2] = {nullptr, nullptr};   // kck x N closure n of each recomputed step (two buffers)
  size_t adj_nseg_cap[2] = {0, 0};

  // UDE closure for n (0: off; else the precision of its network, 32 or 64)
  int32_t ude_prec = 0;
  int32_t ude_nsets = 0;
  void* ude_w = nullptr;            // [ude_nsets][353] weights in that precision
  int32_t* ude_set = nullptr;       // weight set of each cell (lib order), or NULL
  double* ude_nbar = nullptr;       // per-step cotangent of n (N)
  double* ude_wpart = nullptr;      // [ude_nslots][353] weight cotangents per (block, set) slot
  double* ude_wbar = nullptr;       // accumulated weight cotangents [ude_nsets][353]
  int32_t ude_nslots = 0;           // slots, ordered by set then block
  int32_t ude_cpb = 1;              // 64-cell chunks per reverse-mode block
  int32_t* ude_blk_off = nullptr;   // [blocks + 1] first entry of each reverse-mode block in ude_blk_set / _slot
  int32_t* ude_blk_set = nullptr;   // set of each (block, set) pair, a block's sets in order of first appearance
  int32_t* ude_blk_slot = nullptr;  // slot of each (block, set) pair
  int32_t* ude_set_start = nullptr; // [ude_nsets + 1] first slot of each set
  std::vector<void*> ude_allocs;

  // variable Manning-n law (kind 0: off), its per-cell ks and per-step n cotangent
  hg::NLaw nlaw{};
  double* nlaw_ks = nullptr;
  double* nlaw_nbar = nullptr;
  std::vector<void*> nlaw_allocs;

  // the reference's UDE network (choice 0: off), its weights, per-cell ks,
  // friction magnitudes (FlowResistance), per-step output cotangent and the
  // accumulated weight cotangent
  hg::LuxNet lux{};
  int32_t lux_mw = 0;          // register width of the kernels (4 or 16)
  double* lux_w = nullptr;     // P weights (double)
  double* lux_ks = nullptr;
  double* lux_fmag = nullptr;  // N (FlowResistance)
  double* lux_obar = nullptr;  // N
  double* lux_wpart = nullptr; // P x blocks
  double* lux_wbar = nullptr;  // P
  double* lux_dw = nullptr;    // weight tangent of hg_jvp (kLuxMaxP)
  std::vector<void*> lux_allocs;

  // batches: instance boundaries in caller order, and the per-cell time step
  // (lib order) when per-instance steps are set
  std::vector<int64_t> inst_off;
  double* dt_cell = nullptr;  // dt_buf while per-instance steps are set, else NULL
  double* dt_buf = nullptr;

  // forward-mode (JVP) tangents: state (2 x 3N) and parameter fields
  double* jvp = nullptr;

  // Tsit5 work buffers: 7 stages + y, ynew, ytmp (3N each), norm partials
  double* rk = nullptr;
  double* rk_part = nullptr;

  // events attached to timed step dispatches (hg_advance_timed), reused
  std::vector<hipEvent_t> ev_pool;
  hipEvent_t ev_xfer[2] = {nullptr, nullptr};  // chunk events of the pinned host transfers

  // cached graph of kGraphSteps steps (buffer 0 -> ... -> buffer 0) per dt
  hipGraphExec_t gexec = nullptr;
  // cached graphs of the adjoint's segment recompute and reversal (dropped
  // with gexec whenever a launch they captured would change)
  std::map<SegGraphKey, hipGraphExec_t> seg_graphs;
  double gdt = 0.0;
  bool use_graphs = true;  // false (environment HG_GRAPHS=0 at creation): every launch is a plain launch

  std::vector<void*> allocs;
};
