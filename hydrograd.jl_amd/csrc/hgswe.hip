// hgswe.hip -- forward path of libhgswe: context set-up, the fused
// RHS / Euler-step kernel, inlet conveyance, bed data, zone gather, CFL, and
// the C ABI declared in include/hgswe.h.
//
// Reference path (Hydrograd.jl): swe_2d_rhs, semi_discretize_swe_2D.jl:18-277
// -> process_all_boundaries_2d (bc_2D.jl:575-875), compute_inviscid_fluxes
// (:281-434) -> Riemann_2D_Roe (swe_2D_solvers.jl:4-164),
// compute_source_terms (:449-559); custom_ODE_update_cells
// (custom_ODE_solvers.jl:5-33).
//
// Design: one work-item per cell gathers its own faces in the reference's
// local order (no atomics, deterministic, every internal face evaluated from
// both sides exactly like the reference).  Ghost states are formed inline
// because a ghost's internal cell is the cell that owns the boundary face;
// only the inlet conveyance (a sequential sum over an inlet's faces) needs a
// separate tiny kernel per step.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <unistd.h>

#include <condition_variable>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hg_ctx.hpp"
#include "hg_device.hpp"
#include "hg_roe_pair.hpp"
#include "hg_adjoint.hpp"
#include "hg_ude.hpp"
#include "hg_dual.hpp"  // after the double paths: its sqrt / cbrt overloads must stay out of them
#include "hg_nlaw.hpp"
#include "hg_lux.hpp"
#include "hgswe.h"

using hg::Side;
using hg::StepArgs;

static thread_local std::string g_create_err;  // per caller thread, like the contexts

// Scratch of hg_cfl_dt_noalloc: the per-work-group partial minima of the CFL
// reduction (one per 256 cells, so meshes of up to 32768 x 256 = 8,388,608
// cells) followed by its result.  A static device array, so that the entry
// point allocates nothing; each GPU holds its own copy (256 KiB) until the
// process ends, and every call on a GPU shares that GPU's copy.  In host code
// the name is only a stand-in: host code uses the address that
// hipGetSymbolAddress returned on the device (kept per device below).
constexpr int kNoAllocParts = 32768;
static __device__ __align__(256) double scratch[kNoAllocParts + 32];

#define HG_TRY(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                    \
      return HG_ERR_HIP;                                                               \
    }                                                                                  \
  } while (0)

namespace {

// Device address of `scratch` on each device, resolved once per device by
// the first hg_create with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

constexpr int kBlock = 256;
constexpr int kRewarmSteps = 1024;  // graph-replayed steps the autotuner ends with (a multiple of kGraphSteps)
constexpr int kGraphSteps = 16;  // steps per captured graph (even: the graph returns to buffer 0)
constexpr int kXCD = 8;         // MI355X: 8 accelerator complex dies, one L2 each
constexpr int kTileInlets = 64;  // inlet boundaries one tile may touch when the conveyance is fused

template <class T>
int dalloc(hg_ctx* ctx, T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
  if (e != hipSuccess) {
    ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return HG_ERR_HIP;
  }
  ctx->allocs.push_back(*p);
  return HG_OK;
}

// Frees a buffer dalloc gave (NULL: nothing), after the context's streams
// are idle.
template <class T>
int dfree(hg_ctx* ctx, T*& p) {
  if (!p) return HG_OK;
  for (hipStream_t st : {ctx->stream, ctx->stream2})
    if (st && hipStreamSynchronize(st) != hipSuccess) {
      ctx->err = "hipStreamSynchronize before a free failed";
      return HG_ERR_HIP;
    }
  auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), static_cast<void*>(p));
  if (it != ctx->allocs.end()) ctx->allocs.erase(it);
  const hipError_t e = hipFree(p);
  p = nullptr;
  if (e != hipSuccess) {
    ctx->err = std::string("hipFree: ") + hipGetErrorString(e);
    return HG_ERR_HIP;
  }
  return HG_OK;
}

template <class T>
int upload(hg_ctx* ctx, T* dst, const T* src, size_t n) {
  if (n == 0) return HG_OK;
  HG_TRY(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  HG_TRY(hipStreamSynchronize(ctx->stream));
  return HG_OK;
}

// --------------------------------------------------------------- kernels ---


// Inlet conveyance (bc_2D.jl:640-680): total_A_b = sum over the inlet's
// faces, in NodeString order, of ((L^(5/3) h_c) / n_c) * wet_c.  The terms
// are formed in parallel, the sum is a left fold by one lane (the
// reference's sequential order).
__global__ __launch_bounds__(kBlock) void k_inlet_conveyance(hg::InletArgs a) {
  __shared__ double term[kBlock];
  const int b = blockIdx.x;              // inlet over all fields
  const int fld = b / a.nin_m;           // its field
  const int bl = b - fld * a.nin_m;      // the inlet within the mesh
  const double* xi = a.xi + int64_t(fld) * a.Nm;
  const double* nf = a.n + int64_t(fld) * a.Nm;
  const int f0 = a.off[bl], f1 = a.off[bl + 1];
  double acc = 0.0;
  for (int base = f0; base < f1; base += kBlock) {
    const int f = base + threadIdx.x;
    if (f < f1) {
      const int c = a.cell[f];
      double h = xi[c] + a.hstill[c];
      if (h <= a.hsmall) h = a.hsmall;
      const double w = (h > a.hsmall) ? 1.0 : 0.0;
      term[threadIdx.x] = a.L53[f] * h / nf[c] * w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int cnt = min(kBlock, f1 - base);
      for (int k = 0; k < cnt; ++k) acc = (base + k == f0) ? term[k] : acc + term[k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.totalA[b] = acc;
    if (!(acc > 1e-10)) a.flag[0] = 1;
  }
}

// The new state of a step, stored as a relaxed store at device (agent)
// scope: the same vector store with the device-scope cache policy bit
// (sc1).  The next step's kernel reads these values, from any XCD; on the
// 1M-cell channel the step kernel measured about 1 % faster with it than
// with plain stores (the stored values are the same).
__device__ __forceinline__ void store_state(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Fused RHS (+ Euler step and dry clamp when STEP).  One work-item per cell.
template <bool STEP>
__global__ __launch_bounds__(kBlock) void k_swe_step(StepArgs a0) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;  // over the K fields' cells
  const int f = a0.K > 1 ? ig / a0.N : 0;
  const int i = ig - f * a0.N;
  if (f >= a0.K || i >= a0.N) return;
  const StepArgs a = hg::field_view(a0, f);
  const double hsmall = a.hsmall;
  const double xi = a.xi[i], hs = a.hstill[i];
  const double qx0 = a.qx[i], qy0 = a.qy[i];
  double h = xi + hs, qx = qx0, qy = qy0;  // semi_discretize_swe_2D.jl:101-106
  if (h <= hsmall) {
    h = hsmall;
    qx = 0.0;
    qy = 0.0;
  }
  const double zb = a.zb[i];
  const double nman = a.n[i];
  const Side Ls{xi, hs, h, qx, qy, zb};
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
  const int d = a.deg[i];
  for (int k = 0; k < d; ++k) {
    const int s = k * a.N + i;
    const int j = a.nbr[s];
    const double nx = a.fnx[s], ny = a.fny[s], L = a.flen[s];
    Side Rs;
    if (j >= 0) {
      const double xj = a.xi[j], hsj = a.hstill[j];
      double hj = xj + hsj, qxj = a.qx[j], qyj = a.qy[j];
      if (hj <= hsmall) {
        hj = hsmall;
        qxj = 0.0;
        qyj = 0.0;
      }
      Rs = Side{xj, hsj, hj, qxj, qyj, a.zb[j]};
    } else {  // ghost (bc_2D.jl:640-834), internal cell = this cell
      const int gi = ~j;
      double hg, qxg, qyg;
      switch (a.g_kind[gi]) {
        case HG_BC_INLETQ: {
          const int b = a.g_bc[gi];
          const double w = (h > hsmall) ? 1.0 : 0.0;
          const double v = a.inletQ[b] / a.totalA[b] * a.g_L23[gi] / nman;
          hg = h;
          qxg = -h * v * nx * w;
          qyg = -h * v * ny * w;
          break;
        }
        case HG_BC_EXITH: {
          const double hx = a.exitWSE[a.g_bc[gi]] - zb;
          hg = (hx > hsmall) ? hx : hsmall;
          qxg = qx;
          qyg = qy;
          break;
        }
        case HG_BC_WALL:
          hg = h;
          qxg = -qx;
          qyg = -qy;
          break;
        default: {  // HG_BC_SYMM
          const double vdn = qx * nx + qy * ny;
          hg = h;
          qxg = qx - 2.0 * vdn * nx;
          qyg = qy - 2.0 * vdn * ny;
          break;
        }
      }
      const double hsg = a.g_hstill[gi];
      Rs = Side{hg - hsg, hsg, hg, qxg, qyg, a.g_zb[gi]};
    }
    double f0, f1, f2;
    hg::roe_flux(Ls, Rs, nx, ny, a.g, a.half_g, hsmall, f0, f1, f2);
    acc0 = acc0 + f0 * L;
    acc1 = acc1 + f1 * L;
    acc2 = acc2 + f2 * L;
  }
  const double A = a.area[i];
  const double r0 = -acc0 / A, r1 = -acc1 / A, r2 = -acc2 / A;
  double s1 = 0.0, s2 = 0.0;
  if (h > hsmall) {  // :463-475
    double fx, fy;
    if (a.fmag)
      hg::friction_mag(a.fmag[i], qx, qy, fx, fy);
    else
      hg::friction(a.g, a.kn2, hsmall, h, qx, qy, nman, fx, fy);
    s1 = a.g * xi * a.S0x[i] - fx;
    s2 = a.g * xi * a.S0y[i] - fy;
  }
  const double d0 = r0 + 0.0, d1 = r1 + s1, d2 = r2 + s2;  // :264
  if (STEP) {  // custom_ODE_solvers.jl:16-26
    const double dti = a.dtc ? a.dtc[i] : a.dt;  // per-instance step of a batch, when set
    double xn = xi + dti * d0;
    double qxn = qx0 + dti * d1;
    double qyn = qy0 + dti * d2;
    const bool dry = xn < hsmall;
    if (dry) {
      xn = hsmall;
      qxn = 0.0;
      qyn = 0.0;
    }
    if (a.clamp) a.clamp[i] = dry ? 1 : 0;
    if (!(isfinite(xn) && isfinite(qxn) && isfinite(qyn))) a.flag[1] = 1;  // checked after the stepping call
    store_state(a.xo + i, xn);
    store_state(a.qxo + i, qxn);
    store_state(a.qyo + i, qyn);
  } else {
    a.xo[i] = d0;
    a.qxo[i] = d1;
    a.qyo[i] = d2;
  }
}

// Tiled face-once step (production path).  One work-group per tile of T
// consecutive cells:
//  1. every cell's state and derived quantities (u, v, sqrt(h+eps),
//     pressure) go to LDS; tiles that own inlet-Q ghosts also form the
//     inlet conveyance (sequential left fold over the inlet's faces);
//  2. work items: an intra-tile face is evaluated once and yields both
//     sides' fluxes (roe_pair<true>, bit-identical to the per-side
//     evaluation); cross-tile and boundary faces are evaluated from the
//     tile's side only; each result F * L lands in the side's LDS slot;
//  3. each cell sums its slots in the reference's local face order, adds
//     the source terms and applies the Euler update and dry clamp.
// T (cells per tile) and MD (slots per cell) are compile-time constants so
// every LDS address is a base register plus an immediate offset.
// Work item it of a tile whose items before ix0 carry a ref (see ItemTable).
__device__ __forceinline__ hg::ItemRec load_item(const hg::ItemTable& tb, int it, int ix0) {
  hg::ItemRec r;
  r.code = tb.code[it];
  r.ref = it < ix0 ? tb.ref[it] : 0;
  const hg::ItemGeo g = tb.geo[it];
  r.nx = g.nx;
  r.ny = g.ny;
  r.len = g.len;
  return r;
}

// This work-group's tile t and field fld (the launch covers every tile of
// each of K fields; see k_swe_tile).
__device__ __forceinline__ void tile_of_block(int K, int xcd_chunk, int& t, int& fld) {
  const int nt = int(gridDim.x) / K;
  const int Gc = xcd_chunk > 0 ? xcd_chunk : 1;
  const int blk = int(blockIdx.x);
  const int W = kXCD * Gc * K;
  const int w = blk / W, rr = blk - w * W;
  const int tw = w * kXCD * Gc;
  if (tw + kXCD * Gc <= nt) {
    const int q = rr / kXCD;
    t = tw + (rr % kXCD) * Gc + q % Gc;
    fld = q / Gc;
  } else {
    const int rem = nt - tw;
    t = tw + rr % rem;
    fld = rr / rem;
  }
}

template <bool STEP, int T, int MD>
__global__ __launch_bounds__(T) void k_swe_tile(hg::TileArgs a0) {
  extern __shared__ double lds[];
  // Work-groups are dealt round-robin to the 8 XCDs (block b -> XCD b % 8)
  // in launch order; the launch covers nt tiles of each of K fields.  Each
  // window of 8 G' consecutive Morton-ordered tiles (G' = max(G, 1), G =
  // xcd_chunk) and its K fields is split so that XCD x takes G' consecutive
  // tiles of it for every field: cross-tile neighbours then tend to share
  // that XCD's L2, and the K fields of a tile read its mesh data from that
  // L2.  A final partial window takes its tiles in turn for each field.
  // With K = 1, G = 0 is the identity.  Bijection on [0, K nt).
  int t, fld;
  tile_of_block(a0.K, a0.xcd_chunk, t, fld);
  const hg::TileArgs a = hg::field_view(a0, fld);
  const int li = threadIdx.x;
  const int i = t * T + li;
  const bool valid = i < a.N;
  double* C = lds;               // 10 x T cell quantities
  double* FL = lds + 10 * T;     // 3 x MD x T face contributions
  double* tA = FL + 3 * MD * T;  // conveyance of the tile's inlets (kTileInlets)
  const double hsmall = a.hsmall;

  double xi = 0.0, qx0 = 0.0, qy0 = 0.0, hs_i = 0.0, zb_i = 0.0;
  double n_i = 0.0, S0x_i = 0.0, S0y_i = 0.0, A_i = 1.0;
  int d_i = 0;
  // Wave priority: raised while a wave issues a phase's loads, lowered for
  // its arithmetic, so that the waves of the CU's other tiles get their
  // loads out ahead of a long flux phase (2.7 % off the 1M-cell step).
  __builtin_amdgcn_s_setprio(3);
  if (valid) {
    xi = a.xi[i];
    qx0 = a.qx[i];
    qy0 = a.qy[i];
    hs_i = a.hstill[i];
    zb_i = a.zb[i];
    n_i = a.n[i];
    S0x_i = a.S0x[i];
    S0y_i = a.S0y[i];
    A_i = a.area[i];
    d_i = a.deg ? a.deg[i] : MD;
  }
  hg::CellQ me{};
  if (valid) {
    me = hg::make_cellq(xi, hs_i, qx0, qy0, zb_i, a.half_g, hsmall);
    C[0 * T + li] = me.xi; C[1 * T + li] = me.hs; C[2 * T + li] = me.h; C[3 * T + li] = me.hu;
    C[4 * T + li] = me.hv; C[5 * T + li] = me.zb; C[6 * T + li] = me.u; C[7 * T + li] = me.v;
    C[8 * T + li] = me.s; C[9 * T + li] = me.p;
  }
  __builtin_amdgcn_s_setprio(0);
  if (a.n_inlet > 0 && li < 64) {  // wave 0: inlet conveyance (bc_2D.jl:640-680)
    const int s0 = a.tile_in_off[t], s1 = a.tile_in_off[t + 1];
    for (int sl = s0; sl < s1; ++sl) {
      const int b = a.tile_in_list[sl];
      const int f0 = a.in_off[b], f1 = a.in_off[b + 1];
      double acc = 0.0;
      for (int base = f0; base < f1; base += 64) {
        const int f = base + li;
        double term = 0.0;
        if (f < f1) {
          const int c = a.in_cell[f];
          double h = a.xi[c] + a.hstill[c];
          if (h <= hsmall) h = hsmall;
          const double w = (h > hsmall) ? 1.0 : 0.0;
          term = a.in_L53[f] * h / a.n[c] * w;
        }
        const int cnt = min(64, f1 - base);
        for (int k = 0; k < cnt; ++k) {
          const double tk = __shfl(term, k, 64);
          acc = (base + k == f0) ? tk : acc + tk;
        }
      }
      if (li == 0) {
        tA[sl - s0] = acc;
        if (a.totalA_out) a.totalA_out[b] = acc;  // every tile of inlet b writes the same value
        if (!(acc > 1e-10)) a.flag[0] = 1;
      }
    }
  }
  __syncthreads();

  const int it0 = a.items.off[t], it1 = a.items.off[t + 1], ix0 = a.items.x0[t];
  for (int it = it0 + li; it < it1; it += T) {
    __builtin_amdgcn_s_setprio(3);
    const hg::ItemRec rec = load_item(a.items, it, ix0);
    const int code = rec.code;
    const int ia = code & 1023;
    const int ka = (code >> 10) & 15;
    const int type = (code >> 28) & 3;
    double jxi = 0.0, jhs = 0.0, jqx = 0.0, jqy = 0.0, jzb = 0.0;
    if (type == 1) {  // the cross-tile neighbour's raw state, read at the raised priority too
      const int j = rec.ref;
      jxi = a.xi[j]; jhs = a.hstill[j]; jqx = a.qx[j]; jqy = a.qy[j]; jzb = a.zb[j];
    }
    __builtin_amdgcn_s_setprio(0);
    const double nx = rec.nx, ny = rec.ny, L = rec.len;
    hg::CellQ A;
    A.xi = C[0 * T + ia]; A.hs = C[1 * T + ia]; A.h = C[2 * T + ia]; A.hu = C[3 * T + ia];
    A.hv = C[4 * T + ia]; A.zb = C[5 * T + ia]; A.u = C[6 * T + ia]; A.v = C[7 * T + ia];
    A.s = C[8 * T + ia]; A.p = C[9 * T + ia];
    // Only the right-hand state depends on the item type; the flux itself
    // is one code path for all items (no divergent duplicates of it).
    hg::CellQ B;
    const int ib = (code >> 14) & 1023;
    const int kb = (code >> 24) & 15;
    if (type == 0) {
      B.xi = C[0 * T + ib]; B.hs = C[1 * T + ib]; B.h = C[2 * T + ib]; B.hu = C[3 * T + ib];
      B.hv = C[4 * T + ib]; B.zb = C[5 * T + ib]; B.u = C[6 * T + ib]; B.v = C[7 * T + ib];
      B.s = C[8 * T + ib]; B.p = C[9 * T + ib];
    } else if (type == 1) {
      B = hg::make_cellq(jxi, jhs, jqx, jqy, jzb, a.half_g, hsmall);
    } else {  // ghost (bc_2D.jl:640-834); its internal cell is A
      const int gi = rec.ref;
      double hg_, qxg, qyg;
      switch (a.g_kind[gi]) {
        case HG_BC_INLETQ: {
          const int b = a.g_bc[gi];
          const double w = (A.h > hsmall) ? 1.0 : 0.0;
          const double tot = (a.n_inlet > 0) ? tA[(code >> 14) & (kTileInlets - 1)] : a.totalA[b];
          const double v = a.inletQ[b] / tot * a.g_L23[gi] / a.n[t * T + ia];
          hg_ = A.h;
          qxg = -A.h * v * nx * w;
          qyg = -A.h * v * ny * w;
          break;
        }
        case HG_BC_EXITH: {
          const double hx = a.exitWSE[a.g_bc[gi]] - A.zb;
          hg_ = (hx > hsmall) ? hx : hsmall;
          qxg = A.hu;
          qyg = A.hv;
          break;
        }
        case HG_BC_WALL:
          hg_ = A.h;
          qxg = -A.hu;
          qyg = -A.hv;
          break;
        default: {
          const double vdn = A.hu * nx + A.hv * ny;
          hg_ = A.h;
          qxg = A.hu - 2.0 * vdn * nx;
          qyg = A.hv - 2.0 * vdn * ny;
          break;
        }
      }
      B.hs = a.g_hstill[gi];
      B.xi = hg_ - B.hs;
      B.h = hg_;
      B.hu = qxg;
      B.hv = qyg;
      B.zb = a.g_zb[gi];
      B.u = qxg / hg_;
      B.v = qyg / hg_;
      B.s = sqrt(hg_ + hg::EPS);
      B.p = a.half_g * ((B.xi + hg::EPS) * (B.xi + hg::EPS) + 2.0 * B.xi * B.hs);
    }
    double fa[3], fb[3];
    hg::roe_pair<true>(A, B, nx, ny, a.g, a.half_g, hsmall, fa, fb);
    if (type == 0)
      for (int c = 0; c < 3; ++c) FL[(c * MD + kb) * T + ib] = fb[c] * L;
    for (int c = 0; c < 3; ++c) FL[(c * MD + ka) * T + ia] = fa[c] * L;
  }
  __syncthreads();

  if (!valid) return;
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
  for (int k = 0; k < MD; ++k) {  // local face order (semi_discretize_swe_2D.jl:293-387)
    if (k < d_i) {
      acc0 = acc0 + FL[(0 * MD + k) * T + li];
      acc1 = acc1 + FL[(1 * MD + k) * T + li];
      acc2 = acc2 + FL[(2 * MD + k) * T + li];
    }
  }
  const double r0 = -acc0 / A_i, r1 = -acc1 / A_i, r2 = -acc2 / A_i;
  double s1 = 0.0, s2 = 0.0;
  const double h = me.h;
  if (h > hsmall) {
    double fx, fy;
    if (a.fmag)
      hg::friction_mag(a.fmag[i], me.hu, me.hv, fx, fy);
    else
      hg::friction(a.g, a.kn2, hsmall, h, me.hu, me.hv, n_i, fx, fy);
    s1 = a.g * xi * S0x_i - fx;
    s2 = a.g * xi * S0y_i - fy;
  }
  const double d0 = r0 + 0.0, d1 = r1 + s1, d2 = r2 + s2;
  if (STEP) {
    const double dti = a.dtc ? a.dtc[i] : a.dt;  // per-instance step of a batch, when set
    double xn = xi + dti * d0;
    double qxn = qx0 + dti * d1;
    double qyn = qy0 + dti * d2;
    const bool dry = xn < hsmall;
    if (dry) {
      xn = hsmall;
      qxn = 0.0;
      qyn = 0.0;
    }
    if (a.clamp) a.clamp[i] = dry ? 1 : 0;
    if (!(isfinite(xn) && isfinite(qxn) && isfinite(qyn))) a.flag[1] = 1;  // checked after the stepping call
    store_state(a.xo + i, xn);
    store_state(a.qxo + i, qxn);
    store_state(a.qyo + i, qyn);
  } else {
    a.xo[i] = d0;
    a.qxo[i] = d1;
    a.qyo[i] = d2;
  }
}

// update_bed_data (process_bed_2D.jl:46-66): zb_ghost and S0 from zb.
__global__ __launch_bounds__(kBlock) void k_bed_data(int N, int G, const double* zb, const int32_t* deg,
                                                    const int32_t* nbr, const double* fnx, const double* fny,
                                                    const double* flen, const double* area, const int32_t* g_cell,
                                                    double* S0x, double* S0y, double* g_zb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < G) g_zb[i] = zb[g_cell[i]];
  if (i >= N) return;
  const double z = zb[i];
  double gx = 0.0, gy = 0.0;
  for (int k = 0; k < deg[i]; ++k) {
    const int s = k * N + i;
    const int j = nbr[s];
    const double zf = (j >= 0) ? (z + zb[j]) * 0.5 : z;
    gx = gx + fnx[s] * zf * flen[s];
    gy = gy + fny[s] * zf * flen[s];
  }
  S0x[i] = -(gx / area[i]);
  S0y[i] = -(gy / area[i]);
}

// n_cell = params[zone] (process_ManningN_2D.jl:71-98).
__global__ __launch_bounds__(kBlock) void k_zone_gather(int N, const int32_t* zone, const double* zn, double* n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) n[i] = zn[zone[i]];
}

// CFL: per-block minimum of (2A/P)/(|u| + sqrt(g h)) over wet cells.
__global__ __launch_bounds__(kBlock) void k_cfl(int N, int Nm, const double* xi, const double* qx, const double* qy,
                                               const double* hstill, const int32_t* deg, const double* flen,
                                               const double* area, double g, double hsmall, double* part) {
  __shared__ double red[kBlock];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // over all fields; the mesh cell is i % Nm
  double v = INFINITY;
  if (i < N) {
    const int m = i % Nm;
    const double h = xi[i] + hstill[m];
    if (h > hsmall) {
      double P = 0.0;
      for (int k = 0; k < deg[m]; ++k) P = P + flen[k * Nm + m];
      const double u = qx[i] / h, w = qy[i] / h;
      const double spd = sqrt(u * u + w * w) + sqrt(g * h);
      v = ((2.0 * area[m]) / P) / spd;
    }
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out[0] = scale * min(part[0 .. n-1]), one work-group (the last stage of
// the CFL reduction on the device).
__global__ __launch_bounds__(kBlock) void k_min_parts(int n, const double* part, double scale, double* out) {
  __shared__ double red[kBlock];
  double v = INFINITY;
  for (int j = threadIdx.x; j < n; j += kBlock) v = fmin(v, part[j]);
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = scale * red[0];
}

// Caller order -> library order: d_c[i] = src[c N + perm[i]] for the
// ncomp (1 or 3) components.
__global__ __launch_bounds__(kBlock) void k_perm_in(int64_t N, int ncomp, const int32_t* __restrict__ perm,
                                                   const double* __restrict__ src, double* d0, double* d1,
                                                   double* d2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t p = perm[i];
  d0[i] = src[p];
  if (ncomp > 1) {
    d1[i] = src[N + p];
    d2[i] = src[2 * N + p];
  }
}

// Library order -> caller order: dst[c N + perm[i]] = s_c[i].
__global__ __launch_bounds__(kBlock) void k_perm_out(int64_t N, int ncomp, const int32_t* __restrict__ perm,
                                                    const double* s0, const double* s1, const double* s2,
                                                    double* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t p = perm[i];
  dst[p] = s0[i];
  if (ncomp > 1) {
    dst[N + p] = s1[i];
    dst[2 * N + p] = s2[i];
  }
}

// Water-surface residuals at the probe cells (library order):
// r_k = (xi + h_still) + zb - target_k (swe_2D_inversion.jl:387-399).
__global__ __launch_bounds__(kBlock) void k_probe_resid(int n, int Nm, const int32_t* cell, const double* target,
                                                       const double* xi, const double* hstill, const double* zb,
                                                       double* r) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int c = cell[k];  // over all fields; the mesh cell is c % Nm
  r[k] = xi[c] + hstill[c % Nm] + zb[c % Nm] - target[k];
}

// lam[cell[k]] = val[k] over distinct cells.
__global__ __launch_bounds__(kBlock) void k_seed(int n, const int32_t* cell, const double* val, double* lam) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) lam[cell[k]] = val[k];
}

// Plain streaming copy (16 B per lane per iteration): the reference kernel
// of known traffic used to calibrate rocprof's FETCH_SIZE / WRITE_SIZE and to
// measure the achievable HBM bandwidth.
__global__ __launch_bounds__(kBlock) void k_copy_probe(int64_t n2, const double2* __restrict__ src,
                                                      double2* __restrict__ dst) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n2; k += int64_t(gridDim.x) * blockDim.x)
    dst[k] = src[k];
}

// ------------------------------------------------------------- adjoint ----
// Reverse-mode adjoint of one RHS evaluation (STEP = false: the cotangent of
// dQ/dt is given) or of one Euler step with its dry clamp (STEP = true: the
// cotangent of the step output is given; cells the clamp reset pass
// nothing back).  One work-item per cell gathers, deterministically, the
// cotangent that flows into its own state:
//   * an internal face (i | j) enters RHS_i as -F L / A_i and, through the
//     exact mirror F(j | i, -n) = -F(i | j, n), RHS_j as +F L / A_j; the cell
//     takes the left-state VJP of F(i | j, n) with the combined cotangent
//     c = -(L / A_i) ybar_i + (L / A_j) ybar_j;
//   * a boundary face also routes the ghost-state VJP back through the
//     ghost's boundary rule (bc_2D.jl:640-834);
//   * the source term (bed slope, Manning friction) is the cell's own.
// Derivatives with respect to n, the exit-H bed term, S0 and the inlet
// ghost velocity are accumulated / written for the follow-up kernels.
// (One work-item per (cell, face) slot with a second pass for the sums was
// slower at every size: the slot values round-trip through memory.)
struct CellAdj {  // one cell's quantities the adjoint passes need
  hg::CellQ me;
  bool wet;
  double xi, zb, n, A, lt0, lt1, lt2, y0, y1, y2;
};

template <bool STEP>
__device__ __forceinline__ CellAdj cell_adj(const hg::AdjArgs& a, int i) {
  CellAdj c;
  c.xi = a.xi[i];
  c.zb = a.zb[i];
  c.n = a.n[i];
  c.A = a.area[i];
  c.me = hg::make_cellq(c.xi, a.hstill[i], a.qx[i], a.qy[i], c.zb, a.half_g, a.hsmall);
  c.wet = c.me.h > a.hsmall;
  c.lt0 = a.lx[i];
  c.lt1 = a.lqx[i];
  c.lt2 = a.lqy[i];
  if (STEP && a.mask && a.mask[i]) c.lt0 = c.lt1 = c.lt2 = 0.0;
  const double dti = a.dtc ? a.dtc[i] : a.dt;
  c.y0 = STEP ? dti * c.lt0 : c.lt0;
  c.y1 = STEP ? dti * c.lt1 : c.lt1;
  c.y2 = STEP ? dti * c.lt2 : c.lt2;
  return c;
}

// A boundary face's ghost state (bc_2D.jl:640-834) formed from its cell's
// flux inputs `me`, and what the routing of its cotangent back to the cell
// needs (the ghost's kind; inlet-Q: the ghost velocity v and wet flag w).
struct GhostAdj {
  int kind, bc;
  double v, w;
};

__device__ __forceinline__ hg::CellQ adj_ghost_state(const hg::AdjArgs& a, const hg::CellQ& me, double n_i, int gi,
                                                     double nx, double ny, GhostAdj& gs) {
  const double hsmall = a.hsmall;
  gs.kind = a.g_kind[gi];
  gs.bc = a.g_bc[gi];
  gs.v = 0.0;
  gs.w = 0.0;
  double hgh, qxg, qyg;
  switch (gs.kind) {
    case HG_BC_INLETQ: {
      gs.w = (me.h > hsmall) ? 1.0 : 0.0;
      gs.v = a.inletQ[gs.bc] / a.totalA[gs.bc] * a.g_L23[gi] / n_i;
      hgh = me.h;
      qxg = -me.h * gs.v * nx * gs.w;
      qyg = -me.h * gs.v * ny * gs.w;
      break;
    }
    case HG_BC_EXITH: {
      const double hx = a.exitWSE[gs.bc] - me.zb;
      hgh = (hx > hsmall) ? hx : hsmall;
      qxg = me.hu;
      qyg = me.hv;
      break;
    }
    case HG_BC_WALL:
      hgh = me.h;
      qxg = -me.hu;
      qyg = -me.hv;
      break;
    default: {
      const double vdn = me.hu * nx + me.hv * ny;
      hgh = me.h;
      qxg = me.hu - 2.0 * vdn * nx;
      qyg = me.hv - 2.0 * vdn * ny;
      break;
    }
  }
  hg::CellQ B;
  B.hs = a.g_hstill[gi];
  B.xi = hgh - B.hs;
  B.h = hgh;
  B.hu = qxg;
  B.hv = qyg;
  B.zb = a.g_zb[gi];
  B.u = qxg / hgh;
  B.v = qyg / hgh;
  B.s = sqrt(hgh + hg::EPS);
  B.p = a.half_g * ((B.xi + hg::EPS) * (B.xi + hg::EPS) + 2.0 * B.xi * B.hs);
  return B;
}

// Routes the ghost's cotangent Bb back through its boundary rule: adds to
// the cell's U, and its n / exit-H bed cotangents; writes the inlet ghost's
// velocity cotangent for the inlet conveyance adjoint.
__device__ __forceinline__ void adj_ghost_route(const hg::AdjArgs& a, const hg::CellQ& me, double n_i, int gi,
                                                const GhostAdj& gs, const hg::SideBar& Bb, double nx, double ny,
                                                hg::SideBar& U, double& nb, double& zbb) {
  const double hgb = Bb.h + Bb.xi;  // xi_g = h_g - hstill_g
  switch (gs.kind) {
    case HG_BC_INLETQ: {  // h_g = h, q_g = -h v n w, v = (Q / A) L^(2/3) / n
      const double v = gs.v, w = gs.w;
      U.h += hgb + Bb.hu * (-v * nx * w) + Bb.hv * (-v * ny * w);
      const double vb = Bb.hu * (-me.h * nx * w) + Bb.hv * (-me.h * ny * w);
      nb += vb * (-v / n_i);
      a.vbar[gi] = vb;
      break;
    }
    case HG_BC_EXITH: {  // h_g = max(h_small, WSE - zb_cell), q_g = q
      if (a.exitWSE[gs.bc] - me.zb > a.hsmall) zbb -= hgb;
      U.hu += Bb.hu;
      U.hv += Bb.hv;
      break;
    }
    case HG_BC_WALL:
      U.h += hgb;
      U.hu -= Bb.hu;
      U.hv -= Bb.hv;
      break;
    default: {
      const double vdnb = -2.0 * (Bb.hu * nx + Bb.hv * ny);
      U.h += hgb;
      U.hu += Bb.hu + vdnb * nx;
      U.hv += Bb.hv + vdnb * ny;
      break;
    }
  }
}

// The VJP of the boundary face in slot s (cell i = s % N, local face
// k = s / N) seen from cell i: its contribution to the cotangent of the
// cell's state (U), of n and of the exit-H bed, with the ghost's dependence
// on the cell routed back through its boundary rule.
template <bool STEP>
__device__ __forceinline__ void adj_ghost_face(const hg::AdjArgs& a, const CellAdj& ca, int64_t s, hg::SideBar& U,
                                               double& nb, double& zbb) {
  const int gi = ~a.nbr[s];
  const double nx = a.fnx[s], ny = a.fny[s], L = a.flen[s];
  const double fi = -(L / ca.A);
  nb = 0.0;
  zbb = 0.0;
  GhostAdj gs;
  const hg::CellQ B = adj_ghost_state(a, ca.me, ca.n, gi, nx, ny, gs);
  const double c[3] = {fi * ca.y0, fi * ca.y1, fi * ca.y2};
  hg::SideBar Bb;
  hg::roe_vjp(ca.me, B, nx, ny, a.g, a.half_g, a.hsmall, c, U, Bb);
  adj_ghost_route(a, ca.me, ca.n, gi, gs, Bb, nx, ny, U, nb, zbb);
}

template <bool STEP>
__device__ __forceinline__ void adj_cell_finish(const hg::AdjArgs& a, int i, const CellAdj& ca, double xib,
                                                double hb, double qxb, double qyb, double nb, double zbb);

// Reverse mode of one RHS evaluation or one step, one work-item per cell,
// over the cell's faces in local face order: of an internal face, the part
// of its VJP that falls on the cell (the flux cotangent combines both
// cells' output cotangents, since the neighbour's flux through the face is
// minus this one), accumulated over the cell's raw and derived quantities
// and folded once; of a boundary face, the VJP with the ghost's boundary
// rule routed back to the cell.  Then the source VJP, the clamp and the
// identity part of the step.
template <bool STEP>
__global__ __launch_bounds__(kBlock) void k_adj_cell(hg::AdjArgs a0) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;  // over the K fields' cells
  const int fld = a0.K > 1 ? ig / a0.N : 0;
  const int i = ig - fld * a0.N;
  if (fld >= a0.K || i >= a0.N) return;
  const hg::AdjArgs a = hg::field_view(a0, fld);
  const CellAdj ca = cell_adj<STEP>(a, i);
  hg::ExtBar e;
  hg::ext_zero(e);
  double gxi = 0.0, gh = 0.0, gqx = 0.0, gqy = 0.0, nb = 0.0, zbb = 0.0;  // the boundary faces' sums
  const int d = a.deg[i];
  const double hsmall = a.hsmall;
  const double rA = 1.0 / ca.A;
  for (int k = 0; k < d; ++k) {  // local face order
    const int64_t s = int64_t(k) * a.N + i;
    const int j = a.nbr[s];
    if (j < 0) {
      hg::SideBar U;
      double fn, fz;
      adj_ghost_face<STEP>(a, ca, s, U, fn, fz);
      gxi += U.xi;
      gh += U.h;
      gqx += U.hu;
      gqy += U.hv;
      nb += fn;
      zbb += fz;
      continue;
    }
    const double nx = a.fnx[s], ny = a.fny[s], L = a.flen[s];
    const hg::CellQ B = hg::make_cellq(a.xi[j], a.hstill[j], a.qx[j], a.qy[j], a.zb[j], a.half_g, hsmall);
    double m0 = a.lx[j], m1 = a.lqx[j], m2 = a.lqy[j];
    if (STEP && a.mask && a.mask[j]) m0 = m1 = m2 = 0.0;
    const double dtj = a.dtc ? a.dtc[j] : a.dt;
    const double fj = L / a.area[j], fi = -(L * rA);
    const double c[3] = {fi * ca.y0 + fj * (STEP ? dtj * m0 : m0), fi * ca.y1 + fj * (STEP ? dtj * m1 : m1),
                         fi * ca.y2 + fj * (STEP ? dtj * m2 : m2)};
    hg::roe_vjp_a(ca.me, B, nx, ny, a.g, a.half_g, hsmall, c, e);
  }
  const hg::SideBar U = hg::ext_to_side(e, ca.me, a.half_g);
  adj_cell_finish<STEP>(a, i, ca, U.xi + gxi, U.h + gh, U.hu + gqx, U.hv + gqy, nb, zbb);
}

// What the finish reads of one cell besides its flux inputs and output
// cotangent: n, friction magnitude, S0, and the sums it adds to.
struct FinishIn {
  double n, fmag, S0x, S0y, fbar, nbar, zbbar, S0bx, S0by;
};

__device__ __forceinline__ FinishIn adj_finish_load(const hg::AdjArgs& a, int i) {
  FinishIn f;
  f.n = a.n[i];
  f.fmag = a.fmag ? a.fmag[i] : 0.0;
  f.S0x = a.S0x[i];
  f.S0y = a.S0y[i];
  f.fbar = (a.fmag && !a.fbar_set) ? a.fbar[i] : 0.0;
  f.nbar = (a.nbar && !a.nbar_set) ? a.nbar[i] : 0.0;
  f.zbbar = a.zbbar ? a.zbbar[i] : 0.0;
  f.S0bx = a.S0bar ? a.S0bar[i] : 0.0;
  f.S0by = a.S0bar ? a.S0bar[a.N + i] : 0.0;
  return f;
}

// The source VJP, the clamp and the identity part of the step, and the
// writes of one cell, given its summed face cotangents and its FinishIn.
template <bool STEP>
__device__ __forceinline__ void adj_finish_store(const hg::AdjArgs& a, int i, const CellAdj& ca, const FinishIn& f,
                                                 double xib, double hb, double qxb, double qyb, double nb,
                                                 double zbb) {
  const hg::CellQ& me = ca.me;
  double S0xb = 0.0, S0yb = 0.0;
  if (a.fmag) {
    double mb = 0.0;
    if (ca.wet)
      hg::source_vjp_mag(a.g, ca.xi, me.hu, me.hv, f.fmag, f.S0x, f.S0y, ca.y1, ca.y2, xib, qxb, qyb, mb, S0xb,
                         S0yb);
    // this step's cotangent of the friction magnitude: set (the sum with 0.0
    // gives the bits of an add into a cleared array), or summed over the steps
    if (a.fbar_set)
      a.fbar[i] = mb + 0.0;
    else if (ca.wet)
      a.fbar[i] = f.fbar + mb;
  } else if (ca.wet) {
    hg::source_vjp(a.g, a.kn2, a.hsmall, ca.xi, me.h, me.hu, me.hv, f.n, f.S0x, f.S0y, ca.y1, ca.y2, xib, hb, qxb,
                   qyb, nb, S0xb, S0yb);
  }
  // clamped depth / masked discharge are constants (semi_discretize_swe_2D.jl:101-106)
  double ox = xib, oqx = 0.0, oqy = 0.0;
  if (ca.wet) {
    ox += hb;
    oqx = qxb;
    oqy = qyb;
  }
  if (STEP) {
    ox += ca.lt0;
    oqx += ca.lt1;
    oqy += ca.lt2;
  }
  a.ox[i] = ox;
  a.oqx[i] = oqx;
  a.oqy[i] = oqy;
  if (a.nbar) a.nbar[i] = a.nbar_set ? nb + 0.0 : f.nbar + nb;
  if (a.zbbar) a.zbbar[i] = f.zbbar + zbb;
  if (a.S0bar) {  // summed over the steps; the bed-slope transpose is applied once at the end
    a.S0bar[i] = f.S0bx + S0xb;
    a.S0bar[a.N + i] = f.S0by + S0yb;
  }
}

template <bool STEP>
__device__ __forceinline__ void adj_cell_finish(const hg::AdjArgs& a, int i, const CellAdj& ca, double xib,
                                                double hb, double qxb, double qyb, double nb, double zbb) {
  adj_finish_store<STEP>(a, i, ca, adj_finish_load(a, i), xib, hb, qxb, qyb, nb, zbb);
}

// Tiled reverse step (production path on meshes the tiled step kernel
// runs).  One work-group per tile of T cells, over the step kernel's work
// items:
//  1. every cell's flux inputs (make_cellq; the pressure is re-formed from
//     xi and h_still where read), its output cotangent times dt over its
//     area (z = y / A, y zero where the step's clamp fired), 1 / h and
//     0.5 / sqrt(h + eps) go to LDS;
//  2. each item forms one face's VJP: an intra-tile face once for both
//     cells, a cross-tile face (neighbour gathered from memory) for the
//     tile's cell only, a boundary face with its ghost formed from the cell
//     and its cotangent routed back through the boundary rule.  The face
//     cotangent combines both cells' outputs, c = L (z_b - z_a).  An
//     internal face is evaluated from the side whose normal points to +x
//     (or along +y when n_x = 0), so both cells receive the same bits
//     whichever work-group evaluates the face: the results do not depend on
//     the tiling, the numbering or the launch shape.  Each side's
//     cotangent lands in its (cell, local face) LDS slot as (xi + h for a
//     wet cell, xi for a dry one -- the step's clamp drops h-bar there),
//     hu, hv and the boundary face's n or exit-H bed term (a per-cell mask
//     marks exit-H slots);
//  3. each cell sums its slots in local face order and finishes as
//     k_adj_cell does (source VJP, identity part of the step).
// LDS: (13 + 4 MD) doubles and one int per cell, 51 KiB for 256 triangles:
// three work-groups per CU, as many as the registers allow.  (At 53 KiB and
// more per work-group only two of them run on a CU at once on MI355X.)
constexpr int kAdjRows = 13;  // CellQ without h and p (8), z (3), 1 / h, 0.5 / s

size_t adj_tile_lds_bytes(int T, int MD) {
  return size_t(kAdjRows * T + 4 * MD * T) * sizeof(double) + size_t(T) * sizeof(int32_t);
}

// Rows: 0 xi, 1 h_still, 2 hu, 3 hv, 4 zb, 5 u, 6 v, 7 s, 8-10 z, 11 1 / h,
// 12 0.5 / s.  The depth and the pressure are re-formed where read, with
// make_cellq's operations.
template <int T>
__device__ __forceinline__ hg::CellQ lds_cellq(const double* C, int r, double half_g, double hsmall) {
  hg::CellQ q;
  q.xi = C[0 * T + r]; q.hs = C[1 * T + r]; q.hu = C[2 * T + r]; q.hv = C[3 * T + r];
  q.zb = C[4 * T + r]; q.u = C[5 * T + r]; q.v = C[6 * T + r]; q.s = C[7 * T + r];
  double h = q.xi + q.hs;
  if (h <= hsmall) h = hsmall;
  q.h = h;
  q.p = half_g * ((q.xi + hg::EPS) * (q.xi + hg::EPS) + 2.0 * q.xi * q.hs);
  return q;
}

// A cell's flux inputs as the tiled reverse step forms them, for the tile's
// own cells and for cross-tile neighbours alike (so a face's VJP has the same
// bits whichever work-group evaluates it): make_cellq's clamp and pressure,
// with u and v from one reciprocal of h and sqrt(h + eps) and its reciprocal
// from one rsqrt.  rh = 1 / h, r2s = 0.5 / sqrt(h + eps).
__device__ __forceinline__ hg::CellQ adj_cellq(double xi, double hs, double qx, double qy, double zb, double half_g,
                                               double hsmall, double& rh, double& r2s) {
  hg::CellQ c;
  c.xi = xi;
  c.hs = hs;
  double h = xi + hs, hu = qx, hv = qy;
  if (h <= hsmall) {
    h = hsmall;
    hu = 0.0;
    hv = 0.0;
  }
  c.h = h;
  c.hu = hu;
  c.hv = hv;
  c.zb = zb;
  rh = 1.0 / h;
  c.u = hu * rh;
  c.v = hv * rh;
  const double he = h + hg::EPS;
  const double rs = rsqrt(he);
  c.s = he * rs;
  r2s = 0.5 * rs;
  c.p = half_g * ((xi + hg::EPS) * (xi + hg::EPS) + 2.0 * xi * hs);
  return c;
}

// A cross-tile neighbour's raw data as a work item of the tiled reverse step
// reads it.
struct NbrRaw {
  double xi, hs, qx, qy, zb, m0, m1, m2, dt, area;
  uint8_t clamped;
};

template <bool STEP>
__device__ __forceinline__ NbrRaw nbr_load(const hg::AdjArgs& a, int j) {
  NbrRaw r;
  r.xi = a.xi[j]; r.hs = a.hstill[j]; r.qx = a.qx[j]; r.qy = a.qy[j]; r.zb = a.zb[j];
  r.m0 = a.lx[j]; r.m1 = a.lqx[j]; r.m2 = a.lqy[j];
  r.clamped = (STEP && a.mask) ? a.mask[j] : uint8_t(0);
  r.dt = a.dtc ? a.dtc[j] : a.dt;
  r.area = a.area[j];
  return r;
}

// One work item of the tiled reverse step (see k_adj_tile); nr is the
// neighbour of a cross-tile item.
template <bool STEP, int T, int MD>
__device__ __forceinline__ void adj_item(const hg::AdjArgs& a, int t, const double* C, double* S, int32_t* Xm,
                                         const hg::ItemRec& rec, const NbrRaw& nr) {
  const double hsmall = a.hsmall, half_g = a.half_g;
  const int code = rec.code;
  const int ia = code & 1023, ka = (code >> 10) & 15;
  const int ib = (code >> 14) & 1023, kb = (code >> 24) & 15;
  const int type = (code >> 28) & 3;
  const double nx = rec.nx, ny = rec.ny, L = rec.len;
  // P / Q: the evaluation's first / second side (normal from P)
  const bool sw = type != 2 && !(nx > 0.0 || (nx == 0.0 && ny > 0.0));
  const double ex = sw ? -nx : nx, ey = sw ? -ny : ny;
  hg::CellQ P, Q;
  double zP0, zP1, zP2, zQ0, zQ1, zQ2, rhP, rhQ, r2P, r2Q;
  int ip = ia, kp = ka, iq = ia, kq = ka;  // LDS slots of P and Q (type 0); the tile's cell otherwise
  GhostAdj gs{};
  double n_a = 0.0;
  if (type == 0) {
    if (sw) { ip = ib; kp = kb; } else { iq = ib; kq = kb; }
    P = lds_cellq<T>(C, ip, half_g, hsmall);
    Q = lds_cellq<T>(C, iq, half_g, hsmall);
    zP0 = C[8 * T + ip]; zP1 = C[9 * T + ip]; zP2 = C[10 * T + ip];
    zQ0 = C[8 * T + iq]; zQ1 = C[9 * T + iq]; zQ2 = C[10 * T + iq];
    rhP = C[11 * T + ip]; r2P = C[12 * T + ip];
    rhQ = C[11 * T + iq]; r2Q = C[12 * T + iq];
  } else {
    const hg::CellQ Lc = lds_cellq<T>(C, ia, half_g, hsmall);
    const double zL0 = C[8 * T + ia], zL1 = C[9 * T + ia], zL2 = C[10 * T + ia];
    const double rhL = C[11 * T + ia], r2L = C[12 * T + ia];
    hg::CellQ G;
    double zG0 = 0.0, zG1 = 0.0, zG2 = 0.0;
    double rhG, r2G;
    if (type == 1) {  // cross-tile neighbour
      G = adj_cellq(nr.xi, nr.hs, nr.qx, nr.qy, nr.zb, half_g, hsmall, rhG, r2G);
      double m0 = nr.m0, m1 = nr.m1, m2 = nr.m2;
      if (nr.clamped) m0 = m1 = m2 = 0.0;
      const double rAj = 1.0 / nr.area;
      zG0 = (STEP ? nr.dt * m0 : m0) * rAj;
      zG1 = (STEP ? nr.dt * m1 : m1) * rAj;
      zG2 = (STEP ? nr.dt * m2 : m2) * rAj;
    } else {  // ghost (no cotangent of its own)
      n_a = a.n[t * T + ia];
      G = adj_ghost_state(a, Lc, n_a, rec.ref, nx, ny, gs);
      rhG = 1.0 / G.h;
      r2G = 0.5 / G.s;
    }
    P = sw ? G : Lc;
    Q = sw ? Lc : G;
    zP0 = sw ? zG0 : zL0; zP1 = sw ? zG1 : zL1; zP2 = sw ? zG2 : zL2;
    zQ0 = sw ? zL0 : zG0; zQ1 = sw ? zL1 : zG1; zQ2 = sw ? zL2 : zG2;
    rhP = sw ? rhG : rhL; r2P = sw ? r2G : r2L;
    rhQ = sw ? rhL : rhG; r2Q = sw ? r2L : r2G;
  }
  const double c[3] = {L * zQ0 - L * zP0, L * zQ1 - L * zP1, L * zQ2 - L * zP2};
  hg::ExtBar eP, eQ;
  hg::roe_vjp_ext(P, Q, ex, ey, a.g, half_g, hsmall, c, eP, eQ);
  hg::SideBar sP = hg::ext_to_side_r(eP, P, rhP, r2P, half_g);
  const hg::SideBar sQ = hg::ext_to_side_r(eQ, Q, rhQ, r2Q, half_g);
  if (type == 2) {
    double nb = 0.0, zbb = 0.0;
    adj_ghost_route(a, P, n_a, rec.ref, gs, sQ, nx, ny, sP, nb, zbb);
    if (gs.kind == HG_BC_EXITH) {
      S[(3 * MD + ka) * T + ia] = zbb;
      atomicOr(&Xm[ia], 1 << ka);
    } else {
      S[(3 * MD + ka) * T + ia] = nb;
    }
  }
  if (type == 0 || !sw) {
    S[(0 * MD + kp) * T + ip] = (P.h > hsmall) ? sP.xi + sP.h : sP.xi;
    S[(1 * MD + kp) * T + ip] = sP.hu;
    S[(2 * MD + kp) * T + ip] = sP.hv;
  }
  if (type == 0 || sw) {
    S[(0 * MD + kq) * T + iq] = (Q.h > hsmall) ? sQ.xi + sQ.h : sQ.xi;
    S[(1 * MD + kq) * T + iq] = sQ.hu;
    S[(2 * MD + kq) * T + iq] = sQ.hv;
  }
}

template <bool STEP, int T, int MD>
__global__ __launch_bounds__(T, 3) void k_adj_tile(hg::AdjArgs a0) {
  extern __shared__ double lds[];
  int t, fld;
  tile_of_block(a0.K, a0.xcd_chunk, t, fld);
  const hg::AdjArgs a = hg::field_view(a0, fld);
  const int li = threadIdx.x;
  const int i = t * T + li;
  const bool valid = i < a.N;
  double* C = lds;                                         // kAdjRows x T
  double* S = lds + kAdjRows * T;                          // 4 x MD x T: per slot xi(+h), hu, hv, n or zb
  int32_t* Xm = reinterpret_cast<int32_t*>(S + 4 * MD * T);  // T: exit-H slots of each cell
  const double hsmall = a.hsmall;
  // this work-item's first two work items, and the neighbour of a cross-tile
  // first item, are read with the cell data: one round of memory latency
  const int it0 = a.items.off[t], it1 = a.items.off[t + 1], ix0 = a.items.x0[t];
  const int itA = it0 + li, itB = itA + T;
  hg::ItemRec recA{}, recB{};
  NbrRaw nbrA{};
  __builtin_amdgcn_s_setprio(3);  // raised while the loads go out, as in k_swe_tile
  if (itA < it1) {
    recA = load_item(a.items, itA, ix0);
    if (((recA.code >> 28) & 3) == 1) nbrA = nbr_load<STEP>(a, recA.ref);
  }
  if (itB < it1) recB = load_item(a.items, itB, ix0);
  int d_i = 0;
  if (valid) {
    double rh, r2s;
    const hg::CellQ me = adj_cellq(a.xi[i], a.hstill[i], a.qx[i], a.qy[i], a.zb[i], a.half_g, hsmall, rh, r2s);
    double l0 = a.lx[i], l1 = a.lqx[i], l2 = a.lqy[i];
    if (STEP && a.mask && a.mask[i]) l0 = l1 = l2 = 0.0;
    const double dti = a.dtc ? a.dtc[i] : a.dt;
    const double rA = 1.0 / a.area[i];
    C[0 * T + li] = me.xi; C[1 * T + li] = me.hs; C[2 * T + li] = me.hu; C[3 * T + li] = me.hv;
    C[4 * T + li] = me.zb; C[5 * T + li] = me.u; C[6 * T + li] = me.v; C[7 * T + li] = me.s;
    C[8 * T + li] = (STEP ? dti * l0 : l0) * rA;
    C[9 * T + li] = (STEP ? dti * l1 : l1) * rA;
    C[10 * T + li] = (STEP ? dti * l2 : l2) * rA;
    C[11 * T + li] = rh;
    C[12 * T + li] = r2s;
    d_i = a.deg ? a.deg[i] : MD;
    for (int k = 0; k < MD; ++k) S[(3 * MD + k) * T + li] = 0.0;
    Xm[li] = 0;
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();

  if (itA < it1) adj_item<STEP, T, MD>(a, t, C, S, Xm, recA, nbrA);
  for (int it = itB; it < it1; it += T) {
    hg::ItemRec rec;
    __builtin_amdgcn_s_setprio(3);
    if (it == itB)
      rec = recB;
    else
      rec = load_item(a.items, it, ix0);
    NbrRaw nr{};
    if (((rec.code >> 28) & 3) == 1) nr = nbr_load<STEP>(a, rec.ref);
    __builtin_amdgcn_s_setprio(0);
    adj_item<STEP, T, MD>(a, t, C, S, Xm, rec, nr);
  }
  // the finish's own loads, in flight while the other waves finish their items
  FinishIn fin{};
  double lt0 = 0.0, lt1 = 0.0, lt2 = 0.0, dti = 0.0;
  __builtin_amdgcn_s_setprio(3);
  if (valid) {
    fin = adj_finish_load(a, i);
    lt0 = a.lx[i];
    lt1 = a.lqx[i];
    lt2 = a.lqy[i];
    if (STEP && a.mask && a.mask[i]) lt0 = lt1 = lt2 = 0.0;
    dti = a.dtc ? a.dtc[i] : a.dt;
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();

  if (!valid) return;
  double uxh = 0.0, uqx = 0.0, uqy = 0.0, nb = 0.0, zbb = 0.0;
  const int xm = Xm[li];
  for (int k = 0; k < MD; ++k) {  // local face order
    if (k < d_i) {
      uxh += S[(0 * MD + k) * T + li];
      uqx += S[(1 * MD + k) * T + li];
      uqy += S[(2 * MD + k) * T + li];
      const double e = S[(3 * MD + k) * T + li];
      if ((xm >> k) & 1)
        zbb += e;
      else
        nb += e;
    }
  }
  CellAdj ca;
  ca.me = lds_cellq<T>(C, li, a.half_g, hsmall);
  ca.wet = ca.me.h > hsmall;
  ca.xi = ca.me.xi;
  ca.zb = ca.me.zb;
  ca.n = fin.n;
  ca.A = 1.0;  // not used by the finish
  ca.lt0 = lt0;
  ca.lt1 = lt1;
  ca.lt2 = lt2;
  ca.y0 = STEP ? dti * lt0 : lt0;
  ca.y1 = STEP ? dti * lt1 : lt1;
  ca.y2 = STEP ? dti * lt2 : lt2;
  // the slots carry xi + h already (h-bar only where the cell is wet)
  adj_finish_store<STEP>(a, i, ca, fin, uxh, 0.0, uqx, uqy, nb, zbb);
}

// Inlet conveyance adjoint (bc_2D.jl:640-694): v_f = (Q_b / A_b) L^(2/3) / n
// and A_b = sum_f (L^(5/3) h / n) w.  One work-item walks every inlet's
// faces in NodeString order (deterministic; a cell may own several inlet
// faces), adds d/dh of A_b to the state cotangent and accumulates d/dn and
// d/dQ_b.
__global__ __launch_bounds__(kBlock) void k_inlet_adj(int n_inlet, int nin_m, int Nm, int Gm, const int32_t* off,
                                                     const hg::InletFaceRec* rec, const double* xi_all,
                                                     const double* hstill, const double* n_all, const double* inletQ,
                                                     const double* totalA, const double* vbar_all, double hsmall,
                                                     double* ox_all, double* nbar_all, double* Qbar) {
  // Per inlet: the per-face terms are formed in parallel and summed by a
  // fixed tree; each cell's faces of the inlet are applied, in NodeString
  // order, by the work-item of its first face (rec.first; rec.next: the
  // cell's next face in the inlet, or -1).  A cell's faces all belong to it,
  // so its depth and n are read once.  A work-item reads the record, cell data
  // and ghost cotangent of its first face, and for a cell's first face the
  // cell's cotangent values and its next face's record, before the sums; each
  // cotangent value is then read once and written once.  With one
  // work-group, the inlets are taken in order (a cell may sit on two of
  // them); when no cell does, work-group b takes inlet b alone (same
  // arithmetic, same bits).
  static_assert(kBlock == 256, "the tree below is written for 256 work-items");
  __shared__ double sA[kBlock], sQ[kBlock];
  const int tid = threadIdx.x;
  const int b0 = gridDim.x > 1 ? int(blockIdx.x) : 0;
  const int b1 = gridDim.x > 1 ? int(blockIdx.x) + 1 : n_inlet;
  for (int b = b0; b < b1; ++b) {  // inlet over all fields: field fl, inlet bl within the mesh
    const int fl = b / nin_m, bl = b - fl * nin_m;
    const int64_t o = int64_t(fl) * Nm;
    const double* xi = xi_all + o;
    const double* n = n_all + o;
    const double* vbar = vbar_all + int64_t(fl) * Gm;
    double* ox = ox_all + o;
    double* nbar = nbar_all ? nbar_all + o : nullptr;
    const double A = totalA[b], Qb = inletQ[b];
    const int f0 = off[bl], f1 = off[bl + 1];
    const int fmine = f0 + tid;
    const double qbar0 = tid == 0 ? Qbar[b] : 0.0;
    hg::InletFaceRec r0{}, rn{};
    double n0 = 1.0, h0 = 0.0, ox0 = 0.0, nb0 = 0.0;
    double pa = 0.0, pq = 0.0;
    if (fmine < f1) {
      r0 = rec[fmine];
      const int c = r0.cell;
      n0 = n[c];
      h0 = xi[c] + hstill[c];
      if (r0.first) {
        ox0 = ox[c];
        if (nbar) nb0 = nbar[c];
        if (r0.next >= 0) rn = rec[r0.next];
      }
      const double vb = vbar[r0.ghost];
      const double v = Qb / A * r0.L23 / n0;
      pa += vb * (-v / A);
      pq += vb * (r0.L23 / (A * n0));
    }
    for (int f = fmine + kBlock; f < f1; f += kBlock) {
      const hg::InletFaceRec r = rec[f];
      const double nc = n[r.cell];
      const double vb = vbar[r.ghost];
      const double v = Qb / A * r.L23 / nc;
      pa += vb * (-v / A);
      pq += vb * (r.L23 / (A * nc));
    }
    // pairwise tree: level w adds entry t + w into entry t (t < w), w = 128 .. 1;
    // the levels below 64 stay in wave 0 and run as lane exchanges
    sA[tid] = pa;
    sQ[tid] = pq;
    __syncthreads();
    if (tid < 128) {
      sA[tid] += sA[tid + 128];
      sQ[tid] += sQ[tid + 128];
    }
    __syncthreads();
    if (tid < 64) {
      double xa = sA[tid] + sA[tid + 64], xq = sQ[tid] + sQ[tid + 64];
      for (int w = 32; w > 0; w >>= 1) {
        xa = xa + __shfl_xor(xa, w, 64);
        xq = xq + __shfl_xor(xq, w, 64);
      }
      if (tid == 0) {
        sA[0] = xa;
        Qbar[b] = qbar0 + xq;
      }
    }
    __syncthreads();
    const double Ab = sA[0];
    if (fmine < f1 && r0.first && h0 > hsmall) {  // else w = 0 and h clamped: no derivative
      double ov = ox0, nv = nb0;
      ov += Ab * r0.L53 / n0;
      if (nbar) nv += Ab * (-r0.L53 * h0 / (n0 * n0));
      for (int g = r0.next; g >= 0;) {
        const hg::InletFaceRec rg = g == r0.next ? rn : rec[g];
        ov += Ab * rg.L53 / n0;
        if (nbar) nv += Ab * (-rg.L53 * h0 / (n0 * n0));
        g = rg.next;
      }
      ox[r0.cell] = ov;
      if (nbar) nbar[r0.cell] = nv;
    }
    for (int f = fmine + kBlock; f < f1; f += kBlock) {  // inlets of more than 256 faces
      const hg::InletFaceRec r = rec[f];
      if (!r.first) continue;
      const int c = r.cell;
      const double h = xi[c] + hstill[c], nc = n[c];
      if (!(h > hsmall)) continue;
      double ov = ox[c], nv = nbar ? nbar[c] : 0.0;
      for (int g = f; g >= 0;) {
        const hg::InletFaceRec rg = g == f ? r : rec[g];
        ov += Ab * rg.L53 / nc;
        if (nbar) nv += Ab * (-rg.L53 * h / (nc * nc));
        g = rg.next;
      }
      ox[c] = ov;
      if (nbar) nbar[c] = nv;
    }
    __syncthreads();  // sA / sQ reuse; cells shared by two inlets take them in inlet order
  }
}

// Transpose of the Green-Gauss bed slope (process_bed_2D.jl:46-66,
// fvm_schemes_2D.jl:89-167): zbbar_m += sum over its faces of
//   -(w L / A_m) (n . S0bar_m) + [internal] (L / (2 A_j)) (n . S0bar_j),
// w = 1/2 on internal faces, 1 on boundary faces.
__global__ __launch_bounds__(kBlock) void k_bed_adj(int Ntot, int N, const int32_t* deg, const int32_t* nbr,
                                                   const double* fnx, const double* fny, const double* flen,
                                                   const double* area, const double* S0bar_all, double* zbbar_all) {
  const int mg = blockIdx.x * blockDim.x + threadIdx.x;  // over the fields' cells; N cells per field
  if (mg >= Ntot) return;
  const int fl = mg / N, m = mg - fl * N;
  const double* S0bar = S0bar_all + 2 * int64_t(fl) * N;  // field fl: [x: N][y: N]
  double* zbbar = zbbar_all + int64_t(fl) * N;
  const double sx = S0bar[m], sy = S0bar[N + m], Am = area[m];
  double acc = 0.0;
  for (int k = 0; k < deg[m]; ++k) {
    const int s = k * N + m;
    const int j = nbr[s];
    const double nx = fnx[s], ny = fny[s], L = flen[s];
    const double w = (j >= 0) ? 0.5 : 1.0;
    acc += -(w * L / Am) * (nx * sx + ny * sy);
    if (j >= 0) acc += (0.5 * L / area[j]) * (nx * S0bar[j] + ny * S0bar[N + j]);
  }
  zbbar[m] += acc;
}

// ------------------------------------------------------------ host helpers --

uint64_t spread21(uint64_t v) {
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// Morton (Z-order) renumbering of the cell centroids, separately within
// each segment [seg[s], seg[s+1]) so that batched instances stay contiguous.
std::vector<int32_t> morton_perm(const std::vector<int64_t>& seg, const double* xy) {
  const int64_t N = seg.back();
  std::vector<int32_t> perm(N);
  std::vector<std::pair<uint64_t, int32_t>> key;
  for (size_t sg = 0; sg + 1 < seg.size(); ++sg) {
    const int64_t a = seg[sg], b = seg[sg + 1];
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int64_t i = a; i < b; ++i)
      for (int d = 0; d < 2; ++d) {
        lo[d] = std::min(lo[d], xy[2 * i + d]);
        hi[d] = std::max(hi[d], xy[2 * i + d]);
      }
    const double span = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), 1e-300);
    const double scale = double((1u << 21) - 1) / span;
    key.resize(b - a);
    for (int64_t i = a; i < b; ++i) {
      uint64_t qx = uint64_t(std::min(std::max((xy[2 * i] - lo[0]) * scale, 0.0), double((1u << 21) - 1)));
      uint64_t qy = uint64_t(std::min(std::max((xy[2 * i + 1] - lo[1]) * scale, 0.0), double((1u << 21) - 1)));
      key[i - a] = {spread21(qx) | (spread21(qy) << 1), int32_t(i)};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint64_t, int32_t>& x, const std::pair<uint64_t, int32_t>& y) {
                       return x.first < y.first;
                     });
    for (int64_t i = a; i < b; ++i) perm[i] = key[i - a].second;
  }
  return perm;
}

// ---- UDE closure for Manning's n (matrix cores; one wave = 64 cells) ----

// Hidden layers of 64 cells: features F[64][4] (column 3 = 1) -> A1 =
// tanh(F [W1 | b1]^T) -> A2 = tanh(A1 W2^T + b2), each layer as 16-row blocks
// of 16x16x4 matrix-core products (b2 as their initial accumulator); results
// land in LDS row-major [cell][unit].
template <class T>
__device__ __forceinline__ void ude_hidden(const T* __restrict__ W, const T* F, T* A1, T* A2, int l) {
  using M = hg::Mfma<T>;
  const int col = l & 15, kq = l >> 4;
  for (int rb = 0; rb < 4; ++rb) {
    const T av = F[(16 * rb + col) * 4 + kq];                                    // F[16 rb + i][k]
    const T bv = kq < 3 ? W[hg::kUdeW1 + col * 3 + kq] : W[hg::kUdeB1 + col];  // [W1 | b1][j][k]
    typename M::V acc = {T(0), T(0), T(0), T(0)};
    acc = M::op(av, bv, acc);
    for (int r = 0; r < 4; ++r) A1[(16 * rb + M::row(kq, r)) * 16 + col] = M::tanh_(acc[r]);
  }
  hg::wave_sync();
  const T b2 = W[hg::kUdeB2 + col];
  for (int rb = 0; rb < 4; ++rb) {
    typename M::V acc = {b2, b2, b2, b2};
    for (int kc = 0; kc < 4; ++kc) {
      const T av = A1[(16 * rb + col) * 16 + 4 * kc + kq];           // A1[16 rb + i][4 kc + k]
      const T bv = W[hg::kUdeW2 + col * 16 + 4 * kc + kq];           // W2[j][4 kc + k]
      acc = M::op(av, bv, acc);
    }
    for (int r = 0; r < 4; ++r) A2[(16 * rb + M::row(kq, r)) * 16 + col] = M::tanh_(acc[r]);
  }
  hg::wave_sync();
}

// Output layer of cell l: z3 = b3 + sum_j A2[l][j] W3[j].
template <class T>
__device__ __forceinline__ T ude_out(const T* __restrict__ W, const T* A2, int l) {
  T z3 = W[hg::kUdeB3];
  for (int j = 0; j < 16; ++j) z3 = hg::Mfma<T>::fma_(A2[l * 16 + j], W[hg::kUdeW3 + j], z3);
  return z3;
}

// One weight set held in registers, laid out for the lane's share of the
// matrix-core operands (see ude_hidden) and for the output layer.
template <class T>
struct UdeRegs {
  T w1, w2[4], b2, w3[16], b3;
  __device__ __forceinline__ void load(const T* __restrict__ W, int col, int kq) {
    w1 = kq < 3 ? W[hg::kUdeW1 + col * 3 + kq] : W[hg::kUdeB1 + col];  // [W1 | b1]
    for (int kc = 0; kc < 4; ++kc) w2[kc] = W[hg::kUdeW2 + col * 16 + 4 * kc + kq];
    b2 = W[hg::kUdeB2 + col];
    for (int j = 0; j < 16; ++j) w3[j] = W[hg::kUdeW3 + j];
    b3 = W[hg::kUdeB3];
  }
};

// ude_hidden with the weights in registers, for the first ROWS cells (16,
// 32, 48 or 64) of the chunk; each cell's values do not depend on ROWS.
template <class T, int ROWS = 64>
__device__ __forceinline__ void ude_hidden_r(const UdeRegs<T>& R, const T* F, T* A1, T* A2, int l) {
  using M = hg::Mfma<T>;
  const int col = l & 15, kq = l >> 4;
  for (int rb = 0; rb < ROWS / 16; ++rb) {
    typename M::V acc = {T(0), T(0), T(0), T(0)};
    acc = M::op(F[(16 * rb + col) * 4 + kq], R.w1, acc);
    for (int r = 0; r < 4; ++r) A1[(16 * rb + M::row(kq, r)) * 16 + col] = M::tanh_(acc[r]);
  }
  hg::wave_sync();
  for (int rb = 0; rb < ROWS / 16; ++rb) {
    typename M::V acc = {R.b2, R.b2, R.b2, R.b2};
    for (int kc = 0; kc < 4; ++kc) acc = M::op(A1[(16 * rb + col) * 16 + 4 * kc + kq], R.w2[kc], acc);
    for (int r = 0; r < 4; ++r) A2[(16 * rb + M::row(kq, r)) * 16 + col] = M::tanh_(acc[r]);
  }
  hg::wave_sync();
}

// n = 0.01 + softplus(MLP(features)) for every cell.  Each wave walks
// CH-cell chunks (grid-stride; CH = 64, or 32 to spread a small mesh over
// twice the waves) with its weight set in registers; a chunk whose cells
// use several sets takes one pass per set (the matrix-core products need one
// set across the wave).
template <class T, int WPB, int CH = 64>
__global__ __launch_bounds__(64 * WPB) void k_ude_n(hg::UdeArgs a) {
  static_assert(CH == 32 || CH == 64, "chunk of 32 or 64 cells");
  __shared__ T Fs[WPB][CH * 4];
  __shared__ T A1s[WPB][CH * 16];
  __shared__ T A2s[WPB][CH * 16];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int col = l & 15, kq = l >> 4;
  T* F = Fs[w];
  T* A1 = A1s[w];
  T* A2 = A2s[w];
  const int nchunks = (a.N + CH - 1) / CH;
  UdeRegs<T> R;
  int cur = -1;
  for (int ch = blockIdx.x * WPB + w; ch < nchunks; ch += gridDim.x * WPB) {
    const int i = ch * CH + l;
    const bool valid = l < CH && i < a.N;
    double h = 0.0, u = 0.0, v = 0.0, um = 0.0, zb = 0.0;
    bool wet = false;
    int set = 0;
    if (valid) {
      hg::ude_features(a, i, h, u, v, um, wet);
      zb = a.zb[i % a.Nm];
      if (a.set_of_cell) set = a.set_of_cell[i];
    }
    if (l < CH) {  // F has CH rows
      F[l * 4 + 0] = T(h);
      F[l * 4 + 1] = T(um);
      F[l * 4 + 2] = T(zb);
      F[l * 4 + 3] = T(1);  // the bias column of layer 1
    }
    hg::wave_sync();
    uint64_t pending = __ballot(valid);
    while (pending) {
      const int lead = __ffsll((unsigned long long)pending) - 1;
      const int s = __builtin_amdgcn_readfirstlane(__shfl(set, lead, 64));
      const bool mine = valid && set == s;
      pending &= ~uint64_t(__ballot(mine));
      if (s != cur) {
        R.load(static_cast<const T*>(a.w) + size_t(s) * hg::kUdeW, col, kq);
        cur = s;
      }
      ude_hidden_r<T, CH>(R, F, A1, A2, l);
      if (mine) {
        T z3 = R.b3;
        for (int j = 0; j < 16; ++j) z3 = hg::Mfma<T>::fma_(A2[l * 16 + j], R.w3[j], z3);
        a.n[i] = 0.01 + double(hg::softplus<T>(z3));
      }
      hg::wave_sync();  // the next pass rewrites A1, A2
    }
  }
}

// Slot of (this block, set s): the entry of s among the block's sets.
__device__ __forceinline__ int ude_slot(const hg::UdeArgs& a, int s) {
  if (!a.blk_slot) return int(blockIdx.x);
  const int e0 = a.blk_off[blockIdx.x], e1 = a.blk_off[blockIdx.x + 1];
  int slot = a.blk_slot[e0];
  for (int e = e0; e < e1; ++e)
    if (a.blk_set[e] == s) {
      slot = a.blk_slot[e];
      break;
    }
  return slot;
}

// Adds one run of weight cotangents (the matrix-core accumulators of
// k_ude_vjp, row ir = M::row(kq, r), column col) into slot wp; each of the 353
// weights comes from exactly one lane.  All of the lane's slot values are read
// before any is written, so the reads take one memory round trip.
template <class T>
__device__ __forceinline__ void ude_flush(double* __restrict__ wp, const typename hg::Mfma<T>::V& w1,
                                          const typename hg::Mfma<T>::V& w2, const typename hg::Mfma<T>::V& w3,
                                          double b3, int l) {
  using M = hg::Mfma<T>;
  const int col = l & 15, kq = l >> 4;
  int i2[4], i1[4], i3[4];
  double x2[4], x1[4] = {0.0, 0.0, 0.0, 0.0}, x3[4] = {0.0, 0.0, 0.0, 0.0};
  for (int r = 0; r < 4; ++r) {
    const int ir = M::row(kq, r);
    i2[r] = hg::kUdeW2 + ir * 16 + col;                            // W2bar
    i1[r] = col < 3 ? hg::kUdeW1 + ir * 3 + col : hg::kUdeB1 + ir;  // W1bar, b1bar (col < 4)
    i3[r] = col == 0 ? hg::kUdeB2 + ir : hg::kUdeW3 + ir;           // b2bar, W3bar (col < 2)
  }
  for (int r = 0; r < 4; ++r) {
    x2[r] = wp[i2[r]];
    if (col < 4) x1[r] = wp[i1[r]];
    if (col < 2) x3[r] = wp[i3[r]];
  }
  const double xb3 = l == 0 ? wp[hg::kUdeB3] : 0.0;
  for (int r = 0; r < 4; ++r) {
    wp[i2[r]] = x2[r] + double(w2[r]);
    if (col < 4) wp[i1[r]] = x1[r] + double(w1[r]);
    if (col < 2) wp[i3[r]] = x3[r] + double(w3[r]);
  }
  if (l == 0) wp[hg::kUdeB3] = xb3 + b3;
}

// Reverse mode of k_ude_n: given the cotangent of n, adds the cotangents of
// the features to the state cotangent (through the clamp and u = q / h, as
// the forward forms them) and the weight cotangents into per-(block, set)
// slots.  Block b (one wave) takes the 64-cell chunks b cpb .. (b + 1) cpb - 1
// in order, one pass per distinct set of a chunk (order of first appearance,
// as k_ude_n).  The weight cotangents are sums over cells on the matrix cores
// (K = cells); consecutive passes of one set keep adding into the same
// accumulators, which go into the (block, set) slot when the set changes and
// at the end.  A chunk's inputs are read before its passes, its
// state-cotangent values after them.  The backward activations overwrite the forward ones
// once read: Z2 replaces A2, then Z1 replaces A1.
template <class T>
__global__ __launch_bounds__(64, sizeof(T) == 4 ? 3 : 1) void k_ude_vjp(hg::UdeArgs a) {
  using M = hg::Mfma<T>;
  __shared__ T F[64 * 4];
  __shared__ T A1[64 * 16];  // A1, then Z1 = A1bar (1 - A1^2)
  __shared__ T A2[64 * 16];  // A2, then Z2 = z3bar W3 (1 - A2^2)
  __shared__ T Z3[64];
  const int l = threadIdx.x;
  const int col = l & 15, kq = l >> 4;
  const int nchunks = (a.N + 63) / 64;
  const int ch0 = int(blockIdx.x) * a.cpb, ch1 = min(nchunks, ch0 + a.cpb);
  typename M::V w1 = {T(0), T(0), T(0), T(0)}, w2 = w1, w3 = w1;
  double b3 = 0.0;
  int acc_set = -1;  // set of the running accumulators (-1: none yet)
  for (int ch = ch0; ch < ch1; ++ch) {
    const int i = ch * 64 + l;
    const bool valid = i < a.N;
    double h = 0.0, u = 0.0, v = 0.0, um = 0.0, zb = 0.0;
    T nb = T(0);
    bool wet = false;
    int set = 0;
    if (valid) {
      hg::ude_features(a, i, h, u, v, um, wet);
      zb = a.zb[i % a.Nm];
      if (a.set_of_cell) set = a.set_of_cell[i];
      nb = T(a.nbar[i]);
    }
    F[l * 4 + 0] = T(h);
    F[l * 4 + 1] = T(um);
    F[l * 4 + 2] = T(zb);
    F[l * 4 + 3] = T(1);  // the bias column of layer 1
    hg::wave_sync();
    T g0 = T(0), g1 = T(0), g2 = T(0);  // the cell's feature cotangents, from its set's pass
    uint64_t pending = __ballot(valid);
    while (pending) {
      const int lead = __ffsll((unsigned long long)pending) - 1;
      const int s = __builtin_amdgcn_readfirstlane(__shfl(set, lead, 64));  // wave-uniform: scalar weight loads
      const bool mine = valid && set == s;
      pending &= ~uint64_t(__ballot(mine));
      if (s != acc_set) {
        if (acc_set >= 0) ude_flush<T>(a.wpart + size_t(ude_slot(a, acc_set)) * hg::kUdeW, w1, w2, w3, b3, l);
        w1 = typename M::V{T(0), T(0), T(0), T(0)};
        w2 = w1;
        w3 = w1;
        b3 = 0.0;
        acc_set = s;
      }
      const T* W = static_cast<const T*>(a.w) + size_t(s) * hg::kUdeW;
      ude_hidden<T>(W, F, A1, A2, l);
      // output layer and softplus, lane = cell; cells of other sets carry a zero cotangent
      const T z3 = ude_out<T>(W, A2, l);
      const T z3b = mine ? nb * hg::sigmoid<T>(z3) : T(0);
      Z3[l] = z3b;
      hg::wave_sync();
      // Weight cotangents as sums over the cells on the matrix cores (K = cells):
      // w3 column 1: W3bar = A2^T Z3; column 0: b2bar = Z2^T 1; w2: W2bar = Z2^T A1;
      // w1: [W1bar | b1bar] = Z1^T [F | 1]; b3bar = sum of Z3 by a fixed xor tree.
      for (int kc = 0; kc < 16; ++kc) {
        const int c = 4 * kc + kq;  // cell of this lane's K-row
        w3 = M::op(A2[c * 16 + col], col == 1 ? Z3[c] : T(0), w3);
      }
      hg::wave_sync();
      for (int j = 0; j < 16; ++j) {  // Z2 over A2, own row
        const T a2 = A2[l * 16 + j];
        A2[l * 16 + j] = (z3b * W[hg::kUdeW3 + j]) * M::fma_(-a2, a2, T(1));
      }
      hg::wave_sync();
      const T* Z2 = A2;
      for (int kc = 0; kc < 16; ++kc) {
        const int c = 4 * kc + kq;
        const T z2 = Z2[c * 16 + col];
        w2 = M::op(z2, A1[c * 16 + col], w2);
        w3 = M::op(z2, T(col == 0), w3);
      }
      // A1bar = Z2 W2 (rows = cells)
      typename M::V ab[4];
      for (int rb = 0; rb < 4; ++rb) {
        typename M::V acc = {T(0), T(0), T(0), T(0)};
        for (int kc = 0; kc < 4; ++kc) {
          const T av = Z2[(16 * rb + col) * 16 + 4 * kc + kq];    // Z2[16 rb + i][4 kc + k]
          const T bv = W[hg::kUdeW2 + (4 * kc + kq) * 16 + col];  // W2[4 kc + k][j]
          acc = M::op(av, bv, acc);
        }
        ab[rb] = acc;
      }
      hg::wave_sync();
      for (int rb = 0; rb < 4; ++rb)  // Z1 = A1bar (1 - A1^2) over A1
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * rb + M::row(kq, r);
          const T a1 = A1[row * 16 + col];
          A1[row * 16 + col] = ab[rb][r] * M::fma_(-a1, a1, T(1));
        }
      hg::wave_sync();
      const T* Z1 = A1;
      // feature cotangents -> state cotangent (lane = cell)
      T fb0 = T(0), fb1 = T(0), fb2 = T(0);
      for (int j = 0; j < 16; ++j) {
        const T z = Z1[l * 16 + j];
        fb0 = M::fma_(z, W[hg::kUdeW1 + j * 3 + 0], fb0);
        fb1 = M::fma_(z, W[hg::kUdeW1 + j * 3 + 1], fb1);
        fb2 = M::fma_(z, W[hg::kUdeW1 + j * 3 + 2], fb2);
      }
      if (mine) {
        g0 = fb0;
        g1 = fb1;
        g2 = fb2;
      }
      for (int kc = 0; kc < 16; ++kc) {
        const int c = 4 * kc + kq;
        w1 = M::op(Z1[c * 16 + col], col < 3 ? F[c * 4 + col] : T(col == 3), w1);
      }
      double z3s = double(Z3[l]);
      for (int m = 32; m > 0; m >>= 1) z3s += __shfl_xor(z3s, m, 64);
      b3 += z3s;
      hg::wave_sync();  // the next pass rewrites A1, A2 and Z3
    }
    if (valid) {  // all four read before any is written: one round trip
      const double o0 = a.ox[i], o1 = a.oqx[i], o2 = a.oqy[i], oz = a.zbbar ? a.zbbar[i] : 0.0;
      double xib = 0.0, qxb = 0.0, qyb = 0.0;
      if (wet) {  // h = xi + h_still, |u| = sqrt(u^2 + v^2 + eps), u = q_x / h, v = q_y / h
        const double ub = double(g1) * u / um, vb = double(g1) * v / um;
        xib = double(g0) - (ub * u + vb * v) / h;
        qxb = ub / h;
        qyb = vb / h;
      }
      const double zbb = double(g2);
      a.ox[i] = o0 + xib;
      a.oqx[i] = o1 + qxb;
      a.oqy[i] = o2 + qyb;
      if (a.zbbar) a.zbbar[i] = oz + zbb;
    }
  }
  if (acc_set >= 0) ude_flush<T>(a.wpart + size_t(ude_slot(a, acc_set)) * hg::kUdeW, w1, w2, w3, b3, l);
}

// wbar[s][k] += the sum of the slots of weight set s (set_start[s] ..
// set_start[s + 1], or all nslots when set_start is NULL), where slot b of
// weight k sits at wpart[k * kstride + b * sstride]; one work-group per
// (weight k = blockIdx.x, set s = blockIdx.y), strided partial sums then a
// fixed tree (deterministic).
__global__ __launch_bounds__(kBlock) void k_ude_wsum(int nslots, const int32_t* set_start, const double* wpart,
                                                      int64_t kstride, int64_t sstride, double* wbar) {
  __shared__ double part[kBlock];
  const int k = blockIdx.x, set = blockIdx.y, tid = threadIdx.x;
  const int lo = set_start ? set_start[set] : 0, hi = set_start ? set_start[set + 1] : nslots;
  const double* p = wpart + k * kstride;
  double acc = 0.0;
  for (int b = lo + tid; b < hi; b += kBlock) acc += p[b * sstride];
  part[tid] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (tid < w) part[tid] += part[tid + w];
    __syncthreads();
  }
  if (tid == 0) wbar[size_t(set) * gridDim.x + k] += part[0];
}

// ---- variable Manning n: one law evaluation per cell (hg_nlaw.hpp) ----

__device__ __forceinline__ double nlaw_ks(const hg::NLawArgs& a, int i) {
  return a.law.kind == hg::NLAW_H_UMAG_KS ? a.ks[i] : 1.0;
}

// n <- law(features of the state), before every RHS evaluation.
__global__ __launch_bounds__(kBlock) void k_nlaw_n(hg::NLawArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  double h, U;
  hg::nlaw_features<double>(a.law, a.xi[i], a.qx[i], a.qy[i], a.hstill[i], a.hsmall, h, U);
  a.n[i] = hg::nlaw_eval<double, false>(a.law, h, U, nlaw_ks(a, i), nullptr, nullptr, nullptr);
}

// The law at given (h, |U|), with its auxiliary outputs.
__global__ __launch_bounds__(kBlock) void k_nlaw_eval(hg::NLawArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  double hks, f, Re;
  a.n[i] = hg::nlaw_eval<double, true>(a.law, a.h_in[i], a.U_in[i], nlaw_ks(a, i), &hks, &f, &Re);
  if (a.hks) a.hks[i] = hks;
  if (a.f) a.f[i] = f;
  if (a.Re) a.Re[i] = Re;
}

// dn/d(state component s) of cell i: one dual pass seeded on that component.
__device__ __forceinline__ double nlaw_partial(const hg::NLawArgs& a, int i, int s) {
  using hg::Dual1;
  const Dual1 xi(a.xi[i], s == 0 ? 1.0 : 0.0), qx(a.qx[i], s == 1 ? 1.0 : 0.0), qy(a.qy[i], s == 2 ? 1.0 : 0.0);
  Dual1 h, U;
  hg::nlaw_features<Dual1>(a.law, xi, qx, qy, a.hstill[i], a.hsmall, h, U);
  return hg::nlaw_eval<Dual1, false>(a.law, h, U, nlaw_ks(a, i), nullptr, nullptr, nullptr).d;
}

// Reverse mode: state cotangent += nbar * dn/d(xi, q_x, q_y).  The depth-only
// laws do not read the discharge.
__global__ __launch_bounds__(kBlock) void k_nlaw_vjp(hg::NLawArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const double nb = a.nbar[i];
  if (nb == 0.0) return;
  a.ox[i] += nb * nlaw_partial(a, i, 0);
  if (a.law.kind == hg::NLAW_H_UMAG_KS) {
    a.oqx[i] += nb * nlaw_partial(a, i, 1);
    a.oqy[i] += nb * nlaw_partial(a, i, 2);
  }
}

// Forward mode: n and its tangent along the state tangent.
__global__ __launch_bounds__(kBlock) void k_nlaw_jvp(hg::NLawArgs a) {
  using hg::Dual1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const Dual1 xi(a.xi[i], a.dxi[i]), qx(a.qx[i], a.dqx[i]), qy(a.qy[i], a.dqy[i]);
  Dual1 h, U;
  hg::nlaw_features<Dual1>(a.law, xi, qx, qy, a.hstill[i], a.hsmall, h, U);
  const Dual1 n = hg::nlaw_eval<Dual1, false>(a.law, h, U, nlaw_ks(a, i), nullptr, nullptr, nullptr);
  a.n[i] = n.v;
  a.dn[i] = n.d;
}

// ---- the reference's UDE network (hg_lux.hpp) ----

// out <- network(features of the state): n, or the friction magnitude.
template <int MW>
__global__ __launch_bounds__(kBlock) void k_lux_fwd(hg::LuxArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  double f[3];
  hg::lux_features<double>(a.net, a.xi[i], a.qx[i], a.qy[i], a.hstill[i], a.hsmall, hg::lux_ks(a, i), f);
  a.out[i] = hg::lux_eval<MW>(a.net, a.w, f);
}

// Forward mode: the network's output and its tangent along the state
// tangent (dxi, dqx, dqy) and the weight tangent dw (NULL: none), one cell
// per work-item, on dual numbers (the value part is bitwise k_lux_fwd's).
template <int MW>
__global__ __launch_bounds__(kBlock) void k_lux_jvp(hg::LuxArgs a, const double* dxi, const double* dqx,
                                                   const double* dqy, const double* dw, double* dout) {
  using hg::Dual1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  Dual1 f[3];
  hg::lux_features<Dual1>(a.net, Dual1(a.xi[i], dxi[i]), Dual1(a.qx[i], dqx[i]), Dual1(a.qy[i], dqy[i]),
                          a.hstill[i], a.hsmall, hg::lux_ks(a, i), f);
  const Dual1 o = hg::lux_eval<MW, Dual1>(a.net, a.w, f, dw);
  a.out[i] = o.v;
  dout[i] = o.d;
}

// Reverse mode: one wave per work-group, 64 cells at a time (grid-stride).
// Weight cotangents: wave sums into LDS, then one slot per work-group in
// wpart; the state cotangent adds through the features (one dual pass per
// state component the features read).
template <int MW>
__global__ __launch_bounds__(64) void k_lux_vjp(hg::LuxArgs a) {
  __shared__ double acc[hg::kLuxMaxP];
  const int lane = threadIdx.x;
  const int P = a.net.P;
  for (int k = lane; k < P; k += 64) acc[k] = 0.0;
  __syncthreads();
  for (int base = blockIdx.x * 64; base < a.N; base += gridDim.x * 64) {
    const int i = base + lane;
    const bool valid = i < a.N;
    double f[3] = {0.0, 0.0, 0.0};
    double ob = 0.0;
    if (valid) {
      ob = a.obar[i];
      hg::lux_features<double>(a.net, a.xi[i], a.qx[i], a.qy[i], a.hstill[i], a.hsmall, hg::lux_ks(a, i), f);
    }
    double fb[3];
    hg::lux_backward<MW>(a.net, a.w, f, ob, acc, lane, fb);
    if (valid && ob != 0.0) {
      using hg::Dual1;
      const int ns = a.net.choice == hg::LUX_MANNING_H ? 1 : 3;  // ManningN_h reads the depth only
      for (int st = 0; st < ns; ++st) {
        const Dual1 xi(a.xi[i], st == 0 ? 1.0 : 0.0), qx(a.qx[i], st == 1 ? 1.0 : 0.0),
            qy(a.qy[i], st == 2 ? 1.0 : 0.0);
        Dual1 fd[3];
        hg::lux_features<Dual1>(a.net, xi, qx, qy, a.hstill[i], a.hsmall, hg::lux_ks(a, i), fd);
        const double d = fb[0] * fd[0].d + fb[1] * fd[1].d + fb[2] * fd[2].d;
        if (st == 0)
          a.ox[i] += d;
        else if (st == 1)
          a.oqx[i] += d;
        else
          a.oqy[i] += d;
      }
    }
  }
  __syncthreads();
  for (int k = lane; k < P; k += 64) a.wpart[size_t(k) * gridDim.x + blockIdx.x] += acc[k];
}

// Layout check of the 16x16x4 matrix-core products used above: D = A B for
// A (16x4) and B (4x16) given row-major, D returned row-major.
template <class T>
__global__ __launch_bounds__(64) void k_mfma_probe(const T* A, const T* B, T* D) {
  using M = hg::Mfma<T>;
  const int l = threadIdx.x, col = l & 15, kq = l >> 4;
  typename M::V acc = {T(0), T(0), T(0), T(0)};
  acc = M::op(A[col * 4 + kq], B[kq * 16 + col], acc);
  for (int r = 0; r < 4; ++r) D[M::row(kq, r) * 16 + col] = acc[r];
}

// ---- Tsit5: Tsitouras' 5(4) Runge-Kutta pair (Ch. Tsitouras, Comput. Math.
// Appl. 62 (2011) 770-775), first-same-as-last, the adaptive integrator the
// reference's examples run through OrdinaryDiffEq (Tsit5()).
namespace tsit5 {
constexpr double c[7] = {0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0};
constexpr double a[7][6] = {
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.161, 0.0, 0.0, 0.0, 0.0, 0.0},
    {-0.008480655492356989, 0.335480655492357, 0.0, 0.0, 0.0, 0.0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0.0, 0.0, 0.0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0.0, 0.0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0.0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774}};
// b = a[6] (fifth order, b7 = 0); btilde = b - bhat (bhat: the embedded fourth-order weights)
constexpr double btilde[7] = {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                              -0.1447110071732629,     0.5823571654525552,    -0.45808210592918697,
                              0.015151515151515152};
}  // namespace tsit5

struct RkArgs {
  int64_t n;             // 3N
  const double* y;
  double* out;
  const double* k[7];
  double coef[7];
  int nk;
};

// out = y + sum_j coef[j] k[j] (coef already scaled by the step).
__global__ __launch_bounds__(kBlock) void k_rk_combine(RkArgs a) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= a.n) return;
  double s = 0.0;
  for (int j = 0; j < a.nk; ++j) s += a.coef[j] * a.k[j][i];
  a.out[i] = a.y[i] + s;
}

// Partial sums of ((sum_j coef[j] k[j]) / (abstol + reltol max(|y|, |ynew|)))^2
// per work-group (fixed tree), for the RMS error norm.
__global__ __launch_bounds__(kBlock) void k_rk_err(RkArgs a, const double* ynew, double abstol, double reltol,
                                                   double* part) {
  __shared__ double s[kBlock];
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  double v = 0.0;
  if (i < a.n) {
    double e = 0.0;
    for (int j = 0; j < a.nk; ++j) e += a.coef[j] * a.k[j][i];
    const double sc = abstol + reltol * fmax(fabs(a.y[i]), fabs(ynew[i]));
    v = (e / sc) * (e / sc);
  }
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// out[0] = sum of part[0 .. n) by one work-group (fixed tree).
__global__ __launch_bounds__(kBlock) void k_sum(int n, const double* part, double* out) {
  __shared__ double s[kBlock];
  double acc = 0.0;
  for (int b = threadIdx.x; b < n; b += kBlock) acc += part[b];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// ---- forward mode (JVP) of the Euler step: the per-cell step kernel on
// dual numbers (value, one tangent), the device counterpart of
// ForwardDiff.Dual through swe_2d_rhs and custom_ODE_update_cells.
struct JvpArgs {
  StepArgs a;  // values (fields after the active parameter is applied)
  const double* dxi;
  const double* dqx;
  const double* dqy;
  double* dxo;
  double* dqxo;
  double* dqyo;
  const double* dzb;      // tangents of the parameter fields (nullable = 0)
  const double* dn;
  const double* dS0x;
  const double* dS0y;
  const double* dg_zb;
  const double* dinletQ;
  const double* dtotalA;  // tangent of the inlet conveyances of this step
  const double* dfmag;    // tangent of the FlowResistance friction magnitude (with a.fmag)
};

__device__ __forceinline__ hg::Dual1 ldd(const double* v, const double* d, int i) {
  return hg::Dual1(v[i], d ? d[i] : 0.0);
}

// Inlet conveyance on dual numbers, sequential fold in NodeString order.
__global__ void k_inlet_conveyance_jvp(hg::InletArgs a, const double* dxi, const double* dn, double* dtotalA) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.n_inlet) return;
  using hg::Dual1;
  Dual1 acc(0.0);
  for (int f = a.off[b]; f < a.off[b + 1]; ++f) {
    const int c = a.cell[f];
    Dual1 h = ldd(a.xi, dxi, c) + a.hstill[c];
    if (h <= a.hsmall) h = Dual1(a.hsmall);
    const double w = (h > a.hsmall) ? 1.0 : 0.0;
    const Dual1 term = a.L53[f] * h / ldd(a.n, dn, c) * w;
    acc = (f == a.off[b]) ? term : acc + term;
  }
  a.totalA[b] = acc.v;
  dtotalA[b] = acc.d;
  if (!(acc.v > 1e-10)) a.flag[0] = 1;
}

__global__ __launch_bounds__(kBlock) void k_swe_step_jvp(JvpArgs J) {
  using hg::Dual1;
  const StepArgs& a = J.a;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const double hsmall = a.hsmall;
  const Dual1 xi = ldd(a.xi, J.dxi, i);
  const double hs = a.hstill[i];
  const Dual1 qx0 = ldd(a.qx, J.dqx, i), qy0 = ldd(a.qy, J.dqy, i);
  Dual1 h = xi + hs, qx = qx0, qy = qy0;  // semi_discretize_swe_2D.jl:101-106
  if (h <= hsmall) {
    h = Dual1(hsmall);
    qx = Dual1(0.0);
    qy = Dual1(0.0);
  }
  const Dual1 zb = ldd(a.zb, J.dzb, i);
  const Dual1 nman = ldd(a.n, J.dn, i);
  const hg::SideT<Dual1> Ls{xi, Dual1(hs), h, qx, qy, zb};
  Dual1 acc0(0.0), acc1(0.0), acc2(0.0);
  const int d = a.deg[i];
  for (int k = 0; k < d; ++k) {
    const int s = k * a.N + i;
    const int j = a.nbr[s];
    const double nx = a.fnx[s], ny = a.fny[s], L = a.flen[s];
    hg::SideT<Dual1> Rs;
    if (j >= 0) {
      const Dual1 xj = ldd(a.xi, J.dxi, j);
      const double hsj = a.hstill[j];
      Dual1 hj = xj + hsj, qxj = ldd(a.qx, J.dqx, j), qyj = ldd(a.qy, J.dqy, j);
      if (hj <= hsmall) {
        hj = Dual1(hsmall);
        qxj = Dual1(0.0);
        qyj = Dual1(0.0);
      }
      Rs = hg::SideT<Dual1>{xj, Dual1(hsj), hj, qxj, qyj, ldd(a.zb, J.dzb, j)};
    } else {  // ghost (bc_2D.jl:640-834), internal cell = this cell
      const int gi = ~j;
      Dual1 hgh, qxg, qyg;
      switch (a.g_kind[gi]) {
        case HG_BC_INLETQ: {
          const int b = a.g_bc[gi];
          const double w = (h > hsmall) ? 1.0 : 0.0;
          const Dual1 v = ldd(a.inletQ, J.dinletQ, b) / ldd(a.totalA, J.dtotalA, b) * a.g_L23[gi] / nman;
          hgh = h;
          qxg = -h * v * nx * w;
          qyg = -h * v * ny * w;
          break;
        }
        case HG_BC_EXITH: {
          const Dual1 hx = a.exitWSE[a.g_bc[gi]] - zb;
          hgh = (hx > hsmall) ? hx : Dual1(hsmall);
          qxg = qx;
          qyg = qy;
          break;
        }
        case HG_BC_WALL:
          hgh = h;
          qxg = -qx;
          qyg = -qy;
          break;
        default: {  // HG_BC_SYMM
          const Dual1 vdn = qx * nx + qy * ny;
          hgh = h;
          qxg = qx - 2.0 * vdn * nx;
          qyg = qy - 2.0 * vdn * ny;
          break;
        }
      }
      const double hsg = a.g_hstill[gi];
      Rs = hg::SideT<Dual1>{hgh - hsg, Dual1(hsg), hgh, qxg, qyg, ldd(a.g_zb, J.dg_zb, gi)};
    }
    Dual1 f0, f1, f2;
    hg::roe_flux(Ls, Rs, nx, ny, a.g, a.half_g, hsmall, f0, f1, f2);
    acc0 = acc0 + f0 * L;
    acc1 = acc1 + f1 * L;
    acc2 = acc2 + f2 * L;
  }
  const double A = a.area[i];
  const Dual1 r0 = -acc0 / A, r1 = -acc1 / A, r2 = -acc2 / A;
  Dual1 s1(0.0), s2(0.0);
  if (h > hsmall) {  // :463-475
    Dual1 fx, fy;
    if (a.fmag)
      hg::friction_mag(ldd(a.fmag, J.dfmag, i), qx, qy, fx, fy);
    else
      hg::friction(a.g, a.kn2, hsmall, h, qx, qy, nman, fx, fy);
    s1 = a.g * xi * ldd(a.S0x, J.dS0x, i) - fx;
    s2 = a.g * xi * ldd(a.S0y, J.dS0y, i) - fy;
  }
  const Dual1 d0 = r0 + 0.0, d1 = r1 + s1, d2 = r2 + s2;
  const double dti = a.dtc ? a.dtc[i] : a.dt;
  Dual1 xn = xi + dti * d0;  // custom_ODE_solvers.jl:16-26
  Dual1 qxn = qx0 + dti * d1;
  Dual1 qyn = qy0 + dti * d2;
  if (xn < hsmall) {
    xn = Dual1(hsmall);
    qxn = Dual1(0.0);
    qyn = Dual1(0.0);
  }
  a.xo[i] = xn.v;
  a.qxo[i] = qxn.v;
  a.qyo[i] = qyn.v;
  J.dxo[i] = xn.d;
  J.dqxo[i] = qxn.d;
  J.dqyo[i] = qyn.d;
}

// Work items of the tiled kernel (see TileArgs): per tile, cross-tile faces
// first (the kernel prefetches their neighbour states before its first
// barrier), then boundary faces, then intra-tile faces (owned by the side
// with the lower local index), each group ordered by cell.
// Inlet ghosts carry the slot of their inlet in the tile's inlet list in
// code bits 14..19; max_inlets is the longest such list.
struct TileBuild {
  std::vector<int32_t> off, x0, code, ref;
  std::vector<double> nx, ny, len;
  std::vector<int32_t> in_off, in_list;
  int32_t max_inlets = 0;
};

TileBuild build_tiles(int64_t N, int T, const std::vector<int32_t>& deg, const std::vector<int32_t>& nbr,
                      const std::vector<double>& fnx, const std::vector<double>& fny, const std::vector<double>& flen,
                      const std::vector<int32_t>& g_kind, const std::vector<int32_t>& g_bc) {
  TileBuild tb;
  const int64_t ntiles = (N + T - 1) / T;
  tb.off.assign(ntiles + 1, 0);
  tb.x0.assign(ntiles, 0);
  tb.in_off.assign(ntiles + 1, 0);
  std::vector<int64_t> grp[3];
  std::vector<int32_t> local;  // this tile's inlets in order of first appearance
  for (int64_t t = 0; t < ntiles; ++t) {
    const int64_t c0 = t * T, c1 = std::min<int64_t>(N, c0 + T);
    for (auto& v : grp) v.clear();
    local.clear();
    for (int64_t i = c0; i < c1; ++i) {
      const int32_t li = int32_t(i - c0);
      for (int32_t k = 0; k < deg[i]; ++k) {
        const int64_t s = int64_t(k) * N + i;
        const int32_t j = nbr[s];
        int32_t code, ref = 0, type;
        if (j < 0) {
          type = 2;
          ref = ~j;
          code = li | (k << 10) | (type << 28);
          if (g_kind[ref] == HG_BC_INLETQ) {
            int32_t sl = 0;
            while (sl < int32_t(local.size()) && local[sl] != g_bc[ref]) ++sl;
            if (sl == int32_t(local.size())) local.push_back(g_bc[ref]);
            code |= (sl & (kTileInlets - 1)) << 14;
          }
        } else if (j >= c0 && j < c1) {
          const int32_t lj = int32_t(j - c0);
          if (lj < li) continue;  // owned by the other side
          int32_t kb = -1;
          for (int32_t k2 = 0; k2 < deg[j]; ++k2)
            if (nbr[int64_t(k2) * N + j] == int32_t(i)) kb = k2;
          type = 0;
          code = li | (k << 10) | (lj << 14) | (kb << 24) | (type << 28);
        } else {
          type = 1;
          ref = j;
          code = li | (k << 10) | (type << 28);
        }
        grp[type].push_back(int64_t(code) | (int64_t(uint32_t(ref)) << 32));
        (void)s;
      }
    }
    for (int type : {1, 2, 0}) {  // cross-tile first: their neighbour gathers are prefetched
      if (type == 0) tb.x0[t] = int32_t(tb.code.size());
      for (int64_t e : grp[type]) {
        const int32_t code = int32_t(e & 0xffffffff);
        const int32_t ref = int32_t(uint32_t(uint64_t(e) >> 32));
        const int32_t li = code & 1023, k = (code >> 10) & 15;
        const int64_t s = int64_t(k) * N + c0 + li;
        tb.code.push_back(code);
        tb.ref.push_back(ref);
        tb.nx.push_back(fnx[s]);
        tb.ny.push_back(fny[s]);
        tb.len.push_back(flen[s]);
      }
    }
    tb.off[t + 1] = int32_t(tb.code.size());
    tb.in_list.insert(tb.in_list.end(), local.begin(), local.end());
    tb.in_off[t + 1] = int32_t(tb.in_list.size());
    tb.max_inlets = std::max(tb.max_inlets, int32_t(local.size()));
  }
  return tb;
}

// Device copy of a tiling's work items (dalloc'd: freed with the context,
// or by free_items).
int upload_items(hg_ctx* ctx, const TileBuild& tb, hg::ItemTable& d) {
  const size_t ni = tb.code.size();
  std::vector<hg::ItemGeo> geo(ni);
  for (size_t k = 0; k < ni; ++k) geo[k] = hg::ItemGeo{tb.nx[k], tb.ny[k], tb.len[k]};
  d = hg::ItemTable{};
  int rc = dalloc(ctx, &d.off, tb.off.size());
  if (!rc) rc = dalloc(ctx, &d.x0, tb.x0.size());
  if (!rc) rc = dalloc(ctx, &d.code, ni);
  if (!rc) rc = dalloc(ctx, &d.ref, ni);
  if (!rc) rc = dalloc(ctx, &d.geo, ni);
  if (!rc) rc = upload(ctx, d.off, tb.off.data(), tb.off.size());
  if (!rc) rc = upload(ctx, d.x0, tb.x0.data(), tb.x0.size());
  if (!rc) rc = upload(ctx, d.code, tb.code.data(), ni);
  if (!rc) rc = upload(ctx, d.ref, tb.ref.data(), ni);
  if (!rc) rc = upload(ctx, d.geo, geo.data(), ni);
  return rc;
}

int free_items(hg_ctx* ctx, hg::ItemTable& d) {
  int rc = dfree(ctx, d.off);
  if (!rc) rc = dfree(ctx, d.x0);
  if (!rc) rc = dfree(ctx, d.code);
  if (!rc) rc = dfree(ctx, d.ref);
  if (!rc) rc = dfree(ctx, d.geo);
  return rc;
}

size_t tile_lds_bytes(int T, int maxdeg) {
  return size_t(10 * T + 3 * maxdeg * T + kTileInlets) * sizeof(double);
}

hg::TileArgs make_tile_args_p(hg_ctx* ctx, const double* const in[3], double* const out[3], double dt,
                              uint8_t* clamp = nullptr) {
  hg::TileArgs a{};
  a.N = int32_t(ctx->Nm);
  a.K = ctx->K;
  a.nin_m = ctx->nin_m;
  a.nex_m = ctx->nex_m;
  a.T = ctx->tileT;
  a.maxdeg = ctx->tileMD;
  a.n_inlet = ctx->fuse_inlet ? ctx->n_inlet : 0;
  a.xi = in[0];
  a.qx = in[1];
  a.qy = in[2];
  a.xo = out[0];
  a.qxo = out[1];
  a.qyo = out[2];
  a.clamp = clamp;
  a.hstill = ctx->hstill;
  a.zb = ctx->zb;
  a.n = ctx->n;
  a.S0x = ctx->S0x;
  a.S0y = ctx->S0y;
  a.area = ctx->area;
  a.deg = ctx->deg_full ? nullptr : ctx->deg;  // NULL: every cell has tileMD faces
  a.items = ctx->items;
  a.tile_in_off = ctx->tile_in_off;
  a.tile_in_list = ctx->tile_in_list;
  a.xcd_chunk = ctx->xcd_chunk;
  a.g_kind = ctx->g_kind;
  a.g_bc = ctx->g_bc;
  a.g_hstill = ctx->g_hstill;
  a.g_zb = ctx->g_zb;
  a.g_L23 = ctx->g_L23;
  a.inletQ = ctx->inletQ;
  a.exitWSE = ctx->exitWSE;
  a.in_off = ctx->in_off;
  a.in_cell = ctx->in_cell;
  a.in_L53 = ctx->in_L53;
  a.totalA = ctx->totalA;
  a.flag = ctx->flags;
  a.g = ctx->c.g;
  a.half_g = 0.5 * ctx->c.g;
  a.kn2 = ctx->c.k_n * ctx->c.k_n;
  a.hsmall = ctx->c.h_small;
  a.dt = dt;
  a.dtc = ctx->dt_cell;
  a.fmag = ctx->lux.choice == hg::LUX_FLOW_RESISTANCE ? ctx->lux_fmag : nullptr;
  return a;
}

hg::TileArgs make_tile_args(hg_ctx* ctx, int src, int dst, double dt, double* o0 = nullptr, double* o1 = nullptr,
                            double* o2 = nullptr) {
  double* out[3] = {o0 ? o0 : ctx->st[dst][0], o1 ? o1 : ctx->st[dst][1], o2 ? o2 : ctx->st[dst][2]};
  return make_tile_args_p(ctx, ctx->st[src], out, dt);
}

StepArgs make_args_p(hg_ctx* ctx, const double* const in[3], double* const out[3], double dt,
                     uint8_t* clamp = nullptr) {
  StepArgs a{};
  a.flag = ctx->flags;
  a.N = int32_t(ctx->Nm);
  a.K = ctx->K;
  a.nin_m = ctx->nin_m;
  a.nex_m = ctx->nex_m;
  a.maxdeg = ctx->maxdeg;
  a.xi = in[0];
  a.qx = in[1];
  a.qy = in[2];
  a.xo = out[0];
  a.qxo = out[1];
  a.qyo = out[2];
  a.clamp = clamp;
  a.hstill = ctx->hstill;
  a.zb = ctx->zb;
  a.n = ctx->n;
  a.S0x = ctx->S0x;
  a.S0y = ctx->S0y;
  a.area = ctx->area;
  a.deg = ctx->deg;
  a.nbr = ctx->nbr;
  a.fnx = ctx->fnx;
  a.fny = ctx->fny;
  a.flen = ctx->flen;
  a.g_kind = ctx->g_kind;
  a.g_bc = ctx->g_bc;
  a.g_hstill = ctx->g_hstill;
  a.g_zb = ctx->g_zb;
  a.g_L23 = ctx->g_L23;
  a.inletQ = ctx->inletQ;
  a.exitWSE = ctx->exitWSE;
  a.totalA = ctx->totalA;
  a.g = ctx->c.g;
  a.half_g = 0.5 * ctx->c.g;
  a.kn2 = ctx->c.k_n * ctx->c.k_n;
  a.hsmall = ctx->c.h_small;
  a.dt = dt;
  a.dtc = ctx->dt_cell;
  a.fmag = ctx->lux.choice == hg::LUX_FLOW_RESISTANCE ? ctx->lux_fmag : nullptr;
  return a;
}

StepArgs make_args(hg_ctx* ctx, int src, int dst, double dt, double* out0 = nullptr, double* out1 = nullptr,
                   double* out2 = nullptr) {
  double* out[3] = {out0 ? out0 : ctx->st[dst][0], out1 ? out1 : ctx->st[dst][1], out2 ? out2 : ctx->st[dst][2]};
  return make_args_p(ctx, ctx->st[src], out, dt);
}

hg::InletArgs inlet_args_p(hg_ctx* ctx, const double* xi) {
  hg::InletArgs a{};
  a.n_inlet = ctx->n_inlet;
  a.nin_m = ctx->nin_m;
  a.Nm = int32_t(ctx->Nm);
  a.off = ctx->in_off;
  a.cell = ctx->in_cell;
  a.L53 = ctx->in_L53;
  a.xi = xi;
  a.hstill = ctx->hstill;
  a.n = ctx->n;
  a.hsmall = ctx->c.h_small;
  a.totalA = ctx->totalA;
  a.flag = ctx->flags;
  return a;
}

hg::InletArgs inlet_args(hg_ctx* ctx, int src) {
  hg::InletArgs a{};
  a.n_inlet = ctx->n_inlet;
  a.nin_m = ctx->nin_m;
  a.Nm = int32_t(ctx->Nm);
  a.off = ctx->in_off;
  a.cell = ctx->in_cell;
  a.L53 = ctx->in_L53;
  a.xi = ctx->st[src][0];
  a.hstill = ctx->hstill;
  a.n = ctx->n;
  a.hsmall = ctx->c.h_small;
  a.totalA = ctx->totalA;
  a.flag = ctx->flags;
  return a;
}

int grid_of(int64_t n) { return int((n + kBlock - 1) / kBlock); }

// Tile kernel instances: T in {128, 256, 512} cells per tile (one work-item
// per cell), MD = 3 (triangles), 4 (quads or mixed), 8 (up to octagons).
template <int T, int MD>
void launch_tile_tm(hg_ctx* ctx, const hg::TileArgs& a, bool step, hipEvent_t e0, hipEvent_t e1) {
  const size_t lds = tile_lds_bytes(T, MD);
  const dim3 grid(unsigned(ctx->ntiles) * unsigned(ctx->K));  // every tile of every field
  if (e0)  // start / stop events attached to the dispatch itself
    hipExtLaunchKernelGGL((k_swe_tile<true, T, MD>), grid, dim3(T), lds, ctx->stream, e0, e1, 0, a);
  else if (step)
    hipLaunchKernelGGL((k_swe_tile<true, T, MD>), grid, dim3(T), lds, ctx->stream, a);
  else
    hipLaunchKernelGGL((k_swe_tile<false, T, MD>), grid, dim3(T), lds, ctx->stream, a);
}

template <int T>
void launch_tile_t(hg_ctx* ctx, const hg::TileArgs& a, bool step, hipEvent_t e0, hipEvent_t e1) {
  if (ctx->tileMD == 3)
    launch_tile_tm<T, 3>(ctx, a, step, e0, e1);
  else if (ctx->tileMD == 4)
    launch_tile_tm<T, 4>(ctx, a, step, e0, e1);
  else
    launch_tile_tm<T, 8>(ctx, a, step, e0, e1);
}

// Launches the tiled step (or RHS) kernel; with events e0 / e1 (a step
// only) they record the dispatch's own start and end.
void launch_tile(hg_ctx* ctx, const hg::TileArgs& a, bool step, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (ctx->tileT == 128)
    launch_tile_t<128>(ctx, a, step, e0, e1);
  else if (ctx->tileT == 256)
    launch_tile_t<256>(ctx, a, step, e0, e1);
  else
    launch_tile_t<512>(ctx, a, step, e0, e1);
}

hg::UdeArgs ude_args(hg_ctx* ctx, const double* const q[3]) {
  hg::UdeArgs u{};
  u.N = int32_t(ctx->N);
  u.Nm = int32_t(ctx->Nm);
  u.n_sets = ctx->ude_nsets;
  u.xi = q[0];
  u.qx = q[1];
  u.qy = q[2];
  u.hstill = ctx->hstill;
  u.zb = ctx->zb;
  u.set_of_cell = ctx->ude_set;
  u.w = ctx->ude_w;
  u.n = ctx->n;
  u.hsmall = ctx->c.h_small;
  return u;
}

int ude_blocks(const hg_ctx* ctx) { return int((ctx->N + 63) / 64); }

// Chunks per reverse-mode block of the closure: one while the mesh has at
// most 3,072 chunks (the waves the kernel holds at once on 256 CUs, 3 per
// SIMD), else the fewest that keep the blocks within 3,072.  A function of the
// chunk count alone, so a mesh's weight cotangents group its cells the same
// way on every run.
int ude_vjp_cpb(int nchunks) { return std::max(1, (nchunks + 3071) / 3072); }
int ude_vjp_blocks(const hg_ctx* ctx) { return (ude_blocks(ctx) + ctx->ude_cpb - 1) / ctx->ude_cpb; }

// n <- 0.01 + softplus(MLP(features of state q)) when the closure is on.
int enqueue_ude_n(hg_ctx* ctx, const double* const q[3]) {
  if (!ctx->ude_prec) return HG_OK;
  const hg::UdeArgs u = ude_args(ctx, q);
  const int chunks = ude_blocks(ctx);  // 64-cell chunks
  if (ctx->ude_prec == 32 && chunks <= 2048) {
    // a small mesh: 32-cell chunks, twice the waves, each half as long (the
    // same n per cell); still within the 4,096 wave slots the kernel can hold
    const int chunks32 = int((ctx->N + 31) / 32);
    hipLaunchKernelGGL((k_ude_n<float, 4, 32>), dim3((chunks32 + 3) / 4), dim3(256), 0, ctx->stream, u);
  } else if (ctx->ude_prec == 32) {
    const int grid = std::min((chunks + 3) / 4, 4096);
    hipLaunchKernelGGL((k_ude_n<float, 4>), dim3(grid), dim3(256), 0, ctx->stream, u);
  } else {
    const int grid = std::min((chunks + 1) / 2, 8192);
    hipLaunchKernelGGL((k_ude_n<double, 2>), dim3(grid), dim3(128), 0, ctx->stream, u);
  }
  HG_TRY(hipGetLastError());
  return HG_OK;
}

hg::NLawArgs nlaw_args(hg_ctx* ctx, const double* const q[3]) {
  hg::NLawArgs a{};
  a.law = ctx->nlaw;
  a.N = int32_t(ctx->N);
  a.xi = q[0];
  a.qx = q[1];
  a.qy = q[2];
  a.hstill = ctx->hstill;
  a.ks = ctx->nlaw_ks;
  a.hsmall = ctx->c.h_small;
  a.n = ctx->n;
  return a;
}

// n <- law(state q) when a variable-n law is set.
int enqueue_nlaw_n(hg_ctx* ctx, const double* const q[3]) {
  if (ctx->nlaw.kind == hg::NLAW_OFF) return HG_OK;
  hipLaunchKernelGGL(k_nlaw_n, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, nlaw_args(ctx, q));
  HG_TRY(hipGetLastError());
  return HG_OK;
}

hg::LuxArgs lux_args(hg_ctx* ctx, const double* const q[3]) {
  hg::LuxArgs a{};
  a.net = ctx->lux;
  a.N = int32_t(ctx->N);
  a.w = ctx->lux_w;
  a.xi = q[0];
  a.qx = q[1];
  a.qy = q[2];
  a.hstill = ctx->hstill;
  a.ks = ctx->lux_ks;
  a.hsmall = ctx->c.h_small;
  a.out = ctx->lux.choice == hg::LUX_FLOW_RESISTANCE ? ctx->lux_fmag : ctx->n;
  return a;
}

int lux_vjp_blocks(const hg_ctx* ctx) { return int(std::min<int64_t>((ctx->N + 63) / 64, 1024)); }

// n (or the friction magnitude) <- the reference's UDE network at state q.
int enqueue_lux_fwd(hg_ctx* ctx, const double* const q[3]) {
  if (ctx->lux.choice == hg::LUX_OFF) return HG_OK;
  const hg::LuxArgs a = lux_args(ctx, q);
  if (ctx->lux_mw == 4)
    hipLaunchKernelGGL(k_lux_fwd<4>, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL(k_lux_fwd<hg::kLuxMaxW>, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, a);
  HG_TRY(hipGetLastError());
  return HG_OK;
}

// The state-dependent n (or friction) in force: UDE closure, variable-n law
// or the reference's UDE network, if any.
int enqueue_cell_n(hg_ctx* ctx, const double* const q[3]) {
  if (ctx->ude_prec) return enqueue_ude_n(ctx, q);
  if (ctx->lux.choice) return enqueue_lux_fwd(ctx, q);
  return enqueue_nlaw_n(ctx, q);
}

// Enqueue one step (or one RHS evaluation) reading state `src`.
int enqueue_step(hg_ctx* ctx, int src, int dst, double dt, bool step, double* o0 = nullptr, double* o1 = nullptr,
                 double* o2 = nullptr, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  {
    const double* q[3] = {ctx->st[src][0], ctx->st[src][1], ctx->st[src][2]};
    int rc = enqueue_cell_n(ctx, q);
    if (rc) return rc;
  }
  if (ctx->tiled) {
    if (ctx->n_inlet > 0 && !ctx->fuse_inlet)
      hipLaunchKernelGGL(k_inlet_conveyance, dim3(ctx->n_inlet), dim3(kBlock), 0, ctx->stream, inlet_args(ctx, src));
    hg::TileArgs a = make_tile_args(ctx, src, dst, dt, o0, o1, o2);
    launch_tile(ctx, a, step, e0, e1);
    HG_TRY(hipGetLastError());
    return HG_OK;
  }
  if (ctx->n_inlet > 0)
    hipLaunchKernelGGL(k_inlet_conveyance, dim3(ctx->n_inlet), dim3(kBlock), 0, ctx->stream, inlet_args(ctx, src));
  StepArgs a = make_args(ctx, src, dst, dt, o0, o1, o2);
  if (e0 && step)
    hipExtLaunchKernelGGL(k_swe_step<true>, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, e0, e1, 0, a);
  else if (step)
    hipLaunchKernelGGL(k_swe_step<true>, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL(k_swe_step<false>, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, a);
  HG_TRY(hipGetLastError());
  return HG_OK;
}

// One step (or RHS) from explicit state buffers; optional clamp record and
// record of the inlet conveyances of this evaluation (totalA_out).
int enqueue_step_p(hg_ctx* ctx, const double* const in[3], double* const out[3], double dt, bool step,
                   uint8_t* clamp = nullptr, double* totalA_out = nullptr) {
  {
    int rc = enqueue_cell_n(ctx, in);
    if (rc) return rc;
  }
  hg::InletArgs ia = inlet_args_p(ctx, in[0]);
  if (totalA_out) ia.totalA = totalA_out;
  if (ctx->tiled) {
    if (ctx->n_inlet > 0 && !ctx->fuse_inlet)
      hipLaunchKernelGGL(k_inlet_conveyance, dim3(ctx->n_inlet), dim3(kBlock), 0, ctx->stream, ia);
    hg::TileArgs a = make_tile_args_p(ctx, in, out, dt, clamp);
    if (ctx->fuse_inlet)
      a.totalA_out = totalA_out;
    else if (totalA_out)
      a.totalA = totalA_out;
    launch_tile(ctx, a, step);
  } else {
    if (ctx->n_inlet > 0)
      hipLaunchKernelGGL(k_inlet_conveyance, dim3(ctx->n_inlet), dim3(kBlock), 0, ctx->stream, ia);
    StepArgs a = make_args_p(ctx, in, out, dt, clamp);
    if (totalA_out) a.totalA = totalA_out;
    if (step)
      hipLaunchKernelGGL(k_swe_step<true>, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, a);
    else
      hipLaunchKernelGGL(k_swe_step<false>, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, a);
  }
  HG_TRY(hipGetLastError());
  return HG_OK;
}

// ---- adjoint host side
int advance_async(hg_ctx* ctx, double dt, int64_t nsteps);
int advance_timed(hg_ctx* ctx, double dt, int64_t nsteps, double* kernel_ms);
int drop_graph(hg_ctx* ctx);
int ensure_adj_tiles(hg_ctx* ctx);

// Launches the cached graph of `key` on `stream`, capturing it first from
// enqueue() (which enqueues on ctx->stream, pointed at `stream` meanwhile;
// other streams join through events).
template <class F>
int launch_seg_graph(hg_ctx* ctx, hipStream_t stream, const SegGraphKey& key, F&& enqueue) {
  if (!ctx->use_graphs) {  // plain launches
    const hipStream_t saved = ctx->stream;
    ctx->stream = stream;
    const int rc = enqueue();
    ctx->stream = saved;
    return rc;
  }
  auto it = ctx->seg_graphs.find(key);
  if (it == ctx->seg_graphs.end()) {
    hipGraph_t graph = nullptr;
    HG_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    const hipStream_t saved = ctx->stream;
    ctx->stream = stream;
    int rc = enqueue();
    ctx->stream = saved;
    const hipError_t e = hipStreamEndCapture(stream, &graph);
    if (rc) {
      if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) {
      ctx->err = std::string("hipStreamEndCapture (adjoint segment): ") + hipGetErrorString(e);
      return HG_ERR_HIP;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      ctx->err = std::string("hipGraphInstantiate (adjoint segment): ") + hipGetErrorString(ei);
      return HG_ERR_HIP;
    }
    it = ctx->seg_graphs.emplace(key, exec).first;
  }
  HG_TRY(hipGraphLaunch(it->second, stream));
  return HG_OK;
}

int ensure_adj(hg_ctx* ctx, int64_t kck, int64_t nseg) {
  const int64_t N = ctx->N;
  int rc = 0;
  if (!ctx->adj_lam[0]) {
    rc = dalloc(ctx, &ctx->adj_lam[0], 3 * N);
    if (!rc) rc = dalloc(ctx, &ctx->adj_lam[1], 3 * N);
    if (!rc) rc = dalloc(ctx, &ctx->adj_nbar, N);
    if (!rc) rc = dalloc(ctx, &ctx->adj_zbbar, N);
    if (!rc) rc = dalloc(ctx, &ctx->adj_S0bar, 2 * N);
    if (!rc) rc = dalloc(ctx, &ctx->adj_vbar, std::max<int64_t>(ctx->G, 1));
    if (!rc) rc = dalloc(ctx, &ctx->adj_Qbar, std::max<int32_t>(ctx->n_inlet, 1));
    if (rc) return rc;
  }
  // Segment buffers: two (one recomputed while the other is reversed), or
  // one when a single segment covers the run.  Grown when kck grows (the
  // segment graphs that captured the old ones are dropped first).
  const int nbuf = nseg > 1 ? 2 : 1;
  for (int b = 0; b < nbuf && !rc; ++b) {
    if (size_t(kck) > ctx->adj_seg_cap[b]) {
      rc = drop_graph(ctx);
      if (!rc) rc = dfree(ctx, ctx->adj_seg[b]);
      if (!rc) rc = dfree(ctx, ctx->adj_mask[b]);
      if (!rc) rc = dfree(ctx, ctx->adj_totalA[b]);
      if (!rc) rc = dalloc(ctx, &ctx->adj_seg[b], size_t(kck + 1) * 3 * N);
      if (!rc) rc = dalloc(ctx, &ctx->adj_mask[b], size_t(kck) * N);
      if (!rc) rc = dalloc(ctx, &ctx->adj_totalA[b], size_t(kck) * std::max<int32_t>(ctx->n_inlet, 1));
      if (!rc) ctx->adj_seg_cap[b] = size_t(kck);
    }
    if (!rc && ctx->ude_prec && size_t(kck) > ctx->adj_nseg_cap[b]) {  // the closure's n of each step
      rc = drop_graph(ctx);
      if (!rc) rc = dfree(ctx, ctx->adj_nseg[b]);
      if (!rc) rc = dalloc(ctx, &ctx->adj_nseg[b], size_t(kck) * N);
      if (!rc) ctx->adj_nseg_cap[b] = size_t(kck);
    }
  }
  if (rc) return rc;
  if (!ctx->stream2) {  // the recompute stream (default priority) and its events
    HG_TRY(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
    for (hipEvent_t* e : {&ctx->ev_seg[0], &ctx->ev_seg[1], &ctx->ev_rev[0], &ctx->ev_rev[1], &ctx->ev_fwd})
      HG_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }
  if (size_t(nseg) > ctx->adj_ckpt_cap) {
    rc = dfree(ctx, ctx->adj_ckpt);
    if (!rc) rc = dalloc(ctx, &ctx->adj_ckpt, size_t(nseg) * 3 * N);
    if (rc) return rc;
    ctx->adj_ckpt_cap = size_t(nseg);
  }
  if (ctx->tiled) return ensure_adj_tiles(ctx);  // before any capture: it may build and upload
  return HG_OK;
}

hg::AdjArgs make_adj_args(hg_ctx* ctx, const double* const q[3], const double* const lam[3], const uint8_t* mask,
                          double* const out[3], double dt) {
  hg::AdjArgs a{};
  a.N = int32_t(ctx->Nm);
  a.K = ctx->K;
  a.nin_m = ctx->nin_m;
  a.nex_m = ctx->nex_m;
  a.G_m = int32_t(ctx->Gm);
  a.maxdeg = ctx->maxdeg;
  a.xi = q[0]; a.qx = q[1]; a.qy = q[2];
  a.lx = lam[0]; a.lqx = lam[1]; a.lqy = lam[2];
  a.mask = mask;
  a.ox = out[0]; a.oqx = out[1]; a.oqy = out[2];
  a.nbar = ctx->adj_nbar;
  a.zbbar = ctx->adj_want_zb ? ctx->adj_zbbar : nullptr;
  a.S0bar = ctx->adj_want_zb ? ctx->adj_S0bar : nullptr;
  a.vbar = ctx->adj_vbar;
  a.hstill = ctx->hstill;
  a.zb = ctx->zb;
  a.n = ctx->n;
  a.S0x = ctx->S0x;
  a.S0y = ctx->S0y;
  a.area = ctx->area;
  a.deg = ctx->deg;
  a.nbr = ctx->nbr;
  a.fnx = ctx->fnx;
  a.fny = ctx->fny;
  a.flen = ctx->flen;
  a.g_kind = ctx->g_kind;
  a.g_bc = ctx->g_bc;
  a.g_hstill = ctx->g_hstill;
  a.g_zb = ctx->g_zb;
  a.g_L23 = ctx->g_L23;
  a.inletQ = ctx->inletQ;
  a.exitWSE = ctx->exitWSE;
  a.totalA = ctx->totalA;
  a.g = ctx->c.g;
  a.half_g = 0.5 * ctx->c.g;
  a.kn2 = ctx->c.k_n * ctx->c.k_n;
  a.hsmall = ctx->c.h_small;
  a.dt = dt;
  a.dtc = ctx->dt_cell;
  if (ctx->lux.choice == hg::LUX_FLOW_RESISTANCE) {
    a.fmag = ctx->lux_fmag;
    a.fbar = ctx->lux_obar;
  }
  return a;
}

// The reverse step runs tiled wherever the step does.  Its tile size is its
// own: 256 cells, or 128 when 256-cell tiles would give fewer than two
// rounds of three work-groups per CU (its results do not depend on the
// tiling).  It uses the step kernel's work items when they have that tile
// size, else its own copy, built once.
int ensure_adj_tiles(hg_ctx* ctx) {
  const int64_t blocks256 = (ctx->Nm + 255) / 256 * ctx->K;
  int T = ctx->adj_T_req ? ctx->adj_T_req : (blocks256 >= int64_t(2) * 3 * ctx->ncu ? 256 : 128);
  if (adj_tile_lds_bytes(T, ctx->tileMD) > ctx->lds_max) T = 128;
  if (T == ctx->tileT) {
    ctx->adj_T = T;
    ctx->adj_ntiles = ctx->ntiles;
    ctx->adj_items = ctx->items;
    return HG_OK;
  }
  if (T != ctx->adj_own_T) {
    const int64_t N = ctx->Nm;
    TileBuild tb = build_tiles(N, T, ctx->h_deg, ctx->h_nbr, ctx->h_nx, ctx->h_ny, ctx->h_len, ctx->h_gkind,
                               ctx->h_gbc);
    HG_TRY(hipStreamSynchronize(ctx->stream));
    int rc = drop_graph(ctx);
    if (!rc) rc = free_items(ctx, ctx->adj_own);
    if (!rc) rc = upload_items(ctx, tb, ctx->adj_own);
    if (rc) return rc;
    ctx->adj_own_T = T;
    ctx->adj_own_ntiles = int32_t((N + T - 1) / T);
  }
  ctx->adj_T = T;
  ctx->adj_ntiles = ctx->adj_own_ntiles;
  ctx->adj_items = ctx->adj_own;
  return HG_OK;
}

template <int T, int MD>
void launch_adj_tile_tm(hg_ctx* ctx, const hg::AdjArgs& a) {
  const dim3 grid(unsigned(ctx->adj_ntiles) * unsigned(ctx->K));
  hipLaunchKernelGGL((k_adj_tile<true, T, MD>), grid, dim3(T), adj_tile_lds_bytes(T, MD), ctx->stream, a);
}

template <int T>
void launch_adj_tile_t(hg_ctx* ctx, const hg::AdjArgs& a) {
  if (ctx->tileMD == 3)
    launch_adj_tile_tm<T, 3>(ctx, a);
  else if (ctx->tileMD == 4)
    launch_adj_tile_tm<T, 4>(ctx, a);
  else
    launch_adj_tile_tm<T, 8>(ctx, a);
}

void launch_adj_tile(hg_ctx* ctx, hg::AdjArgs a) {
  if (ctx->deg_full) a.deg = nullptr;  // every cell has tileMD faces
  a.items = ctx->adj_items;
  a.xcd_chunk = ctx->xcd_chunk;
  if (ctx->adj_T == 128)
    launch_adj_tile_t<128>(ctx, a);
  else if (ctx->adj_T == 256)
    launch_adj_tile_t<256>(ctx, a);
  else
    launch_adj_tile_t<512>(ctx, a);
}

// Cotangent of one RHS evaluation (step = false) or one step (step = true)
// at state q, given the output cotangent lam; writes the state cotangent to
// out and accumulates into adj_nbar / adj_zbbar / adj_Qbar / adj_S0bar.
// totalA: the inlet conveyances of q when the forward pass recorded them
// (else they are recomputed here).  n_ready: ctx->n already holds the
// closure's n of q (recorded by the recompute), so it is not evaluated again.
int enqueue_adjoint(hg_ctx* ctx, const double* const q[3], const double* const lam[3], const uint8_t* mask,
                    double* const out[3], double dt, bool step, const double* totalA = nullptr,
                    bool n_ready = false) {
  const bool ude = ctx->ude_prec != 0;
  const bool law = ctx->nlaw.kind != hg::NLAW_OFF;
  const bool lux = ctx->lux.choice != hg::LUX_OFF;
  const bool lux_n = lux && ctx->lux.choice != hg::LUX_FLOW_RESISTANCE;  // the network gives n
  // with a state-dependent n, n-bar of this step only
  double* nbar = ude ? ctx->ude_nbar : law ? ctx->nlaw_nbar : lux_n ? ctx->lux_obar
                                                           : (ctx->adj_want_n ? ctx->adj_nbar : nullptr);
  // the closure's, law's or network's n-bar (or friction-magnitude bar) of
  // this step is set by the adjoint kernel, not cleared and added to
  if (ude && !n_ready) {
    int rc = enqueue_ude_n(ctx, q);
    if (rc) return rc;
  }
  if (law) {
    int rc = enqueue_nlaw_n(ctx, q);
    if (rc) return rc;
  }
  if (lux) {
    int rc = enqueue_lux_fwd(ctx, q);
    if (rc) return rc;
  }
  if (ctx->n_inlet > 0 && !totalA) {
    hipLaunchKernelGGL(k_inlet_conveyance, dim3(ctx->n_inlet), dim3(kBlock), 0, ctx->stream, inlet_args_p(ctx, q[0]));
    totalA = ctx->totalA;
  }
  hg::AdjArgs a = make_adj_args(ctx, q, lam, mask, out, dt);
  a.nbar = nbar;
  a.nbar_set = (ude || law || lux_n) ? 1 : 0;
  a.fbar_set = (lux && !lux_n) ? 1 : 0;
  if (totalA) a.totalA = totalA;
  if (step && ctx->tiled)
    launch_adj_tile(ctx, a);
  else if (step)
    hipLaunchKernelGGL(k_adj_cell<true>, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL(k_adj_cell<false>, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, a);
  if (ctx->n_inlet > 0)
    hipLaunchKernelGGL(k_inlet_adj, dim3(ctx->inlets_disjoint ? ctx->n_inlet : 1), dim3(kBlock), 0, ctx->stream,
                       ctx->n_inlet, ctx->nin_m, int(ctx->Nm), int(ctx->Gm), ctx->in_off, ctx->in_rec, q[0],
                       ctx->hstill, ctx->n, ctx->inletQ,
                       a.totalA, ctx->adj_vbar, ctx->c.h_small, out[0], nbar, ctx->adj_Qbar);
  if (ude) {  // through the closure: n(Q) adds to the state cotangent, and the weight cotangent accumulates
    hg::UdeArgs u = ude_args(ctx, q);
    u.nbar = ctx->ude_nbar;
    u.ox = out[0];
    u.oqx = out[1];
    u.oqy = out[2];
    u.zbbar = ctx->adj_zbbar;
    u.wpart = ctx->ude_wpart;
    u.blk_off = ctx->ude_blk_off;
    u.blk_set = ctx->ude_blk_set;
    u.blk_slot = ctx->ude_blk_slot;
    u.nslots = ctx->ude_nslots;
    u.cpb = ctx->ude_cpb;
    if (ctx->ude_prec == 32)
      hipLaunchKernelGGL(k_ude_vjp<float>, dim3(ude_vjp_blocks(ctx)), dim3(64), 0, ctx->stream, u);
    else
      hipLaunchKernelGGL(k_ude_vjp<double>, dim3(ude_vjp_blocks(ctx)), dim3(64), 0, ctx->stream, u);
  }
  if (law) {  // through the law: n(Q) adds to the state cotangent
    hg::NLawArgs u = nlaw_args(ctx, q);
    u.nbar = ctx->nlaw_nbar;
    u.ox = out[0];
    u.oqx = out[1];
    u.oqy = out[2];
    hipLaunchKernelGGL(k_nlaw_vjp, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, u);
  }
  if (lux) {  // through the network: its output adds to the state cotangent; weight cotangents accumulate
    hg::LuxArgs u = lux_args(ctx, q);
    u.obar = ctx->lux_obar;
    u.ox = out[0];
    u.oqx = out[1];
    u.oqy = out[2];
    u.wpart = ctx->lux_wpart;
    const int nb = lux_vjp_blocks(ctx);
    if (ctx->lux_mw == 4)
      hipLaunchKernelGGL(k_lux_vjp<4>, dim3(nb), dim3(64), 0, ctx->stream, u);
    else
      hipLaunchKernelGGL(k_lux_vjp<hg::kLuxMaxW>, dim3(nb), dim3(64), 0, ctx->stream, u);
  }
  HG_TRY(hipGetLastError());
  return HG_OK;
}

// After the last reverse step: the bed-slope transpose of the summed S0
// cotangent (the map zb -> S0 does not change over the steps), and the
// weight cotangents of the closure or network: each reverse-step block
// summed its share over the steps into its own slot; the slots are summed
// here by a fixed tree.
int finish_adjoint(hg_ctx* ctx) {
  if (ctx->adj_want_zb)
    hipLaunchKernelGGL(k_bed_adj, dim3(grid_of(ctx->N)), dim3(kBlock), 0, ctx->stream, int(ctx->N), int(ctx->Nm),
                     ctx->deg, ctx->nbr, ctx->fnx, ctx->fny, ctx->flen, ctx->area, ctx->adj_S0bar, ctx->adj_zbbar);
  if (ctx->ude_prec)
    hipLaunchKernelGGL(k_ude_wsum, dim3(hg::kUdeW, ctx->ude_nsets), dim3(kBlock), 0, ctx->stream, ctx->ude_nslots,
                       ctx->ude_set_start, ctx->ude_wpart, int64_t(1), int64_t(hg::kUdeW), ctx->ude_wbar);
  if (ctx->lux.choice)
    hipLaunchKernelGGL(k_ude_wsum, dim3(ctx->lux.P), dim3(kBlock), 0, ctx->stream, lux_vjp_blocks(ctx), nullptr,
                       ctx->lux_wpart, int64_t(lux_vjp_blocks(ctx)), int64_t(1), ctx->lux_wbar);
  HG_TRY(hipGetLastError());
  return HG_OK;
}

int zero_adj_accumulators(hg_ctx* ctx) {
  if (ctx->ude_prec) {
    HG_TRY(hipMemsetAsync(ctx->ude_wbar, 0, size_t(ctx->ude_nsets) * hg::kUdeW * sizeof(double), ctx->stream));
    HG_TRY(hipMemsetAsync(ctx->ude_wpart, 0, size_t(hg::kUdeW) * ctx->ude_nslots * sizeof(double), ctx->stream));
  }
  if (ctx->lux.choice) {
    HG_TRY(hipMemsetAsync(ctx->lux_wbar, 0, ctx->lux.P * sizeof(double), ctx->stream));
    HG_TRY(hipMemsetAsync(ctx->lux_wpart, 0, size_t(ctx->lux.P) * lux_vjp_blocks(ctx) * sizeof(double),
                          ctx->stream));
  }
  if (ctx->adj_want_zb) {
    HG_TRY(hipMemsetAsync(ctx->adj_S0bar, 0, 2 * ctx->N * sizeof(double), ctx->stream));
    HG_TRY(hipMemsetAsync(ctx->adj_zbbar, 0, ctx->N * sizeof(double), ctx->stream));
  }
  if (ctx->adj_want_n) HG_TRY(hipMemsetAsync(ctx->adj_nbar, 0, ctx->N * sizeof(double), ctx->stream));
  HG_TRY(hipMemsetAsync(ctx->adj_Qbar, 0, std::max<int32_t>(ctx->n_inlet, 1) * sizeof(double), ctx->stream));
  HG_TRY(hipMemsetAsync(ctx->adj_vbar, 0, std::max<int64_t>(ctx->G, 1) * sizeof(double), ctx->stream));
  return HG_OK;
}

int perm_out(hg_ctx* ctx, const double* const s[3], int ncomp, double* dst);

// Per-cell device array (library order) -> host array (caller order).
int cell_out(hg_ctx* ctx, const double* dsrc, double* hdst) {
  const double* const s[3] = {dsrc, dsrc, dsrc};
  return perm_out(ctx, s, 1, hdst);
}

double* seg_state(hg_ctx* ctx, int64_t t, int c, int b = 0) { return ctx->adj_seg[b] + (t * 3 + c) * ctx->N; }

// The closure's n of step t of segment buffer b (recorded by the recompute).
double* seg_n(hg_ctx* ctx, int64_t t, int b) { return ctx->adj_nseg[b] + t * ctx->N; }

// Inlet conveyances of step t of the segment being reversed (nullptr without inlets).
double* seg_totalA(hg_ctx* ctx, int64_t t, int b = 0) {
  return ctx->n_inlet > 0 ? ctx->adj_totalA[b] + t * ctx->n_inlet : nullptr;
}

// Forward nsteps from the device state, storing a checkpoint every kck steps.
int recompute_segment(hg_ctx* ctx, double dt, int64_t nsteps, int64_t kck, int64_t sgi, int b, hipStream_t stream);

// The forward sweep: a checkpoint every kck steps.  Its last segment is run
// as the reverse sweep's recompute would run it (recording every state, the
// clamps and the conveyances into segment buffer 0), so the reversal starts
// without recomputing it; with kck >= nsteps nothing is recomputed.  The
// final state ends in state slot 0.
int forward_checkpointed(hg_ctx* ctx, double dt, int64_t nsteps, int64_t kck) {
  const int64_t N = ctx->N;
  const int64_t nseg = (nsteps + kck - 1) / kck;
  for (int64_t sgi = 0; sgi < nseg; ++sgi) {
    for (int c = 0; c < 3; ++c)
      HG_TRY(hipMemcpyAsync(ctx->adj_ckpt + (sgi * 3 + c) * N, ctx->st[ctx->cur][c], N * sizeof(double),
                            hipMemcpyDeviceToDevice, ctx->stream));
    if (sgi + 1 < nseg) {
      int rc = advance_async(ctx, dt, kck);
      if (rc) return rc;
    }
  }
  int rc = recompute_segment(ctx, dt, nsteps, kck, nseg - 1, 0, ctx->stream);
  if (rc) return rc;
  const int64_t ns = nsteps - (nseg - 1) * kck;
  for (int c = 0; c < 3; ++c)
    HG_TRY(hipMemcpyAsync(ctx->st[0][c], seg_state(ctx, ns, c, 0), N * sizeof(double), hipMemcpyDeviceToDevice,
                          ctx->stream));
  ctx->cur = 0;
  return HG_OK;
}

// Recompute segment sgi from its checkpoint into segment buffer b, recording
// the clamp and the inlet conveyances of every step; enqueued on `stream`.
int recompute_segment(hg_ctx* ctx, double dt, int64_t nsteps, int64_t kck, int64_t sgi, int b, hipStream_t stream) {
  const int64_t N = ctx->N;
  const int64_t t0 = sgi * kck, ns = std::min<int64_t>(kck, nsteps - t0);
  const hipStream_t main = ctx->stream;
  ctx->stream = stream;  // every enqueue below goes to `stream`
  int rc = HG_OK;
  for (int c = 0; c < 3 && !rc; ++c)
    if (hipMemcpyAsync(seg_state(ctx, 0, c, b), ctx->adj_ckpt + (sgi * 3 + c) * N, N * sizeof(double),
                       hipMemcpyDeviceToDevice, stream) != hipSuccess)
      rc = HG_ERR_HIP;
  ctx->stream = main;
  if (rc) return rc;
  // the ns steps from the segment's first state: one cached graph per (b, ns, dt)
  return launch_seg_graph(ctx, stream, SegGraphKey{0, b, 0, ns, dt, 0}, [&] {
    double* const n_field = ctx->n;
    int r = HG_OK;
    for (int64_t t = 0; t < ns && !r; ++t) {
      const double* in[3] = {seg_state(ctx, t, 0, b), seg_state(ctx, t, 1, b), seg_state(ctx, t, 2, b)};
      double* out[3] = {seg_state(ctx, t + 1, 0, b), seg_state(ctx, t + 1, 1, b), seg_state(ctx, t + 1, 2, b)};
      if (ctx->ude_prec) ctx->n = seg_n(ctx, t, b);  // the closure writes this step's n where the reversal reads it
      r = enqueue_step_p(ctx, in, out, dt, true, ctx->adj_mask[b] + t * N, seg_totalA(ctx, t, b));
    }
    ctx->n = n_field;
    return r;
  });
}

// Reverse sweep over the checkpointed trajectory; lam (device, 3 x N) holds
// the cotangent of the final state on entry and of the initial state on exit.
// Segment s - 1 is recomputed on the second stream into the other segment
// buffer while segment s is reversed on the context stream (events order
// the two); with a state-dependent n (closure, law or network), whose
// kernels rewrite the shared n every step, the two run in turn on one stream.
int reverse_sweep(hg_ctx* ctx, double dt, int64_t nsteps, int64_t kck, int& lam_cur) {
  const int64_t N = ctx->N;
  const int64_t nseg = (nsteps + kck - 1) / kck;
  const bool overlap = !ctx->ude_prec && ctx->nlaw.kind == hg::NLAW_OFF && ctx->lux.choice == hg::LUX_OFF;
  const hipStream_t rs = overlap ? ctx->stream2 : ctx->stream;
  if (overlap) {  // the recompute stream starts after everything already on the context stream
    HG_TRY(hipEventRecord(ctx->ev_fwd, ctx->stream));
    HG_TRY(hipStreamWaitEvent(ctx->stream2, ctx->ev_fwd, 0));
  }
  // the last segment is already in buffer 0 (recorded by the forward sweep)
  int rc = HG_OK;
  if (overlap) HG_TRY(hipEventRecord(ctx->ev_seg[0], ctx->stream));
  for (int64_t sgi = nseg - 1; sgi >= 0; --sgi) {
    const int b = int((nseg - 1 - sgi) & 1);
    if (overlap && sgi > 0) {  // the previous segment, into the other buffer once its last reversal is done
      if (sgi < nseg - 1) HG_TRY(hipStreamWaitEvent(rs, ctx->ev_rev[1 - b], 0));
      rc = recompute_segment(ctx, dt, nsteps, kck, sgi - 1, 1 - b, rs);
      if (rc) return rc;
      HG_TRY(hipEventRecord(ctx->ev_seg[1 - b], rs));
    }
    if (overlap) HG_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_seg[b], 0));
    const int64_t t0 = sgi * kck, ns = std::min<int64_t>(kck, nsteps - t0);
    // the segment's reversal: one cached graph per (b, starting cotangent buffer, ns, dt)
    const int lam0 = lam_cur;
    rc = launch_seg_graph(ctx, ctx->stream, SegGraphKey{1, b, lam0, ns, dt, (ctx->adj_want_n ? 1 : 0) | (ctx->adj_want_zb ? 2 : 0)}, [&] {
      int lc = lam0, r = HG_OK;
      for (int64_t t = ns - 1; t >= 0 && !r; --t) {
        const double* q[3] = {seg_state(ctx, t, 0, b), seg_state(ctx, t, 1, b), seg_state(ctx, t, 2, b)};
        double* L = ctx->adj_lam[lc];
        double* O = ctx->adj_lam[1 - lc];
        const double* lam[3] = {L, L + N, L + 2 * N};
        double* out[3] = {O, O + N, O + 2 * N};
        double* const n_field = ctx->n;
        if (ctx->ude_prec) ctx->n = seg_n(ctx, t, b);  // the closure's n of this state, from the recompute
        r = enqueue_adjoint(ctx, q, lam, ctx->adj_mask[b] + t * N, out, dt, true, seg_totalA(ctx, t, b),
                            ctx->ude_prec != 0);
        ctx->n = n_field;
        lc = 1 - lc;
      }
      return r;
    });
    if (rc) return rc;
    lam_cur = (ns & 1) ? 1 - lam0 : lam0;
    if (overlap) {
      HG_TRY(hipEventRecord(ctx->ev_rev[b], ctx->stream));
    } else if (sgi > 0) {
      rc = recompute_segment(ctx, dt, nsteps, kck, sgi - 1, 1 - b, rs);
      if (rc) return rc;
    }
  }
  return HG_OK;
}

// The error flags live in pinned, coherent host memory that the kernels
// write only on an error: clearing and reading them takes no copy on the
// stream.  Every call that clears them has no kernel of its own in flight
// (each entry point ends synchronized), and the stream is synchronized
// before they are read.
int clear_flags(hg_ctx* ctx) {
  for (int k = 0; k < 4; ++k) ctx->flags_host[k] = 0;
  return HG_OK;
}

int check_flags(hg_ctx* ctx) {
  HG_TRY(hipStreamSynchronize(ctx->stream));
  int32_t f[4];
  for (int k = 0; k < 4; ++k) f[k] = ctx->flags_host[k];
  if (f[0]) {
    ctx->err = "Total cross-sectional conveyance for an inlet-q boundary is not positive (<= 1e-10)";
    return HG_ERR_CONVEYANCE;
  }
  if (f[1]) {
    ctx->err = "non-finite value (NaN or Inf) in the state";
    return HG_ERR_NONFINITE;
  }
  return HG_OK;
}

// Finiteness check of the device state (one reduction, at the end of a call).

// Host 3N state (caller order) -> device state `slot` (library order).
// The staging buffer and device permutation of the transfers below
// (allocated on first use).
// ---- host <-> device copies of caller arrays.  A caller's array is
// ordinary (pageable) host memory, which the DMA engine reads or writes only
// through the runtime's own staging, slowly for device -> host.  Instead the
// array moves in 4 MiB chunks through two pinned buffers: host threads copy
// one chunk between the caller's array and a pinned buffer while the DMA
// engine moves the other.  The pinned buffers and the threads are the
// process's, shared by every context (one transfer at a time).

constexpr size_t kXferChunk = size_t(1) << 19;  // doubles per chunk (4 MiB)

// memcpy split over a fixed set of host threads (the caller takes part 0).
struct HostCopyPool {
  std::vector<std::thread> th;
  std::mutex m;
  std::condition_variable go, done;
  uint64_t gen = 0;
  int left = 0, parts = 1;
  char* dst = nullptr;
  const char* src = nullptr;
  size_t bytes = 0;
  explicit HostCopyPool(int workers) : parts(workers + 1) {
    for (int k = 1; k <= workers; ++k) th.emplace_back([this, k] { worker(k); });
  }
  void part(int p) const {
    const size_t per = ((bytes + size_t(parts) - 1) / size_t(parts) + 63) & ~size_t(63);
    const size_t a = std::min(bytes, per * size_t(p)), b = std::min(bytes, a + per);
    if (b > a) std::memcpy(dst + a, src + a, b - a);
  }
  void worker(int p) {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      go.wait(lk, [&] { return gen != seen; });
      seen = gen;
      lk.unlock();
      part(p);
      lk.lock();
      if (--left == 0) done.notify_one();
    }
  }
  void copy(void* d, const void* s, size_t n) {
    {
      std::lock_guard<std::mutex> g(m);
      dst = static_cast<char*>(d);
      src = static_cast<const char*>(s);
      bytes = n;
      left = parts - 1;
      ++gen;
    }
    go.notify_all();
    part(0);
    std::unique_lock<std::mutex> lk(m);
    done.wait(lk, [&] { return left == 0; });
  }
};

struct HostXfer {
  std::mutex m;            // one transfer at a time
  double* pin[2] = {nullptr, nullptr};
  HostCopyPool* pool = nullptr;  // threads live until the process ends
  pid_t pid = 0;                 // the process that made them (a forked child makes its own)
  bool ok = false;
};

HostXfer& host_xfer() {
  static HostXfer x;
  return x;
}

// The pinned buffers and threads, made on first use; false if pinned memory
// is not available (the plain copies are used then).
bool host_xfer_ready(HostXfer& x) {
  if (x.ok && x.pid == getpid()) return true;
  if (x.ok) {  // a fork: the parent's threads did not come along (its pool is left as it is)
    x.pool = nullptr;
    x.ok = false;
  }
  if (!x.pin[0] && hipHostMalloc(reinterpret_cast<void**>(&x.pin[0]), kXferChunk * sizeof(double)) != hipSuccess)
    x.pin[0] = nullptr;
  if (!x.pin[1] && hipHostMalloc(reinterpret_cast<void**>(&x.pin[1]), kXferChunk * sizeof(double)) != hipSuccess)
    x.pin[1] = nullptr;
  if (!x.pin[0] || !x.pin[1]) return false;
  if (!x.pool) {
    const unsigned hw = std::thread::hardware_concurrency();
    x.pool = new HostCopyPool(int(std::min(7u, hw > 1 ? hw / 2 : 0u)));
  }
  x.pid = getpid();
  x.ok = true;
  return true;
}

int ensure_xfer_events(hg_ctx* ctx) {
  for (hipEvent_t& e : ctx->ev_xfer)
    if (!e) HG_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return HG_OK;
}

// n doubles: caller's host array src -> device dst (ctx->stream; returns
// with the copies enqueued after the last host chunk was staged).
int host_to_device(hg_ctx* ctx, double* dst, const double* src, size_t n) {
  HostXfer& x = host_xfer();
  std::lock_guard<std::mutex> g(x.m);
  if (!host_xfer_ready(x)) {
    HG_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HG_TRY(hipStreamSynchronize(ctx->stream));
    return HG_OK;
  }
  if (int rc = ensure_xfer_events(ctx)) return rc;
  for (size_t off = 0, k = 0; off < n; off += kXferChunk, ++k) {
    const int b = int(k & 1);
    const size_t len = std::min(kXferChunk, n - off);
    if (k >= 2) HG_TRY(hipEventSynchronize(ctx->ev_xfer[b]));  // the DMA that last read pin[b]
    x.pool->copy(x.pin[b], src + off, len * sizeof(double));
    HG_TRY(hipMemcpyAsync(dst + off, x.pin[b], len * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HG_TRY(hipEventRecord(ctx->ev_xfer[b], ctx->stream));
  }
  HG_TRY(hipStreamSynchronize(ctx->stream));  // the pinned buffers are free for the next transfer
  return HG_OK;
}

// n doubles: device src -> caller's host array dst, after the work already
// on ctx->stream; synchronous.
int device_to_host(hg_ctx* ctx, double* dst, const double* src, size_t n) {
  HostXfer& x = host_xfer();
  std::lock_guard<std::mutex> g(x.m);
  if (!host_xfer_ready(x)) {
    HG_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HG_TRY(hipStreamSynchronize(ctx->stream));
    return HG_OK;
  }
  if (int rc = ensure_xfer_events(ctx)) return rc;
  const size_t nch = (n + kXferChunk - 1) / kXferChunk;
  auto issue = [&](size_t k) -> int {
    const int b = int(k & 1);
    const size_t off = k * kXferChunk, len = std::min(kXferChunk, n - off);
    HG_TRY(hipMemcpyAsync(x.pin[b], src + off, len * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HG_TRY(hipEventRecord(ctx->ev_xfer[b], ctx->stream));
    return HG_OK;
  };
  for (size_t k = 0; k < std::min<size_t>(2, nch); ++k)
    if (int rc = issue(k)) return rc;
  for (size_t k = 0; k < nch; ++k) {
    const int b = int(k & 1);
    const size_t off = k * kXferChunk, len = std::min(kXferChunk, n - off);
    HG_TRY(hipEventSynchronize(ctx->ev_xfer[b]));
    x.pool->copy(dst + off, x.pin[b], len * sizeof(double));
    if (k + 2 < nch)
      if (int rc = issue(k + 2)) return rc;
  }
  return HG_OK;
}

int ensure_xfer(hg_ctx* ctx) {
  if (ctx->stage) return HG_OK;
  int rc = dalloc(ctx, &ctx->d_perm, size_t(ctx->N));
  if (!rc) rc = upload(ctx, ctx->d_perm, ctx->perm.data(), size_t(ctx->N));
  if (!rc) rc = dalloc(ctx, &ctx->stage, size_t(3) * ctx->N);
  return rc;
}

// Host array of ncomp x N values in caller order -> device arrays d[c] in
// library order: one contiguous copy, then the permutation on the device.
int perm_in(hg_ctx* ctx, const double* src, int ncomp, double* const d[3]) {
  const int64_t N = ctx->N;
  int rc = ensure_xfer(ctx);
  if (rc) return rc;
  rc = host_to_device(ctx, ctx->stage, src, size_t(ncomp) * N);
  if (rc) return rc;
  hipLaunchKernelGGL(k_perm_in, dim3(grid_of(N)), dim3(kBlock), 0, ctx->stream, N, ncomp, ctx->d_perm, ctx->stage,
                     d[0], d[ncomp > 1 ? 1 : 0], d[ncomp > 1 ? 2 : 0]);
  HG_TRY(hipGetLastError());
  HG_TRY(hipStreamSynchronize(ctx->stream));
  return HG_OK;
}

// Device arrays s[c] in library order -> host array in caller order.
int perm_out(hg_ctx* ctx, const double* const s[3], int ncomp, double* dst) {
  const int64_t N = ctx->N;
  int rc = ensure_xfer(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_perm_out, dim3(grid_of(N)), dim3(kBlock), 0, ctx->stream, N, ncomp, ctx->d_perm, s[0],
                     s[ncomp > 1 ? 1 : 0], s[ncomp > 1 ? 2 : 0], ctx->stage);
  HG_TRY(hipGetLastError());
  return device_to_host(ctx, dst, ctx->stage, size_t(ncomp) * N);
}

int state_in(hg_ctx* ctx, const double* Q, int slot) {
  double* const d[3] = {ctx->st[slot][0], ctx->st[slot][1], ctx->st[slot][2]};
  return perm_in(ctx, Q, 3, d);
}

int state_out(hg_ctx* ctx, double* const src[3], double* Q) {
  const double* const s[3] = {src[0], src[1], src[2]};
  return perm_out(ctx, s, 3, Q);
}

int cell_in(hg_ctx* ctx, double* dst, const double* src) {  // per-cell host array -> device (lib order)
  double* const d[3] = {dst, dst, dst};
  return perm_in(ctx, src, 1, d);
}

// A per-cell array of the mesh (field 0's part of a caller array over all
// fields) -> device, library order: the renumbering of field 0 is the mesh's.
int mesh_cell_in(hg_ctx* ctx, double* dst, const double* src) {
  if (ctx->K == 1) return cell_in(ctx, dst, src);
  const int64_t Nm = ctx->Nm;
  HG_TRY(hipMemcpyAsync(ctx->stage, src, size_t(Nm) * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_perm_in, dim3(grid_of(Nm)), dim3(kBlock), 0, ctx->stream, Nm, 1, ctx->d_perm, ctx->stage, dst,
                     dst, dst);
  HG_TRY(hipGetLastError());
  HG_TRY(hipStreamSynchronize(ctx->stream));
  return HG_OK;
}

// Working copies of zb / n / S0 / zb_ghost / inletQ <- the values given to
// hg_set_fields (p_extra is immutable in the reference; params override it
// per call only).
int restore_base(hg_ctx* ctx) {
  const int64_t Nm = ctx->Nm;  // the mesh's bed data (shared by the fields); n per field
  HG_TRY(hipMemcpyAsync(ctx->zb, ctx->zb0, Nm * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  HG_TRY(hipMemcpyAsync(ctx->n, ctx->n0, ctx->N * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  HG_TRY(hipMemcpyAsync(ctx->S0x, ctx->S0x0, Nm * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  HG_TRY(hipMemcpyAsync(ctx->S0y, ctx->S0y0, Nm * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  if (ctx->Gm)
    HG_TRY(hipMemcpyAsync(ctx->g_zb, ctx->g_zb0, ctx->Gm * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  if (ctx->n_inlet)
    HG_TRY(hipMemcpyAsync(ctx->inletQ, ctx->inletQ0, ctx->n_inlet * sizeof(double), hipMemcpyDeviceToDevice,
                          ctx->stream));
  HG_TRY(hipStreamSynchronize(ctx->stream));
  return HG_OK;
}

// Apply params_vector per the active parameter (semi_discretize_swe_2D.jl:114-199).
int apply_params(hg_ctx* ctx, const double* params, int64_t np) {
  const int64_t N = ctx->N;
  int rc0 = restore_base(ctx);
  if (rc0) return rc0;
  switch (ctx->active) {
    case HG_PARAM_NONE:
      break;
    case HG_PARAM_ZB: {
      if (np != N || !params) { ctx->err = "HG_PARAM_ZB expects n_params == n_cells"; return HG_ERR_ARG; }
      int rc = cell_in(ctx, ctx->zb, params);
      if (rc) return rc;
      hipLaunchKernelGGL(k_bed_data, dim3(grid_of(std::max<int64_t>(N, ctx->G))), dim3(kBlock), 0, ctx->stream,
                         int(N), int(ctx->G), ctx->zb, ctx->deg, ctx->nbr, ctx->fnx, ctx->fny, ctx->flen, ctx->area,
                         ctx->g_cell, ctx->S0x, ctx->S0y, ctx->g_zb);
      HG_TRY(hipGetLastError());
      break;
    }
    case HG_PARAM_N_ZONE: {
      if (np != ctx->n_zones || !params) { ctx->err = "HG_PARAM_N_ZONE expects n_params == n_zones"; return HG_ERR_ARG; }
      int rc = upload(ctx, ctx->scratch, params, size_t(np));
      if (rc) return rc;
      hipLaunchKernelGGL(k_zone_gather, dim3(grid_of(N)), dim3(kBlock), 0, ctx->stream, int(N), ctx->zone,
                         ctx->scratch, ctx->n);
      HG_TRY(hipGetLastError());
      break;
    }
    case HG_PARAM_N_CELL: {
      if (np != N || !params) { ctx->err = "HG_PARAM_N_CELL expects n_params == n_cells"; return HG_ERR_ARG; }
      int rc = cell_in(ctx, ctx->n, params);
      if (rc) return rc;
      break;
    }
    case HG_PARAM_Q: {
      if (np != ctx->n_inlet || !params) { ctx->err = "HG_PARAM_Q expects n_params == n_inlet"; return HG_ERR_ARG; }
      int rc = upload(ctx, ctx->inletQ, params, size_t(np));
      if (rc) return rc;
      break;
    }
    case HG_PARAM_UDE: {
      if (np != ctx->lux.P || !params) { ctx->err = "HG_PARAM_UDE expects the network's weights"; return HG_ERR_ARG; }
      for (int64_t k = 0; k < np; ++k)
        if (!std::isfinite(params[k])) { ctx->err = "HG_PARAM_UDE: non-finite weight"; return HG_ERR_ARG; }
      int rc = upload(ctx, ctx->lux_w, params, size_t(np));
      if (rc) return rc;
      break;
    }
    default:
      ctx->err = "unknown active parameter";
      return HG_ERR_ARG;
  }
  HG_TRY(hipStreamSynchronize(ctx->stream));
  return HG_OK;
}

int drop_graph(hg_ctx* ctx) {
  if (ctx->gexec || !ctx->seg_graphs.empty()) {
    HG_TRY(hipStreamSynchronize(ctx->stream));  // no replay of them may still be running
    if (ctx->stream2) HG_TRY(hipStreamSynchronize(ctx->stream2));
  }
  if (ctx->gexec) {
    HG_TRY(hipGraphExecDestroy(ctx->gexec));
    ctx->gexec = nullptr;
  }
  for (auto& kv : ctx->seg_graphs) HG_TRY(hipGraphExecDestroy(kv.second));
  ctx->seg_graphs.clear();
  return HG_OK;
}


// (Re)build the tiled kernel's work items for tiles of T cells from the
// host copies of the face arrays.  Any T gives bit-identical results: every
// face flux is formed exactly as the per-side evaluation forms it.
int setup_tiles(hg_ctx* ctx, int T) {
  if (T != 128 && T != 256 && T != 512) {
    ctx->err = "tile size must be 128, 256 or 512 cells";
    return HG_ERR_ARG;
  }
  if (tile_lds_bytes(T, ctx->tileMD) > ctx->lds_max) {
    ctx->err = "a tile of " + std::to_string(T) + " cells needs " + std::to_string(tile_lds_bytes(T, ctx->tileMD)) +
               " bytes of LDS; the device allows " + std::to_string(ctx->lds_max);
    return HG_ERR_ARG;
  }
  const int64_t N = ctx->Nm;  // tiles of the mesh; every field of it uses them
  const int32_t maxdeg = ctx->maxdeg;
  const std::vector<int32_t>& h_deg = ctx->h_deg;
  TileBuild tb = build_tiles(N, T, h_deg, ctx->h_nbr, ctx->h_nx, ctx->h_ny, ctx->h_len, ctx->h_gkind, ctx->h_gbc);
  const int64_t ntiles = (N + T - 1) / T;
  {  // every (cell, local face) slot must receive exactly one flux, and only
     // the items before x0 (cross-tile and boundary) may need their ref
    std::vector<uint8_t> hits(size_t(maxdeg) * N, 0);
    bool ok = true;
    for (int64_t t = 0; t < ntiles && ok; ++t)
      for (int32_t it = tb.off[t]; it < tb.off[t + 1]; ++it) {
        const int32_t code = tb.code[it];
        const int32_t ia = code & 1023, ka = (code >> 10) & 15, ib = (code >> 14) & 1023, kb = (code >> 24) & 15;
        const int32_t type = (code >> 28) & 3;
        const int64_t ca = t * T + ia;
        if (ia >= T || ca >= N || ka >= h_deg[ca]) { ok = false; break; }
        if ((type == 0) != (it >= tb.x0[t])) { ok = false; break; }
        hits[size_t(ka) * N + ca]++;
        if (type == 0) {
          const int64_t cb = t * T + ib;
          if (ib >= T || cb >= N || kb >= h_deg[cb]) { ok = false; break; }
          hits[size_t(kb) * N + cb]++;
        }
      }
    for (int64_t i = 0; i < N && ok; ++i)
      for (int32_t k = 0; k < maxdeg; ++k)
        if (hits[size_t(k) * N + i] != (k < h_deg[i] ? 1 : 0)) { ok = false; break; }
    if (!ok) {
      ctx->err = "internal: tiled work items do not cover every face slot exactly once";
      return HG_ERR_MESH;
    }
  }
  // release the previous tiling (and the graph that captured it)
  HG_TRY(hipStreamSynchronize(ctx->stream));
  int rc = drop_graph(ctx);
  if (rc) return rc;
  for (void* p : ctx->tile_allocs) {
    ctx->allocs.erase(std::find(ctx->allocs.begin(), ctx->allocs.end(), p));
    HG_TRY(hipFree(p));
  }
  ctx->tile_allocs.clear();
  const size_t nalloc0 = ctx->allocs.size();
  rc = upload_items(ctx, tb, ctx->items);
  if (!rc) rc = dalloc(ctx, &ctx->tile_in_off, tb.in_off.size());
  if (!rc) rc = dalloc(ctx, &ctx->tile_in_list, tb.in_list.size());
  ctx->tile_allocs.assign(ctx->allocs.begin() + nalloc0, ctx->allocs.end());
  if (!rc) rc = upload(ctx, ctx->tile_in_off, tb.in_off.data(), tb.in_off.size());
  if (!rc) rc = upload(ctx, ctx->tile_in_list, tb.in_list.data(), tb.in_list.size());
  if (rc) return rc;
  ctx->tileT = T;
  ctx->ntiles = int32_t(ntiles);
  ctx->fuse_inlet = ctx->n_inlet > 0 && tb.max_inlets <= kTileInlets;
  return HG_OK;
}

// The cached graph of kGraphSteps steps for dt.  (A longer graph saves
// about 0.1 us per step of graph boundaries, but rocprofv3's kernel tracing
// replays a 128-step graph at half speed, so profiles would not describe
// the unprofiled run.)
int ensure_graph(hg_ctx* ctx, double dt) {
  if (!ctx->use_graphs || (ctx->gexec && ctx->gdt == dt)) return HG_OK;
  if (drop_graph(ctx)) return HG_ERR_HIP;
  hipGraph_t graph;
  HG_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int k = 0; k < kGraphSteps && !rc; ++k) rc = enqueue_step(ctx, k & 1, (k + 1) & 1, dt, true);
  hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  if (rc) return rc;
  if (e != hipSuccess) {
    ctx->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
    return HG_ERR_HIP;
  }
  HG_TRY(hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0));
  HG_TRY(hipGraphDestroy(graph));
  ctx->gdt = dt;
  return HG_OK;
}

// Advance the device state nsteps (no host sync at the end).
int advance_async(hg_ctx* ctx, double dt, int64_t nsteps) {
  if (nsteps <= 0) return HG_OK;
  if (ctx->cur == 1) {
    int rc = enqueue_step(ctx, 1, 0, dt, true);
    if (rc) return rc;
    ctx->cur = 0;
    --nsteps;
  }
  if (ctx->use_graphs && nsteps >= kGraphSteps) {
    int rc = ensure_graph(ctx, dt);
    if (rc) return rc;
    for (; nsteps >= kGraphSteps; nsteps -= kGraphSteps) HG_TRY(hipGraphLaunch(ctx->gexec, ctx->stream));
  }
  for (; nsteps > 0; --nsteps) {  // remainder, single launches
    int rc = enqueue_step(ctx, ctx->cur, 1 - ctx->cur, dt, true);
    if (rc) return rc;
    ctx->cur = 1 - ctx->cur;
  }
  return HG_OK;
}

// nsteps single-launch steps, each step kernel with its own start / stop
// events on its dispatch; *kernel_ms = their mean duration.  Synchronous.
// Entry points that launch kernels over the fields refuse to run before
// hg_set_fields has uploaded them (the kernels would read null pointers).
int require_fields(hg_ctx* ctx, const char* who) {
  if (ctx->fields_set) return HG_OK;
  ctx->err = std::string(who) + ": hg_set_fields has not been called";
  return HG_ERR_STATE;
}

// Entry points a context of several fields on one mesh does not provide.
int refuse_fields(hg_ctx* ctx, const char* who) {
  if (ctx->K == 1) return HG_OK;
  ctx->err = std::string(who) + ": not available for a context of several fields on one mesh";
  return HG_ERR_STATE;
}

int advance_timed(hg_ctx* ctx, double dt, int64_t nsteps, double* kernel_ms) {
  std::vector<hipEvent_t>& ev = ctx->ev_pool;  // grown once (at least 2 x 2048), reused by later calls
  int rc = HG_OK;
  const size_t need = size_t(2 * std::max<int64_t>(nsteps, 2048));
  while (!rc && ev.size() < need) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) rc = HG_ERR_HIP;
    else ev.push_back(e);
  }
  for (int64_t k = 0; k < nsteps && !rc; ++k) {
    rc = enqueue_step(ctx, ctx->cur, 1 - ctx->cur, dt, true, nullptr, nullptr, nullptr, ev[2 * k], ev[2 * k + 1]);
    ctx->cur = 1 - ctx->cur;
  }
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = HG_ERR_HIP;
  double sum = 0.0;
  for (int64_t k = 0; k < nsteps && !rc; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) != hipSuccess) rc = HG_ERR_HIP;
    sum += ms;
  }
  if (!rc && kernel_ms) *kernel_ms = nsteps > 0 ? sum / double(nsteps) : 0.0;
  return rc;
}

}  // namespace

// ================================================================ C ABI ====
extern "C" {

const char* hg_info(void) {
  static const std::string s = [] {  // formed once (thread-safe static initialisation)
    int dev = 0;
    hipDeviceProp_t p{};
    char buf[512];
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
      snprintf(buf, sizeof(buf), "libhgswe: %s (%s), %d CUs, %.1f GB HBM; FP64, no FMA contraction", p.name,
               p.gcnArchName, p.multiProcessorCount, double(p.totalGlobalMem) / 1e9);
    else
      snprintf(buf, sizeof(buf), "libhgswe: no HIP device");
    return std::string(buf);
  }();
  return s.c_str();
}

const char* hg_last_error(const hg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

}  // extern "C"

namespace {
// Context creation for one mesh; seg lists the cell offsets of batched
// instances ({0, N} for a single mesh), within which cells are renumbered.
int create_impl(hg_ctx** out, const hg_mesh* m, const hg_bcs* b, const hg_consts* c, int32_t flags,
                std::vector<int64_t> seg, int32_t K = 1) {
  if (!out || !m || !b || !c) {
    g_create_err = "hg_create: NULL argument";
    return HG_ERR_ARG;
  }
  *out = nullptr;
  const int64_t N = m->n_cells, G = b->n_ghosts;
  if (N <= 0 || N > (int64_t(1) << 30) || G < 0) {
    g_create_err = "hg_create: bad cell / ghost count";
    return HG_ERR_ARG;
  }
  if (!m->face_off || !m->face_nbr || !m->face_nx || !m->face_ny || !m->face_len || !m->area ||
      (G > 0 && (!b->ghost_slot || !b->ghost_kind || !b->ghost_bc)) ||
      (b->n_inlet > 0 && (!b->inlet_off || !b->inlet_ghost)) || b->n_inlet < 0 || b->n_exit < 0) {
    g_create_err = "hg_create: NULL array";
    return HG_ERR_ARG;
  }
  // ---- validate topology
  if (m->face_off[0] != 0) { g_create_err = "face_off[0] != 0"; return HG_ERR_MESH; }
  int32_t maxdeg = 0, mindeg = 16;
  for (int64_t i = 0; i < N; ++i) {
    const int32_t d = m->face_off[i + 1] - m->face_off[i];
    if (d < 3 || d > 16) { g_create_err = "cell " + std::to_string(i) + " has " + std::to_string(d) + " faces"; return HG_ERR_MESH; }
    maxdeg = std::max(maxdeg, d);
    mindeg = std::min(mindeg, d);
    if (!(m->area[i] > 0.0)) { g_create_err = "cell " + std::to_string(i) + " has non-positive area"; return HG_ERR_MESH; }
  }
  const int64_t M = m->face_off[N];
  if (int64_t(maxdeg) * N >= (int64_t(1) << 31)) {  // ELL slot k * N + i must fit an int
    g_create_err = "mesh too large for 32-bit face slots";
    return HG_ERR_ARG;
  }
  std::vector<int32_t> ghost_seen(G, -1);
  for (int64_t i = 0; i < N; ++i)
    for (int32_t s = m->face_off[i]; s < m->face_off[i + 1]; ++s) {
      const int32_t j = m->face_nbr[s];
      if (j >= 0) {
        bool back = false;
        if (j >= N || j == i) { g_create_err = "face_nbr out of range at slot " + std::to_string(s); return HG_ERR_MESH; }
        int nback = 0;
        for (int32_t t = m->face_off[j]; t < m->face_off[j + 1]; ++t) nback += (m->face_nbr[t] == i);
        back = nback == 1;
        if (!back) { g_create_err = "neighbour relation not symmetric (or cells share several faces) at slot " + std::to_string(s); return HG_ERR_MESH; }
      } else {
        const int64_t gi = ~int64_t(j);
        if (gi >= G || ghost_seen[gi] >= 0) { g_create_err = "bad ghost reference at slot " + std::to_string(s); return HG_ERR_MESH; }
        ghost_seen[gi] = s;
      }
    }
  for (int64_t gi = 0; gi < G; ++gi) {
    if (ghost_seen[gi] < 0 || b->ghost_slot[gi] != ghost_seen[gi]) { g_create_err = "ghost_slot mismatch for ghost " + std::to_string(gi); return HG_ERR_MESH; }
    const int32_t k = b->ghost_kind[gi];
    if (k < 0 || k > 3) { g_create_err = "bad ghost_kind"; return HG_ERR_MESH; }
    if ((k == HG_BC_INLETQ && (b->ghost_bc[gi] < 0 || b->ghost_bc[gi] >= b->n_inlet)) ||
        (k == HG_BC_EXITH && (b->ghost_bc[gi] < 0 || b->ghost_bc[gi] >= b->n_exit))) {
      g_create_err = "bad ghost_bc"; return HG_ERR_MESH;
    }
  }
  if (b->n_inlet > 0) {
    if (b->inlet_off[0] != 0) { g_create_err = "inlet_off[0] != 0"; return HG_ERR_MESH; }
    for (int bi = 0; bi < b->n_inlet; ++bi)
      for (int32_t f = b->inlet_off[bi]; f < b->inlet_off[bi + 1]; ++f) {
        const int32_t gi = b->inlet_ghost[f];
        if (gi < 0 || gi >= G || b->ghost_kind[gi] != HG_BC_INLETQ || b->ghost_bc[gi] != bi) {
          g_create_err = "inlet_ghost inconsistent with ghost_kind / ghost_bc"; return HG_ERR_MESH;
        }
      }
  }

  hg_ctx* ctx = new hg_ctx();
  {  // HG_GRAPHS=0: no hipGraphs (plain launches throughout; for tools that do not follow graph launches)
    const char* e = std::getenv("HG_GRAPHS");
    ctx->use_graphs = !(e && std::strcmp(e, "0") == 0);
  }
  // K fields of this mesh (1 for a plain or batched context): the counts
  // below cover every field; the geometry is the mesh's (N cells, G ghosts)
  ctx->K = K;
  ctx->Nm = N;
  ctx->Gm = G;
  ctx->nin_m = b->n_inlet;
  ctx->nex_m = b->n_exit;
  ctx->N = int64_t(K) * N;
  ctx->G = int64_t(K) * G;
  ctx->maxdeg = maxdeg;
  ctx->c = *c;
  ctx->n_inlet = K * b->n_inlet;
  ctx->n_exit = K * b->n_exit;
  ctx->n_inlet_faces = b->n_inlet ? b->inlet_off[b->n_inlet] : 0;  // the mesh's inlet faces
  auto fail = [&](int rc) {
    g_create_err = ctx->err;
    hg_destroy(ctx);
    return rc;
  };
  if (hipGetDevice(&ctx->device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    ctx->err = "no HIP device / stream";
    return fail(HG_ERR_HIP);
  }
  {  // the static scratch's address on this device (current now), once per device
    if (ctx->device < 0 || ctx->device >= kMaxDevices) {
      ctx->err = "device index beyond the per-device scratch table";
      return fail(HG_ERR_HIP);
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[ctx->device]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return fail(HG_ERR_HIP);
      }
      g_scratch_dev[ctx->device] = static_cast<double*>(p);
    }
    ctx->noalloc_scratch = g_scratch_dev[ctx->device];
  }
  // ---- renumbering
  if (seg.empty()) seg = {0, N};
  ctx->inst_off = seg;
  if ((flags & HG_REORDER) && m->centroid_xy) {
    ctx->perm = morton_perm(seg, m->centroid_xy);
  } else {
    ctx->perm.resize(N);
    std::iota(ctx->perm.begin(), ctx->perm.end(), 0);
  }
  ctx->iperm.resize(N);
  for (int64_t i = 0; i < N; ++i) ctx->iperm[ctx->perm[i]] = int32_t(i);

  // ---- ELL face arrays in library order
  const int64_t E = int64_t(maxdeg) * N;
  std::vector<int32_t> h_deg(N), h_nbr(E, 0);
  std::vector<double> h_nx(E, 0.0), h_ny(E, 0.0), h_len(E, 0.0), h_area(N);
  std::vector<int32_t> g_cell(G), g_kind(G), g_bc(G);
  std::vector<double> g_L23(G, 0.0);
  for (int64_t li = 0; li < N; ++li) {
    const int32_t oi = ctx->perm[li];
    const int32_t s0 = m->face_off[oi], d = m->face_off[oi + 1] - s0;
    h_deg[li] = d;
    h_area[li] = m->area[oi];
    for (int32_t k = 0; k < d; ++k) {
      const int32_t j = m->face_nbr[s0 + k];
      h_nbr[k * N + li] = (j >= 0) ? ctx->iperm[j] : j;
      h_nx[k * N + li] = m->face_nx[s0 + k];
      h_ny[k * N + li] = m->face_ny[s0 + k];
      h_len[k * N + li] = m->face_len[s0 + k];
    }
  }
  // slot -> owning cell (caller order)
  std::vector<int32_t> slot_cell(M);
  for (int64_t i = 0; i < N; ++i)
    for (int32_t s = m->face_off[i]; s < m->face_off[i + 1]; ++s) slot_cell[s] = int32_t(i);
  for (int64_t gi = 0; gi < G; ++gi) {
    g_cell[gi] = ctx->iperm[slot_cell[b->ghost_slot[gi]]];
    g_kind[gi] = b->ghost_kind[gi];
    g_bc[gi] = b->ghost_bc[gi];
    if (g_kind[gi] == HG_BC_INLETQ) g_L23[gi] = std::pow(m->face_len[b->ghost_slot[gi]], 2.0 / 3.0);
  }
  std::vector<int32_t> in_off(b->n_inlet + 1, 0), in_cell(ctx->n_inlet_faces);
  std::vector<double> in_L53(ctx->n_inlet_faces);
  for (int bi = 0; bi <= b->n_inlet && b->n_inlet > 0; ++bi) in_off[bi] = b->inlet_off[bi];
  for (int32_t f = 0; f < ctx->n_inlet_faces; ++f) {
    const int32_t gi = b->inlet_ghost[f];
    in_cell[f] = g_cell[gi];
    in_L53[f] = std::pow(m->face_len[b->ghost_slot[gi]], 5.0 / 3.0);
  }

  // ---- device allocation (state-like arrays for every field, mesh data once)
  const int64_t NT = ctx->N;
  int rc = 0;
  for (int s = 0; s < 2 && !rc; ++s)
    for (int k = 0; k < 3 && !rc; ++k) rc = dalloc(ctx, &ctx->st[s][k], NT);
  double** cellbufs[] = {&ctx->hstill, &ctx->zb, &ctx->S0x, &ctx->S0y, &ctx->area, &ctx->zb0, &ctx->S0x0, &ctx->S0y0};
  for (double** p : cellbufs)
    if (!rc) rc = dalloc(ctx, p, N);
  if (!rc) rc = dalloc(ctx, &ctx->n, NT);
  if (!rc) rc = dalloc(ctx, &ctx->n0, NT);
  if (!rc) rc = dalloc(ctx, &ctx->deg, N);
  if (!rc) rc = dalloc(ctx, &ctx->nbr, E);
  if (!rc) rc = dalloc(ctx, &ctx->fnx, E);
  if (!rc) rc = dalloc(ctx, &ctx->fny, E);
  if (!rc) rc = dalloc(ctx, &ctx->flen, E);
  if (!rc) rc = dalloc(ctx, &ctx->g_kind, G);
  if (!rc) rc = dalloc(ctx, &ctx->g_bc, G);
  if (!rc) rc = dalloc(ctx, &ctx->g_cell, G);
  if (!rc) rc = dalloc(ctx, &ctx->g_hstill, G);
  if (!rc) rc = dalloc(ctx, &ctx->g_zb, G);
  if (!rc) rc = dalloc(ctx, &ctx->g_zb0, G);
  if (!rc) rc = dalloc(ctx, &ctx->g_L23, G);
  if (!rc) rc = dalloc(ctx, &ctx->in_off, b->n_inlet + 1);
  if (!rc) rc = dalloc(ctx, &ctx->in_cell, ctx->n_inlet_faces);
  if (!rc) rc = dalloc(ctx, &ctx->in_L53, ctx->n_inlet_faces);
  if (!rc) rc = dalloc(ctx, &ctx->in_rec, ctx->n_inlet_faces);
  if (!rc) rc = dalloc(ctx, &ctx->inletQ, ctx->n_inlet);
  if (!rc) rc = dalloc(ctx, &ctx->inletQ0, ctx->n_inlet);
  if (!rc) rc = dalloc(ctx, &ctx->exitWSE, ctx->n_exit);
  if (!rc) rc = dalloc(ctx, &ctx->totalA, ctx->n_inlet);
  if (!rc) {
    void* hp = nullptr;
    if (hipHostMalloc(&hp, 4 * sizeof(int32_t), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
      ctx->err = "hipHostMalloc (error flags) failed";
      rc = HG_ERR_HIP;
    } else {
      ctx->flags_host = static_cast<volatile int32_t*>(hp);
      void* dp = nullptr;
      if (hipHostGetDevicePointer(&dp, hp, 0) != hipSuccess) {
        ctx->err = "hipHostGetDevicePointer (error flags) failed";
        rc = HG_ERR_HIP;
      }
      ctx->flags = static_cast<int32_t*>(dp);
    }
  }
  if (!rc) rc = dalloc(ctx, &ctx->zone, N);
  ctx->scratch_n = size_t(std::max<int64_t>(3 * NT, 4096));
  if (!rc) rc = dalloc(ctx, &ctx->scratch, ctx->scratch_n);
  if (!rc) rc = dalloc(ctx, &ctx->red, size_t(grid_of(NT) + 64));
  if (!rc && K > 1) {  // caller order <-> library order of every field: the mesh's renumbering, per field
    std::vector<int32_t> p(NT), ip(NT);
    for (int64_t f = 0; f < K; ++f)
      for (int64_t i = 0; i < N; ++i) {
        p[f * N + i] = int32_t(f * N + ctx->perm[i]);
        ip[f * N + i] = int32_t(f * N + ctx->iperm[i]);
      }
    ctx->perm.swap(p);
    ctx->iperm.swap(ip);
    ctx->inst_off.resize(K + 1);
    for (int64_t f = 0; f <= K; ++f) ctx->inst_off[f] = f * N;
  }
  if (!rc) rc = ensure_xfer(ctx);  // host transfers never allocate after set-up
  if (rc) return fail(rc);

  if (!rc) rc = upload(ctx, ctx->deg, h_deg.data(), N);
  if (!rc) rc = upload(ctx, ctx->nbr, h_nbr.data(), E);
  if (!rc) rc = upload(ctx, ctx->fnx, h_nx.data(), E);
  if (!rc) rc = upload(ctx, ctx->fny, h_ny.data(), E);
  if (!rc) rc = upload(ctx, ctx->flen, h_len.data(), E);
  if (!rc) rc = upload(ctx, ctx->area, h_area.data(), N);
  if (!rc) rc = upload(ctx, ctx->g_kind, g_kind.data(), G);
  if (!rc) rc = upload(ctx, ctx->g_bc, g_bc.data(), G);
  if (!rc) rc = upload(ctx, ctx->g_cell, g_cell.data(), G);
  if (!rc) rc = upload(ctx, ctx->g_L23, g_L23.data(), G);
  if (!rc) rc = upload(ctx, ctx->in_off, in_off.data(), b->n_inlet + 1);
  if (!rc) rc = upload(ctx, ctx->in_cell, in_cell.data(), ctx->n_inlet_faces);
  if (!rc) rc = upload(ctx, ctx->in_L53, in_L53.data(), ctx->n_inlet_faces);
  if (!rc) {  // per inlet, the faces of each cell in NodeString order (see k_inlet_adj)
    std::vector<int32_t> next(std::max<int32_t>(ctx->n_inlet_faces, 1), -1), first(next.size(), 1);
    for (int bi = 0; bi < b->n_inlet; ++bi) {
      std::unordered_map<int32_t, int32_t> last;  // cell -> its latest face in this inlet
      for (int32_t f = in_off[bi]; f < in_off[bi + 1]; ++f) {
        auto it = last.find(in_cell[f]);
        if (it == last.end()) {
          last.emplace(in_cell[f], f);
        } else {
          next[it->second] = f;
          first[f] = 0;
          it->second = f;
        }
      }
    }
    std::vector<hg::InletFaceRec> recs(std::max<int32_t>(ctx->n_inlet_faces, 1));
    for (int32_t f = 0; f < ctx->n_inlet_faces; ++f) {
      const int32_t gi = b->inlet_ghost[f];
      recs[f] = hg::InletFaceRec{gi, in_cell[f], first[f], next[f], g_L23[gi], in_L53[f]};
    }
    rc = upload(ctx, ctx->in_rec, recs.data(), ctx->n_inlet_faces);
    std::unordered_map<int32_t, int32_t> owner;  // cell -> the inlet it sits on
    ctx->inlets_disjoint = true;
    for (int bi = 0; bi < b->n_inlet && ctx->inlets_disjoint; ++bi)
      for (int32_t f = in_off[bi]; f < in_off[bi + 1]; ++f) {
        auto it = owner.emplace(in_cell[f], bi).first;
        if (it->second != bi) ctx->inlets_disjoint = false;
      }
  }
  if (!rc) rc = clear_flags(ctx);
  if (rc) return fail(rc);

  // ---- tiled face-once kernel data
  // The tiled kernel's LDS slots are sized for at most 8 faces per cell;
  // meshes with larger polygons run the per-cell kernel.
  ctx->h_deg = std::move(h_deg);
  ctx->h_nbr = std::move(h_nbr);
  ctx->h_nx = std::move(h_nx);
  ctx->h_ny = std::move(h_ny);
  ctx->h_len = std::move(h_len);
  ctx->h_gkind = std::move(g_kind);
  ctx->h_gbc = std::move(g_bc);
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) == hipSuccess && v > 0)
      ctx->lds_max = size_t(v);
  }
  ctx->tiled = !(flags & HG_CELL_KERNEL) && maxdeg <= 8;
  if (ctx->tiled) {
    ctx->tileMD = maxdeg <= 3 ? 3 : (maxdeg <= 4 ? 4 : 8);
    ctx->deg_full = mindeg == ctx->tileMD;  // all triangles or all quadrilaterals
    // 256-cell tiles, unless that leaves fewer than two work-groups per CU
    // (a small mesh fills the chip better with 128-cell tiles)
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || ncu <= 0)
      ncu = 256;
    ctx->ncu = ncu;
    const bool small = ctx->N < int64_t(2) * ncu * 256;
    const int T = (!small && tile_lds_bytes(256, ctx->tileMD) <= ctx->lds_max) ? 256 : 128;
    rc = setup_tiles(ctx, T);
    if (rc) return fail(rc);
  }
  *out = ctx;
  return HG_OK;
}
}  // namespace

extern "C" {

int hg_create(hg_ctx** out, const hg_mesh* m, const hg_bcs* b, const hg_consts* c, int32_t flags) {
  return create_impl(out, m, b, c, flags, {});
}

int hg_fields_create(hg_ctx** out, int32_t n_fields, const hg_mesh* m, const hg_bcs* b, const hg_consts* c,
                     int32_t flags) {
  if (!out || !m || n_fields < 1) {
    g_create_err = "hg_fields_create: NULL argument or no fields";
    return HG_ERR_ARG;
  }
  if (int64_t(n_fields) * m->n_cells > (int64_t(1) << 30)) {
    g_create_err = "hg_fields_create: more than 2^30 cells over all fields";
    return HG_ERR_ARG;
  }
  return create_impl(out, m, b, c, flags, {}, n_fields);
}

int hg_batch_create(hg_ctx** out, int32_t n_inst, const hg_mesh* meshes, const hg_bcs* bcs, const hg_consts* c,
                    int32_t flags) {
  if (!out || !meshes || !bcs || !c || n_inst <= 0) {
    g_create_err = "hg_batch_create: NULL argument or no instances";
    return HG_ERR_ARG;
  }
  *out = nullptr;
  auto bad = [](int32_t inst, const std::string& what) {
    g_create_err = "hg_batch_create: instance " + std::to_string(inst) + ": " + what;
    return HG_ERR_MESH;
  };
  // ---- sizes, and every index checked against its own instance before it is offset
  int64_t N = 0, M = 0, G = 0, nIF = 0;
  int64_t nI = 0, nE = 0;
  bool cen = true;
  for (int32_t r = 0; r < n_inst; ++r) {
    const hg_mesh& m = meshes[r];
    const hg_bcs& b = bcs[r];
    if (m.n_cells <= 0 || b.n_ghosts < 0 || b.n_inlet < 0 || b.n_exit < 0) return bad(r, "bad counts");
    if (!m.face_off || !m.face_nbr || !m.face_nx || !m.face_ny || !m.face_len || !m.area ||
        (b.n_ghosts > 0 && (!b.ghost_slot || !b.ghost_kind || !b.ghost_bc)) ||
        (b.n_inlet > 0 && (!b.inlet_off || !b.inlet_ghost)))
      return bad(r, "NULL array");
    if (m.face_off[0] != 0) return bad(r, "face_off[0] != 0");
    for (int64_t i = 0; i < m.n_cells; ++i)
      if (m.face_off[i + 1] < m.face_off[i]) return bad(r, "face_off not monotone");
    const int64_t Mi = m.face_off[m.n_cells];
    for (int64_t s = 0; s < Mi; ++s) {
      const int32_t j = m.face_nbr[s];
      if (j >= 0 ? j >= m.n_cells : ~int64_t(j) >= b.n_ghosts) return bad(r, "face_nbr out of range");
    }
    for (int64_t g = 0; g < b.n_ghosts; ++g) {
      if (b.ghost_slot[g] < 0 || b.ghost_slot[g] >= Mi) return bad(r, "ghost_slot out of range");
      const int32_t k = b.ghost_kind[g], bi = b.ghost_bc[g];
      if ((k == HG_BC_INLETQ && (bi < 0 || bi >= b.n_inlet)) || (k == HG_BC_EXITH && (bi < 0 || bi >= b.n_exit)))
        return bad(r, "ghost_bc out of range");
    }
    if (b.n_inlet > 0) {
      if (b.inlet_off[0] != 0) return bad(r, "inlet_off[0] != 0");
      for (int32_t k = 0; k < b.n_inlet; ++k)
        if (b.inlet_off[k + 1] < b.inlet_off[k]) return bad(r, "inlet_off not monotone");
      for (int32_t f = 0; f < b.inlet_off[b.n_inlet]; ++f)
        if (b.inlet_ghost[f] < 0 || b.inlet_ghost[f] >= b.n_ghosts) return bad(r, "inlet_ghost out of range");
      nIF += b.inlet_off[b.n_inlet];
    }
    cen = cen && m.centroid_xy;
    N += m.n_cells;
    M += Mi;
    G += b.n_ghosts;
    nI += b.n_inlet;
    nE += b.n_exit;
  }
  if (N > (int64_t(1) << 30) || M >= (int64_t(1) << 31) || nI >= (int64_t(1) << 31) || nE >= (int64_t(1) << 31)) {
    g_create_err = "hg_batch_create: batch too large for 32-bit indices";
    return HG_ERR_ARG;
  }
  // ---- concatenation: instance r's cells, slots, ghosts, inlets and exits follow those of r-1
  std::vector<int32_t> off(N + 1, 0), nbr(M), gslot(G), gkind(G), gbc(G), ioff(nI + 1, 0), ighost(nIF);
  std::vector<double> nx(M), ny(M), len(M), area(N), xy(cen ? 2 * N : 0);
  std::vector<int64_t> seg(1, 0);
  int64_t cN = 0, cM = 0, cG = 0, cI = 0, cE = 0, cF = 0;
  for (int32_t r = 0; r < n_inst; ++r) {
    const hg_mesh& m = meshes[r];
    const hg_bcs& b = bcs[r];
    const int64_t Mi = m.face_off[m.n_cells];
    for (int64_t i = 0; i < m.n_cells; ++i) {
      off[cN + i + 1] = int32_t(cM + m.face_off[i + 1]);
      area[cN + i] = m.area[i];
      if (cen) {
        xy[2 * (cN + i)] = m.centroid_xy[2 * i];
        xy[2 * (cN + i) + 1] = m.centroid_xy[2 * i + 1];
      }
    }
    for (int64_t s = 0; s < Mi; ++s) {
      const int32_t j = m.face_nbr[s];
      nbr[cM + s] = j >= 0 ? int32_t(j + cN) : ~int32_t(~j + cG);
      nx[cM + s] = m.face_nx[s];
      ny[cM + s] = m.face_ny[s];
      len[cM + s] = m.face_len[s];
    }
    for (int64_t g = 0; g < b.n_ghosts; ++g) {
      gslot[cG + g] = int32_t(b.ghost_slot[g] + cM);
      gkind[cG + g] = b.ghost_kind[g];
      const int32_t bi = b.ghost_bc[g];
      gbc[cG + g] = b.ghost_kind[g] == HG_BC_INLETQ ? int32_t(bi + cI)
                    : b.ghost_kind[g] == HG_BC_EXITH ? int32_t(bi + cE) : bi;
    }
    for (int32_t k = 0; k < b.n_inlet; ++k) ioff[cI + k + 1] = int32_t(cF + b.inlet_off[k + 1]);
    const int64_t nf = b.n_inlet > 0 ? b.inlet_off[b.n_inlet] : 0;
    for (int64_t f = 0; f < nf; ++f) ighost[cF + f] = int32_t(b.inlet_ghost[f] + cG);
    cN += m.n_cells;
    cM += Mi;
    cG += b.n_ghosts;
    cI += b.n_inlet;
    cE += b.n_exit;
    cF += nf;
    seg.push_back(cN);
  }
  hg_mesh mm{N, off.data(), nbr.data(), nx.data(), ny.data(), len.data(), area.data(), cen ? xy.data() : nullptr};
  hg_bcs bb{G, gslot.data(), gkind.data(), gbc.data(), int32_t(nI), int32_t(nE), ioff.data(), ighost.data()};
  return create_impl(out, &mm, &bb, c, flags, seg);
}

int hg_destroy(hg_ctx* ctx) {
  if (!ctx) return HG_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  for (auto& kv : ctx->seg_graphs) (void)hipGraphExecDestroy(kv.second);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : {ctx->ev_seg[0], ctx->ev_seg[1], ctx->ev_rev[0], ctx->ev_rev[1], ctx->ev_fwd, ctx->ev_xfer[0],
                       ctx->ev_xfer[1]})
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream2) {
    (void)hipStreamSynchronize(ctx->stream2);
    (void)hipStreamDestroy(ctx->stream2);
  }
  for (void* p : ctx->allocs) (void)hipFree(p);
  if (ctx->flags_host) (void)hipHostFree(const_cast<int32_t*>(ctx->flags_host));
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return HG_OK;
}

int hg_set_fields(hg_ctx* ctx, const double* hstill, const double* hstill_ghost, const double* zb,
                  const double* zb_ghost, const double* S0x, const double* S0y, const double* n,
                  const double* inletQ, const double* exitWSE) {
  if (!ctx) return HG_ERR_ARG;
  ctx->adj_pending = false;  // a recorded forward sweep no longer matches
  if (!hstill || !zb || !S0x || !S0y || !n || (ctx->G && (!hstill_ghost || !zb_ghost)) ||
      (ctx->n_inlet && !inletQ) || (ctx->n_exit && !exitWSE)) {
    ctx->err = "hg_set_fields: NULL array";
    return HG_ERR_ARG;
  }
  if (ctx->K > 1) {  // fields of one mesh: its per-cell constants are stored once
    const int64_t Nm = ctx->Nm, Gm = ctx->Gm;
    for (int64_t f = 1; f < ctx->K; ++f) {
      const bool same = !std::memcmp(hstill + f * Nm, hstill, Nm * sizeof(double)) &&
                        !std::memcmp(zb + f * Nm, zb, Nm * sizeof(double)) &&
                        !std::memcmp(S0x + f * Nm, S0x, Nm * sizeof(double)) &&
                        !std::memcmp(S0y + f * Nm, S0y, Nm * sizeof(double)) &&
                        (!Gm || (!std::memcmp(hstill_ghost + f * Gm, hstill_ghost, Gm * sizeof(double)) &&
                                 !std::memcmp(zb_ghost + f * Gm, zb_ghost, Gm * sizeof(double))));
      if (!same) {
        ctx->err = "hg_set_fields: the fields of one mesh must share h_still, zb and S0 (field " + std::to_string(f) +
                   " differs from field 0)";
        return HG_ERR_ARG;
      }
    }
  }
  int rc = 0;
  if (!rc) rc = mesh_cell_in(ctx, ctx->hstill, hstill);
  if (!rc) rc = mesh_cell_in(ctx, ctx->zb0, zb);
  if (!rc) rc = mesh_cell_in(ctx, ctx->S0x0, S0x);
  if (!rc) rc = mesh_cell_in(ctx, ctx->S0y0, S0y);
  if (!rc) rc = cell_in(ctx, ctx->n0, n);
  if (!rc) rc = upload(ctx, ctx->g_hstill, hstill_ghost, ctx->Gm);
  if (!rc) rc = upload(ctx, ctx->g_zb0, zb_ghost, ctx->Gm);
  if (!rc) rc = upload(ctx, ctx->inletQ0, inletQ, ctx->n_inlet);
  if (!rc) rc = upload(ctx, ctx->exitWSE, exitWSE, ctx->n_exit);
  if (rc) return rc;
  ctx->fields_set = true;
  return restore_base(ctx);
}

int hg_set_active_param(hg_ctx* ctx, int32_t kind, const int32_t* zone_of_cell, int32_t n_zones) {
  if (!ctx) return HG_ERR_ARG;
  ctx->adj_pending = false;  // a recorded forward sweep no longer matches
  if ((kind == HG_PARAM_ZB || kind == HG_PARAM_N_ZONE || kind == HG_PARAM_UDE) &&
      refuse_fields(ctx, "hg_set_active_param (zb, zone n or network weights)"))
    return HG_ERR_STATE;
  if (kind < HG_PARAM_NONE || kind > HG_PARAM_UDE) {
    ctx->err = "unknown active parameter kind";
    return HG_ERR_ARG;
  }
  if (ctx->ude_prec && (kind == HG_PARAM_N_ZONE || kind == HG_PARAM_N_CELL)) {
    ctx->err = "Manning n comes from the UDE closure; it cannot also be the active parameter";
    return HG_ERR_STATE;
  }
  if (ctx->nlaw.kind != hg::NLAW_OFF && (kind == HG_PARAM_N_ZONE || kind == HG_PARAM_N_CELL)) {
    ctx->err = "Manning n comes from the variable-n law; it cannot also be the active parameter";
    return HG_ERR_STATE;
  }
  if (ctx->lux.choice == hg::LUX_MANNING_H || ctx->lux.choice == hg::LUX_MANNING_H_UMAG_KS) {
    if (kind == HG_PARAM_N_ZONE || kind == HG_PARAM_N_CELL) {
      ctx->err = "Manning n comes from the UDE network; it cannot also be the active parameter";
      return HG_ERR_STATE;
    }
  }
  if (kind == HG_PARAM_UDE && ctx->lux.choice == hg::LUX_OFF) {
    ctx->err = "HG_PARAM_UDE needs the UDE network (hg_set_lux_ude)";
    return HG_ERR_STATE;
  }
  if (kind == HG_PARAM_N_ZONE) {
    if (!zone_of_cell || n_zones <= 0 || size_t(n_zones) > ctx->scratch_n) {
      ctx->err = "HG_PARAM_N_ZONE needs zone_of_cell and n_zones > 0";
      return HG_ERR_ARG;
    }
    std::vector<int32_t> z(ctx->N);
    for (int64_t i = 0; i < ctx->N; ++i) {
      const int32_t v = zone_of_cell[ctx->perm[i]];
      if (v < 0 || v >= n_zones) {
        ctx->err = "zone_of_cell out of range";
        return HG_ERR_ARG;
      }
      z[i] = v;
    }
    int rc = upload(ctx, ctx->zone, z.data(), ctx->N);
    if (rc) return rc;
    ctx->h_zone = z;
    ctx->n_zones = n_zones;
  }
  ctx->active = kind;
  return HG_OK;
}

int hg_set_params(hg_ctx* ctx, const double* params, int64_t n_params) {
  if (!ctx) return HG_ERR_ARG;
  ctx->adj_pending = false;  // a recorded forward sweep no longer matches
  if (!ctx->fields_set) {
    ctx->err = "hg_set_fields has not been called";
    return HG_ERR_STATE;
  }
  return apply_params(ctx, params, n_params);
}

int hg_rhs(hg_ctx* ctx, const double* Q, const double* params, int64_t n_params, double* dQdt) {
  if (!ctx || !Q || !dQdt) return ctx ? (ctx->err = "hg_rhs: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  int rc = hg_set_params(ctx, params, n_params);
  if (!rc) rc = state_in(ctx, Q, 0);
  if (rc) return rc;
  ctx->cur = 0;
  if (!rc) rc = clear_flags(ctx);
  double* o[3] = {ctx->scratch, ctx->scratch + ctx->N, ctx->scratch + 2 * ctx->N};
  if (!rc) rc = enqueue_step(ctx, 0, 1, 0.0, false, o[0], o[1], o[2]);
  if (!rc) rc = state_out(ctx, o, dQdt);
  if (!rc) rc = check_flags(ctx);
  return rc;
}

int hg_state_upload(hg_ctx* ctx, const double* Q) {
  if (!ctx || !Q) return HG_ERR_ARG;
  int rc = state_in(ctx, Q, 0);
  if (!rc) ctx->cur = 0;
  return rc;
}

int hg_state_download(hg_ctx* ctx, double* Q) {
  if (!ctx || !Q) return HG_ERR_ARG;
  return state_out(ctx, ctx->st[ctx->cur], Q);
}

int hg_advance(hg_ctx* ctx, double dt, int64_t nsteps) {
  if (!ctx) return HG_ERR_ARG;
  if (!ctx->fields_set) {
    ctx->err = "hg_set_fields has not been called";
    return HG_ERR_STATE;
  }
  int rc = clear_flags(ctx);
  if (!rc) rc = advance_async(ctx, dt, nsteps);
  if (!rc) rc = check_flags(ctx);
  return rc;
}

int hg_advance_region(hg_ctx* ctx, double dt, int64_t nsteps, double* gpu_ms_per_step) {
  if (!ctx || nsteps <= 0) return HG_ERR_ARG;
  if (!ctx->fields_set) {
    ctx->err = "hg_set_fields has not been called";
    return HG_ERR_STATE;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = clear_flags(ctx);
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = HG_ERR_HIP;
  if (!rc && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = HG_ERR_HIP;
  if (!rc) rc = advance_async(ctx, dt, nsteps);
  if (!rc && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = HG_ERR_HIP;
  if (!rc) rc = check_flags(ctx);  // synchronizes the stream
  float ms = 0.f;
  if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = HG_ERR_HIP;
  if (!rc && gpu_ms_per_step) *gpu_ms_per_step = double(ms) / double(nsteps);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

int hg_prepare_advance(hg_ctx* ctx, double dt) {
  if (!ctx || !(dt > 0.0)) return HG_ERR_ARG;
  if (require_fields(ctx, "hg_prepare_advance")) return HG_ERR_STATE;
  int rc = ensure_graph(ctx, dt);
  if (!rc) rc = advance_timed(ctx, dt, 0, nullptr);  // the dispatch events
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = HG_ERR_HIP;
  return rc;
}

int hg_advance_timed(hg_ctx* ctx, double dt, int64_t nsteps, double* kernel_ms) {
  if (!ctx || nsteps < 0) return HG_ERR_ARG;
  if (!ctx->fields_set) {
    ctx->err = "hg_set_fields has not been called";
    return HG_ERR_STATE;
  }
  int rc = clear_flags(ctx);
  if (!rc) rc = advance_timed(ctx, dt, nsteps, kernel_ms);
  if (!rc) rc = check_flags(ctx);
  return rc;
}

int hg_step(hg_ctx* ctx, double* Q, const double* params, int64_t n_params, double dt, int64_t nsteps) {
  if (!ctx || !Q) return HG_ERR_ARG;
  int rc = hg_set_params(ctx, params, n_params);
  if (!rc) rc = hg_state_upload(ctx, Q);
  if (!rc) rc = hg_advance(ctx, dt, nsteps);
  if (!rc) rc = hg_state_download(ctx, Q);
  return rc;
}

int hg_cfl_dt(hg_ctx* ctx, const double* Q, double cfl, double* dt) {
  if (!ctx || !Q || !dt) return HG_ERR_ARG;
  if (!ctx->fields_set) {
    ctx->err = "hg_set_fields has not been called";
    return HG_ERR_STATE;
  }
  int rc = state_in(ctx, Q, 0);
  if (rc) return rc;
  ctx->cur = 0;
  const int nb = grid_of(ctx->N);
  hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(kBlock), 0, ctx->stream, int(ctx->N), int(ctx->Nm), ctx->st[0][0], ctx->st[0][1],
                     ctx->st[0][2], ctx->hstill, ctx->deg, ctx->flen, ctx->area, ctx->c.g, ctx->c.h_small, ctx->red);
  HG_TRY(hipGetLastError());
  std::vector<double> part(nb);
  HG_TRY(hipMemcpyAsync(part.data(), ctx->red, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HG_TRY(hipStreamSynchronize(ctx->stream));
  double mn = INFINITY;
  for (double v : part) mn = std::min(mn, v);
  *dt = cfl * mn;
  return HG_OK;
}

int hg_cfl_dt_noalloc(hg_ctx* ctx, const double* Q, double cfl, double* dt) {
  if (!ctx || !Q || !dt) return HG_ERR_ARG;
  if (require_fields(ctx, "hg_cfl_dt_noalloc")) return HG_ERR_STATE;
  const int64_t N = ctx->N;
  const int nb = grid_of(N);
  if (nb > kNoAllocParts) {
    ctx->err = "hg_cfl_dt_noalloc: " + std::to_string(N) + " cells need " + std::to_string(nb) +
               " partial minima; the static scratch holds " + std::to_string(kNoAllocParts) +
               " (meshes of up to " + std::to_string(int64_t(kNoAllocParts) * kBlock) + " cells)";
    return HG_ERR_ARG;
  }
  double* const sc = ctx->noalloc_scratch;
  if (!sc || !ctx->stage || !ctx->d_perm) {
    ctx->err = "hg_cfl_dt_noalloc: context set-up incomplete";
    return HG_ERR_STATE;
  }
  // Q -> the context's transfer buffer -> state slot 0 in library order
  // (as hg_cfl_dt); partial minima in scratch[0, nb), the result in
  // scratch[kNoAllocParts].  Every slot read is written first by this call.
  HG_TRY(hipMemcpyAsync(ctx->stage, Q, size_t(3) * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_perm_in, dim3(grid_of(N)), dim3(kBlock), 0, ctx->stream, N, 3, ctx->d_perm, ctx->stage,
                     ctx->st[0][0], ctx->st[0][1], ctx->st[0][2]);
  ctx->cur = 0;
  hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(kBlock), 0, ctx->stream, int(N), int(ctx->Nm), ctx->st[0][0], ctx->st[0][1],
                     ctx->st[0][2], ctx->hstill, ctx->deg, ctx->flen, ctx->area, ctx->c.g, ctx->c.h_small, sc);
  hipLaunchKernelGGL(k_min_parts, dim3(1), dim3(kBlock), 0, ctx->stream, nb, sc, cfl, sc + kNoAllocParts);
  HG_TRY(hipGetLastError());
  HG_TRY(hipMemcpyAsync(dt, sc + kNoAllocParts, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HG_TRY(hipStreamSynchronize(ctx->stream));
  return HG_OK;
}

int hg_set_ude(hg_ctx* ctx, int32_t n_sets, const double* weights, const int32_t* set_of_cell,
               int32_t precision) {
  if (!ctx) return HG_ERR_ARG;
  ctx->adj_pending = false;  // a recorded forward sweep no longer matches
  HG_TRY(hipStreamSynchronize(ctx->stream));
  if (drop_graph(ctx)) return HG_ERR_HIP;  // captured steps may include (or lack) the closure
  for (void* p : ctx->ude_allocs) {
    ctx->allocs.erase(std::find(ctx->allocs.begin(), ctx->allocs.end(), p));
    HG_TRY(hipFree(p));
  }
  ctx->ude_allocs.clear();
  ctx->ude_prec = 0;
  ctx->ude_nsets = 0;
  ctx->ude_w = nullptr;
  ctx->ude_set = nullptr;
  ctx->ude_nbar = ctx->ude_wpart = ctx->ude_wbar = nullptr;
  ctx->ude_blk_off = ctx->ude_blk_set = ctx->ude_blk_slot = ctx->ude_set_start = nullptr;
  ctx->ude_cpb = 1;
  ctx->ude_nslots = 0;
  if (n_sets == 0) {  // closure off: back to the per-cell n of hg_set_fields
    HG_TRY(hipMemcpyAsync(ctx->n, ctx->n0, ctx->N * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HG_TRY(hipStreamSynchronize(ctx->stream));
    return HG_OK;
  }
  if (n_sets < 0 || n_sets > 65535 || !weights || (precision != 32 && precision != 64)) {
    ctx->err = "hg_set_ude: need 0 <= n_sets <= 65535, weights and precision 32 or 64";
    return HG_ERR_ARG;
  }
  if (ctx->active == HG_PARAM_N_ZONE || ctx->active == HG_PARAM_N_CELL) {
    ctx->err = "hg_set_ude: Manning n is the active parameter";
    return HG_ERR_STATE;
  }
  if (ctx->nlaw.kind != hg::NLAW_OFF) {
    ctx->err = "hg_set_ude: a variable-n law is set";
    return HG_ERR_STATE;
  }
  if (ctx->lux.choice) {
    ctx->err = "hg_set_ude: the UDE network is on";
    return HG_ERR_STATE;
  }
  const int64_t N = ctx->N;
  std::vector<int32_t> set;
  if (set_of_cell) {
    set.resize(N);
    for (int64_t i = 0; i < N; ++i) {
      const int32_t v = set_of_cell[ctx->perm[i]];
      if (v < 0 || v >= n_sets) {
        ctx->err = "hg_set_ude: set_of_cell out of range";
        return HG_ERR_ARG;
      }
      set[i] = v;
    }
  }
  for (int64_t k = 0; k < int64_t(n_sets) * hg::kUdeW; ++k)
    if (!std::isfinite(weights[k])) {
      ctx->err = "hg_set_ude: non-finite weight";
      return HG_ERR_ARG;
    }
  const size_t nw = size_t(n_sets) * hg::kUdeW;
  const size_t nalloc0 = ctx->allocs.size();
  int rc = HG_OK;
  if (precision == 32) {
    std::vector<float> w(nw);
    for (size_t k = 0; k < nw; ++k) w[k] = float(weights[k]);
    float* d = nullptr;
    rc = dalloc(ctx, &d, nw);
    if (!rc) rc = upload(ctx, d, w.data(), nw);
    ctx->ude_w = d;
  } else {
    double* d = nullptr;
    rc = dalloc(ctx, &d, nw);
    if (!rc) rc = upload(ctx, d, weights, nw);
    ctx->ude_w = d;
  }
  if (!rc && set_of_cell) {
    rc = dalloc(ctx, &ctx->ude_set, N);
    if (!rc) rc = upload(ctx, ctx->ude_set, set.data(), N);
  }
  if (!rc) rc = dalloc(ctx, &ctx->ude_nbar, N);
  // Reverse-mode weight slots: one per (reverse-mode block, set used in it),
  // ordered by set and then block, so that each set's cotangent sums a
  // contiguous run.  A block takes cpb consecutive 64-cell chunks.
  ctx->ude_cpb = ude_vjp_cpb(ude_blocks(ctx));
  const int nb = ude_vjp_blocks(ctx);
  const int64_t cells_pb = int64_t(ctx->ude_cpb) * 64;
  std::vector<int32_t> blk_off(size_t(nb) + 1, 0), blk_set, set_start(size_t(n_sets) + 1, 0);
  for (int b = 0; b < nb; ++b) {
    const int64_t i1 = std::min<int64_t>(N, int64_t(b + 1) * cells_pb);
    for (int64_t i = int64_t(b) * cells_pb; i < i1; ++i) {
      const int32_t v = set_of_cell ? set[i] : 0;
      if (std::find(blk_set.begin() + blk_off[b], blk_set.end(), v) == blk_set.end()) blk_set.push_back(v);
    }
    blk_off[b + 1] = int32_t(blk_set.size());
  }
  for (int32_t v : blk_set) ++set_start[size_t(v) + 1];
  for (int32_t v = 0; v < n_sets; ++v) set_start[size_t(v) + 1] += set_start[v];
  std::vector<int32_t> blk_slot(blk_set.size()), next(set_start.begin(), set_start.end() - 1);
  for (size_t e = 0; e < blk_set.size(); ++e) blk_slot[e] = next[blk_set[e]]++;  // blocks in order within a set
  const int32_t nslots = int32_t(blk_set.size());
  if (!rc) rc = dalloc(ctx, &ctx->ude_wpart, size_t(nslots) * hg::kUdeW);
  if (!rc) rc = dalloc(ctx, &ctx->ude_wbar, nw);
  if (!rc) rc = dalloc(ctx, &ctx->ude_blk_off, blk_off.size());
  if (!rc) rc = upload(ctx, ctx->ude_blk_off, blk_off.data(), blk_off.size());
  if (!rc) rc = dalloc(ctx, &ctx->ude_blk_set, blk_set.size());
  if (!rc) rc = upload(ctx, ctx->ude_blk_set, blk_set.data(), blk_set.size());
  if (!rc) rc = dalloc(ctx, &ctx->ude_blk_slot, blk_slot.size());
  if (!rc) rc = upload(ctx, ctx->ude_blk_slot, blk_slot.data(), blk_slot.size());
  if (!rc) rc = dalloc(ctx, &ctx->ude_set_start, set_start.size());
  if (!rc) rc = upload(ctx, ctx->ude_set_start, set_start.data(), set_start.size());
  ctx->ude_allocs.assign(ctx->allocs.begin() + nalloc0, ctx->allocs.end());
  if (rc) return rc;
  HG_TRY(hipMemsetAsync(ctx->ude_wbar, 0, nw * sizeof(double), ctx->stream));
  HG_TRY(hipStreamSynchronize(ctx->stream));
  ctx->ude_nslots = nslots;
  ctx->ude_nsets = n_sets;
  ctx->ude_prec = precision;
  return HG_OK;
}

int hg_ude_n(hg_ctx* ctx, const double* Q, double* n) {
  if (!ctx || !Q || !n) return ctx ? (ctx->err = "hg_ude_n: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  if (require_fields(ctx, "hg_ude_n")) return HG_ERR_STATE;
  if (!ctx->ude_prec) {
    ctx->err = "hg_ude_n: the UDE closure is off";
    return HG_ERR_STATE;
  }
  int rc = state_in(ctx, Q, 0);
  if (rc) return rc;
  ctx->cur = 0;
  const double* q[3] = {ctx->st[0][0], ctx->st[0][1], ctx->st[0][2]};
  rc = enqueue_ude_n(ctx, q);
  if (!rc) rc = cell_out(ctx, ctx->n, n);
  return rc;
}

int hg_ude_grad(hg_ctx* ctx, double* w_bar) {
  if (!ctx || !w_bar) return HG_ERR_ARG;
  if (!ctx->ude_prec) {
    ctx->err = "hg_ude_grad: the UDE closure is off";
    return HG_ERR_STATE;
  }
  HG_TRY(hipMemcpyAsync(w_bar, ctx->ude_wbar, size_t(ctx->ude_nsets) * hg::kUdeW * sizeof(double),
                        hipMemcpyDeviceToHost, ctx->stream));
  HG_TRY(hipStreamSynchronize(ctx->stream));
  return HG_OK;
}

int hg_set_n_law(hg_ctx* ctx, int32_t kind, const double* params, const double* ks) {
  if (!ctx) return HG_ERR_ARG;
  ctx->adj_pending = false;  // a recorded forward sweep no longer matches
  if (kind != HG_NLAW_OFF && refuse_fields(ctx, "hg_set_n_law")) return HG_ERR_STATE;
  if (kind < HG_NLAW_OFF || kind > HG_NLAW_H_UMAG_KS) {
    ctx->err = "hg_set_n_law: unknown law";
    return HG_ERR_ARG;
  }
  HG_TRY(hipStreamSynchronize(ctx->stream));
  if (drop_graph(ctx)) return HG_ERR_HIP;  // captured steps may include (or lack) the law
  for (void* p : ctx->nlaw_allocs) {
    ctx->allocs.erase(std::find(ctx->allocs.begin(), ctx->allocs.end(), p));
    HG_TRY(hipFree(p));
  }
  ctx->nlaw_allocs.clear();
  ctx->nlaw = hg::NLaw{};
  ctx->nlaw_ks = nullptr;
  ctx->nlaw_nbar = nullptr;
  if (kind == HG_NLAW_OFF) {  // back to the per-cell n of hg_set_fields
    HG_TRY(hipMemcpyAsync(ctx->n, ctx->n0, ctx->N * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HG_TRY(hipStreamSynchronize(ctx->stream));
    return HG_OK;
  }
  if (ctx->ude_prec) {
    ctx->err = "hg_set_n_law: the UDE closure is on";
    return HG_ERR_STATE;
  }
  if (ctx->lux.choice) {
    ctx->err = "hg_set_n_law: the UDE network is on";
    return HG_ERR_STATE;
  }
  if (ctx->active == HG_PARAM_N_ZONE || ctx->active == HG_PARAM_N_CELL) {
    ctx->err = "hg_set_n_law: Manning n is the active parameter";
    return HG_ERR_STATE;
  }
  hg::NLaw L{};
  L.kind = kind;
  if (kind == HG_NLAW_H_UMAG_KS) {
    if (!ks) {
      ctx->err = "hg_set_n_law: h_Umag_ks needs ks per cell";
      return HG_ERR_ARG;
    }
  } else {
    if (!params) {
      ctx->err = "hg_set_n_law: params {n_lower, n_upper, k, h_mid} missing";
      return HG_ERR_ARG;
    }
    L.lo = params[0];
    L.hi = params[1];
    L.k = params[2];
    L.hmid = params[3];
    if (!(L.k > 0.0)) {  // the laws' @assert (process_ManningN_2D.jl:140, 155, 171)
      ctx->err = "hg_set_n_law: k must be a positive scaling factor";
      return HG_ERR_ARG;
    }
    if (kind == HG_NLAW_SIGMOID && !(L.hmid > 0.0)) {  // :172
      ctx->err = "hg_set_n_law: h_mid must be a positive water depth";
      return HG_ERR_ARG;
    }
  }
  const size_t nalloc0 = ctx->allocs.size();
  int rc = dalloc(ctx, &ctx->nlaw_nbar, ctx->N);
  if (!rc && ks) {
    rc = dalloc(ctx, &ctx->nlaw_ks, ctx->N);
    if (!rc) rc = cell_in(ctx, ctx->nlaw_ks, ks);
  }
  ctx->nlaw_allocs.assign(ctx->allocs.begin() + nalloc0, ctx->allocs.end());
  if (rc) return rc;
  ctx->nlaw = L;
  return HG_OK;
}

int hg_n_law_eval(hg_ctx* ctx, const double* h, const double* Umag, double* n, double* h_ks, double* f,
                  double* Re) {
  if (!ctx || !h || !Umag) return ctx ? (ctx->err = "hg_n_law_eval: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  if (ctx->nlaw.kind == hg::NLAW_OFF) {
    ctx->err = "hg_n_law_eval: no variable-n law is set";
    return HG_ERR_STATE;
  }
  const int64_t N = ctx->N;
  double* buf = nullptr;  // n, h/ks, f, Re
  int rc = dalloc(ctx, &buf, size_t(4) * N);
  if (rc) return rc;
  ctx->nlaw_allocs.push_back(buf);  // freed with the law
  if (!rc) rc = cell_in(ctx, ctx->scratch, h);
  if (!rc) rc = cell_in(ctx, ctx->scratch + N, Umag);
  if (rc) return rc;
  hg::NLawArgs a = nlaw_args(ctx, ctx->st[0]);
  a.h_in = ctx->scratch;
  a.U_in = ctx->scratch + N;
  a.n = buf;
  a.hks = buf + N;
  a.f = buf + 2 * N;
  a.Re = buf + 3 * N;
  hipLaunchKernelGGL(k_nlaw_eval, dim3(grid_of(N)), dim3(kBlock), 0, ctx->stream, a);
  HG_TRY(hipGetLastError());
  double* outs[4] = {n, h_ks, f, Re};
  for (int k = 0; k < 4 && !rc; ++k)
    if (outs[k]) rc = cell_out(ctx, buf + k * N, outs[k]);
  ctx->nlaw_allocs.pop_back();
  ctx->allocs.erase(std::find(ctx->allocs.begin(), ctx->allocs.end(), static_cast<void*>(buf)));
  if (hipFree(buf) != hipSuccess && !rc) rc = HG_ERR_HIP;
  return rc;
}

int hg_cell_n(hg_ctx* ctx, const double* Q, double* n) {
  if (!ctx || !Q || !n) return ctx ? (ctx->err = "hg_cell_n: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  if (require_fields(ctx, "hg_cell_n")) return HG_ERR_STATE;
  int rc = state_in(ctx, Q, 0);
  if (rc) return rc;
  ctx->cur = 0;
  const double* q[3] = {ctx->st[0][0], ctx->st[0][1], ctx->st[0][2]};
  rc = enqueue_cell_n(ctx, q);
  if (!rc) rc = cell_out(ctx, ctx->n, n);
  return rc;
}

// Validates a network description and lays out its weights; NULL when valid.
const char* lux_build(const hg_lux_net* in, hg::LuxNet& L) {
  L = hg::LuxNet{};
  if (!in) return "NULL network";
  if (in->choice < HG_LUX_MANNING_H || in->choice > HG_LUX_FLOW_RESISTANCE) return "unknown UDE choice";
  if (in->n_hidden < 1 || in->n_hidden > hg::kLuxMaxLayers) return "need 1..4 hidden layers";
  L.choice = in->choice;
  L.n_in = in->choice == HG_LUX_MANNING_H ? 1 : in->choice == HG_LUX_MANNING_H_UMAG_KS ? 3 : 2;
  L.n_hidden = in->n_hidden;
  int32_t cur = 0, win = L.n_in;
  for (int l = 0; l < L.n_hidden; ++l) {
    const int32_t w = in->width[l];
    if (w < 1 || w > hg::kLuxMaxW) return "hidden widths must be 1..16";
    if (in->act[l] < HG_ACT_RELU || in->act[l] > HG_ACT_SOFTPLUS) return "unknown activation";
    L.width[l] = w;
    L.act[l] = in->act[l];
    L.off_W[l] = cur;
    cur += w * win;
    L.off_b[l] = cur;
    cur += w;
    L.off_g[l] = cur;
    cur += w;
    L.off_be[l] = cur;
    cur += w;
    win = w;
  }
  L.off_Wo = cur;
  cur += win;
  L.off_bo = cur;
  cur += 1;
  L.P = cur;
  if (L.P > hg::kLuxMaxP) return "too many weights (at most 2048)";
  L.lo = in->output_bounds[0];
  L.hi = in->output_bounds[1];
  const double* b[3] = {in->h_bounds, in->u_bounds, in->ks_bounds};
  for (int k = 0; k < L.n_in; ++k) {
    L.in_lo[k] = b[k][0];
    L.in_hi[k] = b[k][1];
    if (!(L.in_hi[k] != L.in_lo[k]) || !std::isfinite(L.in_lo[k]) || !std::isfinite(L.in_hi[k]))
      return "input bounds must be finite and distinct";
  }
  if (L.choice == HG_LUX_FLOW_RESISTANCE) {  // |q| bounds from h and velocity bounds (process_FlowResistance_UDE.jl:14)
    L.in_lo[1] = in->h_bounds[0] * in->u_bounds[0];
    L.in_hi[1] = in->h_bounds[1] * in->u_bounds[1];
    if (!(L.in_hi[1] != L.in_lo[1])) return "h and velocity bounds give an empty |q| range";
  }
  L.eps = in->ln_epsilon > 0.0 ? in->ln_epsilon : hg::kLuxLayerNormEps;
  return nullptr;
}

int64_t hg_lux_n_params(const hg_lux_net* net) {
  hg::LuxNet L;
  return lux_build(net, L) ? -1 : int64_t(L.P);
}

int hg_set_lux_ude(hg_ctx* ctx, const hg_lux_net* net, const double* params, const double* ks) {
  if (!ctx) return HG_ERR_ARG;
  ctx->adj_pending = false;  // a recorded forward sweep no longer matches
  if (net && refuse_fields(ctx, "hg_set_lux_ude")) return HG_ERR_STATE;
  HG_TRY(hipStreamSynchronize(ctx->stream));
  if (drop_graph(ctx)) return HG_ERR_HIP;  // captured steps may include (or lack) the network
  for (void* p : ctx->lux_allocs) {
    ctx->allocs.erase(std::find(ctx->allocs.begin(), ctx->allocs.end(), p));
    HG_TRY(hipFree(p));
  }
  ctx->lux_allocs.clear();
  ctx->lux = hg::LuxNet{};
  ctx->lux_w = ctx->lux_ks = ctx->lux_fmag = ctx->lux_obar = ctx->lux_wpart = ctx->lux_wbar = nullptr;
  ctx->lux_mw = 0;
  if (ctx->active == HG_PARAM_UDE) ctx->active = HG_PARAM_NONE;
  if (!net) {  // off: back to the per-cell n of hg_set_fields
    HG_TRY(hipMemcpyAsync(ctx->n, ctx->n0, ctx->N * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HG_TRY(hipStreamSynchronize(ctx->stream));
    return HG_OK;
  }
  hg::LuxNet L;
  const char* why = lux_build(net, L);
  if (why) {
    ctx->err = std::string("hg_set_lux_ude: ") + why;
    return HG_ERR_ARG;
  }
  if (!params || (L.choice == HG_LUX_MANNING_H_UMAG_KS && !ks)) {
    ctx->err = "hg_set_lux_ude: weights (and ks for ManningN_h_Umag_ks) required";
    return HG_ERR_ARG;
  }
  for (int32_t k = 0; k < L.P; ++k)
    if (!std::isfinite(params[k])) {
      ctx->err = "hg_set_lux_ude: non-finite weight";
      return HG_ERR_ARG;
    }
  if (ctx->ude_prec || ctx->nlaw.kind != hg::NLAW_OFF) {
    ctx->err = "hg_set_lux_ude: the MLP closure or a variable-n law is on";
    return HG_ERR_STATE;
  }
  if (L.choice != HG_LUX_FLOW_RESISTANCE && (ctx->active == HG_PARAM_N_ZONE || ctx->active == HG_PARAM_N_CELL)) {
    ctx->err = "hg_set_lux_ude: Manning n is the active parameter";
    return HG_ERR_STATE;
  }
  const int64_t N = ctx->N;
  const size_t nalloc0 = ctx->allocs.size();
  int rc = dalloc(ctx, &ctx->lux_w, size_t(L.P));
  if (!rc) rc = upload(ctx, ctx->lux_w, params, size_t(L.P));
  if (!rc && L.choice == HG_LUX_MANNING_H_UMAG_KS) {
    rc = dalloc(ctx, &ctx->lux_ks, N);
    if (!rc) rc = cell_in(ctx, ctx->lux_ks, ks);
  }
  if (!rc && L.choice == HG_LUX_FLOW_RESISTANCE) rc = dalloc(ctx, &ctx->lux_fmag, N);
  if (!rc) rc = dalloc(ctx, &ctx->lux_obar, N);
  if (!rc) rc = dalloc(ctx, &ctx->lux_wpart, size_t(L.P) * lux_vjp_blocks(ctx));
  if (!rc) rc = dalloc(ctx, &ctx->lux_wbar, size_t(L.P));
  ctx->lux_allocs.assign(ctx->allocs.begin() + nalloc0, ctx->allocs.end());
  if (rc) {
    ctx->lux_w = ctx->lux_ks = ctx->lux_fmag = ctx->lux_obar = ctx->lux_wpart = ctx->lux_wbar = nullptr;
    return rc;
  }
  HG_TRY(hipMemsetAsync(ctx->lux_wbar, 0, size_t(L.P) * sizeof(double), ctx->stream));
  HG_TRY(hipStreamSynchronize(ctx->stream));
  int maxw = 0;
  for (int l = 0; l < L.n_hidden; ++l) maxw = std::max(maxw, int(L.width[l]));
  ctx->lux_mw = maxw <= 4 ? 4 : hg::kLuxMaxW;
  ctx->lux = L;
  return HG_OK;
}

int hg_lux_eval(hg_ctx* ctx, const double* Q, double* out) {
  if (!ctx || !Q || !out) return ctx ? (ctx->err = "hg_lux_eval: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  if (require_fields(ctx, "hg_lux_eval")) return HG_ERR_STATE;
  if (!ctx->lux.choice) {
    ctx->err = "hg_lux_eval: the UDE network is off";
    return HG_ERR_STATE;
  }
  int rc = state_in(ctx, Q, 0);
  if (rc) return rc;
  ctx->cur = 0;
  const double* q[3] = {ctx->st[0][0], ctx->st[0][1], ctx->st[0][2]};
  rc = enqueue_lux_fwd(ctx, q);
  if (!rc) rc = cell_out(ctx, ctx->lux.choice == hg::LUX_FLOW_RESISTANCE ? ctx->lux_fmag : ctx->n, out);
  return rc;
}

int hg_lux_grad(hg_ctx* ctx, double* params_bar) {
  if (!ctx || !params_bar) return HG_ERR_ARG;
  if (!ctx->lux.choice) {
    ctx->err = "hg_lux_grad: the UDE network is off";
    return HG_ERR_STATE;
  }
  HG_TRY(hipMemcpyAsync(params_bar, ctx->lux_wbar, size_t(ctx->lux.P) * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  HG_TRY(hipStreamSynchronize(ctx->stream));
  return HG_OK;
}

int hg_mfma_probe(int32_t precision, const double* A, const double* B, double* D) {
  if (!A || !B || !D || (precision != 32 && precision != 64)) return HG_ERR_ARG;
  const size_t bytes = 512 * sizeof(double);
  void *dA = nullptr, *dB = nullptr, *dD = nullptr;
  if (hipMalloc(&dA, bytes) != hipSuccess || hipMalloc(&dB, bytes) != hipSuccess || hipMalloc(&dD, bytes) != hipSuccess) {
    (void)hipFree(dA);
    (void)hipFree(dB);
    return HG_ERR_HIP;
  }
  hipError_t e = hipSuccess;
  if (precision == 32) {
    float a[64], b[64], d[256];
    for (int k = 0; k < 64; ++k) {
      a[k] = float(A[k]);
      b[k] = float(B[k]);
    }
    if (e == hipSuccess) e = hipMemcpy(dA, a, sizeof(a), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, b, sizeof(b), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_mfma_probe<float>, dim3(1), dim3(64), 0, 0, static_cast<const float*>(dA),
                         static_cast<const float*>(dB), static_cast<float*>(dD));
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(d, dD, sizeof(d), hipMemcpyDeviceToHost);
    for (int k = 0; k < 256 && e == hipSuccess; ++k) D[k] = d[k];
  } else {
    if (e == hipSuccess) e = hipMemcpy(dA, A, 64 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, B, 64 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_mfma_probe<double>, dim3(1), dim3(64), 0, 0, static_cast<const double*>(dA),
                         static_cast<const double*>(dB), static_cast<double*>(dD));
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(D, dD, 256 * sizeof(double), hipMemcpyDeviceToHost);
  }
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dD);
  return e == hipSuccess ? HG_OK : HG_ERR_HIP;
}

int hg_jvp(hg_ctx* ctx, const double* Q0, const double* dQ0, const double* params, int64_t n_params,
           const double* dparams, double dt, int64_t nsteps, double* QN, double* dQN) {
  if (!ctx || !Q0 || !QN || !dQN) return ctx ? (ctx->err = "hg_jvp: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  if (refuse_fields(ctx, "hg_jvp")) return HG_ERR_STATE;
  if (nsteps < 0) {
    ctx->err = "hg_jvp: nsteps must be >= 0";
    return HG_ERR_ARG;
  }
  if (ctx->ude_prec) {
    ctx->err = "hg_jvp: forward mode through the UDE closure is not provided";
    return HG_ERR_STATE;
  }
  const int64_t N = ctx->N, G = ctx->G;
  int rc = hg_set_params(ctx, params, n_params);
  if (rc) return rc;
  if (!ctx->jvp) {  // tangents: state (2 x 3N), dn, dzb, dS0x, dS0y (N each), dg_zb (G), dinletQ, dtotalA
    const size_t nI = size_t(std::max<int32_t>(ctx->n_inlet, 1));
    rc = dalloc(ctx, &ctx->jvp, size_t(10) * N + size_t(std::max<int64_t>(G, 1)) + 2 * nI);
    if (rc) return rc;
  }
  double* T0 = ctx->jvp;
  double* T1 = ctx->jvp + 3 * N;
  double* dn = ctx->jvp + 6 * N;
  double* dzb = dn + N;
  double* dS0x = dzb + N;
  double* dS0y = dS0x + N;
  double* dgzb = dS0y + N;
  double* dQin = dgzb + std::max<int64_t>(G, 1);
  double* dtot = dQin + std::max<int32_t>(ctx->n_inlet, 1);
  // tangents of the parameter fields (semi_discretize_swe_2D.jl:114-199)
  const double* pn = nullptr;
  const double* pzb = nullptr;
  const double* pQ = nullptr;
  if (dparams && ctx->active != HG_PARAM_NONE) {
    switch (ctx->active) {
      case HG_PARAM_N_ZONE: {
        std::vector<double> v(N);
        for (int64_t i = 0; i < N; ++i) v[i] = dparams[ctx->h_zone[i]];
        rc = upload(ctx, dn, v.data(), N);
        pn = dn;
        break;
      }
      case HG_PARAM_N_CELL:
        rc = cell_in(ctx, dn, dparams);
        pn = dn;
        break;
      case HG_PARAM_ZB:  // bed data are linear in zb: their tangent is the same map of dzb
        rc = cell_in(ctx, dzb, dparams);
        if (!rc) {
          hipLaunchKernelGGL(k_bed_data, dim3(grid_of(std::max<int64_t>(N, G))), dim3(kBlock), 0, ctx->stream,
                             int(N), int(G), dzb, ctx->deg, ctx->nbr, ctx->fnx, ctx->fny, ctx->flen, ctx->area,
                             ctx->g_cell, dS0x, dS0y, dgzb);
          if (hipGetLastError() != hipSuccess) rc = HG_ERR_HIP;
        }
        pzb = dzb;
        break;
      case HG_PARAM_Q:
        rc = upload(ctx, dQin, dparams, size_t(ctx->n_inlet));
        pQ = dQin;
        break;
      default:
        break;
    }
    if (rc) return rc;
  }
  rc = state_in(ctx, Q0, 0);
  if (!rc) {
    if (dQ0) {
      double* dst[3] = {T0, T0 + N, T0 + 2 * N};
      for (int c = 0; c < 3 && !rc; ++c) rc = cell_in(ctx, dst[c], dQ0 + c * N);
    } else if (hipMemsetAsync(T0, 0, 3 * N * sizeof(double), ctx->stream) != hipSuccess) {
      rc = HG_ERR_HIP;
    }
  }
  // the network's weight tangent, when its weights are the active parameter
  const bool lux = ctx->lux.choice != hg::LUX_OFF;
  const bool lux_fr = ctx->lux.choice == hg::LUX_FLOW_RESISTANCE;
  const double* pw = nullptr;
  if (!rc && lux && dparams && ctx->active == HG_PARAM_UDE) {
    if (!ctx->lux_dw) rc = dalloc(ctx, &ctx->lux_dw, size_t(hg::kLuxMaxP));
    if (!rc) rc = upload(ctx, ctx->lux_dw, dparams, size_t(ctx->lux.P));
    pw = ctx->lux_dw;
  }
  if (!rc) rc = clear_flags(ctx);
  if (rc) return rc;
  int cur = 0;
  double* Tb[2] = {T0, T1};
  const bool law = ctx->nlaw.kind != hg::NLAW_OFF;
  if (law || (lux && !lux_fr)) pn = dn;  // the law's or network's tangent of n, rewritten every step
  double* dfm = lux_fr ? dn : nullptr;     // the network's tangent of the friction magnitude
  for (int64_t t = 0; t < nsteps; ++t) {
    const double* in[3] = {ctx->st[cur][0], ctx->st[cur][1], ctx->st[cur][2]};
    double* out[3] = {ctx->st[1 - cur][0], ctx->st[1 - cur][1], ctx->st[1 - cur][2]};
    if (law) {
      hg::NLawArgs u = nlaw_args(ctx, in);
      u.dxi = Tb[cur];
      u.dqx = Tb[cur] + N;
      u.dqy = Tb[cur] + 2 * N;
      u.dn = dn;
      hipLaunchKernelGGL(k_nlaw_jvp, dim3(grid_of(N)), dim3(kBlock), 0, ctx->stream, u);
    }
    if (lux) {  // n (or the friction magnitude) and its tangent from the network on dual numbers
      const hg::LuxArgs u = lux_args(ctx, in);
      if (ctx->lux_mw == 4)
        hipLaunchKernelGGL(k_lux_jvp<4>, dim3(grid_of(N)), dim3(kBlock), 0, ctx->stream, u, Tb[cur], Tb[cur] + N,
                           Tb[cur] + 2 * N, pw, dn);
      else
        hipLaunchKernelGGL(k_lux_jvp<hg::kLuxMaxW>, dim3(grid_of(N)), dim3(kBlock), 0, ctx->stream, u, Tb[cur],
                           Tb[cur] + N, Tb[cur] + 2 * N, pw, dn);
    }
    if (ctx->n_inlet > 0)
      hipLaunchKernelGGL(k_inlet_conveyance_jvp, dim3((ctx->n_inlet + 63) / 64), dim3(64), 0, ctx->stream,
                         inlet_args_p(ctx, in[0]), Tb[cur], pn, dtot);
    JvpArgs J{};
    J.a = make_args_p(ctx, in, out, dt, nullptr);
    J.dxi = Tb[cur];
    J.dqx = Tb[cur] + N;
    J.dqy = Tb[cur] + 2 * N;
    J.dxo = Tb[1 - cur];
    J.dqxo = Tb[1 - cur] + N;
    J.dqyo = Tb[1 - cur] + 2 * N;
    J.dzb = pzb;
    J.dn = pn;
    J.dS0x = pzb ? dS0x : nullptr;
    J.dS0y = pzb ? dS0y : nullptr;
    J.dg_zb = pzb ? dgzb : nullptr;
    J.dinletQ = pQ;
    J.dtotalA = dtot;
    J.dfmag = dfm;
    hipLaunchKernelGGL(k_swe_step_jvp, dim3(grid_of(N)), dim3(kBlock), 0, ctx->stream, J);
    HG_TRY(hipGetLastError());
    cur = 1 - cur;
  }
  double* src[3] = {ctx->st[cur][0], ctx->st[cur][1], ctx->st[cur][2]};
  rc = state_out(ctx, src, QN);
  double* dsrc[3] = {Tb[cur], Tb[cur] + N, Tb[cur] + 2 * N};
  if (!rc) rc = state_out(ctx, dsrc, dQN);
  ctx->cur = cur;
  if (!rc) rc = check_flags(ctx);
  return rc;
}

int hg_tsit5_tableau(double* c, double* a, double* btilde) {
  for (int i = 0; i < 7; ++i) {
    if (c) c[i] = tsit5::c[i];
    if (btilde) btilde[i] = tsit5::btilde[i];
    for (int j = 0; j < 6; ++j)
      if (a) a[i * 6 + j] = tsit5::a[i][j];
  }
  return HG_OK;
}

int hg_tsit5_solve(hg_ctx* ctx, const double* Q0, const double* params, int64_t n_params, double t0, double t1,
                   int64_t n_save, double abstol, double reltol, double dt0, double* Qsave, int64_t* stats) {
  if (!ctx || !Q0 || !Qsave) return ctx ? (ctx->err = "hg_tsit5_solve: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  if (refuse_fields(ctx, "hg_tsit5_solve")) return HG_ERR_STATE;
  if (!(t1 > t0) || n_save <= 0 || !(abstol > 0.0) || !(reltol >= 0.0) || !(dt0 > 0.0)) {
    ctx->err = "hg_tsit5_solve: need t1 > t0, n_save > 0, abstol > 0, reltol >= 0, dt0 > 0";
    return HG_ERR_ARG;
  }
  const int64_t N = ctx->N, n3 = 3 * N;
  int rc = hg_set_params(ctx, params, n_params);
  if (rc) return rc;
  if (!ctx->rk) {
    rc = dalloc(ctx, &ctx->rk, size_t(10) * n3);
    if (!rc) rc = dalloc(ctx, &ctx->rk_part, size_t((n3 + kBlock - 1) / kBlock) + 1);
    if (rc) return rc;
  }
  double* K[7];
  for (int j = 0; j < 7; ++j) K[j] = ctx->rk + j * n3;
  double* y = ctx->rk + 7 * n3;
  double* ynew = ctx->rk + 8 * n3;
  double* ytmp = ctx->rk + 9 * n3;
  double* errv = ctx->rk_part + (n3 + kBlock - 1) / kBlock;
  const int grid = int((n3 + kBlock - 1) / kBlock);
  auto trip = [N](double* p, const double** c) {
    c[0] = p;
    c[1] = p + N;
    c[2] = p + 2 * N;
  };
  auto rhs = [&](const double* in, double* out) {
    const double* i3[3];
    trip(const_cast<double*>(in), i3);
    double* o3[3] = {out, out + N, out + 2 * N};
    return enqueue_step_p(ctx, i3, o3, 0.0, false);
  };
  auto combine = [&](const double* base, double* out, double h, const double* w, int nk) {
    RkArgs r{};
    r.n = n3;
    r.y = base;
    r.out = out;
    r.nk = nk;
    for (int j = 0; j < nk; ++j) {
      r.k[j] = K[j];
      r.coef[j] = h * w[j];
    }
    hipLaunchKernelGGL(k_rk_combine, dim3(grid), dim3(kBlock), 0, ctx->stream, r);
  };
  auto save = [&](int64_t idx) {
    double* src[3] = {y, y + N, y + 2 * N};
    return state_out(ctx, src, Qsave + idx * n3);
  };
  rc = state_in(ctx, Q0, 0);
  for (int c3 = 0; c3 < 3 && !rc; ++c3)
    if (hipMemcpyAsync(y + c3 * N, ctx->st[0][c3], N * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream) !=
        hipSuccess)
      rc = HG_ERR_HIP;
  if (!rc) rc = clear_flags(ctx);
  if (!rc) rc = save(0);
  if (!rc) rc = rhs(y, K[0]);
  if (rc) return rc;
  int64_t accepted = 0, rejected = 0, nrhs = 1;
  double t = t0, dt = dt0;
  const int64_t max_steps = 100000000;
  for (int64_t si = 1; si <= n_save && !rc; ++si) {
    const double tt = (si == n_save) ? t1 : t0 + (t1 - t0) * double(si) / double(n_save);
    while (t < tt && !rc) {
      if (accepted + rejected >= max_steps) {
        ctx->err = "hg_tsit5_solve: too many steps";
        return HG_ERR_STATE;
      }
      const bool last = dt >= (tt - t) * (1.0 - 1e-12);
      const double h = last ? tt - t : dt;
      for (int st = 1; st < 7 && !rc; ++st) {  // stages 2..7; stage 7's input is the new solution
        double* dst = (st == 6) ? ynew : ytmp;
        combine(y, dst, h, tsit5::a[st], st);
        rc = rhs(dst, K[st]);
      }
      nrhs += 6;
      if (rc) break;
      RkArgs r{};
      r.n = n3;
      r.y = y;
      r.nk = 7;
      for (int j = 0; j < 7; ++j) {
        r.k[j] = K[j];
        r.coef[j] = h * tsit5::btilde[j];
      }
      hipLaunchKernelGGL(k_rk_err, dim3(grid), dim3(kBlock), 0, ctx->stream, r, ynew, abstol, reltol,
                         ctx->rk_part);
      hipLaunchKernelGGL(k_sum, dim3(1), dim3(kBlock), 0, ctx->stream, grid, ctx->rk_part, errv);
      double ss = 0.0;
      if (hipMemcpyAsync(&ss, errv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        ctx->err = "hg_tsit5_solve: HIP error";
        return HG_ERR_HIP;
      }
      const double err = std::sqrt(ss / double(n3));
      if (!std::isfinite(err)) {
        ctx->err = "hg_tsit5_solve: non-finite error estimate";
        return HG_ERR_NONFINITE;
      }
      const double fac = err > 0.0 ? 0.9 * std::pow(err, -0.2) : 10.0;
      if (err <= 1.0) {  // accept: the new solution and its derivative (first same as last)
        t = last ? tt : t + h;
        std::swap(y, ynew);
        std::swap(K[0], K[6]);
        ++accepted;
        const double hn = h * std::min(10.0, std::max(0.2, fac));
        if (!last || hn < dt) dt = hn;
      } else {
        ++rejected;
        dt = h * std::max(0.2, fac);
      }
      if (dt < 1e-14 * (t1 - t0)) {
        ctx->err = "hg_tsit5_solve: step size underflow";
        return HG_ERR_STATE;
      }
    }
    if (!rc) rc = save(si);
  }
  if (!rc) rc = check_flags(ctx);
  if (stats) {
    stats[0] = accepted;
    stats[1] = rejected;
    stats[2] = nrhs;
  }
  return rc;
}

int hg_probe_copy(hg_ctx* ctx, int64_t nbytes, int32_t reps, double* gbps) {
  if (!ctx || nbytes < 16 || reps <= 0) return HG_ERR_ARG;
  const int64_t n2 = nbytes / 16;
  double2 *src = nullptr, *dst = nullptr;
  HG_TRY(hipMalloc(reinterpret_cast<void**>(&src), n2 * 16));
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&dst), n2 * 16);
  if (e != hipSuccess) {
    (void)hipFree(src);
    ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return HG_ERR_HIP;
  }
  int rc = HG_OK;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int grid = 256 * 8 * 4;  // 4 work-groups per SIMD
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(kBlock), 0, ctx->stream, n2, src, dst);  // warm-up
  (void)hipEventRecord(e0, ctx->stream);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(kBlock), 0, ctx->stream, n2, src, dst);
  (void)hipEventRecord(e1, ctx->stream);
  if (hipEventSynchronize(e1) != hipSuccess) rc = HG_ERR_HIP;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (gbps) *gbps = 2.0 * double(n2) * 16.0 * reps / (double(ms) * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(dst);
  if (rc) ctx->err = "hg_probe_copy: kernel failed";
  return rc;
}

int hg_time_advance(hg_ctx* ctx, double dt, int64_t nsteps, int32_t reps, double* ms_per_step,
                    double* ms_step_kernel) {
  if (!ctx || nsteps <= 0 || reps <= 0) return HG_ERR_ARG;
  if (require_fields(ctx, "hg_time_advance")) return HG_ERR_STATE;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = clear_flags(ctx);
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = HG_ERR_HIP;
  if (!rc) rc = advance_async(ctx, dt, kGraphSteps);  // warm-up (graph instantiation)
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = HG_ERR_HIP;
  if (!rc && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = HG_ERR_HIP;
  for (int r = 0; r < reps && !rc; ++r) rc = advance_async(ctx, dt, nsteps);
  if (!rc && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = HG_ERR_HIP;
  if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = HG_ERR_HIP;
  float ms = 0.f;
  if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = HG_ERR_HIP;
  if (!rc && ms_per_step) *ms_per_step = double(ms) / double(reps * nsteps);
  // the step kernel alone, continuing from the current state: events
  // attached to each of 200 single dispatches
  if (!rc && ms_step_kernel) rc = advance_timed(ctx, dt, 200, ms_step_kernel);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc) rc = check_flags(ctx);
  return rc;
}

int hg_batch_set_dt(hg_ctx* ctx, const double* dt_inst) {
  if (!ctx) return HG_ERR_ARG;
  ctx->adj_pending = false;  // a recorded forward sweep no longer matches
  if (dt_inst && refuse_fields(ctx, "hg_batch_set_dt")) return HG_ERR_STATE;
  HG_TRY(hipStreamSynchronize(ctx->stream));
  if (drop_graph(ctx)) return HG_ERR_HIP;  // captured launches hold the previous step pointer
  if (!dt_inst) {
    ctx->dt_cell = nullptr;  // the buffer stays owned by the context
    return HG_OK;
  }
  const size_t ninst = ctx->inst_off.size() - 1;
  for (size_t r = 0; r < ninst; ++r)
    if (!(dt_inst[r] > 0.0) || !std::isfinite(dt_inst[r])) {
      ctx->err = "hg_batch_set_dt: time steps must be positive and finite";
      return HG_ERR_ARG;
    }
  std::vector<double> v(ctx->N);
  for (int64_t i = 0; i < ctx->N; ++i) {
    const int64_t c = ctx->perm[i];
    const size_t r = size_t(std::upper_bound(ctx->inst_off.begin(), ctx->inst_off.end(), c) - ctx->inst_off.begin()) - 1;
    v[i] = dt_inst[r];
  }
  if (!ctx->dt_buf) {
    int rc = dalloc(ctx, &ctx->dt_buf, ctx->N);
    if (rc) return rc;
  }
  int rc = upload(ctx, ctx->dt_buf, v.data(), ctx->N);
  if (rc) return rc;
  ctx->dt_cell = ctx->dt_buf;
  return HG_OK;
}

int hg_last_checkpoint_interval(const hg_ctx* ctx, int64_t* kck) {
  if (!ctx || !kck) return HG_ERR_ARG;
  *kck = ctx->last_kck;
  return HG_OK;
}

int hg_set_adjoint_tile(hg_ctx* ctx, int32_t tile_cells) {
  if (!ctx) return HG_ERR_ARG;
  if (tile_cells != 0 && tile_cells != 128 && tile_cells != 256 && tile_cells != 512) {
    ctx->err = "hg_set_adjoint_tile: tile size must be 0 (automatic), 128, 256 or 512 cells";
    return HG_ERR_ARG;
  }
  if (!ctx->tiled) {
    ctx->err = "hg_set_adjoint_tile: this context runs the per-cell kernels";
    return HG_ERR_STATE;
  }
  ctx->adj_T_req = tile_cells;
  return drop_graph(ctx);  // captured reversals hold the old tiling; the next adjoint call re-tiles
}

int hg_set_tile_config(hg_ctx* ctx, int32_t tile_cells, int32_t xcd_chunk) {
  if (!ctx) return HG_ERR_ARG;
  if (!ctx->tiled) {
    ctx->err = "hg_set_tile_config: this context runs the per-cell kernel";
    return HG_ERR_STATE;
  }
  if (xcd_chunk < 0 || xcd_chunk > 4096) {
    ctx->err = "hg_set_tile_config: xcd_chunk must be in [0, 4096]";
    return HG_ERR_ARG;
  }
  ctx->xcd_chunk = xcd_chunk;
  if (tile_cells != ctx->tileT) return setup_tiles(ctx, tile_cells);
  return drop_graph(ctx);  // the captured launches hold the old mapping
}

int hg_get_tile_config(hg_ctx* ctx, int32_t* tile_cells, int32_t* xcd_chunk) {
  if (!ctx) return HG_ERR_ARG;
  if (tile_cells) *tile_cells = ctx->tiled ? ctx->tileT : 0;
  if (xcd_chunk) *xcd_chunk = ctx->xcd_chunk;
  return HG_OK;
}

int hg_autotune(hg_ctx* ctx, double dt, int64_t nsteps, int32_t* tile_cells, int32_t* xcd_chunk,
                double* ms_per_step) {
  if (!ctx || !(dt > 0.0) || nsteps <= 0) {
    if (ctx) ctx->err = "hg_autotune: need dt > 0 and nsteps > 0";
    return HG_ERR_ARG;
  }
  if (!ctx->fields_set) {
    ctx->err = "hg_autotune: fields not set";
    return HG_ERR_STATE;
  }
  const int64_t N = ctx->N;
  double best = INFINITY;
  int32_t bestT = ctx->tileT, bestX = ctx->xcd_chunk;
  if (ctx->tiled) {
    // keep the device state: every candidate starts from it, and it is restored at the end
    for (int k = 0; k < 3; ++k)
      HG_TRY(hipMemcpyAsync(ctx->scratch + k * N, ctx->st[ctx->cur][k], N * sizeof(double), hipMemcpyDeviceToDevice,
                            ctx->stream));
    int rc = HG_OK;
    // The device clocks settle over several hundred steps: warm up first, so
    // that candidates timed early are not slowed by a cold device; then two
    // passes over the candidates, each keeping its faster pass.
    {
      double warm = 0.0;
      ctx->cur = 0;
      for (int k = 0; k < 3 && !rc; ++k)
        if (hipMemcpyAsync(ctx->st[0][k], ctx->scratch + k * N, N * sizeof(double), hipMemcpyDeviceToDevice,
                           ctx->stream) != hipSuccess)
          rc = HG_ERR_HIP;
      if (!rc) rc = clear_flags(ctx);
      if (!rc) rc = advance_timed(ctx, dt, 500, &warm);
      if (!rc) rc = check_flags(ctx);
    }
    std::vector<std::pair<int, int>> shapes;
    for (int T : {128, 256, 512})
      if (tile_lds_bytes(T, ctx->tileMD) <= ctx->lds_max)
        for (int x : {0, 1, 2, 4, 8, 16, 64}) shapes.emplace_back(T, x);
    std::vector<double> t_best(shapes.size(), INFINITY);
    for (int pass = 0; pass < 2 && !rc; ++pass)
      for (size_t c = 0; c < shapes.size() && !rc; ++c) {
        rc = hg_set_tile_config(ctx, shapes[c].first, shapes[c].second);
        for (int k = 0; k < 3 && !rc; ++k)
          if (hipMemcpyAsync(ctx->st[0][k], ctx->scratch + k * N, N * sizeof(double), hipMemcpyDeviceToDevice,
                             ctx->stream) != hipSuccess)
            rc = HG_ERR_HIP;
        ctx->cur = 0;
        if (!rc) rc = clear_flags(ctx);
        double warm = 0.0, ms = 0.0;  // step kernel alone: events on each dispatch, after a short warm-up
        if (!rc) rc = advance_timed(ctx, dt, 16, &warm);
        if (!rc) rc = advance_timed(ctx, dt, nsteps, &ms);
        if (!rc) rc = check_flags(ctx);
        if (!rc) t_best[c] = std::min(t_best[c], ms);
      }
    for (size_t c = 0; c < shapes.size() && !rc; ++c)
      if (t_best[c] < best) {
        best = t_best[c];
        bestT = shapes[c].first;
        bestX = shapes[c].second;
      }
    if (!rc) rc = hg_set_tile_config(ctx, bestT, bestX);
    // Re-warm with the kept shape: switching back may rebuild the tiling on
    // the host while the device idles, and a caller that times right after
    // the tuner should find the device as warm as the candidates found it
    // and the stepping graph for dt already instantiated.
    for (int k = 0; k < 3 && !rc; ++k)
      if (hipMemcpyAsync(ctx->st[0][k], ctx->scratch + k * N, N * sizeof(double), hipMemcpyDeviceToDevice,
                         ctx->stream) != hipSuccess)
        rc = HG_ERR_HIP;
    ctx->cur = 0;
    if (!rc) rc = clear_flags(ctx);
    if (!rc) rc = ensure_graph(ctx, dt);
    if (!rc) rc = advance_async(ctx, dt, kRewarmSteps);
    if (!rc) rc = check_flags(ctx);
    for (int k = 0; k < 3 && !rc; ++k)
      if (hipMemcpyAsync(ctx->st[0][k], ctx->scratch + k * N, N * sizeof(double), hipMemcpyDeviceToDevice,
                         ctx->stream) != hipSuccess)
        rc = HG_ERR_HIP;
    ctx->cur = 0;
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = HG_ERR_HIP;
    if (rc) {
      if (ctx->err.empty()) ctx->err = "hg_autotune: HIP error";
      return rc;
    }
  }
  if (tile_cells) *tile_cells = ctx->tiled ? bestT : 0;
  if (xcd_chunk) *xcd_chunk = bestX;
  if (ms_per_step) *ms_per_step = ctx->tiled ? best : 0.0;
  return HG_OK;
}

// ------------------------------------------------------------ adjoint ABI --
static int adj_upload_lam(hg_ctx* ctx, const double* hbar, double* dst) {
  const int64_t N = ctx->N;
  double* const d[3] = {dst, dst + N, dst + 2 * N};
  return perm_in(ctx, hbar, 3, d);
}

static int adj_download(hg_ctx* ctx, const double* lam_dev, double* Q_bar, double* n_bar, double* zb_bar,
                        double* inletQ_bar) {
  const int64_t N = ctx->N;
  int rc = 0;
  if (Q_bar) {
    double* src[3] = {const_cast<double*>(lam_dev), const_cast<double*>(lam_dev) + N,
                      const_cast<double*>(lam_dev) + 2 * N};
    rc = state_out(ctx, src, Q_bar);
  }
  if (!rc && n_bar) rc = cell_out(ctx, ctx->adj_nbar, n_bar);
  if (!rc && zb_bar) rc = cell_out(ctx, ctx->adj_zbbar, zb_bar);
  if (!rc && inletQ_bar && ctx->n_inlet) {
    HG_TRY(hipMemcpyAsync(inletQ_bar, ctx->adj_Qbar, ctx->n_inlet * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    HG_TRY(hipStreamSynchronize(ctx->stream));
  }
  return rc;
}

int hg_rhs_vjp(hg_ctx* ctx, const double* Q, const double* params, int64_t n_params, const double* dQdt_bar,
               double* Q_bar, double* n_bar, double* zb_bar, double* inletQ_bar) {
  if (!ctx || !Q || !dQdt_bar) return ctx ? (ctx->err = "hg_rhs_vjp: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  const int64_t N = ctx->N;
  ctx->adj_want_n = n_bar != nullptr;  // cotangents nobody asked for are not accumulated
  ctx->adj_want_zb = zb_bar != nullptr;
  int rc = hg_set_params(ctx, params, n_params);
  if (!rc) rc = ensure_adj(ctx, 1, 1);
  if (!rc) rc = state_in(ctx, Q, 0);
  if (!rc) rc = adj_upload_lam(ctx, dQdt_bar, ctx->adj_lam[0]);
  if (!rc) rc = zero_adj_accumulators(ctx);
  if (!rc) rc = clear_flags(ctx);
  if (rc) return rc;
  ctx->cur = 0;
  const double* lam[3] = {ctx->adj_lam[0], ctx->adj_lam[0] + N, ctx->adj_lam[0] + 2 * N};
  double* out[3] = {ctx->adj_lam[1], ctx->adj_lam[1] + N, ctx->adj_lam[1] + 2 * N};
  rc = enqueue_adjoint(ctx, ctx->st[0], lam, nullptr, out, 1.0, false);
  if (!rc) rc = finish_adjoint(ctx);
  if (!rc) rc = adj_download(ctx, ctx->adj_lam[1], Q_bar, n_bar, zb_bar, inletQ_bar);
  if (!rc) rc = check_flags(ctx);
  return rc;
}

// The checkpoint interval the library picks for kck <= 0: the whole run as
// one stored segment (nothing recomputed) when its states, clamp flags and
// recorded conveyances (and the closure's n) take at most half of the device
// memory free for them (the context's segment buffers count as free), else
// ceil(sqrt(nsteps)) (two segment buffers and nsteps / kck checkpoints:
// about 3 sqrt(nsteps) states).
int64_t auto_kck(hg_ctx* ctx, int64_t nsteps) {
  const double N = double(ctx->N);
  const double per_step = 24.0 * N + N + 8.0 * std::max<int32_t>(ctx->n_inlet, 1) + (ctx->ude_prec ? 8.0 * N : 0.0);
  size_t fr = 0, tot = 0;
  double avail = 0.0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess) avail = double(fr);
  for (int b = 0; b < 2; ++b) avail += double(ctx->adj_seg_cap[b]) * per_step;
  if (double(nsteps) * per_step + 24.0 * N <= 0.5 * avail) return nsteps;
  return std::max<int64_t>(1, int64_t(std::ceil(std::sqrt(double(nsteps)))));
}

static int adj_run(hg_ctx* ctx, const double* Q0, const double* params, int64_t n_params, double dt, int64_t nsteps,
                   int64_t kck) {
  if (nsteps <= 0 || kck <= 0) {
    ctx->err = "nsteps and the checkpoint interval must be positive";
    return HG_ERR_ARG;
  }
  const int64_t nseg = (nsteps + kck - 1) / kck;
  ctx->last_kck = kck;
  ctx->adj_pending = false;
  int rc = hg_set_params(ctx, params, n_params);
  if (!rc) rc = ensure_adj(ctx, kck, nseg);
  if (!rc) rc = hg_state_upload(ctx, Q0);
  if (!rc) rc = clear_flags(ctx);
  if (!rc) rc = forward_checkpointed(ctx, dt, nsteps, kck);
  return rc;
}

int hg_adjoint(hg_ctx* ctx, const double* Q0, const double* params, int64_t n_params, double dt, int64_t nsteps,
               int64_t kck, const double* QN_bar, double* Q0_bar, double* n_bar, double* zb_bar,
               double* inletQ_bar) {
  if (!ctx || !Q0 || !QN_bar) return ctx ? (ctx->err = "hg_adjoint: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  ctx->adj_want_n = n_bar != nullptr;  // cotangents nobody asked for are not accumulated
  ctx->adj_want_zb = zb_bar != nullptr;
  if (kck <= 0 && nsteps > 0) kck = auto_kck(ctx, nsteps);
  int rc = adj_run(ctx, Q0, params, n_params, dt, nsteps, kck);
  if (!rc) rc = adj_upload_lam(ctx, QN_bar, ctx->adj_lam[0]);
  if (!rc) rc = zero_adj_accumulators(ctx);
  int cur = 0;
  if (!rc) rc = reverse_sweep(ctx, dt, nsteps, kck, cur);
  if (!rc) rc = finish_adjoint(ctx);
  if (!rc) rc = adj_download(ctx, ctx->adj_lam[cur], Q0_bar, n_bar, zb_bar, inletQ_bar);
  if (!rc) rc = check_flags(ctx);
  return rc;
}

int hg_adjoint_forward(hg_ctx* ctx, const double* Q0, const double* params, int64_t n_params, double dt,
                       int64_t nsteps, int64_t kck, double* QN) {
  if (!ctx || !Q0) return ctx ? (ctx->err = "hg_adjoint_forward: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  if (kck <= 0 && nsteps > 0) kck = auto_kck(ctx, nsteps);
  int rc = adj_run(ctx, Q0, params, n_params, dt, nsteps, kck);
  if (!rc) rc = check_flags(ctx);
  if (!rc && QN) rc = hg_state_download(ctx, QN);
  if (rc) return rc;
  ctx->adj_pending = true;
  ctx->pend_dt = dt;
  ctx->pend_nsteps = nsteps;
  ctx->pend_kck = kck;
  return HG_OK;
}

int hg_adjoint_reverse(hg_ctx* ctx, const double* QN_bar, double* Q0_bar, double* n_bar, double* zb_bar,
                       double* inletQ_bar) {
  if (!ctx || !QN_bar) return ctx ? (ctx->err = "hg_adjoint_reverse: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  if (!ctx->adj_pending) {
    ctx->err = "hg_adjoint_reverse: no forward sweep recorded (hg_adjoint_forward must come right before, with no "
               "gradient call or field / parameter change in between)";
    return HG_ERR_STATE;
  }
  ctx->adj_pending = false;  // the reversal recomputes into the segment buffers the forward sweep recorded
  ctx->adj_want_n = n_bar != nullptr;
  ctx->adj_want_zb = zb_bar != nullptr;
  int rc = adj_upload_lam(ctx, QN_bar, ctx->adj_lam[0]);
  if (!rc) rc = zero_adj_accumulators(ctx);
  int cur = 0;
  if (!rc) rc = reverse_sweep(ctx, ctx->pend_dt, ctx->pend_nsteps, ctx->pend_kck, cur);
  if (!rc) rc = finish_adjoint(ctx);
  if (!rc) rc = adj_download(ctx, ctx->adj_lam[cur], Q0_bar, n_bar, zb_bar, inletQ_bar);
  if (!rc) rc = check_flags(ctx);
  return rc;
}

int hg_grad_probes(hg_ctx* ctx, const double* Q0, const double* params, int64_t n_params, double dt, int64_t nsteps,
                   int64_t kck, int64_t n_probe, const int32_t* probe, const double* target, double* loss,
                   double* n_bar, double* zb_bar, double* inletQ_bar) {
  if (!ctx || !Q0 || (n_probe > 0 && (!probe || !target)) || !loss)
    return ctx ? (ctx->err = "hg_grad_probes: NULL array", HG_ERR_ARG) : HG_ERR_ARG;
  const int64_t N = ctx->N;
  ctx->adj_want_n = n_bar != nullptr;  // cotangents nobody asked for are not accumulated
  ctx->adj_want_zb = zb_bar != nullptr;
  if (kck <= 0 && nsteps > 0) kck = auto_kck(ctx, nsteps);
  for (int64_t k = 0; k < n_probe; ++k)
    if (probe[k] < 0 || probe[k] >= N) {
      ctx->err = "probe cell out of range";
      return HG_ERR_ARG;
    }
  int rc = adj_run(ctx, Q0, params, n_params, dt, nsteps, kck);
  if (rc) return rc;
  // WSE = (xi + h_still) + zb at the probes (swe_2D_inversion.jl:387-399),
  // formed on the device: only the probe cells travel
  if (n_probe > ctx->probe_cap) {
    rc = dalloc(ctx, &ctx->probe_i, size_t(n_probe));
    if (!rc) rc = dalloc(ctx, &ctx->probe_d, size_t(3) * n_probe);
    if (rc) return rc;
    ctx->probe_cap = n_probe;
  }
  const int64_t cap = ctx->probe_cap;
  std::vector<int32_t> pc(n_probe);
  for (int64_t k = 0; k < n_probe; ++k) pc[k] = ctx->iperm[probe[k]];
  std::vector<double> r(n_probe);
  if (n_probe > 0) {
    rc = upload(ctx, ctx->probe_i, pc.data(), size_t(n_probe));
    if (!rc) rc = upload(ctx, ctx->probe_d, target, size_t(n_probe));
    if (rc) return rc;
    hipLaunchKernelGGL(k_probe_resid, dim3(grid_of(n_probe)), dim3(kBlock), 0, ctx->stream, int(n_probe),
                       int(ctx->Nm), ctx->probe_i, ctx->probe_d, ctx->st[ctx->cur][0], ctx->hstill, ctx->zb,
                       ctx->probe_d + cap);
    HG_TRY(hipGetLastError());
    HG_TRY(hipMemcpyAsync(r.data(), ctx->probe_d + cap, n_probe * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    HG_TRY(hipStreamSynchronize(ctx->stream));
  }
  // the loss, and its cotangent of the final state: 2 r at each probe cell
  // (summed over repeated probes, in probe order), zero elsewhere
  double L = 0.0;
  std::vector<int32_t> sc;
  std::vector<double> sv;
  std::unordered_map<int32_t, size_t> at;
  for (int64_t k = 0; k < n_probe; ++k) {
    L = L + r[k] * r[k];
    auto ins = at.emplace(pc[k], sc.size());
    if (ins.second) {
      sc.push_back(pc[k]);
      sv.push_back(0.0);
    }
    sv[ins.first->second] += 2.0 * r[k];
  }
  *loss = L;
  HG_TRY(hipMemsetAsync(ctx->adj_lam[0], 0, size_t(3) * N * sizeof(double), ctx->stream));
  if (!sc.empty()) {
    rc = upload(ctx, ctx->probe_i, sc.data(), sc.size());
    if (!rc) rc = upload(ctx, ctx->probe_d + 2 * cap, sv.data(), sv.size());
    if (rc) return rc;
    hipLaunchKernelGGL(k_seed, dim3(grid_of(int64_t(sc.size()))), dim3(kBlock), 0, ctx->stream, int(sc.size()),
                       ctx->probe_i, ctx->probe_d + 2 * cap, ctx->adj_lam[0]);
    HG_TRY(hipGetLastError());
  }
  rc = zero_adj_accumulators(ctx);
  int cur = 0;
  if (!rc) rc = reverse_sweep(ctx, dt, nsteps, kck, cur);
  if (!rc) rc = finish_adjoint(ctx);
  if (!rc) rc = adj_download(ctx, ctx->adj_lam[cur], nullptr, n_bar, zb_bar, inletQ_bar);
  if (!rc) rc = check_flags(ctx);
  if (!rc && zb_bar)  // the loss reads zb directly
    for (int64_t k = 0; k < n_probe; ++k) zb_bar[probe[k]] += 2.0 * r[k];
  return rc;
}

}  // extern "C"
