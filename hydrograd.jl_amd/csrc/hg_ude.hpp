// hg_ude.hpp -- universal-differential-equation closure for Manning's n:
// n_i = 0.01 + softplus(MLP(h_i, |u|_i, zb_i)), MLP = 3 -> 16 -> 16 -> 1 with
// tanh hidden layers (353 weights), evaluated on the matrix cores.
//
// The reference's UDE path (update_ManningN_UDE, process_ManningN_2D.jl:
// 216-273, Lux model of process_UDE.jl:2-65) evaluates a per-cell MLP of the
// local flow state once per RHS call and uses its output as ManningN_cells;
// this closure has that role with the network above.
//
// Weight layout of one set (353 values): W1[16][3] (row = hidden unit),
// b1[16], W2[16][16] (W2[j][k]: unit j of layer 2 from unit k of layer 1),
// b2[16], W3[16], b3.
//
// Features of cell i (from the state, as the flux sees it):
//   h  = xi + h_still, clamped to h_small (then q = 0);
//   |u| = sqrt(u^2 + v^2 + eps), u = q_x / h, v = q_y / h;
//   zb.
//
// Matrix-core layout (v_mfma_{f32,f64}_16x16x4): for D = A (16x4) * B (4x16),
// lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and receives, in
// element r = 0..3, D[row(l / 16, r)][l % 16] with row = 4 (l / 16) + r for
// f32 and row = l / 16 + 4 r for f64 (checked by hg_mfma_probe).
#pragma once
#include <hip/hip_runtime.h>

#include "hg_device.hpp"

namespace hg {

constexpr int kUdeW = 353;
constexpr int kUdeW1 = 0, kUdeB1 = 48, kUdeW2 = 64, kUdeB2 = 320, kUdeW3 = 336, kUdeB3 = 352;

template <class T>
struct Mfma;

template <>
struct Mfma<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int kq, int r) { return 4 * kq + r; }
  // The f32 network's own products and sums fuse (one rounding): its values
  // are checked against a float64 evaluation to f32 tolerance, not bit for
  // bit against the reference's Float64 arithmetic, which the solver keeps.
  static __device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
  // tanh in f32 as 1 - 2 / (e^(2|x|) + 1) with the sign of x, from the
  // hardware exp and reciprocal: with every step rounded as f32, the exp off
  // by up to three ulp and the reciprocal by one, the absolute error against
  // float64 tanh stays below 5e-7 (largest near 0, where 1 - 2r cancels;
  // below 1e-7 for |x| >= 0.5 with correctly rounded steps), a few ulp of an
  // f32 activation, which the next layer's f32 sums carry at the same
  // magnitude.  A Taylor branch for |x| < 0.4 that kept the
  // relative error to 4 ulp cost 6 % of the 32-set closure gradient and left
  // the closure's n against a float64 evaluation unchanged (1.0e-6).  The
  // closure's 32 activations per cell dominate its forward and reverse kernels.
  static __device__ __forceinline__ float tanh_(float x) {
    const float e = __expf(2.0f * fabsf(x));
    return copysignf(fma_(-2.0f, __builtin_amdgcn_rcpf(e + 1.0f), 1.0f), x);
  }
  static __device__ __forceinline__ float exp_(float x) { return expf(x); }
  static __device__ __forceinline__ float log1p_(float x) { return log1pf(x); }
};

template <>
struct Mfma<double> {
  typedef double V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int kq, int r) { return kq + 4 * r; }
  static __device__ __forceinline__ double fma_(double a, double b, double c) { return a * b + c; }
  static __device__ __forceinline__ double tanh_(double x) { return tanh(x); }
  static __device__ __forceinline__ double exp_(double x) { return exp(x); }
  static __device__ __forceinline__ double log1p_(double x) { return log1p(x); }
};

template <class T>
__device__ __forceinline__ T softplus(T x) {
  return x > T(20) ? x : Mfma<T>::log1p_(Mfma<T>::exp_(x));
}

template <class T>
__device__ __forceinline__ T sigmoid(T x) {
  return T(1) / (T(1) + Mfma<T>::exp_(-x));
}

// LDS exchange between lanes of one wave: stores before, loads after.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct UdeArgs {
  int32_t N;   // cells over all fields
  int32_t Nm;  // cells per field: h_still and zb of cell i are at i % Nm
  int32_t n_sets;
  const double* xi;
  const double* qx;
  const double* qy;
  const double* hstill;
  const double* zb;
  const int32_t* set_of_cell;  // NULL: every cell uses set 0
  const void* w;               // [n_sets][kUdeW] of T
  double* n;                   // out: Manning n per cell
  // reverse mode
  const double* nbar;          // cotangent of n
  double* ox;                  // state cotangent, += d/d(xi, q_x, q_y)
  double* oqx;
  double* oqy;
  double* zbbar;               // optional, += d/d zb
  double* wpart;               // [nslots][kUdeW] weight cotangents, one slot per (block, set) pair
  const int32_t* blk_off;      // entries of block b: blk_off[b] .. blk_off[b + 1] - 1, one per set
  const int32_t* blk_set;      //   its cells use (order of first appearance): that set and its
  const int32_t* blk_slot;     //   slot; all NULL: slot = block (one set)
  int32_t nslots;
  int32_t cpb;                 // 64-cell chunks per reverse-mode block
  double g_unused;
  double hsmall;
};

// Features of one cell (double), see the header comment.
__device__ __forceinline__ void ude_features(const UdeArgs& a, int i, double& h, double& u, double& v, double& um,
                                             bool& wet) {
  const double xi = a.xi[i];
  h = xi + a.hstill[i % a.Nm];
  double hu = a.qx[i], hv = a.qy[i];
  wet = h > a.hsmall;
  if (!wet) {
    h = a.hsmall;
    hu = 0.0;
    hv = 0.0;
  }
  u = hu / h;
  v = hv / h;
  um = sqrt(u * u + v * v + EPS);
}

}  // namespace hg
