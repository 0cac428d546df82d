// hg_lux.hpp -- the reference's UDE network on the device, in FP64.
//
// create_NN_model (process_UDE.jl:2-65) builds a Lux chain
//   [Dense(in, w_l, act_l) -> LayerNorm(w_l)] per hidden layer l,
//   Dense(w_L, 1), then x -> lo + (hi - lo) sigmoid(x)  (output bounds),
// and the RHS feeds it inputs normalised to [-1, 1] by bounds,
//   x_norm = 2 (x - x_lo) / (x_hi - x_lo) - 1,
// per UDE_choice:
//   ManningN_h          h                     -> n   (process_ManningN_2D.jl:216-273)
//   ManningN_h_Umag_ks  h, |U| = |q| / h, ks  -> n
//   FlowResistance      h, |q|_eps            -> friction magnitude
//                       (process_FlowResistance_UDE.jl:3-46, |q| bounds
//                        h_lo U_lo .. h_hi U_hi; semi_discretize_swe_2D.jl:521-531)
// with h the clamped depth and q the masked discharge of the RHS.
//
// Layer semantics are Lux's and NNlib's published ones: Dense y = act(W x + b)
// with W (out x in); LayerNorm over the units of one cell, uncorrected
// variance, y = (a - mean) / sqrt(var + eps) * scale + bias, eps = 1f-5 (as
// Float64); NNlib's relu, leakyrelu (slope 0.01), sigmoid
// (exp(-|x|) form), tanh and softplus (log1p(exp(-|x|)) + relu(x)).
//
// One cell per lane, every layer in registers (widths up to MW, a compile-
// time bound; runtime widths are predicates).  Reverse mode recomputes each
// layer's input from the cell's features instead of storing activations, and
// sums weight cotangents per wave (fixed xor tree) into per-block slots.
#pragma once
#include <hip/hip_runtime.h>

#include "hg_ctx.hpp"
#include "hg_dual.hpp"
#include "hg_nlaw.hpp"

namespace hg {

constexpr double kLuxLayerNormEps = 9.999999747378752e-06;  // Float64(1f-5), Lux's LayerNorm default

__device__ __forceinline__ double nn_sigmoid(double x) {  // NNlib.sigmoid
  const double t = ::exp(-::fabs(x));
  return x >= 0.0 ? 1.0 / (1.0 + t) : t / (1.0 + t);
}

__device__ __forceinline__ double act_f(int k, double x) {
  switch (k) {
    case ACT_RELU: return x < 0.0 ? 0.0 : x;
    case ACT_LEAKYRELU: return x > 0.0 ? x : 0.01 * x;
    case ACT_SIGMOID: return nn_sigmoid(x);
    case ACT_TANH: return ::tanh(x);
    default: return ::log1p(::exp(-::fabs(x))) + (x < 0.0 ? 0.0 : x);  // softplus
  }
}

// d act / dz at z, given a = act(z) (NNlib's derivative rules).
__device__ __forceinline__ double act_d(int k, double z, double a) {
  switch (k) {
    case ACT_RELU: return z > 0.0 ? 1.0 : 0.0;
    case ACT_LEAKYRELU: return z > 0.0 ? 1.0 : 0.01;
    case ACT_SIGMOID: return a * (1.0 - a);
    case ACT_TANH: return 1.0 - a * a;
    default: return nn_sigmoid(z);
  }
}

// Normalised inputs of one cell (features f[0..n_in-1]).
template <class T>
__device__ __forceinline__ void lux_features(const LuxNet& L, const T& xi, const T& qx, const T& qy, double hs,
                                             double hsmall, double ks, T f[3]) {
  T h = xi + hs;
  T hu = qx, hv = qy;
  if (h <= hsmall) {
    h = T(hsmall);
    hu = T(0.0);
    hv = T(0.0);
  }
  f[0] = 2.0 * (h - L.in_lo[0]) / (L.in_hi[0] - L.in_lo[0]) - 1.0;
  f[1] = T(0.0);
  f[2] = T(0.0);
  if (L.choice == LUX_MANNING_H_UMAG_KS) {
    const T u = hu / h, v = hv / h;
    const T U = l_sqrt(u * u + v * v);
    f[1] = 2.0 * (U - L.in_lo[1]) / (L.in_hi[1] - L.in_lo[1]) - 1.0;
    f[2] = T(2.0 * (ks - L.in_lo[2]) / (L.in_hi[2] - L.in_lo[2]) - 1.0);
  } else if (L.choice == LUX_FLOW_RESISTANCE) {
    const T qm = l_sqrt(hu * hu + hv * hv + EPS);
    f[1] = 2.0 * (qm - L.in_lo[1]) / (L.in_hi[1] - L.in_lo[1]) - 1.0;
  }
}

// Scalars of the forward pass: double, or Dual1 for forward mode (value and
// one tangent; the tangent of weight k is dP[k], none when dP is NULL).  The
// value part of a Dual1 pass is computed by the same operations in the same
// order as the double pass, so both give the same bits.
template <class T>
__device__ __forceinline__ T lux_w(const double* __restrict__ P, const double* dP, int k);
template <>
__device__ __forceinline__ double lux_w<double>(const double* __restrict__ P, const double*, int k) { return P[k]; }
template <>
__device__ __forceinline__ Dual1 lux_w<Dual1>(const double* __restrict__ P, const double* dP, int k) {
  return Dual1(P[k], dP ? dP[k] : 0.0);
}
__device__ __forceinline__ double act_t(int k, double z) { return act_f(k, z); }
__device__ __forceinline__ Dual1 act_t(int k, const Dual1& z) {
  const double a = act_f(k, z.v);
  return Dual1(a, act_d(k, z.v, a) * z.d);
}
__device__ __forceinline__ double sigmoid_t(double x) { return nn_sigmoid(x); }
__device__ __forceinline__ Dual1 sigmoid_t(const Dual1& x) {
  const double s = nn_sigmoid(x.v);
  return Dual1(s, s * (1.0 - s) * x.d);
}
__device__ __forceinline__ double sqrt_t(double x) { return ::sqrt(x); }
__device__ __forceinline__ Dual1 sqrt_t(const Dual1& x) { return hg::sqrt(x); }

// Hidden layer l on input x (width win): pre-activation z, activation a,
// normalised xh = (a - mean) / s.
template <int MW, class T = double>
__device__ __forceinline__ void lux_layer(const LuxNet& L, const double* __restrict__ P, int l, const T (&x)[MW],
                                          int win, T (&z)[MW], T (&a)[MW], T (&xh)[MW], T& s,
                                          const double* dP = nullptr) {
  const int w = L.width[l];
  const int act = L.act[l];
  const int oW = L.off_W[l], ob = L.off_b[l];
  T mu = 0.0;
#pragma unroll
  for (int j = 0; j < MW; ++j) {
    if (j < w) {
      T acc = 0.0;
#pragma unroll
      for (int k = 0; k < MW; ++k)
        if (k < win)
          acc = (k == 0) ? lux_w<T>(P, dP, oW + j) * x[0] : acc + lux_w<T>(P, dP, oW + j + k * w) * x[k];
      z[j] = acc + lux_w<T>(P, dP, ob + j);
      a[j] = act_t(act, z[j]);
      mu = (j == 0) ? a[0] : mu + a[j];
    }
  }
  mu = mu / double(w);
  T var = 0.0;
#pragma unroll
  for (int j = 0; j < MW; ++j) {
    if (j < w) {
      const T d = a[j] - mu;
      var = (j == 0) ? d * d : var + d * d;
    }
  }
  var = var / double(w);
  s = sqrt_t(var + L.eps);
#pragma unroll
  for (int j = 0; j < MW; ++j)
    if (j < w) xh[j] = (a[j] - mu) / s;
}

// Input of hidden layer `upto` (the last hidden output when upto == n_hidden).
template <int MW, class T = double>
__device__ __forceinline__ int lux_forward_to(const LuxNet& L, const double* __restrict__ P, const T (&f)[3],
                                              int upto, T (&x)[MW], const double* dP = nullptr) {
#pragma unroll
  for (int k = 0; k < MW; ++k) x[k] = k < 3 ? f[k < 3 ? k : 0] : T(0.0);
  int win = L.n_in;
  for (int l = 0; l < upto; ++l) {
    T z[MW], a[MW], xh[MW], s;
    lux_layer<MW, T>(L, P, l, x, win, z, a, xh, s, dP);
    const int w = L.width[l];
    const int og = L.off_g[l], obe = L.off_be[l];
#pragma unroll
    for (int j = 0; j < MW; ++j)
      if (j < w) x[j] = xh[j] * lux_w<T>(P, dP, og + j) + lux_w<T>(P, dP, obe + j);
    win = w;
  }
  return win;
}

// Output layer and bounds: returns lo + (hi - lo) sigmoid(Wo x + bo).
template <int MW, class T = double>
__device__ __forceinline__ T lux_output(const LuxNet& L, const double* __restrict__ P, const T (&x)[MW], int win,
                                        T& sg, const double* dP = nullptr) {
  T acc = 0.0;
#pragma unroll
  for (int k = 0; k < MW; ++k)
    if (k < win)
      acc = (k == 0) ? lux_w<T>(P, dP, L.off_Wo) * x[0] : acc + lux_w<T>(P, dP, L.off_Wo + k) * x[k];
  sg = sigmoid_t(acc + lux_w<T>(P, dP, L.off_bo));
  return L.lo + (L.hi - L.lo) * sg;
}

template <int MW, class T = double>
__device__ __forceinline__ T lux_eval(const LuxNet& L, const double* __restrict__ P, const T (&f)[3],
                                      const double* dP = nullptr) {
  T x[MW];
  const int win = lux_forward_to<MW, T>(L, P, f, L.n_hidden, x, dP);
  T sg;
  return lux_output<MW, T>(L, P, x, win, sg, dP);
}

// Sum over the 64 lanes of a wave (fixed xor tree: every lane gets the total).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Reverse mode of one cell: given the output cotangent ob, adds the weight
// cotangents (summed over the wave; lane 0 adds into acc) and returns the
// feature cotangents fb.  Every lane of the wave must call it (lanes without
// a cell pass ob = 0).
template <int MW>
__device__ __forceinline__ void lux_backward(const LuxNet& L, const double* __restrict__ P, const double (&f)[3],
                                             double ob, double* acc, int lane, double (&fb)[3]) {
  auto add = [&](int k, double v) {
    v = wave_sum(v);
    if (lane == 0) acc[k] += v;
  };
  double x[MW];
  const int wL = lux_forward_to<MW>(L, P, f, L.n_hidden, x);
  double sg;
  lux_output<MW>(L, P, x, wL, sg);
  const double oo = ob * (L.hi - L.lo) * (sg * (1.0 - sg));
  const double* Wo = P + L.off_Wo;
  double xb[MW];
#pragma unroll
  for (int k = 0; k < MW; ++k) {
    if (k < wL) add(L.off_Wo + k, oo * x[k]);
    xb[k] = k < wL ? oo * Wo[k < wL ? k : 0] : 0.0;
  }
  add(L.off_bo, oo);
  for (int l = L.n_hidden - 1; l >= 0; --l) {
    double xin[MW];
    const int wi = lux_forward_to<MW>(L, P, f, l, xin);
    double z[MW], a[MW], xh[MW], s;
    lux_layer<MW>(L, P, l, xin, wi, z, a, xh, s);
    const int w = L.width[l];
    const double* g = P + L.off_g[l];
    const double* W = P + L.off_W[l];
    double gg[MW];
    double m1 = 0.0, m2 = 0.0;
#pragma unroll
    for (int j = 0; j < MW; ++j) {
      if (j < w) {
        add(L.off_g[l] + j, xb[j] * xh[j]);
        add(L.off_be[l] + j, xb[j]);
        gg[j] = xb[j] * g[j];
        m1 += gg[j];
        m2 += gg[j] * xh[j];
      }
    }
    m1 = m1 / double(w);
    m2 = m2 / double(w);
    double zb[MW];
#pragma unroll
    for (int j = 0; j < MW; ++j) {
      zb[j] = 0.0;
      if (j < w) {
        zb[j] = (gg[j] - m1 - xh[j] * m2) / s * act_d(L.act[l], z[j], a[j]);
#pragma unroll
        for (int k = 0; k < MW; ++k)
          if (k < wi) add(L.off_W[l] + j + k * w, zb[j] * xin[k]);
        add(L.off_b[l] + j, zb[j]);
      }
    }
#pragma unroll
    for (int k = 0; k < MW; ++k) {
      double t = 0.0;
      if (k < wi) {
#pragma unroll
        for (int j = 0; j < MW; ++j)
          if (j < w) t += W[j + k * w] * zb[j];
      }
      xb[k] = t;
    }
  }
  fb[0] = xb[0];
  fb[1] = MW > 1 ? xb[MW > 1 ? 1 : 0] : 0.0;
  fb[2] = MW > 2 ? xb[MW > 2 ? 2 : 0] : 0.0;
}

struct LuxArgs {
  LuxNet net;
  int32_t N;
  const double* w;       // P weights
  const double* xi;      // state (lib order)
  const double* qx;
  const double* qy;
  const double* hstill;
  const double* ks;      // per cell (ManningN_h_Umag_ks), else unused
  double hsmall;
  double* out;           // forward: n or friction magnitude per cell
  const double* obar;    // reverse: cotangent of the output per cell
  double* ox;            // += state cotangent
  double* oqx;
  double* oqy;
  double* wpart;         // [P][gridDim.x] per-block weight cotangents
};

__device__ __forceinline__ double lux_ks(const LuxArgs& a, int i) {
  return a.net.choice == LUX_MANNING_H_UMAG_KS ? a.ks[i] : 0.0;
}

}  // namespace hg
