// hg_nlaw.hpp -- the forward simulation's variable Manning n
// ("ManningN_option": "variable"): n is re-evaluated from the state at every
// RHS evaluation (semi_discretize_swe_2D.jl:129-135) by one of the laws of
// create_manning_function (process_ManningN_2D.jl:122-134):
//
//   power_law  n = n_lo + (n_hi - n_lo) (h + eps)^-k                  :154-165
//   sigmoid    n = n_lo + (n_hi - n_lo) / (1 + exp(k (h - h_mid)))   :170-182
//   inverse    n = n_lo + (n_hi - n_lo) / (1 + k h)                   :139-150
//   h_Umag_ks  Cheng (2008) friction factor of h, |U| and ks          :186-213
//              Re = |U| h / nu, alpha = 1 / (1 + (Re/850)^9),
//              beta = 1 / (1 + (Re / (160 h/ks))^2),
//              f = 1 / ((Re/24)^alpha (1.8 log10(Re/2.1))^(2(1-alpha)beta)
//                       (2 log10(11.8 h/ks))^(2(1-alpha)(1-beta))),
//              n = sqrt(f/8) h^(1/6) / sqrt(9.81)
//
// Inputs inside the RHS (semi_discretize_swe_2D.jl:101-106, 130-134):
// h = xi + h_still clamped to h_small (the discharge zeroed there),
// u = q_x / h, v = q_y / h, |U| = sqrt(u^2 + v^2) (no smoothing).
//
// Every law is written once, templated on the scalar: double for the RHS,
// Dual1 for the forward-mode tangent and for the state cotangent of reverse
// mode (one dual pass per state component).  Included after hg_dual.hpp.
#pragma once
#include <hip/hip_runtime.h>

#include "hg_ctx.hpp"
#include "hg_dual.hpp"

namespace hg {

// Elementary functions on double and on Dual1.  A zero tangent stays zero
// (no 0 * inf from a derivative that is infinite at the primal point).
__device__ __forceinline__ double l_sqrt(double x) { return ::sqrt(x); }
__device__ __forceinline__ Dual1 l_sqrt(const Dual1& x) {
  const double r = ::sqrt(x.v);
  return Dual1(r, x.d == 0.0 ? 0.0 : x.d / (2.0 * r));
}
__device__ __forceinline__ double l_exp(double x) { return ::exp(x); }
__device__ __forceinline__ Dual1 l_exp(const Dual1& x) {
  const double r = ::exp(x.v);
  return Dual1(r, x.d * r);
}
__device__ __forceinline__ double l_log10(double x) { return ::log10(x); }
__device__ __forceinline__ Dual1 l_log10(const Dual1& x) {
  return Dual1(::log10(x.v), x.d == 0.0 ? 0.0 : x.d / (x.v * 2.302585092994046));
}
__device__ __forceinline__ double l_pow(double x, double y) { return ::pow(x, y); }
__device__ __forceinline__ Dual1 l_pow(const Dual1& x, double y) {
  const double r = ::pow(x.v, y);
  return Dual1(r, x.d == 0.0 ? 0.0 : x.d * (y * ::pow(x.v, y - 1.0)));
}
__device__ __forceinline__ Dual1 l_pow(const Dual1& x, const Dual1& y) {
  const double r = ::pow(x.v, y.v);
  double d = x.d == 0.0 ? 0.0 : x.d * (y.v * ::pow(x.v, y.v - 1.0));
  if (y.d != 0.0) d = d + y.d * (r * ::log(x.v));
  return Dual1(r, d);
}

// The law at depth h and velocity magnitude U; AUX also returns h/ks, f, Re
// (zero for the depth-only laws, like the reference's outputs).
template <class T, bool AUX>
__device__ __forceinline__ T nlaw_eval(const NLaw& L, const T& h, const T& U, double ks, T* hks, T* fo, T* Reo) {
  if (L.kind != NLAW_H_UMAG_KS) {
    if (AUX) *hks = *fo = *Reo = T(0.0);
    if (L.kind == NLAW_POWER) return L.lo + (L.hi - L.lo) * l_pow(h + EPS, -L.k);
    if (L.kind == NLAW_SIGMOID) return L.lo + (L.hi - L.lo) / (1.0 + l_exp(L.k * (h - L.hmid)));
    return L.lo + (L.hi - L.lo) / (1.0 + L.k * h);  // NLAW_INVERSE
  }
  const T Re = U * h / 1.0e-6;
  const T h_ks = h / ks;
  const T alpha = 1.0 / (1.0 + l_pow(Re / 850.0, 9.0));
  const T r = Re / (h_ks * 160.0);
  const T beta = 1.0 / (1.0 + r * r);
  const T part1 = l_pow(Re / 24.0, alpha);
  const T part2 = l_pow(1.8 * l_log10(Re / 2.1), 2.0 * (1.0 - alpha) * beta);
  const T part3 = l_pow(2.0 * l_log10(11.8 * h_ks), 2.0 * (1.0 - alpha) * (1.0 - beta));
  const T f = 1.0 / (part1 * part2 * part3);
  if (AUX) {
    *hks = h_ks;
    *fo = f;
    *Reo = Re;
  }
  return l_sqrt(f / 8.0) * l_pow(h, 1.0 / 6.0) / ::sqrt(9.81);
}

// The law's inputs from one cell's state, as the RHS forms them.
template <class T>
__device__ __forceinline__ void nlaw_features(const NLaw& L, const T& xi, const T& qx, const T& qy, double hs,
                                              double hsmall, T& h, T& U) {
  h = xi + hs;
  T hu = qx, hv = qy;
  if (h <= hsmall) {
    h = T(hsmall);
    hu = T(0.0);
    hv = T(0.0);
  }
  if (L.kind == NLAW_H_UMAG_KS) {
    const T u = hu / h, v = hv / h;
    U = l_sqrt(u * u + v * v);
  } else {
    U = T(0.0);
  }
}

struct NLawArgs {
  NLaw law;
  int32_t N;
  const double* xi;      // state (lib order)
  const double* qx;
  const double* qy;
  const double* hstill;
  const double* ks;      // per cell (h_Umag_ks), else unused
  double hsmall;
  double* n;             // out: n per cell
  // evaluation at given (h, |U|)
  const double* h_in;
  const double* U_in;
  double* hks;           // nullable outputs
  double* f;
  double* Re;
  // reverse mode: state cotangent += nbar * dn/d(xi, q_x, q_y)
  const double* nbar;
  double* ox;
  double* oqx;
  double* oqy;
  // forward mode: state tangent in, tangent of n out
  const double* dxi;
  const double* dqx;
  const double* dqy;
  double* dn;
};

}  // namespace hg
