// hg_device.hpp -- device-side arithmetic shared by the forward and adjoint
// kernels of libhgswe (gfx950).
//
// Every expression keeps Hydrograd.jl's evaluation order (left-to-right, no
// fused multiply-add: the library is compiled with -ffp-contract=off), so a
// forward evaluation is bit-identical to the reference's Float64 arithmetic
// except for pow(), whose last bit may differ between math libraries.
// Multiplications by 0.5 replace the reference's divisions by 2.0 (exact).
#pragma once
#include <hip/hip_runtime.h>

namespace hg {

constexpr double EPS = 2.220446049250313e-16;  // eps(Float64)

template <class T>
struct SideT {  // one side of a face: the inputs Riemann_2D_Roe reads
  T xi, hs, h, hu, hv, zb;
};
using Side = SideT<double>;

// Riemann_2D_Roe (swe_2D_solvers.jl:4-164).  The 3x3 products of
// :103-118 are written out; entries that multiply by the literal 0.0 / 1.0
// of R and |Lambda| are folded (identical up to the sign of an exact zero).
// T is double on the forward path and a dual number for the device JVP.
template <class T>
__host__ __device__ __forceinline__ void roe_flux(SideT<T> L, SideT<T> R, double nx, double ny, double g,
                                                  double half_g, double hmin, T& f0, T& f1, T& f2) {
  if (L.h <= hmin && R.h <= hmin) {  // :16-22
    f0 = 0.0; f1 = 0.0; f2 = 0.0;
    return;
  } else if ((R.h <= hmin) && (L.h + L.zb) < (R.zb + hmin)) {  // :23-36
    R.h = L.h; R.hu = -L.hu; R.hv = -L.hv;
  } else if ((L.h <= hmin) && (R.h + R.zb) < (L.zb + hmin)) {  // :39-51
    L.h = R.h; L.hu = -R.hu; L.hv = -R.hv;
  } else if (L.h <= hmin) {  // :54-64
    const T u = R.hu / R.h, v = R.hv / R.h;
    const T pr = half_g * ((R.h + EPS) * (R.h + EPS));
    f0 = R.hu * nx + R.hv * ny;
    f1 = (R.hu * u + pr) * nx + R.hu * v * ny;
    f2 = (R.hv * u) * nx + (R.hv * v + pr) * ny;
    return;
  } else if (R.h <= hmin) {  // :65-75
    const T u = L.hu / L.h, v = L.hv / L.h;
    const T pr = half_g * ((L.h + EPS) * (L.h + EPS));
    f0 = L.hu * nx + L.hv * ny;
    f1 = (L.hu * u + pr) * nx + L.hu * v * ny;
    f2 = (L.hv * u) * nx + (L.hv * v + pr) * ny;
    return;
  }
  const T uL = L.hu / L.h, vL = L.hv / L.h;
  const T uR = R.hu / R.h, vR = R.hv / R.h;
  const T sL = sqrt(L.h + EPS), sR = sqrt(R.h + EPS);
  const T hRoe = (L.h + R.h) * 0.5;
  const T sLR = sL + sR;
  const T uRoe = (sL * uL + sR * uR) / sLR;
  const T vRoe = (sL * vL + sR * vR) / sLR;
  const T unRoe = uRoe * nx + vRoe * ny;
  const T cRoe = sqrt(g * hRoe + EPS);
  const T o = 0.5 / cRoe;
  const T d0 = R.xi - L.xi, d1 = R.hu - L.hu, d2 = R.hv - L.hv;
  // L_mat * dQ
  const T w0 = (-(uRoe * ny - vRoe * nx)) * d0 + ny * d1 + (-nx) * d2;
  const T w1 = (unRoe * o + 0.5) * d0 + (-nx * o) * d1 + (-ny * o) * d2;
  const T w2 = (-unRoe * o + 0.5) * d0 + (nx * o) * d1 + (ny * o) * d2;
  // |Lambda| (smooth_abs)
  const T a0 = sqrt(unRoe * unRoe + EPS) * w0;
  const T um = unRoe - cRoe, up = unRoe + cRoe;
  const T a1 = sqrt(um * um + EPS) * w1;
  const T a2 = sqrt(up * up + EPS) * w2;
  // R_mat * (...)
  const T r0 = a1 + a2;
  const T r1 = ny * a0 + (uRoe - cRoe * nx) * a1 + (uRoe + cRoe * nx) * a2;
  const T r2 = (-nx) * a0 + (vRoe - cRoe * ny) * a1 + (vRoe + cRoe * ny) * a2;
  // physical fluxes, pre-balanced pressure 0.5 g (xi^2 + 2 xi hs)
  const T pL = half_g * ((L.xi + EPS) * (L.xi + EPS) + 2.0 * L.xi * L.hs);
  const T pR = half_g * ((R.xi + EPS) * (R.xi + EPS) + 2.0 * R.xi * R.hs);
  const T fL0 = L.hu * nx + L.hv * ny;
  const T fL1 = (L.hu * uL + pL) * nx + L.hu * vL * ny;
  const T fL2 = (L.hv * uL) * nx + (L.hv * vL + pL) * ny;
  const T fR0 = R.hu * nx + R.hv * ny;
  const T fR1 = (R.hu * uR + pR) * nx + R.hu * vR * ny;
  const T fR2 = (R.hv * uR) * nx + (R.hv * vR + pR) * ny;
  f0 = (fL0 + fR0 - r0) * 0.5;
  f1 = (fL1 + fR1 - r1) * 0.5;
  f2 = (fL2 + fR2 - r2) * 0.5;
}

// x^(7/3) for the Manning term as (x * x) * cbrt(x): within a few ulp of
// the correctly rounded power (the reference's pow(x, 7.0/3.0) carries an
// exponent 1.5e-16 off 7/3, a relative 1.5e-16 * |ln x| on its own), at a
// sixth of the instruction count of a general double pow.
template <class T>
__host__ __device__ __forceinline__ T pow_7_3(T x) { return (x * x) * cbrt(x); }

// Manning friction (semi_discretize_swe_2D.jl:544-547); h is the clamped
// depth, q the masked discharge.  Dividing by k_n^2 = 1 is exact and skipped.
template <class T>
__host__ __device__ __forceinline__ void friction(double g, double kn2, double hsmall, T h, T qx, T qy, T n, T& fx,
                                                  T& fy) {
  T c = g * (n * n);
  if (kn2 != 1.0) c = c / kn2;
  c = c / pow_7_3(h + hsmall);
  const T s = sqrt(qx * qx + qy * qy + EPS);
  fx = c * s * qx;
  fy = c * s * qy;
}

// FlowResistance friction (semi_discretize_swe_2D.jl:521-531): the UDE's
// friction magnitude m along the discharge, m q / sqrt(q_x^2 + q_y^2 + eps).
template <class T>
__host__ __device__ __forceinline__ void friction_mag(T m, T qx, T qy, T& fx, T& fy) {
  const T s = sqrt(qx * qx + qy * qy + EPS);
  fx = m * qx / s;
  fy = m * qy / s;
}

}  // namespace hg
