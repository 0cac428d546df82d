// hg_roe_pair.hpp -- face-once Roe flux that returns the flux seen from
// BOTH cells of an internal face, each bit-identical to the reference's
// per-cell evaluation (Riemann_2D_Roe, swe_2D_solvers.jl:4-164, called
// once from each side by compute_inviscid_fluxes, semi_discretize_swe_2D.jl:286-419).
//
// Why two values from one evaluation are exact: evaluating the face from
// cell B (left = B, right = A, normal -n) negates n and dQ exactly, so
// unRoe, w0 and every physical flux negate exactly, w1 <-> -w2 and
// lambda1 <-> lambda2 swap exactly, and uRoe / vRoe / cRoe are unchanged
// (IEEE addition is commutative, multiplication and subtraction are
// sign-symmetric).  Only the final rows of R * (Lambda * w) for the two
// momentum components are summed in a different order:
//   A side: (T0 + T1) + T2,  B side: -((T0 + T2) + T1).
// Both sums are formed here, so each side receives exactly the value the
// reference computes for it.  The branch tests are mirror images of each
// other (branches 2/3 and 4/5 swap), so the same holds on wet/dry faces.
#pragma once
#include "hg_device.hpp"

namespace hg {

// Cell state as the flux reads it, plus per-cell derived quantities that
// the reference recomputes identically for every face of the cell:
// u = hu / h, v = hv / h, s = sqrt(h + eps), p = 0.5 g (xi^2 + 2 xi h_still)
// with the eps-smoothed square of smooth_pow.
struct CellQ {
  double xi, hs, h, hu, hv, zb, u, v, s, p;
};

__host__ __device__ __forceinline__ CellQ make_cellq(double xi, double hs, double qx, double qy, double zb, double half_g,
                                            double hsmall) {
  CellQ c;
  c.xi = xi;
  c.hs = hs;
  double h = xi + hs, hu = qx, hv = qy;
  if (h <= hsmall) {  // semi_discretize_swe_2D.jl:101-106
    h = hsmall;
    hu = 0.0;
    hv = 0.0;
  }
  c.h = h;
  c.hu = hu;
  c.hv = hv;
  c.zb = zb;
  c.u = hu / h;
  c.v = hv / h;
  c.s = sqrt(h + EPS);
  c.p = half_g * ((xi + EPS) * (xi + EPS) + 2.0 * xi * hs);
  return c;
}

// Flux through the face (A | B), normal n pointing from A to B.
// fa = flux seen from A; when BOTH, fb = flux seen from B (normal -n).
template <bool BOTH>
__host__ __device__ __forceinline__ void roe_pair(const CellQ& A, const CellQ& B, double nx, double ny, double g,
                                         double half_g, double hmin, double fa[3], double fb[3]) {
  if (A.h <= hmin && B.h <= hmin) {  // :16-22
    fa[0] = fa[1] = fa[2] = 0.0;
    if (BOTH) fb[0] = fb[1] = fb[2] = 0.0;
    return;
  }
  double hA = A.h, huA = A.hu, hvA = A.hv, uA = A.u, vA = A.v, sA = A.s;
  double hB = B.h, huB = B.hu, hvB = B.hv, uB = B.u, vB = B.v, sB = B.s;
  if ((B.h <= hmin) && (A.h + A.zb) < (B.zb + hmin)) {  // :23-36, B := mirror of A (xi_B, h_still_B kept)
    hB = A.h; huB = -A.hu; hvB = -A.hv; uB = -A.u; vB = -A.v; sB = A.s;
  } else if ((A.h <= hmin) && (B.h + B.zb) < (A.zb + hmin)) {  // :39-51
    hA = B.h; huA = -B.hu; hvA = -B.hv; uA = -B.u; vA = -B.v; sA = B.s;
  } else if (A.h <= hmin || B.h <= hmin) {  // :54-75, physical flux of the wet side, h^2 pressure
    const bool left_dry = A.h <= hmin;
    const double hw = left_dry ? B.h : A.h;
    const double huw = left_dry ? B.hu : A.hu, hvw = left_dry ? B.hv : A.hv;
    const double uw = left_dry ? B.u : A.u, vw = left_dry ? B.v : A.v;
    const double pr = half_g * ((hw + EPS) * (hw + EPS));
    const double f0 = huw * nx + hvw * ny;
    const double f1 = (huw * uw + pr) * nx + huw * vw * ny;
    const double f2 = (hvw * uw) * nx + (hvw * vw + pr) * ny;
    fa[0] = f0; fa[1] = f1; fa[2] = f2;
    if (BOTH) { fb[0] = -f0; fb[1] = -f1; fb[2] = -f2; }
    return;
  }
  const double hRoe = (hA + hB) * 0.5;
  const double sAB = sA + sB;
  const double uRoe = (sA * uA + sB * uB) / sAB;
  const double vRoe = (sA * vA + sB * vB) / sAB;
  const double unRoe = uRoe * nx + vRoe * ny;
  const double cRoe = sqrt(g * hRoe + EPS);
  const double o = 0.5 / cRoe;
  const double d0 = B.xi - A.xi, d1 = huB - huA, d2 = hvB - hvA;
  const double w0 = (-(uRoe * ny - vRoe * nx)) * d0 + ny * d1 + (-nx) * d2;
  const double w1 = (unRoe * o + 0.5) * d0 + (-nx * o) * d1 + (-ny * o) * d2;
  const double w2 = (-unRoe * o + 0.5) * d0 + (nx * o) * d1 + (ny * o) * d2;
  const double a0 = sqrt(unRoe * unRoe + EPS) * w0;
  const double um = unRoe - cRoe, up = unRoe + cRoe;
  const double a1 = sqrt(um * um + EPS) * w1;
  const double a2 = sqrt(up * up + EPS) * w2;
  const double r0 = a1 + a2;
  const double T0 = ny * a0, T1 = (uRoe - cRoe * nx) * a1, T2 = (uRoe + cRoe * nx) * a2;
  const double U0 = (-nx) * a0, U1 = (vRoe - cRoe * ny) * a1, U2 = (vRoe + cRoe * ny) * a2;
  // physical fluxes (the pressures of both sides are unchanged by the
  // mirror branches: they depend on xi and h_still only)
  const double fA0 = huA * nx + hvA * ny;
  const double fA1 = (huA * uA + A.p) * nx + huA * vA * ny;
  const double fA2 = (hvA * uA) * nx + (hvA * vA + A.p) * ny;
  const double fB0 = huB * nx + hvB * ny;
  const double fB1 = (huB * uB + B.p) * nx + huB * vB * ny;
  const double fB2 = (hvB * uB) * nx + (hvB * vB + B.p) * ny;
  const double s0 = fA0 + fB0, s1 = fA1 + fB1, s2 = fA2 + fB2;
  fa[0] = (s0 - r0) * 0.5;
  fa[1] = (s1 - ((T0 + T1) + T2)) * 0.5;
  fa[2] = (s2 - ((U0 + U1) + U2)) * 0.5;
  if (BOTH) {
    fb[0] = -fa[0];
    fb[1] = -((s1 - ((T0 + T2) + T1)) * 0.5);
    fb[2] = -((s2 - ((U0 + U2) + U1)) * 0.5);
  }
}

}  // namespace hg
