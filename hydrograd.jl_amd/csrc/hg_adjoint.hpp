// hg_adjoint.hpp -- hand-written reverse-mode derivatives (vector-Jacobian
// products) of the flux and source terms of the shallow-water RHS.
//
// Semantics follow forward-mode / reverse-mode AD of the reference
// (ForwardDiff / Zygote through swe_2d_rhs): branch tests carry no
// derivative, the derivative follows the taken branch, clamped depths and
// masked discharges are constants, and a branch that copies a mirrored state
// (swe_2D_solvers.jl:23-51) routes the copy's cotangent back to its source.
// Derivatives are accurate, not bit-reproductions of an AD tape: repeated
// divisions by one denominator use its reciprocal.
#pragma once
#include "hg_roe_pair.hpp"

namespace hg {

// Cotangents of one side's raw flux inputs.
struct SideBar {
  double xi, h, hu, hv;
};

// Cotangents of one side including its derived quantities (u, v, s, p).
struct ExtBar {
  double xi, h, hu, hv, u, v, s, p;
};

__host__ __device__ __forceinline__ void ext_zero(ExtBar& e) { e.xi = e.h = e.hu = e.hv = e.u = e.v = e.s = e.p = 0.0; }

// u = hu/h, v = hv/h, s = sqrt(h + eps), p = 0.5 g ((xi+eps)^2 + 2 xi hs):
// fold the derived cotangents into (xi, h, hu, hv).
// rh = 1 / h and r2s = 0.5 / s are the side's own reciprocals (a tiled
// kernel forms them once per cell).
__host__ __device__ __forceinline__ SideBar ext_to_side_r(const ExtBar& e, const CellQ& c, double rh, double r2s,
                                                          double half_g) {
  SideBar s;
  s.hu = e.hu + e.u * rh;
  s.hv = e.hv + e.v * rh;
  s.h = e.h - (e.u * c.u + e.v * c.v) * rh + e.s * r2s;
  s.xi = e.xi + e.p * half_g * (2.0 * (c.xi + EPS) + 2.0 * c.hs);
  return s;
}

__host__ __device__ __forceinline__ SideBar ext_to_side(const ExtBar& e, const CellQ& c, double half_g) {
  return ext_to_side_r(e, c, 1.0 / c.h, 0.5 / c.s, half_g);
}

// Physical flux of one side (hu, hv, u, v, p), reverse.
__host__ __device__ __forceinline__ void phys_vjp(double hu, double hv, double u, double v, double nx, double ny,
                                                  const double fb[3], ExtBar& e) {
  const double un = nx * u + ny * v;
  e.hu += fb[0] * nx + fb[1] * un;
  e.hv += fb[0] * ny + fb[2] * un;
  e.u += fb[1] * nx * hu + fb[2] * nx * hv;
  e.v += fb[1] * ny * hu + fb[2] * ny * hv;
  e.p += fb[1] * nx + fb[2] * ny;
}

// VJP of the flux F(A | B, n) seen from A (Riemann_2D_Roe,
// swe_2D_solvers.jl:4-164) with cotangent Fb: cotangents of both sides' raw
// and derived quantities (unfolded; ext_to_side folds them).  A and B are
// the flux inputs (clamped depth, masked discharge, derived u, v, s, p) as
// make_cellq forms them.  Every field of ea and eb is assigned.
__host__ __device__ __forceinline__ void roe_vjp_ext(const CellQ& A, const CellQ& B, double nx, double ny, double g,
                                                     double half_g, double hmin, const double Fb[3], ExtBar& ea,
                                                     ExtBar& eb) {
  if (A.h <= hmin && B.h <= hmin) {
    ext_zero(ea);
    ext_zero(eb);
    return;
  }
  int mode = 0;  // 0 general, 1 B mirrored from A, 2 A mirrored from B
  double hA = A.h, huA = A.hu, hvA = A.hv, uA = A.u, vA = A.v, sA = A.s;
  double hB = B.h, huB = B.hu, hvB = B.hv, uB = B.u, vB = B.v, sB = B.s;
  if ((B.h <= hmin) && (A.h + A.zb) < (B.zb + hmin)) {
    mode = 1;
    hB = A.h; huB = -A.hu; hvB = -A.hv; uB = -A.u; vB = -A.v; sB = A.s;
  } else if ((A.h <= hmin) && (B.h + B.zb) < (A.zb + hmin)) {
    mode = 2;
    hA = B.h; huA = -B.hu; hvA = -B.hv; uA = -B.u; vA = -B.v; sA = B.s;
  } else if (A.h <= hmin || B.h <= hmin) {  // physical flux of the wet side, h^2 pressure
    ext_zero(ea);
    ext_zero(eb);
    const bool left_dry = A.h <= hmin;
    const CellQ& W = left_dry ? B : A;
    ExtBar& ew = left_dry ? eb : ea;
    const double un = nx * W.u + ny * W.v;
    ew.hu = Fb[0] * nx + Fb[1] * un;
    ew.hv = Fb[0] * ny + Fb[2] * un;
    const double t = Fb[1] * W.hu + Fb[2] * W.hv;
    ew.u = t * nx;
    ew.v = t * ny;
    ew.h = (Fb[1] * nx + Fb[2] * ny) * half_g * 2.0 * (W.h + EPS);
    return;
  }
  // ---- forward recompute (general branch).  sqrt(x) and 1 / sqrt(x) come
  // from one reciprocal square root: x r and r, r = 1 / sqrt(x).
  const double hRoe = (hA + hB) * 0.5;
  const double rsAB = 1.0 / (sA + sB);
  const double uRoe = (sA * uA + sB * uB) * rsAB;
  const double vRoe = (sA * vA + sB * vB) * rsAB;
  const double un = uRoe * nx + vRoe * ny;
  const double xc = g * hRoe + EPS;
  const double rc = rsqrt(xc);
  const double c = xc * rc;
  const double o = 0.5 * rc;
  const double d0 = B.xi - A.xi, d1 = huB - huA, d2 = hvB - hvA;
  const double k0 = -(uRoe * ny - vRoe * nx);
  const double w0 = k0 * d0 + ny * d1 + (-nx) * d2;
  const double w1 = (un * o + 0.5) * d0 + (-nx * o) * d1 + (-ny * o) * d2;
  const double w2 = (-un * o + 0.5) * d0 + (nx * o) * d1 + (ny * o) * d2;
  const double um = un - c, up = un + c;
  const double x0 = un * un + EPS, x1 = um * um + EPS, x2 = up * up + EPS;
  const double r0 = rsqrt(x0), r1 = rsqrt(x1), r2 = rsqrt(x2);
  const double l0 = x0 * r0, l1 = x1 * r1, l2 = x2 * r2;
  // ---- reverse: F = ((fA + fB) - r) / 2, r = R (|Lambda| w)
  const double hf0 = 0.5 * Fb[0], hf1 = 0.5 * Fb[1], hf2 = 0.5 * Fb[2];
  const double ab0 = hf2 * nx - hf1 * ny;
  const double ab1 = -(hf0 + hf1 * (uRoe - c * nx) + hf2 * (vRoe - c * ny));
  const double ab2 = -(hf0 + hf1 * (uRoe + c * nx) + hf2 * (vRoe + c * ny));
  const double a12 = l1 * w1 + l2 * w2;
  double cb = (hf1 * nx + hf2 * ny) * (l1 * w1 - l2 * w2);
  const double q1b = ab1 * w1 * (um * r1), q2b = ab2 * w2 * (up * r2);  // d l_k: cotangent times l_k'
  const double w0b = ab0 * l0, w1b = ab1 * l1, w2b = ab2 * l2;
  double unb = ab0 * w0 * (un * r0) + q1b + q2b;
  cb += q2b - q1b;
  const double k0b = w0b * d0;
  const double wd = w2b - w1b;
  const double d0b = w0b * k0 + w1b * (un * o + 0.5) + w2b * (-un * o + 0.5);
  const double d1b = w0b * ny + wd * nx * o;
  const double d2b = -w0b * nx + wd * ny * o;
  unb -= wd * o * d0;
  cb += wd * (un * d0 - nx * d1 - ny * d2) * o * rc;
  const double uRb = -hf1 * a12 - k0b * ny + unb * nx;
  const double vRb = -hf2 * a12 + k0b * nx + unb * ny;
  const double hRb2 = 0.5 * (cb * g * (0.5 * rc));  // half of hRoe's cotangent reaches each side's h
  const double pb = hf1 * nx + hf2 * ny;
  // the flux states' cotangents (d = B - A; xi is never mirrored)
  const double unA = nx * uA + ny * vA, tA = hf1 * huA + hf2 * hvA;
  const double unB = nx * uB + ny * vB, tB = hf1 * huB + hf2 * hvB;
  ea.xi = -d0b;
  ea.p = pb;
  ea.hu = hf0 * nx + hf1 * unA - d1b;
  ea.hv = hf0 * ny + hf2 * unA - d2b;
  ea.u = tA * nx + uRb * sA * rsAB;
  ea.v = tA * ny + vRb * sA * rsAB;
  ea.s = (uRb * (uA - uRoe) + vRb * (vA - vRoe)) * rsAB;
  ea.h = hRb2;
  eb.xi = d0b;
  eb.p = pb;
  eb.hu = hf0 * nx + hf1 * unB + d1b;
  eb.hv = hf0 * ny + hf2 * unB + d2b;
  eb.u = tB * nx + uRb * sB * rsAB;
  eb.v = tB * ny + vRb * sB * rsAB;
  eb.s = (uRb * (uB - uRoe) + vRb * (vB - vRoe)) * rsAB;
  eb.h = hRb2;
  // ---- route a mirrored copy back to its source side
  if (mode == 1) {  // hB = hA, huB = -huA, hvB = -hvA, uB = -uA, vB = -vA, sB = sA
    ea.h += eb.h; ea.hu -= eb.hu; ea.hv -= eb.hv; ea.u -= eb.u; ea.v -= eb.v; ea.s += eb.s;
    eb.h = eb.hu = eb.hv = eb.u = eb.v = eb.s = 0.0;
  } else if (mode == 2) {
    eb.h += ea.h; eb.hu -= ea.hu; eb.hv -= ea.hv; eb.u -= ea.u; eb.v -= ea.v; eb.s += ea.s;
    ea.h = ea.hu = ea.hv = ea.u = ea.v = ea.s = 0.0;
  }
}

// The same, folded: cotangents of both sides' (xi, h, hu, hv).
__host__ __device__ inline void roe_vjp(const CellQ& A, const CellQ& B, double nx, double ny, double g, double half_g,
                                        double hmin, const double Fb[3], SideBar& Ab, SideBar& Bb) {
  ExtBar ea, eb;
  roe_vjp_ext(A, B, nx, ny, g, half_g, hmin, Fb, ea, eb);
  Ab = ext_to_side(ea, A, half_g);
  Bb = ext_to_side(eb, B, half_g);
}

// The part of roe_vjp that falls on A, for an internal face whose other
// side's cotangent the neighbour forms from its own side: the cotangents of
// A's raw and derived quantities are ADDED to e (unfolded; ext_to_side folds
// a cell's sum once).  Same branches and routing as roe_vjp; B's own
// cotangent is never formed.
__host__ __device__ __forceinline__ void roe_vjp_a(const CellQ& A, const CellQ& B, double nx, double ny, double g,
                                                   double half_g, double hmin, const double Fb[3], ExtBar& e) {
  if (A.h <= hmin && B.h <= hmin) return;
  int mode = 0;  // 0 general, 1 B mirrored from A, 2 A mirrored from B
  double hA = A.h, huA = A.hu, hvA = A.hv, uA = A.u, vA = A.v, sA = A.s;
  double hB = B.h, huB = B.hu, hvB = B.hv, uB = B.u, vB = B.v, sB = B.s;
  if ((B.h <= hmin) && (A.h + A.zb) < (B.zb + hmin)) {
    mode = 1;
    hB = A.h; huB = -A.hu; hvB = -A.hv; uB = -A.u; vB = -A.v; sB = A.s;
  } else if ((A.h <= hmin) && (B.h + B.zb) < (A.zb + hmin)) {
    mode = 2;
    hA = B.h; huA = -B.hu; hvA = -B.hv; uA = -B.u; vA = -B.v; sA = B.s;
  } else if (A.h <= hmin) {  // B's physical flux alone: nothing reaches A
    return;
  } else if (B.h <= hmin) {  // A's physical flux with the h^2 pressure
    const double un = nx * A.u + ny * A.v;
    e.hu += Fb[0] * nx + Fb[1] * un;
    e.hv += Fb[0] * ny + Fb[2] * un;
    const double t = Fb[1] * A.hu + Fb[2] * A.hv;
    e.u += t * nx;
    e.v += t * ny;
    e.h += (Fb[1] * nx + Fb[2] * ny) * half_g * 2.0 * (A.h + EPS);
    return;
  }
  // ---- forward recompute (general branch)
  const double hRoe = (hA + hB) * 0.5;
  const double rsAB = 1.0 / (sA + sB);
  const double uRoe = (sA * uA + sB * uB) * rsAB;
  const double vRoe = (sA * vA + sB * vB) * rsAB;
  const double un = uRoe * nx + vRoe * ny;
  const double c = sqrt(g * hRoe + EPS);
  const double rc = 1.0 / c;
  const double o = 0.5 * rc;
  const double d0 = B.xi - A.xi, d1 = huB - huA, d2 = hvB - hvA;
  const double k0 = -(uRoe * ny - vRoe * nx);
  const double w0 = k0 * d0 + ny * d1 + (-nx) * d2;
  const double w1 = (un * o + 0.5) * d0 + (-nx * o) * d1 + (-ny * o) * d2;
  const double w2 = (-un * o + 0.5) * d0 + (nx * o) * d1 + (ny * o) * d2;
  const double um = un - c, up = un + c;
  const double l0 = sqrt(un * un + EPS), l1 = sqrt(um * um + EPS), l2 = sqrt(up * up + EPS);
  // ---- reverse: F = ((fA + fB) - r) / 2, r = R (|Lambda| w)
  const double hf0 = 0.5 * Fb[0], hf1 = 0.5 * Fb[1], hf2 = 0.5 * Fb[2];
  const double ab0 = hf2 * nx - hf1 * ny;
  const double ab1 = -(hf0 + hf1 * (uRoe - c * nx) + hf2 * (vRoe - c * ny));
  const double ab2 = -(hf0 + hf1 * (uRoe + c * nx) + hf2 * (vRoe + c * ny));
  const double a12 = l1 * w1 + l2 * w2;
  double cb = (hf1 * nx + hf2 * ny) * (l1 * w1 - l2 * w2);
  const double q1b = ab1 * w1 * (um / l1), q2b = ab2 * w2 * (up / l2);  // d l_k: cotangent times l_k'
  const double w0b = ab0 * l0, w1b = ab1 * l1, w2b = ab2 * l2;
  double unb = ab0 * w0 * (un / l0) + q1b + q2b;
  cb += q2b - q1b;
  const double k0b = w0b * d0;
  const double wd = w2b - w1b;
  const double d0b = w0b * k0 + w1b * (un * o + 0.5) + w2b * (-un * o + 0.5);
  const double d1b = w0b * ny + wd * nx * o;
  const double d2b = -w0b * nx + wd * ny * o;
  unb -= wd * o * d0;
  cb += wd * (un * d0 - nx * d1 - ny * d2) * o * rc;
  const double uRb = -hf1 * a12 - k0b * ny + unb * nx;
  const double vRb = -hf2 * a12 + k0b * nx + unb * ny;
  const double hRb = cb * g * (0.5 * rc);
  const double pb = hf1 * nx + hf2 * ny;
  e.xi -= d0b;
  e.p += pb;
  if (mode == 2) return;  // everything else reaches B, whose state A's mirror copies
  const double unA = nx * uA + ny * vA;
  const double tA = hf1 * huA + hf2 * hvA;
  e.hu += hf0 * nx + hf1 * unA - d1b;
  e.hv += hf0 * ny + hf2 * unA - d2b;
  e.u += tA * nx + uRb * sA * rsAB;
  e.v += tA * ny + vRb * sA * rsAB;
  e.s += (uRb * (uA - uRoe) + vRb * (vA - vRoe)) * rsAB;
  e.h += 0.5 * hRb;
  if (mode == 1) {  // B's copy of A (hB = hA, qB = -qA, uB = -uA, sB = sA) routes back to A
    const double unB = -unA;
    const double tB = -tA;
    e.hu -= hf0 * nx + hf1 * unB + d1b;
    e.hv -= hf0 * ny + hf2 * unB + d2b;
    e.u -= tB * nx + uRb * sB * rsAB;
    e.v -= tB * ny + vRb * sB * rsAB;
    e.s += (uRb * (uB - uRoe) + vRb * (vB - vRoe)) * rsAB;
    e.h += 0.5 * hRb;
  }
}

// VJP of the momentum source of one wet cell (semi_discretize_swe_2D.jl:
// 474-475, 544-547): src = g xi S0 - friction.  y1, y2 are the cotangents of
// the x / y momentum RHS; outputs add to the cotangents of xi, h, hu, hv,
// n, S0x, S0y.
__host__ __device__ __forceinline__ void source_vjp(double g, double kn2, double hsmall, double xi, double h,
                                                    double qx, double qy, double n, double S0x, double S0y, double y1,
                                                    double y2, double& xib, double& hb, double& qxb, double& qyb,
                                                    double& nb, double& S0xb, double& S0yb) {
  xib += g * S0x * y1 + g * S0y * y2;
  S0xb = g * xi * y1;
  S0yb = g * xi * y2;
  const double fxb = -y1, fyb = -y2;
  const double x = h + hsmall;
  const double p73 = pow_7_3(x);
  const double rp = 1.0 / (kn2 * p73);
  const double cc = g * (n * n) * rp;
  const double q2 = qx * qx + qy * qy + EPS;
  const double rs = rsqrt(q2);  // 1 / |q|_eps, and |q|_eps = q2 rs
  const double s = q2 * rs;
  const double cb = fxb * s * qx + fyb * s * qy;
  const double sb = fxb * cc * qx + fyb * cc * qy;
  qxb += fxb * cc * s + sb * qx * rs;
  qyb += fyb * cc * s + sb * qy * rs;
  nb += cb * 2.0 * g * n * rp;
  hb += cb * (-7.0 / 3.0) * cc / x;
}

// VJP of the momentum source with FlowResistance friction m q / |q|_eps
// (semi_discretize_swe_2D.jl:474-475, 521-531); mb += d/dm.  No depth term.
__host__ __device__ __forceinline__ void source_vjp_mag(double g, double xi, double qx, double qy, double m,
                                                        double S0x, double S0y, double y1, double y2, double& xib,
                                                        double& qxb, double& qyb, double& mb, double& S0xb,
                                                        double& S0yb) {
  xib += g * S0x * y1 + g * S0y * y2;
  S0xb = g * xi * y1;
  S0yb = g * xi * y2;
  const double fxb = -y1, fyb = -y2;
  const double rs = rsqrt(qx * qx + qy * qy + EPS);
  const double dq = fxb * qx + fyb * qy;
  mb += dq * rs;
  const double sb = -dq * m * rs * rs;  // through |q|_eps
  qxb += fxb * m * rs + sb * qx * rs;
  qyb += fyb * m * rs + sb * qy * rs;
}

}  // namespace hg
