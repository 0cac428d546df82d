// hg_oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of Hydrograd.jl's 2-D shallow-water right-hand side and
// explicit Euler step, used as the checker for the HIP library.  Only the
// test-suite, the smoke check and the CPU-baseline leg of the benchmark may
// load this code; the product library never links or calls it.
//
// Every function follows one reference function (paths relative to the
// Hydrograd.jl repository root) and keeps its evaluation order:
// left-to-right association, no fused multiply-add, the same branch tests,
// eps(Float64)-smoothed primitives, and the reference's strong-zero Bool*x.
//
// The scalar type T is double, Dual<K> (forward-mode derivative with K
// tangent directions) or Rev (reverse mode on a tape).  Dual follows
// ForwardDiff's "active branch" semantics: comparisons and branch tests read
// the value part only, and a branch that substitutes a constant carries a
// zero tangent; Rev follows the same branches.
//
// Parity status of each function is tracked in the test-suite
// (tests/test_oracle_*.py); see the header comment of each function for
// the closed forms and fixtures that pin it.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- scalars --
constexpr double EPS = 2.220446049250313e-16;  // eps(Float64) = 2^-52

template <int K>
struct Dual {
  double v;
  double d[K];
  Dual() : v(0.0) { for (int k = 0; k < K; ++k) d[k] = 0.0; }
  Dual(double x) : v(x) { for (int k = 0; k < K; ++k) d[k] = 0.0; }  // NOLINT
};

inline double val(double x) { return x; }
template <int K> inline double val(const Dual<K>& x) { return x.v; }

template <int K> inline Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r; r.v = -a.v; for (int k = 0; k < K; ++k) r.d[k] = -a.d[k]; return r;
}
template <int K> inline Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r; r.v = a.v + b.v; for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k]; return r;
}
template <int K> inline Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r; r.v = a.v - b.v; for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k]; return r;
}
template <int K> inline Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r; r.v = a.v * b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K> inline Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r; r.v = a.v / b.v;
  for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
  return r;
}
template <int K> inline Dual<K> operator+(const Dual<K>& a, double b) { return a + Dual<K>(b); }
template <int K> inline Dual<K> operator+(double a, const Dual<K>& b) { return Dual<K>(a) + b; }
template <int K> inline Dual<K> operator-(const Dual<K>& a, double b) { return a - Dual<K>(b); }
template <int K> inline Dual<K> operator-(double a, const Dual<K>& b) { return Dual<K>(a) - b; }
template <int K> inline Dual<K> operator*(const Dual<K>& a, double b) { return a * Dual<K>(b); }
template <int K> inline Dual<K> operator*(double a, const Dual<K>& b) { return Dual<K>(a) * b; }
template <int K> inline Dual<K> operator/(const Dual<K>& a, double b) { return a / Dual<K>(b); }
template <int K> inline Dual<K> operator/(double a, const Dual<K>& b) { return Dual<K>(a) / b; }

inline double dsqrt(double x) { return std::sqrt(x); }
template <int K> inline Dual<K> dsqrt(const Dual<K>& a) {
  Dual<K> r; r.v = std::sqrt(a.v);
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / (2.0 * r.v);
  return r;
}
// x^p for a constant real exponent p (Julia ^(Float64, Float64)).
inline double dpow(double x, double p) { return std::pow(x, p); }
template <int K> inline Dual<K> dpow(const Dual<K>& a, double p) {
  Dual<K> r; r.v = std::pow(a.v, p);
  const double dd = p * std::pow(a.v, p - 1.0);
  for (int k = 0; k < K; ++k) r.d[k] = dd * a.d[k];
  return r;
}

// Reverse mode by operator overloading (the role Zygote plays in the
// reference's inversion, swe_2D_inversion.jl:247-483): every operation with
// an active operand appends one record to a tape -- the tape indices of its
// operands (-1: a constant) and its partial derivatives with respect to them
// -- and the reverse sweep walks the records backwards, adding
// adjoint * partial into each operand.  Branch tests read the value only, as
// for Dual, so a Rev run differentiates the same branches a Dual run does.
// The tape is per thread; a computation recorded on it must run on one
// thread (the OpenMP loops over cells with one thread).
struct Tape {
  struct Rec {
    int32_t a, b;    // operand records, -1: a constant
    double da, db;   // partials of the result with respect to them
  };
  std::vector<Rec> rec;
  std::vector<double> adj;
  size_t size() const { return rec.size(); }
  void clear() { rec.clear(); }
  int32_t push(int32_t ia, double pa, int32_t ib, double pb) {
    if (rec.size() >= size_t(INT32_MAX)) throw std::runtime_error("reverse-mode tape: more than 2^31 records");
    rec.push_back(Rec{ia, ib, pa, pb});
    return int32_t(rec.size() - 1);
  }
  // adj[k] for every record given the output seeds already placed in adj
  void sweep() {
    for (size_t k = rec.size(); k-- > 0;) {
      const double w = adj[k];
      if (w == 0.0) continue;
      const Rec& r = rec[k];
      if (r.a >= 0) adj[r.a] += r.da * w;
      if (r.b >= 0) adj[r.b] += r.db * w;
    }
  }
};
inline Tape& tape() { static thread_local Tape t; return t; }

struct Rev {
  double v;
  int32_t i;  // tape record, -1 for a constant
  Rev() : v(0.0), i(-1) {}
  Rev(double x) : v(x), i(-1) {}  // NOLINT
  static Rev input(double x) { Rev r(x); r.i = tape().push(-1, 0.0, -1, 0.0); return r; }
};
inline Rev rev_op(double v, int32_t ia, double pa, int32_t ib, double pb) {
  Rev r(v);
  if (ia >= 0 || ib >= 0) r.i = tape().push(ia, pa, ib, pb);
  return r;
}
inline double val(const Rev& x) { return x.v; }
inline Rev operator-(const Rev& a) { return rev_op(-a.v, a.i, -1.0, -1, 0.0); }
inline Rev operator+(const Rev& a, const Rev& b) { return rev_op(a.v + b.v, a.i, 1.0, b.i, 1.0); }
inline Rev operator-(const Rev& a, const Rev& b) { return rev_op(a.v - b.v, a.i, 1.0, b.i, -1.0); }
inline Rev operator*(const Rev& a, const Rev& b) { return rev_op(a.v * b.v, a.i, b.v, b.i, a.v); }
inline Rev operator/(const Rev& a, const Rev& b) {
  const double r = a.v / b.v;
  return rev_op(r, a.i, 1.0 / b.v, b.i, -r / b.v);
}
inline Rev operator+(const Rev& a, double b) { return a + Rev(b); }
inline Rev operator+(double a, const Rev& b) { return Rev(a) + b; }
inline Rev operator-(const Rev& a, double b) { return a - Rev(b); }
inline Rev operator-(double a, const Rev& b) { return Rev(a) - b; }
inline Rev operator*(const Rev& a, double b) { return a * Rev(b); }
inline Rev operator*(double a, const Rev& b) { return Rev(a) * b; }
inline Rev operator/(const Rev& a, double b) { return a / Rev(b); }
inline Rev operator/(double a, const Rev& b) { return Rev(a) / b; }
inline Rev dsqrt(const Rev& a) {
  const double r = std::sqrt(a.v);
  return rev_op(r, a.i, 1.0 / (2.0 * r), -1, 0.0);
}
inline Rev dpow(const Rev& a, double p) {
  return rev_op(std::pow(a.v, p), a.i, p * std::pow(a.v, p - 1.0), -1, 0.0);
}

// Julia `Bool * x`: ifelse(b, x, copysign(zero(x), x)) -- a strong zero.
template <class T> inline T bool_times(bool b, const T& x) {
  return b ? x : T(std::copysign(0.0, val(x)));
}
// Julia max(a, b) on reals; for Dual the derivative follows the larger value
// (ties keep the first argument, which is the constant in every call here).
template <class T> inline T jmax(const T& a, const T& b) { return (val(b) > val(a)) ? b : a; }

// ------------------------------------------- smooth_functions.jl:4-52 ------
// smooth_abs(x) = sqrt(x^2 + eps)            (smooth_functions.jl:4-12)
template <class T> inline T smooth_abs(const T& x) { return dsqrt(x * x + EPS); }
// smooth_sqrt(x) = sqrt(x + eps)            (smooth_functions.jl:39-44)
template <class T> inline T smooth_sqrt(const T& x) { return dsqrt(x + EPS); }
// smooth_pow(x, 2) = (x + eps)^2; Julia's ^(Float64, Int) with n = 2 returns
// the correctly rounded square, i.e. one multiply  (smooth_functions.jl:47-52)
template <class T> inline T smooth_pow2(const T& x) { T y = x + EPS; return y * y; }

// ------------------------------------------------------------ mesh data ----
// Flattened form of mesh_2D + BoundaryConditions2D (mesh_2D.jl:2-71,
// bc_2D.jl:3-47): cells 0..N-1 in the reference's cell order, each cell's
// faces in the reference's local order (edge node_k -> node_{k+1}).
enum BcKind : int32_t { BC_INLETQ = 0, BC_EXITH = 1, BC_WALL = 2, BC_SYMM = 3 };

struct Mesh {
  int64_t N = 0, G = 0;
  const int32_t* face_off = nullptr;  // [N+1]
  const int32_t* face_nbr = nullptr;  // [M]: >=0 cell, <0 ghost ~g
  const double* face_nx = nullptr;    // [M] cell_normals[i][k]
  const double* face_ny = nullptr;
  const double* face_len = nullptr;   // [M] face_lengths[cellFacesList[i,k]]
  const double* area = nullptr;       // [N] cell_areas
  // boundaries
  const int32_t* ghost_cell = nullptr;  // [G] boundaryFaceID_to_internalCellID
  const int32_t* ghost_kind = nullptr;  // [G] BcKind
  const int32_t* ghost_bc = nullptr;    // [G] index of its inlet / exit boundary
  const double* ghost_nx = nullptr;     // [G] face_direction .* face_normals[faceID]
  const double* ghost_ny = nullptr;
  const double* ghost_len = nullptr;    // [G] face_lengths[faceID]
  int32_t n_inlet = 0, n_exit = 0;
  const int32_t* inlet_off = nullptr;    // [n_inlet+1]
  const int32_t* inlet_ghost = nullptr;  // ghosts of each inlet in NodeString order
};

struct Consts {
  double g = 9.81, k_n = 1.0, h_small = 1e-3;
};

// The RHS's parameter bundle: the fields of SWE2D_Extra_Parameters that the
// hot path reads (application_commons.jl:7-44).
template <class T>
struct Fields {
  std::vector<T> hstill, hstill_ghost, zb, zb_ghost, S0x, S0y, n;
  std::vector<T> inletQ, exitWSE;
  std::vector<T> fmag;  // FlowResistance UDE: friction magnitude per cell (empty: Manning friction)
};

template <class T>
struct Ghosts {
  std::vector<T> h, qx, qy;
};

// ----------------------------------------- swe_2D_solvers.jl:4-164 ---------
// Riemann_2D_Roe: Roe flux in the face-normal frame with the pre-balanced
// pressure 0.5 g (xi^2 + 2 xi h_still) and the five dry-bed branches.
// zb_face / S0_on_face are accepted by the reference but unused.
template <class T>
std::array<T, 3> riemann_2d_roe(const T& xiL, const T& hstillL, T hL, T huL, T hvL, const T& zb_L,
                                const T& xiR, const T& hstillR, T hR, T huR, T hvR, const T& zb_R,
                                double g, double nx, double ny, double hmin) {
  if (val(hL) <= hmin && val(hR) <= hmin) {  // both dry (:16-22)
    return {T(0.0), T(0.0), T(0.0)};
  } else if ((val(hL + zb_L) < val(zb_R + hmin)) && (val(hR) <= hmin)) {  // :23-36
    hR = hL;
    huR = -huL;
    hvR = -hvL;
  } else if ((val(hR + zb_R) < val(zb_L + hmin)) && (val(hL) <= hmin)) {  // :39-51
    hL = hR;
    huL = -huR;
    hvL = -hvR;
  } else if (val(hL) <= hmin) {  // left dry (:54-64)
    T h_flux = huR * nx + hvR * ny;
    T hu_flux = (huR * (huR / hR) + 0.5 * g * smooth_pow2(hR)) * nx + huR * (hvR / hR) * ny;
    T hv_flux = (hvR * (huR / hR)) * nx + (hvR * (hvR / hR) + 0.5 * g * smooth_pow2(hR)) * ny;
    return {h_flux, hu_flux, hv_flux};
  } else if (val(hR) <= hmin) {  // right dry (:65-75)
    T h_flux = huL * nx + hvL * ny;
    T hu_flux = (huL * (huL / hL) + 0.5 * g * smooth_pow2(hL)) * nx + huL * (hvL / hL) * ny;
    T hv_flux = (hvL * (huL / hL)) * nx + (hvL * (hvL / hL) + 0.5 * g * smooth_pow2(hL)) * ny;
    return {h_flux, hu_flux, hv_flux};
  }

  // velocities (:79-82); unL / unR / SL / SR of :85-100 are unused
  T uL = huL / hL;
  T vL = hvL / hL;
  T uR = huR / hR;
  T vR = hvR / hR;

  // Roe averages (:89-96)
  T sqrt_hL = smooth_sqrt(hL);
  T sqrt_hR = smooth_sqrt(hR);
  T hRoe = (hL + hR) / 2.0;
  T uRoe = (sqrt_hL * uL + sqrt_hR * uR) / (sqrt_hL + sqrt_hR);
  T vRoe = (sqrt_hL * vL + sqrt_hR * vR) / (sqrt_hL + sqrt_hR);
  T unRoe = uRoe * nx + vRoe * ny;
  T cRoe = smooth_sqrt(g * hRoe);
  T over_two_cRoe = 1.0 / 2.0 / cRoe;

  // R, L, |Lambda| (:103-113); StaticArrays' 3x3 mat-vec expands to
  // (a11*b1 + a12*b2) + a13*b3 with separate multiplies and adds.
  T R[3][3] = {{T(0.0), T(1.0), T(1.0)},
               {T(ny), uRoe - cRoe * nx, uRoe + cRoe * nx},
               {T(-nx), vRoe - cRoe * ny, vRoe + cRoe * ny}};
  T L[3][3] = {{-(uRoe * ny - vRoe * nx), T(ny), T(-nx)},
               {unRoe * over_two_cRoe + 0.5, -nx * over_two_cRoe, -ny * over_two_cRoe},
               {-unRoe * over_two_cRoe + 0.5, nx * over_two_cRoe, ny * over_two_cRoe}};
  T lam[3] = {smooth_abs(unRoe), smooth_abs(unRoe - cRoe), smooth_abs(unRoe + cRoe)};
  T dQ[3] = {xiR - xiL, huR - huL, hvR - hvL};  // (:115)

  T w[3], lw[3], absA_dQ[3];
  for (int r = 0; r < 3; ++r) w[r] = L[r][0] * dQ[0] + L[r][1] * dQ[1] + L[r][2] * dQ[2];
  // absLamda is a full 3x3 matrix with zero off-diagonals (:111-113)
  for (int r = 0; r < 3; ++r) {
    T z0 = (r == 0) ? lam[0] : T(0.0);
    T z1 = (r == 1) ? lam[1] : T(0.0);
    T z2 = (r == 2) ? lam[2] : T(0.0);
    lw[r] = z0 * w[0] + z1 * w[1] + z2 * w[2];
  }
  for (int r = 0; r < 3; ++r) absA_dQ[r] = R[r][0] * lw[0] + R[r][1] * lw[1] + R[r][2] * lw[2];

  // physical fluxes (:121-127)
  T xi_flux_L = huL * nx + hvL * ny;
  T hu_flux_L = (huL * uL + 0.5 * g * (smooth_pow2(xiL) + 2.0 * xiL * hstillL)) * nx + huL * vL * ny;
  T hv_flux_L = (hvL * uL) * nx + (hvL * vL + 0.5 * g * (smooth_pow2(xiL) + 2.0 * xiL * hstillL)) * ny;

  T xi_flux_R = huR * nx + hvR * ny;
  T hu_flux_R = (huR * uR + 0.5 * g * (smooth_pow2(xiR) + 2.0 * xiR * hstillR)) * nx + huR * vR * ny;
  T hv_flux_R = (hvR * uR) * nx + (hvR * vR + 0.5 * g * (smooth_pow2(xiR) + 2.0 * xiR * hstillR)) * ny;

  // Roe flux (:131-133); the bed-slope term of :149-151 is disabled
  T xi_flux = (xi_flux_L + xi_flux_R - absA_dQ[0]) / 2.0;
  T hu_flux = (hu_flux_L + hu_flux_R - absA_dQ[1]) / 2.0;
  T hv_flux = (hv_flux_L + hv_flux_R - absA_dQ[2]) / 2.0;
  (void)zb_L; (void)zb_R;
  return {xi_flux, hu_flux, hv_flux};
}

// ------------------------------------------------ bc_2D.jl:575-875 ---------
// process_all_boundaries_2d: ghost h, q_x, q_y for INLET-Q, EXIT-H, WALL,
// SYMM, written straight into ghost order (the inverse-permutation gather of
// update_1d_array, misc_tools.jl:4-18, is folded into the indexing).
// h / q_x / q_y are the clamped depth and masked discharges.
template <class T>
void process_all_boundaries_2d(const Mesh& m, const Consts& c, const std::vector<T>& h,
                               const std::vector<T>& q_x, const std::vector<T>& q_y,
                               const std::vector<T>& ManningN_cells, const std::vector<T>& zb_cells,
                               const std::vector<T>& inletQ_TotalQ, const std::vector<T>& exitH_WSE,
                               Ghosts<T>& out, std::vector<double>* inlet_total_A = nullptr) {
  out.h.assign(m.G, T(0.0));
  out.qx.assign(m.G, T(0.0));
  out.qy.assign(m.G, T(0.0));
  if (inlet_total_A) inlet_total_A->assign(m.n_inlet, 0.0);
  // INLET-Q (:640-730): conveyance split of the total discharge over the
  // boundary's faces in NodeString order.
  for (int b = 0; b < m.n_inlet; ++b) {
    const int32_t f0 = m.inlet_off[b], f1 = m.inlet_off[b + 1];
    T total_A(0.0);
    bool first = true;
    for (int32_t f = f0; f < f1; ++f) {  // sum(... for iFace ...) is a left fold
      const int32_t gi = m.inlet_ghost[f];
      const int32_t ic = m.ghost_cell[gi];
      const double dry_wet = (val(h[ic]) > c.h_small) ? 1.0 : 0.0;
      T term = dpow(m.ghost_len[gi], 5.0 / 3.0) * h[ic] / ManningN_cells[ic] * dry_wet;
      if (first) { total_A = term; first = false; } else { total_A = total_A + term; }
    }
    if (!(val(total_A) > 1e-10))
      throw std::runtime_error("Total cross-sectional conveyance for inlet-q boundary " +
                               std::to_string(b + 1) + " is not positive");
    if (inlet_total_A) (*inlet_total_A)[b] = val(total_A);
    for (int32_t f = f0; f < f1; ++f) {
      const int32_t gi = m.inlet_ghost[f];
      const int32_t ic = m.ghost_cell[gi];
      const double dry_wet = (val(h[ic]) > c.h_small) ? 1.0 : 0.0;
      T velocity_normal = inletQ_TotalQ[b] / total_A * dpow(m.ghost_len[gi], 2.0 / 3.0) / ManningN_cells[ic];
      out.h[gi] = h[ic];
      out.qx[gi] = -h[ic] * velocity_normal * m.ghost_nx[gi] * dry_wet;
      out.qy[gi] = -h[ic] * velocity_normal * m.ghost_ny[gi] * dry_wet;
    }
  }
  for (int64_t gi = 0; gi < m.G; ++gi) {
    const int32_t ic = m.ghost_cell[gi];
    switch (m.ghost_kind[gi]) {
      case BC_INLETQ:
        break;
      case BC_EXITH:  // (:748-773): depth from the exit WSE and the CELL bed
        out.h[gi] = jmax(T(c.h_small), exitH_WSE[m.ghost_bc[gi]] - zb_cells[ic]);
        out.qx[gi] = q_x[ic];
        out.qy[gi] = q_y[ic];
        break;
      case BC_WALL:  // (:777-799): both momentum components reversed
        out.h[gi] = h[ic];
        out.qx[gi] = -q_x[ic];
        out.qy[gi] = -q_y[ic];
        break;
      case BC_SYMM: {  // (:803-834): normal component reflected
        const double nx = m.ghost_nx[gi], ny = m.ghost_ny[gi];
        T v_dot_n = q_x[ic] * nx + q_y[ic] * ny;
        out.h[gi] = h[ic];
        out.qx[gi] = q_x[ic] - 2.0 * v_dot_n * nx;
        out.qy[gi] = q_y[ic] - 2.0 * v_dot_n * ny;
        break;
      }
      default:
        throw std::runtime_error("unknown boundary kind");
    }
  }
}

// ------------------------------- semi_discretize_swe_2D.jl:281-434 ---------
// compute_inviscid_fluxes: per cell, -(sum_k F(cell, nbr_k, n_k) * L_k) / A,
// every internal face evaluated from both sides, faces in local order.
template <class T>
void compute_inviscid_fluxes(const Mesh& m, const Consts& c, const std::vector<T>& xi,
                             const std::vector<T>& hstill, const std::vector<T>& h,
                             const std::vector<T>& q_x, const std::vector<T>& q_y,
                             const std::vector<T>& xi_ghost, const std::vector<T>& hstill_ghost,
                             const Ghosts<T>& gh, const std::vector<T>& zb_cells,
                             const std::vector<T>& zb_ghost, std::vector<T>& out /*3N*/) {
  const int64_t N = m.N;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    T fs[3] = {T(0.0), T(0.0), T(0.0)};  // flux_sum = zeros(3)
    for (int32_t s = m.face_off[i]; s < m.face_off[i + 1]; ++s) {
      const int32_t j = m.face_nbr[s];
      std::array<T, 3> F;
      if (j >= 0) {  // internal face (:301-309)
        F = riemann_2d_roe(xi[i], hstill[i], h[i], q_x[i], q_y[i], zb_cells[i], xi[j], hstill[j], h[j],
                           q_x[j], q_y[j], zb_cells[j], c.g, m.face_nx[s], m.face_ny[s], c.h_small);
      } else {  // boundary face: right state is the ghost (:310-328)
        const int32_t gi = ~j;
        F = riemann_2d_roe(xi[i], hstill[i], h[i], q_x[i], q_y[i], zb_cells[i], xi_ghost[gi],
                           hstill_ghost[gi], gh.h[gi], gh.qx[gi], gh.qy[gi], zb_ghost[gi], c.g,
                           m.face_nx[s], m.face_ny[s], c.h_small);
      }
      const double L = m.face_len[s];
      for (int r = 0; r < 3; ++r) fs[r] = fs[r] + F[r] * L;  // (:387)
    }
    for (int r = 0; r < 3; ++r) out[r * N + i] = -fs[r] / m.area[i];  // (:415, :437-446)
  }
}

// ------------------------------- semi_discretize_swe_2D.jl:449-559 ---------
// compute_source_terms + compute_friction_terms (Manning branch :542-547):
//   friction = g n^2 / k_n^2 / (h + h_small)^(7/3) * smooth_sqrt(q_x^2+q_y^2) * q
//   source   = (h > h_small) * (g xi S0 - friction); mass source 0.
template <class T>
inline void friction_terms(const Consts& c, const T& h, const T& q_x, const T& q_y, const T& n, T& fx,
                           T& fy) {
  fx = c.g * (n * n) / (c.k_n * c.k_n) / dpow(h + c.h_small, 7.0 / 3.0) * smooth_sqrt(q_x * q_x + q_y * q_y) * q_x;
  fy = c.g * (n * n) / (c.k_n * c.k_n) / dpow(h + c.h_small, 7.0 / 3.0) * smooth_sqrt(q_x * q_x + q_y * q_y) * q_y;
}

// FlowResistance branch (:521-531): the UDE gives the friction magnitude m,
//   mag_q = smooth_sqrt(q_x^2 + q_y^2), friction = m * q / mag_q.
// Pinned by closed forms (tests/test_lux_oracle.py): opposite to q, magnitude
// m for |q| >> sqrt(eps) and m |q| / sqrt(eps) for |q| << sqrt(eps), zero at
// q = 0 and on dry cells, linear in m.
template <class T>
inline void friction_terms_mag(const T& m, const T& q_x, const T& q_y, T& fx, T& fy) {
  const T mag_q = smooth_sqrt(q_x * q_x + q_y * q_y);
  fx = m * q_x / mag_q;
  fy = m * q_y / mag_q;
}

template <class T>
void compute_source_terms(const Mesh& m, const Consts& c, const std::vector<T>& xi, const std::vector<T>& h,
                          const std::vector<T>& q_x, const std::vector<T>& q_y, const std::vector<T>& S0x,
                          const std::vector<T>& S0y, const std::vector<T>& ManningN_cells,
                          std::vector<T>& out /*3N*/, const std::vector<T>* fmag = nullptr) {
  const int64_t N = m.N;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    T fx, fy;
    if (fmag)
      friction_terms_mag((*fmag)[i], q_x[i], q_y[i], fx, fy);
    else
      friction_terms(c, h[i], q_x[i], q_y[i], ManningN_cells[i], fx, fy);
    const bool above_small_h = val(h[i]) > c.h_small;
    out[i] = T(0.0);
    out[N + i] = bool_times(above_small_h, c.g * xi[i] * S0x[i] - fx);
    out[2 * N + i] = bool_times(above_small_h, c.g * xi[i] * S0y[i] - fy);
  }
}

// ------------------- process_bed_2D.jl:46-66 + fvm_schemes_2D.jl -----------
// update_bed_data: zb_ghost = zb of the internal cell (fvm_schemes_2D.jl:3-30),
// zb_face = mean of the face's cells or the cell on the boundary (:89-105),
// S0 = -1.0 * Green-Gauss gradient of zb_face (:133-167).
// S0_faces (process_bed_2D.jl:69-107) is not consumed by the RHS.
template <class T>
void update_bed_data(const Mesh& m, const std::vector<T>& zb_cells, std::vector<T>& zb_ghost,
                     std::vector<T>& S0x, std::vector<T>& S0y) {
  const int64_t N = m.N;
  zb_ghost.assign(m.G, T(0.0));
  for (int64_t gi = 0; gi < m.G; ++gi) zb_ghost[gi] = zb_cells[m.ghost_cell[gi]];
  S0x.assign(N, T(0.0));
  S0y.assign(N, T(0.0));
  for (int64_t i = 0; i < N; ++i) {
    T gx(0.0), gy(0.0);
    for (int32_t s = m.face_off[i]; s < m.face_off[i + 1]; ++s) {
      const int32_t j = m.face_nbr[s];
      T zf = (j >= 0) ? (zb_cells[i] + zb_cells[j]) / 2.0 : zb_cells[i];
      gx = gx + m.face_nx[s] * zf * m.face_len[s];
      gy = gy + m.face_ny[s] * zf * m.face_len[s];
    }
    S0x[i] = -1.0 * (gx / m.area[i]);
    S0y[i] = -1.0 * (gy / m.area[i]);
  }
}

// --------------------------------- semi_discretize_swe_2D.jl:18-277 --------
// swe_2d_rhs with the parameter injection of :114-199 resolved by the caller
// (Fields already holds zb / S0 / n / inletQ for the active parameter).
template <class T>
void swe_2d_rhs(const Mesh& m, const Consts& c, const Fields<T>& p, const std::vector<T>& Q,
                std::vector<T>& dQdt, std::vector<double>* inlet_total_A = nullptr) {
  const int64_t N = m.N;
  std::vector<T> xi(Q.begin(), Q.begin() + N);
  std::vector<T> h(N), q_x(N), q_y(N);
  for (int64_t i = 0; i < N; ++i) {  // (:101-106)
    T hh = xi[i] + p.hstill[i];
    const bool dry = val(hh) <= c.h_small;
    h[i] = dry ? T(c.h_small) : hh;
    q_x[i] = dry ? T(0.0) : Q[N + i];
    q_y[i] = dry ? T(0.0) : Q[2 * N + i];
  }
  // process_dry_wet_flags (:212) is result-neutral: its flags are unused.
  Ghosts<T> gh;
  process_all_boundaries_2d(m, c, h, q_x, q_y, p.n, p.zb, p.inletQ, p.exitWSE, gh, inlet_total_A);
  std::vector<T> xi_ghost(m.G);
  for (int64_t gi = 0; gi < m.G; ++gi) xi_ghost[gi] = gh.h[gi] - p.hstill_ghost[gi];  // (:220)

  std::vector<T> inv(3 * N), src(3 * N);
  compute_inviscid_fluxes(m, c, xi, p.hstill, h, q_x, q_y, xi_ghost, p.hstill_ghost, gh, p.zb, p.zb_ghost, inv);
  compute_source_terms(m, c, xi, h, q_x, q_y, p.S0x, p.S0y, p.n, src, p.fmag.empty() ? nullptr : &p.fmag);
  dQdt.resize(3 * N);
  for (int64_t k = 0; k < 3 * N; ++k) dQdt[k] = inv[k] + src[k];  // (:264)
}

// ---------------------------------- custom_ODE_solvers.jl:5-33 -------------
// custom_ODE_update_cells: Q' = Q + dt * dQdt; where Q'_xi < h_small set
// xi = h_small and q = 0 (the clamp tests xi, not h).
template <class T>
void custom_ODE_update_cells(const Mesh& m, const Consts& c, const Fields<T>& p, std::vector<T>& Q,
                             double dt) {
  const int64_t N = m.N;
  std::vector<T> dQdt;
  swe_2d_rhs(m, c, p, Q, dQdt);
  for (int64_t k = 0; k < 3 * N; ++k) Q[k] = Q[k] + dt * dQdt[k];
  for (int64_t i = 0; i < N; ++i) {
    if (val(Q[i]) < c.h_small) {
      Q[i] = T(c.h_small);
      Q[N + i] = T(0.0);
      Q[2 * N + i] = T(0.0);
    }
  }
}

// ------------------------------------------------------------ CFL dt -------
// No reference counterpart (the reference stores CFL but steps with a fixed
// dt, custom_ODE_solvers.jl:39).  Definition used by the build:
//   dt = CFL * min over wet cells of (2 A / P) / (sqrt(u^2+v^2) + sqrt(g h))
// with h the clamped depth, P the perimeter, u = q_x / h.
inline double cfl_dt(const Mesh& m, const Consts& c, const double* hstill, const double* Q, double cfl) {
  const int64_t N = m.N;
  double best = INFINITY;
  for (int64_t i = 0; i < N; ++i) {
    double h = Q[i] + hstill[i];
    if (!(h > c.h_small)) continue;
    double P = 0.0;
    for (int32_t s = m.face_off[i]; s < m.face_off[i + 1]; ++s) P = P + m.face_len[s];
    const double u = Q[N + i] / h, v = Q[2 * N + i] / h;
    const double spd = std::sqrt(u * u + v * v) + std::sqrt(c.g * h);
    const double d = (2.0 * m.area[i]) / P;
    const double dti = d / spd;
    if (dti < best) best = dti;
  }
  return cfl * best;
}

}  // namespace orc
