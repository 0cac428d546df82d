// hg_dual.hpp -- forward-mode dual numbers (one tangent direction) for the
// device JVP of the step: value v and derivative d, with the product,
// quotient and chain rules of ForwardDiff.Dual.  Comparisons look at values
// only, so branches follow the primal path (the semantics ForwardDiff gives
// the reference).
#pragma once
#include <hip/hip_runtime.h>

namespace hg {

struct Dual1 {
  double v, d;
  __host__ __device__ Dual1() : v(0.0), d(0.0) {}
  __host__ __device__ Dual1(double x) : v(x), d(0.0) {}  // NOLINT: constants carry no tangent
  __host__ __device__ Dual1(double x, double dx) : v(x), d(dx) {}
};

__host__ __device__ __forceinline__ double val(double x) { return x; }
__host__ __device__ __forceinline__ double val(const Dual1& x) { return x.v; }

__host__ __device__ __forceinline__ Dual1 operator-(const Dual1& a) { return Dual1(-a.v, -a.d); }
__host__ __device__ __forceinline__ Dual1 operator+(const Dual1& a, const Dual1& b) {
  return Dual1(a.v + b.v, a.d + b.d);
}
__host__ __device__ __forceinline__ Dual1 operator-(const Dual1& a, const Dual1& b) {
  return Dual1(a.v - b.v, a.d - b.d);
}
__host__ __device__ __forceinline__ Dual1 operator*(const Dual1& a, const Dual1& b) {
  return Dual1(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ __forceinline__ Dual1 operator/(const Dual1& a, const Dual1& b) {
  return Dual1(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
__host__ __device__ __forceinline__ Dual1 operator+(const Dual1& a, double b) { return a + Dual1(b); }
__host__ __device__ __forceinline__ Dual1 operator+(double a, const Dual1& b) { return Dual1(a) + b; }
__host__ __device__ __forceinline__ Dual1 operator-(const Dual1& a, double b) { return a - Dual1(b); }
__host__ __device__ __forceinline__ Dual1 operator-(double a, const Dual1& b) { return Dual1(a) - b; }
__host__ __device__ __forceinline__ Dual1 operator*(const Dual1& a, double b) { return a * Dual1(b); }
__host__ __device__ __forceinline__ Dual1 operator*(double a, const Dual1& b) { return Dual1(a) * b; }
__host__ __device__ __forceinline__ Dual1 operator/(const Dual1& a, double b) { return a / Dual1(b); }
__host__ __device__ __forceinline__ Dual1 operator/(double a, const Dual1& b) { return Dual1(a) / b; }

__host__ __device__ __forceinline__ bool operator<(const Dual1& a, const Dual1& b) { return a.v < b.v; }
__host__ __device__ __forceinline__ bool operator<=(const Dual1& a, const Dual1& b) { return a.v <= b.v; }
__host__ __device__ __forceinline__ bool operator>(const Dual1& a, const Dual1& b) { return a.v > b.v; }
__host__ __device__ __forceinline__ bool operator<(const Dual1& a, double b) { return a.v < b; }
__host__ __device__ __forceinline__ bool operator<=(const Dual1& a, double b) { return a.v <= b; }
__host__ __device__ __forceinline__ bool operator>(const Dual1& a, double b) { return a.v > b; }

__host__ __device__ __forceinline__ Dual1 sqrt(const Dual1& a) {
  const double r = ::sqrt(a.v);
  return Dual1(r, a.d / (2.0 * r));
}
__host__ __device__ __forceinline__ Dual1 cbrt(const Dual1& a) {
  const double r = ::cbrt(a.v);
  return Dual1(r, a.d * (r / (3.0 * a.v)));
}

}  // namespace hg
