// Dual numbers a + b eps (eps^2 = 0) for the derivative kernels: one
// derivative direction carried beside every value.  Comparisons and branch
// decisions use the value part, so a dual evaluation takes the same path as
// the plain evaluation it differentiates.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace pimd {

struct Dual {
  double v, d;
  Dual() = default;
  __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
};

__device__ __forceinline__ Dual dmk(double v, double d = 0.0) { return Dual{v, d}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return Dual{a.v * b.v, a.d * b.v + a.v * b.d};
}
__device__ __forceinline__ Dual operator*(double s, Dual a) { return Dual{s * a.v, s * a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double s) { return Dual{s * a.v, s * a.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator/(Dual a, double s) { return Dual{a.v / s, a.d / s}; }
__device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) {
  a.v += b.v;
  a.d += b.d;
  return a;
}
__device__ __forceinline__ Dual &operator-=(Dual &a, Dual b) {
  a.v -= b.v;
  a.d -= b.d;
  return a;
}
__device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
__device__ __forceinline__ bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__device__ __forceinline__ Dual dabs(Dual a) { return a.v < 0.0 ? -a : a; }

__device__ __forceinline__ Dual dexp(Dual a) {
  const double e = exp(a.v);
  return Dual{e, e * a.d};
}
__device__ __forceinline__ Dual dlog(Dual a) { return Dual{log(a.v), a.d / a.v}; }
__device__ __forceinline__ Dual dsqrt(Dual a) {
  const double r = sqrt(a.v);
  return Dual{r, 0.5 * a.d / r};
}
// a^p for a > 0 and a constant exponent p
__device__ __forceinline__ Dual dpow(Dual a, double p) {
  const double r = pow(a.v, p);
  return Dual{r, p * r / a.v * a.d};
}

}  // namespace pimd
