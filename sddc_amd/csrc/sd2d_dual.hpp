// sd2d_dual.hpp - forward-mode dual numbers for the local Jacobians of the
// adjoint.  The pointwise physics (iflux_n, rusanov, bc_ghost_p) is written
// once as templates over the arithmetic type; instantiated with Dual, one
// evaluation gives the derivative of every output along one input
// direction, which the hand-written reverse sweeps of sd2d_adj.hpp contract
// with the output cotangents.  Comparisons (maxima, branch selections)
// act on the values, so the derivative is that of the branch taken, as
// for the reference's own arithmetic.
#pragma once
#include "sd2d_device.hpp"

namespace sd2d {

// the library's own double / float overloads stay visible next to the
// Dual ones inside this namespace
using ::fabs;
using ::fmax;
using ::sqrt;

struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0.0), d(0.0) {}
  __host__ __device__ Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ Dual(double x, double dx) : v(x), d(dx) {}
};

__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b) {
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline Dual &operator+=(Dual &a, Dual b) { a = a + b; return a; }
__host__ __device__ inline Dual &operator-=(Dual &a, Dual b) { a = a - b; return a; }
__host__ __device__ inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
__host__ __device__ inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__host__ __device__ inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }

__host__ __device__ inline Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return Dual(s, a.d / (2.0 * s));
}
__host__ __device__ inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
__host__ __device__ inline Dual fmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }
__device__ inline Dual rcp(Dual a) { return Dual(1.0) / a; }

}  // namespace sd2d
