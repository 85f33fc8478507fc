// gkm_dual.hpp -- forward-mode dual numbers for the device math templates.
//
// Dual{v, d} carries a value and one directional derivative.  Running the
// templated WENO (gkm_math.hpp) and GKS flux (gkm_flux.hpp) on Dual gives the
// exact derivative of the same discrete operations the FP64 step evaluates:
// the tangent-linear step seeds d with the perturbation, the adjoint step
// seeds one input at a time to form vector-Jacobian products.  Comparisons act
// on the value, so branch choices (the WENO maximum, |.|) follow the forward
// evaluation; where the step is not differentiable (a tie in the WENO maximum,
// p_l = p_r in the collision time's jump term) the result is the derivative of
// the branch taken, which can differ from a central difference.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "gkm_math.hpp"

namespace gkm {

struct Dual {
    double v, d;
    __host__ __device__ Dual() : v(0), d(0) {}
    __host__ __device__ Dual(double x) : v(x), d(0) {}
    __host__ __device__ Dual(int x) : v(x), d(0) {}
    __host__ __device__ Dual(double x, double dx) : v(x), d(dx) {}
    __host__ __device__ explicit operator double() const { return v; }
};

__host__ __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__host__ __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__host__ __device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__host__ __device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__host__ __device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
    const double q = g_div(a.v, b.v);
    return {q, (a.d - q * b.d) * g_rcp(b.v)};
}
__host__ __device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__host__ __device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
__host__ __device__ __forceinline__ Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
__host__ __device__ __forceinline__ Dual& operator/=(Dual& a, Dual b) { return a = a / b; }
__host__ __device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
__host__ __device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ __forceinline__ bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__host__ __device__ __forceinline__ bool operator>=(Dual a, Dual b) { return a.v >= b.v; }

__host__ __device__ __forceinline__ Dual sqrt(Dual a)
{
    const double s = ::sqrt(a.v);
    return {s, a.d / (2.0 * s)};
}
__host__ __device__ __forceinline__ Dual exp(Dual a)
{
    const double e = ::exp(a.v);
    return {e, a.d * e};
}
// d/dx erfc(x) = -2/sqrt(pi) exp(-x^2)
__host__ __device__ __forceinline__ Dual erfc(Dual a)
{
    const double two_over_sqrt_pi = 1.1283791670955125738961589031215452;
    return {::erfc(a.v), -two_over_sqrt_pi * ::exp(-a.v * a.v) * a.d};
}
// |x| with the derivative 0 at x = 0 exactly (the central-difference value there)
__host__ __device__ __forceinline__ Dual fabs(Dual a) { return a.v < 0 ? -a : (a.v > 0 ? a : Dual(0.0, 0.0)); }

}  // namespace gkm
