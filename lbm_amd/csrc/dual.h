// Forward-mode (tangent-linear) FP64 arithmetic: a value and its directional derivative.
// Instantiating the forward kernels with this type (codec CodecTan) runs the step and carries
// d(state)/d(theta) along one input direction through the very same code, which gives the
// tangent run the adjoint is checked against.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace lbm {

struct Dual {
    double v, d;
    __host__ __device__ constexpr Dual() : v(0.0), d(0.0) {}
    __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
    __host__ __device__ constexpr Dual(double x, double y) : v(x), d(y) {}
    __host__ __device__ explicit constexpr operator double() const { return v; }
    __host__ __device__ explicit constexpr operator float() const { return (float)v; }
};

__host__ __device__ constexpr Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ constexpr Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ constexpr Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ constexpr Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ constexpr Dual operator/(Dual a, Dual b) {
    return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
__host__ __device__ inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__host__ __device__ inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
__host__ __device__ inline Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
__host__ __device__ constexpr bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ constexpr bool operator<(Dual a, Dual b) { return a.v < b.v; }
__host__ __device__ constexpr bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__host__ __device__ constexpr bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__host__ __device__ constexpr bool operator==(Dual a, Dual b) { return a.v == b.v; }
__host__ __device__ constexpr bool operator!=(Dual a, Dual b) { return a.v != b.v; }
__host__ __device__ inline Dual sqrt(Dual a) {
    const double s = std::sqrt(a.v);
    return Dual(s, a.d / (2.0 * s));
}

template <typename R>
struct is_dual {
    static constexpr bool value = false;
};
template <>
struct is_dual<Dual> {
    static constexpr bool value = true;
};

}  // namespace lbm
