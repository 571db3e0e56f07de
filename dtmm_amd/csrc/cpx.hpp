// Minimal complex arithmetic for host and device code (float and double).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace dtmmx {

template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T re, T im = T(0)) {
  cpx<T> r;
  r.x = re;
  r.y = im;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a) { return mk<T>(-a.x, -a.y); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator*(cpx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator*(T s, cpx<T> a) { return mk<T>(a.x * s, a.y * s); }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> conj(cpx<T> a) { return mk<T>(a.x, -a.y); }
template <typename T>
__host__ __device__ __forceinline__ T norm2(cpx<T> a) { return a.x * a.x + a.y * a.y; }
template <typename T>
__host__ __device__ __forceinline__ bool is_zero(cpx<T> a) { return a.x == T(0) && a.y == T(0); }
template <typename T>
__host__ __device__ __forceinline__ bool eq(cpx<T> a, cpx<T> b) { return a.x == b.x && a.y == b.y; }

// a / b (Smith's algorithm, avoids overflow in |b|^2).
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cdiv(cpx<T> a, cpx<T> b) {
  if (fabs(b.y) <= fabs(b.x)) {
    T r = b.y / b.x;
    T den = b.x + b.y * r;
    return mk<T>((a.x + a.y * r) / den, (a.y - a.x * r) / den);
  } else {
    T r = b.x / b.y;
    T den = b.x * r + b.y;
    return mk<T>((a.x * r + a.y) / den, (a.y * r - a.x) / den);
  }
}

// Principal square root with the C99 branch cut: sqrt(-x + 0i) = +i sqrt(x),
// sqrt(-x - 0i) = -i sqrt(x).
template <typename T>
__host__ __device__ __forceinline__ cpx<T> csqrt(cpx<T> z) {
  T x = z.x, y = z.y;
  if (x == T(0) && y == T(0)) return mk<T>(T(0), y);
  T r = hypot(x, y);
  T t = sqrt((r + fabs(x)) * T(0.5));
  if (x >= T(0)) return mk<T>(t, y / (T(2) * t));
  return mk<T>(fabs(y) / (T(2) * t), copysign(t, y));
}

// exp(i * phase) for a complex phase: exp(-Im) * (cos Re, sin Re).
__device__ __forceinline__ cpx<float> cexpi(cpx<float> ph) {
  float s, c;
  sincosf(ph.x, &s, &c);
  float m = (ph.y == 0.f) ? 1.f : expf(-ph.y);
  return mk<float>(c * m, s * m);
}
__device__ __forceinline__ cpx<double> cexpi(cpx<double> ph) {
  double s, c;
  sincos(ph.x, &s, &c);
  double m = (ph.y == 0.) ? 1. : exp(-ph.y);
  return mk<double>(c * m, s * m);
}

inline cpx<double> cexpi_host(cpx<double> ph) {
  double m = (ph.y == 0.) ? 1. : std::exp(-ph.y);
  return mk<double>(std::cos(ph.x) * m, std::sin(ph.x) * m);
}

template <typename T, typename U>
__host__ __device__ __forceinline__ cpx<T> cast(cpx<U> a) {
  return mk<T>(T(a.x), T(a.y));
}

}  // namespace dtmmx
