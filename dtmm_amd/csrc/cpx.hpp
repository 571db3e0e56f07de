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
__host__ __device__ __forceinline__ T cabs(cpx<T> a) { return hypot(a.x, a.y); }
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

// sin and cos of a float phase: Cody-Waite reduction by pi/2 with pi/2 split
// into three floats (C1 + C2 + C3 is pi/2 to about 5e-15), each step one
// fused multiply-add, so x - k C1 - k C2 - k C3 is rounded once per step
// and the reduced argument is off by about |k| 5e-15 (3e-9 at |x| = 1e6);
// then the Taylor series of sin to r^9 and of cos to r^10 on |r| <= pi/4
// (truncation below 2e-9, under the float rounding of the result), and the
// quadrant of k.  The error against the exact sin and cos of the float x
// stays below 1e-7 for |x| <= 1e6 (dtmmx_host_sincos checks it); the float
// argument itself carries a relative rounding of 6e-8.
__host__ __device__ __forceinline__ void sincos_cw(float x, float* sp, float* cp) {
  const float C1 = 1.5703125f, C2 = 4.837512969970703125e-4f, C3 = 7.549789954891882e-8f;
  const float k = rintf(x * 0.636619772367581343f);
  float r = fmaf(-k, C1, x);
  r = fmaf(-k, C2, r);
  r = fmaf(-k, C3, r);
  const float r2 = r * r;
  float sn = fmaf(r2, 2.7557319224e-6f, -1.9841269841e-4f);  // 1/9!, -1/7!
  sn = fmaf(r2, sn, 8.3333333333e-3f);                        // 1/5!
  sn = fmaf(r2, sn, -1.6666666667e-1f);                       // -1/3!
  sn = fmaf(r2 * r, sn, r);
  float cs = fmaf(r2, -2.7557319224e-7f, 2.4801587302e-5f);   // -1/10!, 1/8!
  cs = fmaf(r2, cs, -1.3888888889e-3f);                       // -1/6!
  cs = fmaf(r2, cs, 4.1666666667e-2f);                        // 1/4!
  cs = fmaf(r2, cs, -0.5f);
  cs = fmaf(r2, cs, 1.0f);
  const int q = int(k) & 3;
  float s = (q & 1) ? cs : sn;
  float c = (q & 1) ? sn : cs;
  if (q & 2) s = -s;
  if ((q + 1) & 2) c = -c;
  *sp = s;
  *cp = c;
}

// Phase factors of the per-layer kernels (exp(i ph) for a complex phase).
__device__ __forceinline__ cpx<float> cexpi_layer(cpx<float> ph) {
  float s, c;
  sincos_cw(ph.x, &s, &c);
  float m = (ph.y == 0.f) ? 1.f : expf(-ph.y);
  return mk<float>(c * m, s * m);
}
__device__ __forceinline__ cpx<double> cexpi_layer(cpx<double> ph) { return cexpi(ph); }

// exp(i x) of a real phase in the per-layer kernels.
__device__ __forceinline__ cpx<float> cexpi_layer_real(float x) {
  float s, c;
  sincos_cw(x, &s, &c);
  return mk<float>(c, s);
}
__device__ __forceinline__ cpx<double> cexpi_layer_real(double x) {
  double s, c;
  sincos(x, &s, &c);
  return mk<double>(c, s);
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
