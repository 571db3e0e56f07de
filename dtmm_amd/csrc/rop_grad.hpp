// Derivatives of the per-pixel layer operator data with respect to the
// director angles (theta, phi) by forward-mode differentiation.
//
// D2 carries a value and its derivatives in two directions; rop_lean_t is
// the lean per-pixel solve of berreman.hpp (rop_lean: eigenvalues alpha and
// the projectors P0, P1 of E diag(exp(i kd alpha)) E^-1) written for any
// scalar type, so that rop_lean_t<D2> with theta seeded (1, 0) and phi
// seeded (0, 1) returns alpha, P0, P1 and their theta and phi derivatives.
// The reference has no gradients; this serves the adjoint transfer.
#pragma once
#include "berreman.hpp"

namespace dtmmx {

struct D2 {
  double v, d0, d1;
  D2() = default;
  __host__ __device__ constexpr D2(double x) : v(x), d0(0.), d1(0.) {}
  __host__ __device__ constexpr D2(double x, double a, double b) : v(x), d0(a), d1(b) {}
};

__host__ __device__ inline D2 dn(double v, double d0 = 0., double d1 = 0.) { return D2(v, d0, d1); }

// The double overloads stay visible next to the D2 ones in this namespace.
using ::copysign;
using ::fabs;
using ::hypot;
using ::sincos;
using ::sqrt;

template <typename S>
__host__ __device__ inline S cst(double x);
template <>
__host__ __device__ inline double cst<double>(double x) {
  return x;
}
template <>
__host__ __device__ inline D2 cst<D2>(double x) {
  return dn(x);
}

__host__ __device__ inline D2 operator+(D2 a, D2 b) { return dn(a.v + b.v, a.d0 + b.d0, a.d1 + b.d1); }
__host__ __device__ inline D2 operator-(D2 a, D2 b) { return dn(a.v - b.v, a.d0 - b.d0, a.d1 - b.d1); }
__host__ __device__ inline D2 operator-(D2 a) { return dn(-a.v, -a.d0, -a.d1); }
__host__ __device__ inline D2 operator*(D2 a, D2 b) {
  return dn(a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1);
}
__host__ __device__ inline D2 operator/(D2 a, D2 b) {
  const double q = a.v / b.v;
  return dn(q, (a.d0 - q * b.d0) / b.v, (a.d1 - q * b.d1) / b.v);
}
__host__ __device__ inline bool operator==(D2 a, D2 b) { return a.v == b.v; }
__host__ __device__ inline bool operator<(D2 a, D2 b) { return a.v < b.v; }
__host__ __device__ inline bool operator<=(D2 a, D2 b) { return a.v <= b.v; }
__host__ __device__ inline bool operator>=(D2 a, D2 b) { return a.v >= b.v; }

__host__ __device__ inline D2 sqrt(D2 a) {
  const double s = ::sqrt(a.v);
  const double f = s > 0. ? 0.5 / s : 0.;
  return dn(s, a.d0 * f, a.d1 * f);
}
__host__ __device__ inline D2 fabs(D2 a) { return a.v < 0. ? -a : a; }
__host__ __device__ inline D2 copysign(D2 a, D2 b) {
  const bool neg = ::copysign(1., b.v) < 0.;
  const D2 m = fabs(a);
  return neg ? -m : m;
}
__host__ __device__ inline D2 hypot(D2 a, D2 b) {
  const double h = ::hypot(a.v, b.v);
  if (h == 0.) return dn(0.);
  return dn(h, (a.v * a.d0 + b.v * b.d0) / h, (a.v * a.d1 + b.v * b.d1) / h);
}
__host__ __device__ inline void sincos(D2 a, D2* s, D2* c) {
  double sv, cv;
  ::sincos(a.v, &sv, &cv);
  *s = dn(sv, cv * a.d0, cv * a.d1);
  *c = dn(cv, -sv * a.d0, -sv * a.d1);
}

__host__ __device__ inline double val(double x) { return x; }
__host__ __device__ inline double val(D2 x) { return x.v; }

template <typename S>
__host__ __device__ inline cpx<S> lift(cd z) {
  return mk<S>(cst<S>(z.x), cst<S>(z.y));
}
template <>
__host__ __device__ inline cpx<double> lift<double>(cd z) {
  return z;
}

template <typename S>
struct RopDataT {
  cpx<S> a[4];
  cpx<S> P0[2][2];
  cpx<S> P1[2][2];
  bool ok4, ok2;
  int kind;
};

template <typename S>
__host__ __device__ inline bool inv2_t(const cpx<S> a[2][2], cpx<S> o[2][2]) {
  const cpx<S> det = a[0][0] * a[1][1] - a[0][1] * a[1][0];
  if (val(det.x) == 0. && val(det.y) == 0.) {
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) o[i][j] = mk<S>(cst<S>(0.), cst<S>(0.));
    return false;
  }
  const cpx<S> id = cdiv(mk<S>(cst<S>(1.), cst<S>(0.)), det);
  o[0][0] = a[1][1] * id;
  o[0][1] = -(a[0][1] * id);
  o[1][0] = -(a[1][0] * id);
  o[1][1] = a[0][0] * id;
  return true;
}

// rop_lean (berreman.hpp) for any scalar S: the same operations in the same
// order; beta, the beam azimuth and eps are constants, theta and phi may
// carry derivatives.
template <typename S>
__host__ __device__ inline void rop_lean_t(double beta, double c0, double s0, double phi0, cd e0,
                                           cd e1, cd e2, S theta, S phi, RopDataT<S>& r) {
  const S zero = cst<S>(0.);
  auto cs = [](double x) { return mk<S>(cst<S>(x), cst<S>(0.)); };
  cpx<S> Fx[4], Fy[4];
  const int kind = material_kind(e0, e1, e2);
  r.kind = kind;
  if (kind == KIND_ISO) {
    const cpx<S> a = lift<S>(csqrt(e0 - mk<double>(beta * beta)));
    r.a[0] = a;
    r.a[1] = -a;
    r.a[2] = a;
    r.a[3] = -a;
    const double h = (val(a.x) == 0. && val(a.y) == 0.) ? 0. : 0.5;
    Fx[0] = Fx[1] = cs(h);
    Fx[2] = Fx[3] = cs(0.);
    Fy[0] = Fy[1] = cs(0.);
    Fy[2] = Fy[3] = cs(h);
  } else if (kind == KIND_UNIAXIAL) {
    S st, ct, sp, cp;
    sincos(theta, &st, &ct);
    sincos(phi - cst<S>(phi0), &sp, &cp);
    const cpx<S> eps11 = lift<S>(e0), delta = lift<S>(e2 - e0);
    if (beta == 0.) {
      const cpx<S> evs = csqrt(eps11);
      const cpx<S> u = eps11 + delta * (ct * ct);
      const cpx<S> w = eps11 * (eps11 + delta);
      const cpx<S> sq = cdiv(csqrt(u * w), u);
      r.a[0] = sq;
      r.a[1] = -sq;
      r.a[2] = evs;
      r.a[3] = -evs;
      Fx[0] = Fx[1] = mk<S>(cp, zero);
      Fy[0] = Fy[1] = mk<S>(sp, zero);
      Fx[2] = Fx[3] = evs * (-sp);
      Fy[2] = Fy[3] = evs * cp;
    } else {
      const S b = cst<S>(beta);
      const cpx<S> ev02 = eps11 - cs(beta * beta);
      const cpx<S> evs = csqrt(ev02);
      const cpx<S> u = eps11 + delta * (ct * ct);
      const S gama = b * cp;
      const cpx<S> v = delta * (gama * cst<S>(2.) * st * ct);
      const cpx<S> w = delta * (gama * gama * st * st) - eps11 * (ev02 + delta);
      const cpx<S> sq = cdiv(csqrt(v * v - u * w * cst<S>(4.)), u) * cst<S>(0.5);
      const cpx<S> vh = cdiv(v, u) * cst<S>(0.5);
      const cpx<S> evpp = -vh + sq, evpm = -vh - sq;
      r.a[0] = evpp;
      r.a[1] = evpm;
      r.a[2] = evs;
      r.a[3] = -evs;
      const S sfst = st * sp, cfst = cp * st, ctbeta = ct * b;
      const cpx<S> ctbe = cdiv(mk<S>(ctbeta, zero), eps11);
      const cpx<S> ev02cfsteps11 = cdiv(ev02 * cfst, eps11);
      Fx[0] = evpp * ctbe - ev02cfsteps11;
      Fy[0] = mk<S>(-sfst, zero);
      Fx[1] = -(evpm * ctbe) + ev02cfsteps11;
      Fy[1] = mk<S>(sfst, zero);
      Fx[2] = evs * sfst;
      Fy[2] = mk<S>(ctbeta, zero) - evs * cfst;
      Fx[3] = evs * (-sfst);
      Fy[3] = evs * cfst + mk<S>(ctbeta, zero);
    }
  } else {
    const double nan = __builtin_nan("");
    for (int i = 0; i < 4; ++i) {
      r.a[i] = mk<S>(cst<S>(nan), cst<S>(nan));
      Fx[i] = Fy[i] = mk<S>(cst<S>(nan), cst<S>(nan));
    }
  }
  const S C0 = cst<S>(c0), S0 = cst<S>(s0);
  cpx<S> Ex[4], Ey[4];
  for (int j = 0; j < 4; ++j) {
    Ex[j] = Fx[j] * C0 - Fy[j] * S0;
    Ey[j] = Fx[j] * S0 + Fy[j] * C0;
  }
  cpx<S> ep[2][2] = {{Ex[0], Ex[2]}, {Ey[0], Ey[2]}};
  cpx<S> en[2][2] = {{Ex[1], Ex[3]}, {Ey[1], Ey[3]}};
  cpx<S> ip[2][2], in[2][2];
  const bool okp = inv2_t(ep, ip);
  const bool okn = inv2_t(en, in);
  r.ok2 = okp;
  r.ok4 = okp && okn;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      r.P0[i][j] = ep[i][0] * ip[0][j];
      r.P1[i][j] = en[i][0] * in[0][j];
    }
}

// Value and theta / phi derivatives of the operator data of one pixel.
struct RopGrad {
  cd a[4], P0[2][2], P1[2][2];        // values
  cd da[2][4], dP0[2][2][2], dP1[2][2][2];  // [direction][...]
  bool ok4, ok2;
  int kind;
};

__host__ __device__ inline cd part(cpx<D2> z, int k) {
  return k == 0 ? mk<double>(z.x.d0, z.y.d0) : mk<double>(z.x.d1, z.y.d1);
}

__host__ __device__ inline void rop_grad(double beta, double c0, double s0, double phi0, cd e0,
                                         cd e1, cd e2, double theta, double phi, RopGrad& g) {
  RopDataT<D2> r;
  rop_lean_t<D2>(beta, c0, s0, phi0, e0, e1, e2, dn(theta, 1., 0.), dn(phi, 0., 1.), r);
  g.ok4 = r.ok4;
  g.ok2 = r.ok2;
  g.kind = r.kind;
  for (int i = 0; i < 4; ++i) {
    g.a[i] = mk<double>(r.a[i].x.v, r.a[i].y.v);
    for (int k = 0; k < 2; ++k) g.da[k][i] = part(r.a[i], k);
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      g.P0[i][j] = mk<double>(r.P0[i][j].x.v, r.P0[i][j].y.v);
      g.P1[i][j] = mk<double>(r.P1[i][j].x.v, r.P1[i][j].y.v);
      for (int k = 0; k < 2; ++k) {
        g.dP0[k][i][j] = part(r.P0[i][j], k);
        g.dP1[k][i][j] = part(r.P1[i][j], k);
      }
    }
}

}  // namespace dtmmx
