// Per-pixel Berreman 4x4 eigen-solution in double precision.
//
// Field vector order is (Ex, Hy, Ey, Hx); mode columns are interleaved
// (+, -, +, -).  The isotropic and uniaxial closed forms restate
// dtmm/tmm.py:123-288 (reference); the beam-azimuth rotation of the rows
// restates dtmm/linalg.py:674-713 (_dotr2m).
#pragma once
#include "cpx.hpp"

namespace dtmmx {

using cd = cpx<double>;

struct Modes {
  cd a[4];     // eigenvalues alpha
  cd F[4][4];  // field matrix, columns = modes
};

enum MaterialKind { KIND_ISO = 0, KIND_UNIAXIAL = 1, KIND_BIAXIAL = 2 };

// Exact-equality branch selection of dtmm/tmm.py:384-389.
__host__ __device__ inline int material_kind(cd e0, cd e1, cd e2) {
  if (eq(e0, e1) && eq(e1, e2)) return KIND_ISO;
  if (eq(e0, e1)) return KIND_UNIAXIAL;
  return KIND_BIAXIAL;
}

// Isotropic medium, columns not normalised; all zero where sqrt(eps - beta^2)
// vanishes (dtmm/tmm.py:123-154).
__host__ __device__ inline void iso_modes(double beta, cd eps, Modes& m) {
  cd aout = csqrt(eps - mk<double>(beta * beta));
  for (int i = 0; i < 4; ++i) {
    m.a[i] = mk<double>(0.);
    for (int j = 0; j < 4; ++j) m.F[i][j] = mk<double>(0.);
  }
  if (is_zero(aout)) return;
  cd gp = cdiv(eps, aout);
  cd gs = -aout;
  m.a[0] = aout;
  m.a[1] = -aout;
  m.a[2] = aout;
  m.a[3] = -aout;
  m.F[0][0] = mk<double>(0.5);
  m.F[0][1] = mk<double>(0.5);
  m.F[1][0] = gp * 0.5;
  m.F[1][1] = gp * -0.5;
  m.F[2][2] = mk<double>(0.5);
  m.F[2][3] = mk<double>(0.5);
  m.F[3][2] = gs * 0.5;
  m.F[3][3] = gs * -0.5;
}

// Uniaxial medium (ordinary eps_o, extraordinary eps_e, optic axis at polar
// angle theta and azimuth phi relative to the plane of incidence), columns
// normalised to unit 2-norm (dtmm/tmm.py:156-288 with the rotation matrix of
// dtmm/rotation.py:176-193).
__host__ __device__ inline void uniaxial_modes(double beta, cd eps11, cd eps33, double theta,
                                               double phi, Modes& m) {
  const double ct = cos(theta), st = sin(theta), cf = cos(phi), sf = sin(phi);
  const double sfst = st * sf, cfst = cf * st;
  cd delta = eps33 - eps11;
  cd evpp, evpm, evs, ev02;
  if (beta == 0.) {
    ev02 = eps11;
    evs = csqrt(ev02);
    cd u = eps11 + delta * (ct * ct);
    cd w = eps11 * (ev02 + delta);
    cd sq = cdiv(csqrt(u * w), u);
    evpp = sq;
    evpm = -sq;
  } else {
    ev02 = eps11 - mk<double>(beta * beta);
    evs = csqrt(ev02);
    cd u = eps11 + delta * (ct * ct);
    double gama = beta * cf;
    cd v = delta * (gama * 2. * st * ct);
    cd w = delta * (gama * gama * st * st) - eps11 * (ev02 + delta);
    cd sq = cdiv(csqrt(v * v - u * w * 4.), u) * 0.5;
    v = cdiv(v, u) * 0.5;
    evpp = -v + sq;
    evpm = -v - sq;
  }
  m.a[0] = evpp;
  m.a[1] = evpm;
  m.a[2] = evs;
  m.a[3] = -evs;
  if (beta == 0.) {
    m.F[0][2] = evs * (-sf);
    m.F[1][2] = eps11 * (-sf);
    m.F[2][2] = evs * cf;
    m.F[3][2] = eps11 * (-cf);
    m.F[0][3] = evs * (-sf);
    m.F[1][3] = eps11 * sf;
    m.F[2][3] = evs * cf;
    m.F[3][3] = eps11 * cf;
    m.F[0][0] = mk<double>(cf);
    m.F[1][0] = evpp * cf;
    m.F[2][0] = mk<double>(sf);
    m.F[3][0] = evpp * (-sf);
    m.F[0][1] = mk<double>(cf);
    m.F[1][1] = evpm * cf;
    m.F[2][1] = mk<double>(sf);
    m.F[3][1] = evpm * (-sf);
  } else {
    const double ctbeta = ct * beta;
    cd ctbetaeps11 = cdiv(mk<double>(ctbeta), eps11);
    cd e11sfst = eps11 * sfst;
    cd evssfst = evs * sfst;
    cd evscfst = evs * cfst;
    cd evsctbeta = evs * ctbeta;
    cd ev02cfst = ev02 * cfst;
    cd ev02cfsteps11 = cdiv(ev02cfst, eps11);
    m.F[0][2] = evssfst;
    m.F[1][2] = e11sfst;
    m.F[2][2] = mk<double>(ctbeta) - evscfst;
    m.F[3][2] = ev02cfst - evsctbeta;
    m.F[0][3] = -evssfst;
    m.F[1][3] = e11sfst;
    m.F[2][3] = evscfst + mk<double>(ctbeta);
    m.F[3][3] = ev02cfst + evsctbeta;
    m.F[0][0] = evpp * ctbetaeps11 - ev02cfsteps11;
    m.F[1][0] = mk<double>(ctbeta) - evpp * cfst;
    m.F[2][0] = mk<double>(-sfst);
    m.F[3][0] = evpp * sfst;
    m.F[0][1] = -(evpm * ctbetaeps11) + ev02cfsteps11;
    m.F[1][1] = evpm * cfst - mk<double>(ctbeta);
    m.F[2][1] = mk<double>(sfst);
    m.F[3][1] = -(evpm * sfst);
  }
  for (int j = 0; j < 4; ++j) {
    double t = 0.;
    for (int i = 0; i < 4; ++i) t += norm2(m.F[i][j]);
    t = sqrt(t);
    for (int i = 0; i < 4; ++i) m.F[i][j] = mk<double>(m.F[i][j].x / t, m.F[i][j].y / t);
  }
}

// Givens rotation G = [c s; -conj(s) c] (c real) with G [a; b] = [r; 0].
__host__ __device__ inline void givens(cd a, cd b, double& c, cd& s) {
  const double aa = cabs(a), ab = cabs(b);
  if (ab == 0.) {
    c = 1.;
    s = mk<double>(0.);
    return;
  }
  if (aa == 0.) {
    c = 0.;
    s = mk<double>(1.);
    return;
  }
  const double nrm = hypot(aa, ab);
  c = aa / nrm;
  s = (a * (1. / aa)) * conj(b) * (1. / nrm);
}

// Eigenvalues and eigenvectors of a general complex 4x4 matrix A (destroyed):
// Householder reduction to Hessenberg form, single-shift complex QR with
// Wilkinson shifts to the Schur form T = Z^H A Z, eigenvectors of T by
// back-substitution, mapped by Z and scaled to unit 2-norm with the largest
// component real and positive (the normalisation of LAPACK zgeev, which
// numpy.linalg.eig calls).  lam[k] is the k-th diagonal entry of T and V[:,k]
// its eigenvector.  Returns false if the iteration does not converge.
__host__ __device__ inline bool eig4(cd A[4][4], cd lam[4], cd V[4][4]) {
  cd Z[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) Z[i][j] = mk<double>(i == j ? 1. : 0.);
  for (int k = 0; k < 2; ++k) {
    cd v[3];
    double xn = 0.;
    for (int i = k + 1; i < 4; ++i) xn += norm2(A[i][k]);
    xn = sqrt(xn);
    const double a0 = cabs(A[k + 1][k]);
    if (xn == 0. || xn == a0) continue;  // column already reduced
    const cd ph = a0 > 0. ? A[k + 1][k] * (1. / a0) : mk<double>(1.);
    const int m = 3 - k;
    for (int i = 0; i < m; ++i) v[i] = A[k + 1 + i][k];
    v[0] = v[0] + ph * xn;
    double vn = 0.;
    for (int i = 0; i < m; ++i) vn += norm2(v[i]);
    const double f = 2. / vn;
    for (int j = 0; j < 4; ++j) {  // A <- H A
      cd t = mk<double>(0.);
      for (int i = 0; i < m; ++i) t = t + conj(v[i]) * A[k + 1 + i][j];
      t = t * f;
      for (int i = 0; i < m; ++i) A[k + 1 + i][j] = A[k + 1 + i][j] - v[i] * t;
    }
    for (int i = 0; i < 4; ++i) {  // A <- A H, Z <- Z H
      cd t = mk<double>(0.), u = mk<double>(0.);
      for (int j = 0; j < m; ++j) {
        t = t + A[i][k + 1 + j] * v[j];
        u = u + Z[i][k + 1 + j] * v[j];
      }
      t = t * f;
      u = u * f;
      for (int j = 0; j < m; ++j) {
        A[i][k + 1 + j] = A[i][k + 1 + j] - t * conj(v[j]);
        Z[i][k + 1 + j] = Z[i][k + 1 + j] - u * conj(v[j]);
      }
    }
    for (int i = k + 2; i < 4; ++i) A[i][k] = mk<double>(0.);
  }
  double hn = 0.;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) hn += cabs(A[i][j]);
  if (hn == 0.) hn = 1.;
  const double eps = 2.220446049250313e-16;
  int hi = 3, it = 0, total = 0;
  while (hi > 0) {
    int l = hi;
    for (; l > 0; --l) {
      double sc = cabs(A[l - 1][l - 1]) + cabs(A[l][l]);
      if (sc == 0.) sc = hn;
      if (cabs(A[l][l - 1]) <= eps * sc) {
        A[l][l - 1] = mk<double>(0.);
        break;
      }
    }
    if (l == hi) {
      --hi;
      it = 0;
      continue;
    }
    if (++total > 200) return false;
    cd mu;
    if (it == 10 || it == 20) {
      mu = A[hi][hi] + mk<double>(0.75 * cabs(A[hi][hi - 1]));  // exceptional shift
    } else {
      const cd a = A[hi - 1][hi - 1], b = A[hi - 1][hi], c = A[hi][hi - 1], d = A[hi][hi];
      const cd hd = (a - d) * 0.5;
      const cd disc = csqrt(hd * hd + b * c);
      const cd m1 = (a + d) * 0.5 + disc, m2 = (a + d) * 0.5 - disc;
      mu = cabs(m1 - d) <= cabs(m2 - d) ? m1 : m2;
    }
    ++it;
    double gc[3];
    cd gs[3];
    for (int k = l; k <= hi; ++k) A[k][k] = A[k][k] - mu;
    for (int k = l; k < hi; ++k) {
      givens(A[k][k], A[k + 1][k], gc[k], gs[k]);
      for (int j = k; j < 4; ++j) {
        const cd x = A[k][j], y = A[k + 1][j];
        A[k][j] = x * gc[k] + gs[k] * y;
        A[k + 1][j] = y * gc[k] - conj(gs[k]) * x;
      }
    }
    for (int k = l; k < hi; ++k) {
      const int r = k + 2 < hi ? k + 2 : hi;
      for (int i = 0; i <= r; ++i) {
        const cd x = A[i][k], y = A[i][k + 1];
        A[i][k] = x * gc[k] + y * conj(gs[k]);
        A[i][k + 1] = y * gc[k] - x * gs[k];
      }
      for (int i = 0; i < 4; ++i) {
        const cd x = Z[i][k], y = Z[i][k + 1];
        Z[i][k] = x * gc[k] + y * conj(gs[k]);
        Z[i][k + 1] = y * gc[k] - x * gs[k];
      }
    }
    for (int k = l; k <= hi; ++k) A[k][k] = A[k][k] + mu;
  }
  const double smin = eps * hn > 1e-300 ? eps * hn : 1e-300;
  for (int k = 0; k < 4; ++k) {
    lam[k] = A[k][k];
    cd x[4];
    for (int i = 0; i < 4; ++i) x[i] = mk<double>(i == k ? 1. : 0.);
    for (int i = k - 1; i >= 0; --i) {
      cd t = mk<double>(0.);
      for (int j = i + 1; j <= k; ++j) t = t + A[i][j] * x[j];
      cd den = A[i][i] - A[k][k];
      if (cabs(den) < smin) den = mk<double>(smin);
      x[i] = -cdiv(t, den);
    }
    cd v[4];
    double vn = 0.;
    int im = 0;
    double vm = -1.;
    for (int i = 0; i < 4; ++i) {
      cd t = mk<double>(0.);
      for (int j = 0; j <= k; ++j) t = t + Z[i][j] * x[j];
      v[i] = t;
      vn += norm2(t);
      if (cabs(t) > vm) {
        vm = cabs(t);
        im = i;
      }
    }
    const cd ph = conj(v[im]) * (1. / (vm * sqrt(vn)));
    for (int i = 0; i < 4; ++i) V[i][k] = v[i] * ph;
  }
  return true;
}

// Biaxial medium (dtmm/tmm.py:395-403): the principal tensor diag(e0, e1, e2)
// rotated by R = R_phi R_theta R_psi (dtmm/rotation.py:125-150, :285-292),
// the Berreman matrix of the rotated tensor (dtmm/tmm.py:99-121), its
// eigen-decomposition (eig4) and the Poynting sort of dtmm/tmm.py:299-369:
// in eigenvalue order, modes with Re(Ex Hy*) - Re(Ey Hx*) >= 0 fill columns
// 0, 2 and the others columns 1, 3; if a direction has more than two modes
// (or the iteration fails) alpha is NaN and F zero.  Columns are not rotated
// by the beam azimuth here.  Returns false on failure.
__host__ __device__ inline bool biaxial_modes(double beta, cd e0, cd e1, cd e2, double psi,
                                              double theta, double phi, Modes& m) {
  double sp, cp, st, ct, sf, cf;
  sincos(psi, &sp, &cp);
  sincos(theta, &st, &ct);
  sincos(phi, &sf, &cf);
  const double R[3][3] = {{ct * cf * cp - sf * sp, -ct * cf * sp - sf * cp, cf * st},
                          {ct * sf * cp + cf * sp, cf * cp - ct * sf * sp, st * sf},
                          {-cp * st, st * sp, ct}};
  const cd d[3] = {e0, e1, e2};
  cd ep[6];
  const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  for (int q = 0; q < 6; ++q) {
    cd t = mk<double>(0.);
    for (int k = 0; k < 3; ++k) t = t + d[k] * (R[pr[q][0]][k] * R[pr[q][1]][k]);
    ep[q] = t;
  }
  const cd e2m = cdiv(mk<double>(1.), ep[2]);
  const cd e4 = ep[4] * e2m, e5 = ep[5] * e2m;
  const cd z = mk<double>(0.);
  const cd l00 = e4 * (-beta), l01 = mk<double>(1.) - e2m * (beta * beta), l02 = e5 * (-beta);
  const cd l10 = ep[0] - ep[4] * e4, l12 = ep[3] - ep[5] * e4;
  const cd l32 = mk<double>(beta * beta) + ep[5] * e5 - ep[1];
  cd L[4][4] = {{l00, l01, l02, z}, {l10, l00, l12, z}, {z, z, z, mk<double>(-1.)}, {-l12, -l02, l32, z}};
  cd lam[4], V[4][4];
  bool ok = eig4(L, lam, V);
  int i = 0, j = 1;
  for (int k = 0; ok && k < 4; ++k) {
    const double p = (V[0][k] * conj(V[1][k])).x - (V[2][k] * conj(V[3][k])).x;
    int& slot = p >= 0. ? i : j;
    if (slot >= 4) {
      ok = false;
      break;
    }
    m.a[slot] = lam[k];
    for (int r = 0; r < 4; ++r) m.F[r][slot] = V[r][k];
    slot += 2;
  }
  if (!ok) {
    const double nan = __builtin_nan("");
    for (int c = 0; c < 4; ++c) {
      m.a[c] = mk<double>(nan, nan);
      for (int r = 0; r < 4; ++r) m.F[r][c] = mk<double>(0.);
    }
  }
  return ok;
}

// Rotate the (Ex, Hy, Ey, Hx) rows by the beam azimuth (c, s) = (cos, sin).
__host__ __device__ inline void rotate_rows(Modes& m, double c, double s) {
  for (int j = 0; j < 4; ++j) {
    cd f0 = m.F[0][j], f1 = m.F[1][j], f2 = m.F[2][j], f3 = m.F[3][j];
    m.F[0][j] = f0 * c - f2 * s;
    m.F[1][j] = f1 * c + f3 * s;
    m.F[2][j] = f0 * s + f2 * c;
    m.F[3][j] = f3 * c - f1 * s;
  }
}

// 2x2 inverse; zeros (and false) where the determinant is exactly zero
// (dtmm/linalg.py:428-457).
__host__ __device__ inline bool inv2(const cd a[2][2], cd o[2][2]) {
  cd det = a[0][0] * a[1][1] - a[0][1] * a[1][0];
  if (is_zero(det)) {
    o[0][0] = o[0][1] = o[1][0] = o[1][1] = mk<double>(0.);
    return false;
  }
  cd id = cdiv(mk<double>(1.), det);
  o[0][0] = a[1][1] * id;
  o[0][1] = -(a[0][1] * id);
  o[1][0] = -(a[1][0] * id);
  o[1][1] = a[0][0] * id;
  return true;
}

// 4x4 inverse by cofactors; zeros where the determinant is exactly zero
// (dtmm/linalg.py:460-518).
__host__ __device__ inline bool inv4(const cd m[4][4], cd o[4][4]) {
  // 2x2 sub-determinants of rows (0,1) and rows (2,3)
  cd s0 = m[0][0] * m[1][1] - m[1][0] * m[0][1];
  cd s1 = m[0][0] * m[1][2] - m[1][0] * m[0][2];
  cd s2 = m[0][0] * m[1][3] - m[1][0] * m[0][3];
  cd s3 = m[0][1] * m[1][2] - m[1][1] * m[0][2];
  cd s4 = m[0][1] * m[1][3] - m[1][1] * m[0][3];
  cd s5 = m[0][2] * m[1][3] - m[1][2] * m[0][3];
  cd c5 = m[2][2] * m[3][3] - m[3][2] * m[2][3];
  cd c4 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
  cd c3 = m[2][1] * m[3][2] - m[3][1] * m[2][2];
  cd c2 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
  cd c1 = m[2][0] * m[3][2] - m[3][0] * m[2][2];
  cd c0 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
  cd det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  if (is_zero(det)) {
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) o[i][j] = mk<double>(0.);
    return false;
  }
  cd id = cdiv(mk<double>(1.), det);
  o[0][0] = (m[1][1] * c5 - m[1][2] * c4 + m[1][3] * c3) * id;
  o[0][1] = (-(m[0][1] * c5) + m[0][2] * c4 - m[0][3] * c3) * id;
  o[0][2] = (m[3][1] * s5 - m[3][2] * s4 + m[3][3] * s3) * id;
  o[0][3] = (-(m[2][1] * s5) + m[2][2] * s4 - m[2][3] * s3) * id;
  o[1][0] = (-(m[1][0] * c5) + m[1][2] * c2 - m[1][3] * c1) * id;
  o[1][1] = (m[0][0] * c5 - m[0][2] * c2 + m[0][3] * c1) * id;
  o[1][2] = (-(m[3][0] * s5) + m[3][2] * s2 - m[3][3] * s1) * id;
  o[1][3] = (m[2][0] * s5 - m[2][2] * s2 + m[2][3] * s1) * id;
  o[2][0] = (m[1][0] * c4 - m[1][1] * c2 + m[1][3] * c0) * id;
  o[2][1] = (-(m[0][0] * c4) + m[0][1] * c2 - m[0][3] * c0) * id;
  o[2][2] = (m[3][0] * s4 - m[3][1] * s2 + m[3][3] * s0) * id;
  o[2][3] = (-(m[2][0] * s4) + m[2][1] * s2 - m[2][3] * s0) * id;
  o[3][0] = (-(m[1][0] * c3) + m[1][1] * c1 - m[1][2] * c0) * id;
  o[3][1] = (m[0][0] * c3 - m[0][1] * c1 + m[0][2] * c0) * id;
  o[3][2] = (-(m[3][0] * s3) + m[3][1] * s1 - m[3][2] * s0) * id;
  o[3][3] = (m[2][0] * s3 - m[2][1] * s1 + m[2][2] * s0) * id;
  return true;
}

__host__ __device__ inline void matmul4(const cd a[4][4], const cd b[4][4], cd o[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      cd s = mk<double>(0.);
      for (int k = 0; k < 4; ++k) s = s + a[i][k] * b[k][j];
      o[i][j] = s;
    }
}

// Eigen data of one pixel for the real-space layer operator
// g <- E diag(exp(i kd alpha)) E^-1 g written in the Epn basis
// (E+x, E-x, E+y, E-y).  With E+ = F[(0,2),(0,2)] and E- = F[(0,2),(1,3)],
//   E+ diag(p0, p2) E+^-1 = p2 I + (p0 - p2) P0,  P0 = E+[:,0] (x) E+^-1[0,:]
//   E- diag(p1, p3) E-^-1 = p3 I + (p1 - p3) P1,  P1 = E-[:,0] (x) E-^-1[0,:]
// ok4: E = blockdiag(E+, E-) is invertible (4x4 method, dtmm/tmm.py:974-997
// with linalg.inv); ok2: E+ is invertible (2x2 method, mode = +1); okb: E-
// is invertible (2x2 method, mode = -1).
struct RopData {
  cd a[4];
  cd P0[2][2];
  cd P1[2][2];
  bool ok4, ok2, okb;
  int kind;
};

__host__ __device__ inline void layer_modes(double beta, double c0, double s0, double phi0, cd e0,
                                            cd e1, cd e2, double psi, double theta, double phi,
                                            Modes& m, int& kind) {
  kind = material_kind(e0, e1, e2);
  if (kind == KIND_ISO) {
    iso_modes(beta, e0, m);
  } else if (kind == KIND_UNIAXIAL) {
    uniaxial_modes(beta, e0, e2, theta, phi - phi0, m);
  } else {
    biaxial_modes(beta, e0, e1, e2, psi, theta, phi - phi0, m);
  }
  rotate_rows(m, c0, s0);
}

__host__ __device__ inline void rop_data(const Modes& m, int kind, RopData& r) {
  r.kind = kind;
  for (int i = 0; i < 4; ++i) r.a[i] = m.a[i];
  cd ep[2][2] = {{m.F[0][0], m.F[0][2]}, {m.F[2][0], m.F[2][2]}};
  cd en[2][2] = {{m.F[0][1], m.F[0][3]}, {m.F[2][1], m.F[2][3]}};
  cd ip[2][2], in[2][2];
  bool okp = inv2(ep, ip);
  bool okn = inv2(en, in);
  r.ok2 = okp;
  r.okb = okn;
  r.ok4 = okp && okn;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      r.P0[i][j] = ep[i][0] * ip[0][j];
      r.P1[i][j] = en[i][0] * in[0][j];
    }
}

}  // namespace dtmmx

namespace dtmmx {

// Lean per-pixel solve for the real-space layer operator.  Only the E rows
// (Ex, Ey) of the field matrix enter E diag(p) E^-1, and that product is
// invariant to the scale of each mode column, so the column normalisation of
// dtmm/tmm.py:278-288 and the H rows are skipped.  Same eigenvalues and the
// same projectors as layer_modes + rop_data (checked against the complex128
// oracle by tests/test_abi.py through dtmmx_host_rop).
// rop_lean in real fp64 arithmetic for a lossless (real eps) isotropic or
// uniaxial voxel whose square-root arguments are all non-negative
// (propagating modes): the same closed forms without complex square roots
// and divisions.  Returns false, with r untouched, where that does not apply.
__host__ __device__ inline bool rop_lean_real(double beta, double c0, double s0, double phi0, cd e0c,
                                              cd e1c, cd e2c, double theta, double phi, RopData& r) {
  if (e0c.y != 0. || e1c.y != 0. || e2c.y != 0. || e0c.x != e1c.x) return false;
  const double eps11 = e0c.x, e2 = e2c.x;
  double a[4], Fx[4], Fy[4];
  int kind;
  if (e2 == eps11) {
    kind = KIND_ISO;
    const double x = eps11 - beta * beta;
    if (!(x >= 0.)) return false;
    const double q = sqrt(x);
    a[0] = q;
    a[1] = -q;
    a[2] = q;
    a[3] = -q;
    const double h = q == 0. ? 0. : 0.5;
    Fx[0] = Fx[1] = h;
    Fx[2] = Fx[3] = 0.;
    Fy[0] = Fy[1] = 0.;
    Fy[2] = Fy[3] = h;
  } else {
    kind = KIND_UNIAXIAL;
    double st, ct, sp, cp;
    sincos(theta, &st, &ct);
    sincos(phi - phi0, &sp, &cp);
    const double delta = e2 - eps11;
    const double u = eps11 + delta * (ct * ct);
    if (beta == 0.) {
      const double w = eps11 * (eps11 + delta);
      if (!(eps11 >= 0.) || u == 0. || !(u * w >= 0.)) return false;
      const double evs = sqrt(eps11);
      const double sq = sqrt(u * w) / u;
      a[0] = sq;
      a[1] = -sq;
      a[2] = evs;
      a[3] = -evs;
      Fx[0] = Fx[1] = cp;
      Fy[0] = Fy[1] = sp;
      Fx[2] = Fx[3] = -evs * sp;
      Fy[2] = Fy[3] = evs * cp;
    } else {
      const double ev02 = eps11 - beta * beta;
      const double gama = beta * cp;
      const double v = delta * (gama * 2. * st * ct);
      const double w = delta * (gama * gama * st * st) - eps11 * (ev02 + delta);
      const double disc = v * v - u * w * 4.;
      if (!(ev02 >= 0.) || u == 0. || eps11 == 0. || !(disc >= 0.)) return false;
      const double evs = sqrt(ev02);
      const double sq = sqrt(disc) / u * 0.5;
      const double vh = v / u * 0.5;
      const double evpp = -vh + sq, evpm = -vh - sq;
      a[0] = evpp;
      a[1] = evpm;
      a[2] = evs;
      a[3] = -evs;
      const double sfst = st * sp, cfst = cp * st, ctbeta = ct * beta;
      const double ctbe = ctbeta / eps11;
      const double ev02cfsteps11 = ev02 * cfst / eps11;
      Fx[0] = evpp * ctbe - ev02cfsteps11;
      Fy[0] = -sfst;
      Fx[1] = -(evpm * ctbe) + ev02cfsteps11;
      Fy[1] = sfst;
      Fx[2] = evs * sfst;
      Fy[2] = ctbeta - evs * cfst;
      Fx[3] = -(evs * sfst);
      Fy[3] = evs * cfst + ctbeta;
    }
  }
  // beam-azimuth rotation of the E rows and the rank-1 projectors of the
  // forward (ep) and backward (en) 2x2 blocks
  double Ex[4], Ey[4];
  for (int j = 0; j < 4; ++j) {
    Ex[j] = Fx[j] * c0 - Fy[j] * s0;
    Ey[j] = Fx[j] * s0 + Fy[j] * c0;
  }
  const double dp = Ex[0] * Ey[2] - Ex[2] * Ey[0];
  const double dn = Ex[1] * Ey[3] - Ex[3] * Ey[1];
  const bool okp = dp != 0., okn = dn != 0.;
  const double ip = okp ? 1. / dp : 0., in = okn ? 1. / dn : 0.;
  // first rows of the inverses: (Ey2, -Ex2) / dp and (Ey3, -Ex3) / dn
  const double ip0 = Ey[2] * ip, ip1 = -(Ex[2] * ip);
  const double in0 = Ey[3] * in, in1 = -(Ex[3] * in);
  r.kind = kind;
  for (int i = 0; i < 4; ++i) r.a[i] = mk<double>(a[i]);
  r.P0[0][0] = mk<double>(Ex[0] * ip0);
  r.P0[0][1] = mk<double>(Ex[0] * ip1);
  r.P0[1][0] = mk<double>(Ey[0] * ip0);
  r.P0[1][1] = mk<double>(Ey[0] * ip1);
  r.P1[0][0] = mk<double>(Ex[1] * in0);
  r.P1[0][1] = mk<double>(Ex[1] * in1);
  r.P1[1][0] = mk<double>(Ey[1] * in0);
  r.P1[1][1] = mk<double>(Ey[1] * in1);
  r.ok2 = okp;
  r.okb = okn;
  r.ok4 = okp && okn;
  return true;
}

__host__ __device__ inline void rop_lean(double beta, double c0, double s0, double phi0, cd e0,
                                         cd e1, cd e2, double theta, double phi, RopData& r) {
  if (rop_lean_real(beta, c0, s0, phi0, e0, e1, e2, theta, phi, r)) return;
  cd Fx[4], Fy[4];  // rows Ex and Ey of F before the beam-azimuth rotation
  const int kind = material_kind(e0, e1, e2);
  r.kind = kind;
  if (kind == KIND_ISO) {
    const cd a = csqrt(e0 - mk<double>(beta * beta));
    r.a[0] = a;
    r.a[1] = -a;
    r.a[2] = a;
    r.a[3] = -a;
    const double h = is_zero(a) ? 0. : 0.5;
    Fx[0] = Fx[1] = mk<double>(h);
    Fx[2] = Fx[3] = mk<double>(0.);
    Fy[0] = Fy[1] = mk<double>(0.);
    Fy[2] = Fy[3] = mk<double>(h);
  } else if (kind == KIND_UNIAXIAL) {
    double st, ct, sp, cp;
    sincos(theta, &st, &ct);
    sincos(phi - phi0, &sp, &cp);
    const cd eps11 = e0, delta = e2 - e0;
    if (beta == 0.) {
      const cd evs = csqrt(eps11);
      const cd u = eps11 + delta * (ct * ct);
      const cd w = eps11 * (eps11 + delta);
      const cd sq = cdiv(csqrt(u * w), u);
      r.a[0] = sq;
      r.a[1] = -sq;
      r.a[2] = evs;
      r.a[3] = -evs;
      Fx[0] = Fx[1] = mk<double>(cp);
      Fy[0] = Fy[1] = mk<double>(sp);
      Fx[2] = Fx[3] = evs * (-sp);
      Fy[2] = Fy[3] = evs * cp;
    } else {
      const cd ev02 = eps11 - mk<double>(beta * beta);
      const cd evs = csqrt(ev02);
      const cd u = eps11 + delta * (ct * ct);
      const double gama = beta * cp;
      const cd v = delta * (gama * 2. * st * ct);
      const cd w = delta * (gama * gama * st * st) - eps11 * (ev02 + delta);
      const cd sq = cdiv(csqrt(v * v - u * w * 4.), u) * 0.5;
      const cd vh = cdiv(v, u) * 0.5;
      const cd evpp = -vh + sq, evpm = -vh - sq;
      r.a[0] = evpp;
      r.a[1] = evpm;
      r.a[2] = evs;
      r.a[3] = -evs;
      const double sfst = st * sp, cfst = cp * st, ctbeta = ct * beta;
      const cd ctbe = cdiv(mk<double>(ctbeta), eps11);
      const cd ev02cfsteps11 = cdiv(ev02 * cfst, eps11);
      Fx[0] = evpp * ctbe - ev02cfsteps11;
      Fy[0] = mk<double>(-sfst);
      Fx[1] = -(evpm * ctbe) + ev02cfsteps11;
      Fy[1] = mk<double>(sfst);
      Fx[2] = evs * sfst;
      Fy[2] = mk<double>(ctbeta) - evs * cfst;
      Fx[3] = evs * (-sfst);
      Fy[3] = evs * cfst + mk<double>(ctbeta);
    }
  } else {
    const double nan = __builtin_nan("");
    for (int i = 0; i < 4; ++i) {
      r.a[i] = mk<double>(nan, nan);
      Fx[i] = Fy[i] = mk<double>(nan, nan);
    }
  }
  // beam-azimuth rotation of the E rows (dtmm/linalg.py:674-713)
  cd Ex[4], Ey[4];
  for (int j = 0; j < 4; ++j) {
    Ex[j] = Fx[j] * c0 - Fy[j] * s0;
    Ey[j] = Fx[j] * s0 + Fy[j] * c0;
  }
  cd ep[2][2] = {{Ex[0], Ex[2]}, {Ey[0], Ey[2]}};
  cd en[2][2] = {{Ex[1], Ex[3]}, {Ey[1], Ey[3]}};
  cd ip[2][2], in[2][2];
  const bool okp = inv2(ep, ip);
  const bool okn = inv2(en, in);
  r.ok2 = okp;
  r.okb = okn;
  r.ok4 = okp && okn;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      r.P0[i][j] = ep[i][0] * ip[0][j];
      r.P1[i][j] = en[i][0] * in[0][j];
    }
}

// Per-pixel operator data of any material: the lean closed forms for
// isotropic and uniaxial voxels, the numeric eigen-decomposition of
// biaxial_modes (rotated by the beam azimuth) for biaxial ones.
__host__ __device__ inline void rop_full(double beta, double c0, double s0, double phi0, cd e0,
                                         cd e1, cd e2, double psi, double theta, double phi,
                                         RopData& r) {
  if (material_kind(e0, e1, e2) != KIND_BIAXIAL) {
    rop_lean(beta, c0, s0, phi0, e0, e1, e2, theta, phi, r);
    return;
  }
  Modes m;
  biaxial_modes(beta, e0, e1, e2, psi, theta, phi - phi0, m);
  rotate_rows(m, c0, s0);
  rop_data(m, KIND_BIAXIAL, r);
}

}  // namespace dtmmx
