// MI355X (gfx950) engine for the diffractive transfer-matrix path.
//
// Per layer the reference (dtmm/propagate_4x4.py:80-101) does
//   F <- D2 . FFT2( E diag(exp(i kd alpha)) E^-1 . IFFT2( D1 . F ))
// with dense per-k 4x4 matrices D1, D2 built from the isotropic effective
// medium (dtmm/matrix.py:109-174).  For an isotropic effective medium the
// product D1(layer j+1) . D2(layer j) is diagonal in the (E+x, E-x, E+y, E-y)
// basis: exp(+-i k0 [d_j/2 (a_j(k)-a_cj) + d_j+1/2 (a_j+1(k)-a_cj+1)]) on the
// forward/backward components, a(k) = sqrt(eps - beta(k)^2), zero where
// beta(k) >= betamax.  The engine therefore keeps the field in that basis
// between layers and runs two kernels per layer:
//   row kernel    (x lines): IFFT_x -> per-pixel layer operator -> FFT_x
//   column kernel (y lines): FFT_y -> diagonal k-space operator -> IFFT_y
// Dense k-space kernels (fp64) convert F -> Epn before the first layer,
// Epn -> F after the last one, and handle boundaries between different
// effective media.  The 2x2 method is the same with the two forward
// components only.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "berreman.hpp"
#include "cpx.hpp"
#include "dtmmx.h"
#include "fft.hpp"

namespace dtmmx {

constexpr int NT = 256;  // threads per work-group for every kernel
constexpr int MAXL = DTMMX_MAX_WAVELENGTHS;
constexpr int MAXB = DTMMX_MAX_BEAMS;

thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(DTMMX_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_));            \
  } while (0)

// Wavelengths and beams, passed by value to every kernel.
struct Beams {
  int nlam;           // wavelengths
  int rays_per_beam;  // consecutive rays sharing a beam
  double k0[MAXL];
  double beta[MAXB];
  double phi[MAXB];
};

__host__ __device__ inline double fftfreq(int i, int n) {
  // numpy.fft.fftfreq(n)[i] = k * (1.0 / n) with k in [0..(n-1)/2, -n/2..-1]
  int k = (i < (n + 1) / 2) ? i : i - n;
  return double(k) * (1.0 / double(n));
}

__host__ __device__ inline int mode_index(int i, int n) {  // |k| of fftfreq
  return (i < (n + 1) / 2) ? i : n - i;
}

// ---------------------------------------------------------------------------
// Row kernel: one work-group transforms NC component lines of one row y of
// a set of planes that share one beam.
// ---------------------------------------------------------------------------

enum { OP_FWD = 0, OP_INV = 1, OP_ROP = 2, OP_DIAG = 3 };

template <typename T>
struct RowArgs {
  cpx<T>* field;
  int H, W;
  int nplanes;           // total planes (rays * wavelengths)
  int planes_per_group;  // planes sharing one beam
  int ppw;               // planes per work-group
  int nchunk;            // work-groups per group per row
  const cpx<T>* tw;      // twiddles of length W
  // layer operator (OP_ROP)
  const double* eps_re;
  const double* eps_im;
  const double* ang;
  size_t comp_stride;  // elements between eps components (L*H*W)
  double d;            // layer thickness
  int method4;         // 1: 4x4 (ok4), 0: 2x2 (ok2)
  int* status;
  Beams bm;
};

template <typename T>
struct RopT {
  cpx<T> a[4];
  cpx<T> P0[4];
  cpx<T> P1[4];
  T ok;
};

template <typename T, int N, int NC, int OP>
__global__ __launch_bounds__(NT) void row_kernel(RowArgs<T> A) {
  constexpr int PPT = (N + NT - 1) / NT;  // pixels per thread in the layer operator
  __shared__ cpx<T> buf[NC * N];
  const int tid = threadIdx.x;
  const int y = blockIdx.x;
  const int g = blockIdx.y / A.nchunk;
  const int ch = blockIdx.y - g * A.nchunk;
  const int q0 = g * A.planes_per_group + ch * A.ppw;
  const int q1 = min(q0 + A.ppw, (g + 1) * A.planes_per_group);
  const size_t plane_elems = size_t(A.H) * N;

  RopT<T> rd[PPT];
  if constexpr (OP == OP_ROP) {
    const int beam = g;  // one group per beam
    const double beta = A.bm.beta[beam];
    const double phi0 = A.bm.phi[beam];
    const double c0 = cos(phi0), s0 = sin(phi0);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int x = tid + k * NT;
      if (x < N) {
        const size_t o = size_t(y) * N + x;
        cd e[3];
        for (int c = 0; c < 3; ++c)
          e[c] = mk<double>(A.eps_re[c * A.comp_stride + o],
                            A.eps_im ? A.eps_im[c * A.comp_stride + o] : 0.);
        const double theta = A.ang[1 * A.comp_stride + o];
        const double phi = A.ang[2 * A.comp_stride + o];
        Modes m;
        int kind;
        layer_modes(beta, c0, s0, phi0, e[0], e[1], e[2], theta, phi, m, kind);
        RopData r;
        rop_data(m, kind, r);
        if (kind == KIND_BIAXIAL && A.status) atomicOr(A.status, 1);
        for (int i = 0; i < 4; ++i) rd[k].a[i] = cast<T>(r.a[i]);
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) {
            rd[k].P0[2 * i + j] = cast<T>(r.P0[i][j]);
            rd[k].P1[2 * i + j] = cast<T>(r.P1[i][j]);
          }
        rd[k].ok = (A.method4 ? r.ok4 : r.ok2) ? T(1) : T(0);
        if (kind == KIND_BIAXIAL) rd[k].ok = T(__builtin_nan(""));
      }
    }
  }

  for (int q = q0; q < q1; ++q) {
    cpx<T>* base = A.field + size_t(q) * NC * plane_elems + size_t(y) * N;
    for (int idx = tid; idx < NC * N; idx += NT) {
      const int c = idx / N;
      const int x = idx - c * N;
      buf[idx] = base[c * plane_elems + x];
    }
    __syncthreads();
    if constexpr (OP == OP_FWD) {
      lds_fft<T, N, NC, NT, -1>(buf, N, A.tw, tid);
    } else if constexpr (OP == OP_INV) {
      lds_fft<T, N, NC, NT, +1>(buf, N, A.tw, tid);
    } else {
      lds_fft<T, N, NC, NT, +1>(buf, N, A.tw, tid);
      const int lam = q % A.bm.nlam;
      const T kd = T(A.bm.k0[lam] * A.d);
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        const int x = tid + k * NT;
        if (x < N) {
          const RopT<T>& r = rd[k];
          cpx<T> p0 = cexpi(r.a[0] * kd);
          cpx<T> p2 = cexpi(r.a[2] * kd);
          cpx<T> dp = (p0 - p2) * r.ok;
          p2 = p2 * r.ok;
          if constexpr (NC == 4) {
            cpx<T> p1 = cexpi(r.a[1] * kd);
            cpx<T> p3 = cexpi(r.a[3] * kd);
            cpx<T> dn = (p1 - p3) * r.ok;
            p3 = p3 * r.ok;
            cpx<T> u0 = buf[x], u1 = buf[N + x], u2 = buf[2 * N + x], u3 = buf[3 * N + x];
            cpx<T> w0 = r.P0[0] * u0 + r.P0[1] * u2;
            cpx<T> w2 = r.P0[2] * u0 + r.P0[3] * u2;
            cpx<T> w1 = r.P1[0] * u1 + r.P1[1] * u3;
            cpx<T> w3 = r.P1[2] * u1 + r.P1[3] * u3;
            buf[x] = p2 * u0 + dp * w0;
            buf[2 * N + x] = p2 * u2 + dp * w2;
            buf[N + x] = p3 * u1 + dn * w1;
            buf[3 * N + x] = p3 * u3 + dn * w3;
          } else {
            cpx<T> u0 = buf[x], u1 = buf[N + x];
            cpx<T> w0 = r.P0[0] * u0 + r.P0[1] * u1;
            cpx<T> w1 = r.P0[2] * u0 + r.P0[3] * u1;
            buf[x] = p2 * u0 + dp * w0;
            buf[N + x] = p2 * u1 + dp * w1;
          }
        }
      }
      __syncthreads();
      lds_fft<T, N, NC, NT, -1>(buf, N, A.tw, tid);
    }
    for (int idx = tid; idx < NC * N; idx += NT) {
      const int c = idx / N;
      const int x = idx - c * N;
      base[c * plane_elems + x] = buf[idx];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Column kernel: one work-group transforms TX adjacent columns of one
// component plane.  OP_DIAG applies the diagonal k-space operator between
// the forward and inverse transforms.
// ---------------------------------------------------------------------------

template <typename T>
struct ColArgs {
  cpx<T>* field;
  int H, W;
  int NC;             // components per plane
  const cpx<T>* tw;   // twiddles of length H
  const int* amax;    // [nlam][H]: largest passing |kx| index per ky row, -1 if none
  double da, db;      // thicknesses of the two half steps
  cd ea, eb;          // isotropic effective eps of the two half steps
  double scale;
  int dense_sign;     // 1: backward components (odd c) get exp(-i phase)
  Beams bm;
};

template <typename T>
__device__ __forceinline__ cpx<T> diag_factor(const ColArgs<T>& A, int lam, int beam, int ky,
                                              int kx, int sign) {
  const int mx = mode_index(kx, A.W), my = mode_index(ky, A.H);
  if (mx > A.amax[lam * A.H + ky]) return mk<T>(T(0), T(0));
  (void)my;
  const T k0 = T(A.bm.k0[lam]);
  const T l = T(2 * M_PI) / k0;
  const T fx = T(mx) / T(A.W), fy = T(my) / T(A.H);
  const T b2 = l * l * (fx * fx + fy * fy);
  const T bc = T(A.bm.beta[beam]);
  cpx<T> ph = mk<T>(T(0), T(0));
  if (A.da != 0.) {
    cpx<T> e = cast<T>(A.ea);
    cpx<T> ak = csqrt(e - mk<T>(b2));
    cpx<T> ac = csqrt(e - mk<T>(bc * bc));
    ph = ph + (ak - ac) * T(0.5 * A.da);
  }
  if (A.db != 0.) {
    cpx<T> e = cast<T>(A.eb);
    cpx<T> ak = csqrt(e - mk<T>(b2));
    cpx<T> ac = csqrt(e - mk<T>(bc * bc));
    ph = ph + (ak - ac) * T(0.5 * A.db);
  }
  ph = ph * (k0 * T(sign));
  return cexpi(ph) * T(A.scale);
}

template <typename T, int N, int TX, int LSTR, int OP>
__global__ __launch_bounds__(NT) void col_kernel(ColArgs<T> A) {
  __shared__ cpx<T> buf[TX * LSTR];
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * TX;
  const int pc = blockIdx.y;
  cpx<T>* base = A.field + size_t(pc) * N * A.W;
  for (int idx = tid; idx < TX * N; idx += NT) {
    const int yy = idx / TX;
    const int xx = idx - yy * TX;
    const int col = x0 + xx;
    buf[xx * LSTR + yy] = (col < A.W) ? base[size_t(yy) * A.W + col] : mk<T>(T(0), T(0));
  }
  __syncthreads();
  if constexpr (OP == OP_FWD) {
    lds_fft<T, N, TX, NT, -1>(buf, LSTR, A.tw, tid);
  } else if constexpr (OP == OP_INV) {
    lds_fft<T, N, TX, NT, +1>(buf, LSTR, A.tw, tid);
  } else {
    lds_fft<T, N, TX, NT, -1>(buf, LSTR, A.tw, tid);
    const int c = pc % A.NC;
    const int q = pc / A.NC;
    const int lam = q % A.bm.nlam;
    const int beam = (q / A.bm.nlam) / A.bm.rays_per_beam;
    const int sign = (A.dense_sign && (c & 1)) ? -1 : 1;
    for (int idx = tid; idx < TX * N; idx += NT) {
      const int xx = idx / N;
      const int ky = idx - xx * N;
      const int col = x0 + xx;
      if (col < A.W) buf[xx * LSTR + ky] = buf[xx * LSTR + ky] * diag_factor(A, lam, beam, ky, col, sign);
    }
    __syncthreads();
    lds_fft<T, N, TX, NT, +1>(buf, LSTR, A.tw, tid);
  }
  for (int idx = tid; idx < TX * N; idx += NT) {
    const int yy = idx / TX;
    const int xx = idx - yy * TX;
    const int col = x0 + xx;
    if (col < A.W) base[size_t(yy) * A.W + col] = buf[xx * LSTR + yy];
  }
}

// ---------------------------------------------------------------------------
// Dense per-k operators (fp64), one thread per (k, plane).
// ---------------------------------------------------------------------------

enum { KOP_HEAD4 = 0, KOP_TAIL4 = 1, KOP_BOUND4 = 2, KOP_PROJ2 = 3 };

template <typename T>
struct KArgs {
  const cpx<T>* src;
  cpx<T>* dst;
  int H, W;
  int nplanes;
  const int* amax;
  int kop;
  double d0, d1;  // thickness of the layer(s) involved
  cd e0, e1;      // effective eps of the layer(s) involved
  cd ein;         // input medium eps (PROJ2)
  double scale;
  Beams bm;
};

// Modes of an isotropic medium at in-plane wave vector (beta, phi), rotated.
__device__ inline void iso_at(double beta, double phi, cd eps, Modes& m) {
  iso_modes(beta, eps, m);
  rotate_rows(m, cos(phi), sin(phi));
}

__device__ inline void emat(const Modes& m, cd E[4][4]) {  // dtmm/tmm.py:990-996
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) E[i][j] = mk<double>(0.);
  E[0][0] = m.F[0][0];
  E[0][2] = m.F[0][2];
  E[2][0] = m.F[2][0];
  E[2][2] = m.F[2][2];
  E[1][1] = m.F[0][1];
  E[1][3] = m.F[0][3];
  E[3][1] = m.F[2][1];
  E[3][3] = m.F[2][3];
}

// X . diag(exp(i kd a)) . Y
__device__ inline void xpy(const cd X[4][4], const cd a[4], double kd, const cd Y[4][4], cd O[4][4]) {
  cd p[4];
  for (int i = 0; i < 4; ++i) p[i] = cexpi(a[i] * kd);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      cd s = mk<double>(0.);
      for (int k = 0; k < 4; ++k) s = s + X[i][k] * p[k] * Y[k][j];
      O[i][j] = s;
    }
}

// Mean-beam correction E_c P(-kd) E_c^-1 (dtmm/matrix.py:132-139).
__device__ inline void correction(double beta, double phi, cd eps, double kd, cd C[4][4]) {
  Modes m;
  iso_at(beta, phi, eps, m);
  cd E[4][4], Ei[4][4];
  emat(m, E);
  inv4(E, Ei);
  xpy(E, m.a, -kd, Ei, C);
}

// First half step E P(kd) F^-1 at wave vector k (dtmm/matrix.py:109-118).
__device__ inline void first_half(double beta, double phi, cd eps, double kd, cd D[4][4]) {
  Modes m;
  iso_at(beta, phi, eps, m);
  cd E[4][4], Fi[4][4];
  emat(m, E);
  inv4(m.F, Fi);
  xpy(E, m.a, kd, Fi, D);
}

// Second half step F P(kd) E^-1 at wave vector k (dtmm/matrix.py:120-130).
__device__ inline void second_half(double beta, double phi, cd eps, double kd, cd D[4][4]) {
  Modes m;
  iso_at(beta, phi, eps, m);
  cd E[4][4], Ei[4][4];
  emat(m, E);
  inv4(E, Ei);
  xpy(m.F, m.a, kd, Ei, D);
}

template <typename T>
__global__ __launch_bounds__(NT) void kspace_kernel(KArgs<T> A) {
  const int k = blockIdx.x * NT + threadIdx.x;
  const int q = blockIdx.y;
  const size_t HW = size_t(A.H) * A.W;
  if (k >= int(HW)) return;
  const int ky = k / A.W, kx = k - ky * A.W;
  const int lam = q % A.bm.nlam;
  const int beam = (q / A.bm.nlam) / A.bm.rays_per_beam;
  const double k0 = A.bm.k0[lam];
  const double bb = A.bm.beta[beam], pb = A.bm.phi[beam];
  const bool pass = mode_index(kx, A.W) <= A.amax[lam * A.H + ky];
  const double fx = fftfreq(kx, A.W), fy = fftfreq(ky, A.H);
  const double beta = (2 * M_PI / k0) * sqrt(fx * fx + fy * fy);
  const double phi = atan2(fy, fx);
  const int nin = (A.kop == KOP_TAIL4 || A.kop == KOP_BOUND4) ? 4 : (A.kop == KOP_PROJ2 ? 4 : 4);
  const int nout = (A.kop == KOP_PROJ2) ? 2 : 4;
  const cpx<T>* s = A.src + size_t(q) * nin * HW + k;
  cpx<T>* o = A.dst + size_t(q) * nout * HW + k;
  cd v[4];
  for (int c = 0; c < nin; ++c) v[c] = cast<double>(s[c * HW]);
  cd M[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) M[i][j] = mk<double>(0.);
  if (pass) {
    if (A.kop == KOP_HEAD4) {
      cd C[4][4], D[4][4];
      correction(bb, pb, A.e0, k0 * A.d0 * 0.5, C);
      first_half(beta, phi, A.e0, k0 * A.d0 * 0.5, D);
      matmul4(C, D, M);
    } else if (A.kop == KOP_TAIL4) {
      cd C[4][4], D[4][4];
      correction(bb, pb, A.e0, k0 * A.d0 * 0.5, C);
      second_half(beta, phi, A.e0, k0 * A.d0 * 0.5, D);
      matmul4(D, C, M);
    } else if (A.kop == KOP_BOUND4) {
      cd C0[4][4], D0[4][4], C1[4][4], D1[4][4], X[4][4], Y[4][4];
      correction(bb, pb, A.e0, k0 * A.d0 * 0.5, C0);
      second_half(beta, phi, A.e0, k0 * A.d0 * 0.5, D0);
      correction(bb, pb, A.e1, k0 * A.d1 * 0.5, C1);
      first_half(beta, phi, A.e1, k0 * A.d1 * 0.5, D1);
      matmul4(D0, C0, X);  // second corrected of layer 0
      matmul4(C1, D1, Y);  // first corrected of layer 1
      matmul4(Y, X, M);
    } else {
      // rows (Ex, Ey) of the forward projection in the input medium
      // (dtmm/diffract.py:242-251) times the first 2x2 half step.
      Modes m;
      iso_at(beta, phi, A.ein, m);
      cd Fi[4][4];
      inv4(m.F, Fi);
      cd P[4][4];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) P[i][j] = m.F[i][0] * Fi[0][j] + m.F[i][2] * Fi[2][j];
      cd ph = mk<double>(0.);
      if (A.d0 != 0.) {
        cd ak = csqrt(A.e0 - mk<double>(beta * beta));
        cd ac = csqrt(A.e0 - mk<double>(bb * bb));
        ph = (ak - ac) * (0.5 * A.d0 * k0);
      }
      cd K = cexpi(ph);
      for (int j = 0; j < 4; ++j) {
        M[0][j] = K * P[0][j];
        M[1][j] = K * P[2][j];
      }
    }
  }
  for (int i = 0; i < nout; ++i) {
    cd r = mk<double>(0.);
    for (int j = 0; j < nin; ++j) r = r + M[i][j] * v[j];
    o[i * HW] = cast<T>(r * A.scale);
  }
}

// (Ex, Ey) -> (Ex, Hy, Ey, Hx) with the mean-beam H/E relation of the output
// medium (dtmm/propagate_2x2.py:81-85, dtmm/tmm.py:1051-1072).
template <typename T>
struct E2HArgs {
  const cpx<T>* src;
  cpx<T>* dst;
  int H, W;
  cd eout;
  Beams bm;
};

template <typename T>
__global__ __launch_bounds__(NT) void e2h_kernel(E2HArgs<T> A) {
  const int k = blockIdx.x * NT + threadIdx.x;
  const int q = blockIdx.y;
  const size_t HW = size_t(A.H) * A.W;
  if (k >= int(HW)) return;
  const int beam = (q / A.bm.nlam) / A.bm.rays_per_beam;
  Modes m;
  iso_modes(A.bm.beta[beam], A.eout, m);
  rotate_rows(m, cos(A.bm.phi[beam]), sin(A.bm.phi[beam]));
  cd Am[2][2] = {{m.F[0][0], m.F[0][2]}, {m.F[2][0], m.F[2][2]}};
  cd Bm[2][2] = {{m.F[1][0], m.F[1][2]}, {m.F[3][0], m.F[3][2]}};
  cd Ai[2][2];
  inv2(Am, Ai);
  cd G[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) G[i][j] = Bm[i][0] * Ai[0][j] + Bm[i][1] * Ai[1][j];
  const cpx<T>* s = A.src + size_t(q) * 2 * HW + k;
  cpx<T>* o = A.dst + size_t(q) * 4 * HW + k;
  cd ex = cast<double>(s[0]), ey = cast<double>(s[HW]);
  o[0] = cast<T>(ex);
  o[HW] = cast<T>(G[0][0] * ex + G[0][1] * ey);
  o[2 * HW] = cast<T>(ey);
  o[3 * HW] = cast<T>(G[1][0] * ex + G[1][1] * ey);
}

// ---------------------------------------------------------------------------
// Host side: size dispatch and the transfer driver.
// ---------------------------------------------------------------------------

template <typename T>
static std::vector<cpx<T>> twiddles(int n) {
  std::vector<cpx<T>> t(n);
  for (int i = 0; i < n; ++i) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)i / (long double)n;
    t[i] = mk<T>(T(std::cos(a)), T(std::sin(a)));
  }
  return t;
}

static bool size_supported(int n) {
  switch (n) {
    case 8: case 12: case 16: case 20: case 24: case 32: case 40: case 48: case 64: case 80:
    case 96: case 128: case 160: case 192: case 256: case 320: case 384: case 512: case 640:
    case 768: case 1024:
      return true;
    default:
      return false;
  }
}

#define DTMMX_SIZES(X)                                                                        \
  X(8) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) \
  X(320) X(384) X(512) X(640) X(768) X(1024)

template <typename T, int NC, int OP>
static int launch_row_n(int n, const RowArgs<T>& a, dim3 grid, hipStream_t s) {
  switch (n) {
#define X(N)                                                                  \
  case N:                                                                     \
    hipLaunchKernelGGL((row_kernel<T, N, NC, OP>), grid, dim3(NT), 0, s, a); \
    break;
    DTMMX_SIZES(X)
#undef X
    default:
      return fail(DTMMX_E_UNSUPPORTED, "unsupported width " + std::to_string(n));
  }
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

template <typename T, int N>
struct ColShape {
  static constexpr int BYTES = 65536;
  static constexpr int CB = int(sizeof(cpx<T>));
  static constexpr int TX0 = BYTES / (N * CB);
  static constexpr int TX = TX0 >= 16 ? 16 : (TX0 >= 8 ? 8 : (TX0 >= 4 ? 4 : (TX0 >= 2 ? 2 : 1)));
  static constexpr int LSTR = (TX * (N + 1) * CB <= BYTES) ? N + 1 : N;
};

template <typename T, int OP>
static int launch_col_n(int n, const ColArgs<T>& a, int nlines, hipStream_t s) {
  switch (n) {
#define X(N)                                                                              \
  case N: {                                                                               \
    using S = ColShape<T, N>;                                                             \
    dim3 grid((a.W + S::TX - 1) / S::TX, nlines);                                         \
    hipLaunchKernelGGL((col_kernel<T, N, S::TX, S::LSTR, OP>), grid, dim3(NT), 0, s, a);  \
    break;                                                                                \
  }
    DTMMX_SIZES(X)
#undef X
    default:
      return fail(DTMMX_E_UNSUPPORTED, "unsupported height " + std::to_string(n));
  }
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

}  // namespace dtmmx

using namespace dtmmx;

struct dtmmx_plan {
  dtmmx_plan_desc d;
  std::vector<double> k0, beta, phi;
  Beams bm;
  int device;
  void* tw_w = nullptr;  // twiddles of width (precision type)
  void* tw_h = nullptr;
  int* amax = nullptr;   // [nlam][H]
};

namespace dtmmx {

// Optional per-launch timing of the per-layer kernels with HIP events on the
// launch stream (class 0: row layer kernel, 1: column k-space kernel).
struct KTimer {
  bool on = false;
  int cap = 0, used = 0;
  std::vector<hipEvent_t> ev;
  std::vector<int> cls;
};
static KTimer g_timer;

static void timer_start(int cls, hipStream_t s) {
  if (!g_timer.on || g_timer.used >= g_timer.cap) return;
  g_timer.cls[g_timer.used] = cls;
  (void)hipEventRecord(g_timer.ev[2 * g_timer.used], s);
}

static void timer_stop(hipStream_t s) {
  if (!g_timer.on || g_timer.used >= g_timer.cap) return;
  (void)hipEventRecord(g_timer.ev[2 * g_timer.used + 1], s);
  ++g_timer.used;
}

template <typename T>
static int row_pass(const dtmmx_plan* P, cpx<T>* field, int nplanes, int nc, int op,
                    hipStream_t s, const RowArgs<T>* proto = nullptr, int ppg = 0) {
  RowArgs<T> a{};
  if (proto) a = *proto;
  a.field = field;
  a.H = P->d.height;
  a.W = P->d.width;
  a.nplanes = nplanes;
  a.tw = (const cpx<T>*)P->tw_w;
  a.bm = P->bm;
  if (op == OP_ROP) {
    a.planes_per_group = ppg;
    const int groups = nplanes / ppg;
    int nchunk = (2048 + a.H * groups - 1) / (a.H * groups);
    if (nchunk < 1) nchunk = 1;
    if (nchunk > ppg) nchunk = ppg;
    a.ppw = (ppg + nchunk - 1) / nchunk;
    a.nchunk = (ppg + a.ppw - 1) / a.ppw;
  } else {
    a.planes_per_group = 1;
    a.ppw = 1;
    a.nchunk = 1;
  }
  dim3 grid(a.H, (nplanes / a.planes_per_group) * a.nchunk);
  const int w = a.W;
  if (nc == 1 && op == OP_FWD) return launch_row_n<T, 1, OP_FWD>(w, a, grid, s);
  if (nc == 1 && op == OP_INV) return launch_row_n<T, 1, OP_INV>(w, a, grid, s);
  if (nc == 2 && op == OP_FWD) return launch_row_n<T, 2, OP_FWD>(w, a, grid, s);
  if (nc == 2 && op == OP_INV) return launch_row_n<T, 2, OP_INV>(w, a, grid, s);
  if (op == OP_ROP) {
    timer_start(0, s);
    int r = nc == 2 ? launch_row_n<T, 2, OP_ROP>(w, a, grid, s)
                    : launch_row_n<T, 4, OP_ROP>(w, a, grid, s);
    timer_stop(s);
    return r;
  }
  if (nc == 4 && op == OP_FWD) return launch_row_n<T, 4, OP_FWD>(w, a, grid, s);
  if (nc == 4 && op == OP_INV) return launch_row_n<T, 4, OP_INV>(w, a, grid, s);
  return fail(DTMMX_E_ARG, "bad row pass");
}

template <typename T>
static int col_pass(const dtmmx_plan* P, cpx<T>* field, int nlines, int nc, int op, hipStream_t s,
                    const ColArgs<T>* proto = nullptr) {
  ColArgs<T> a{};
  if (proto) a = *proto;
  a.field = field;
  a.H = P->d.height;
  a.W = P->d.width;
  a.NC = nc;
  a.tw = (const cpx<T>*)P->tw_h;
  a.amax = P->amax;
  a.bm = P->bm;
  const int h = a.H;
  if (op == OP_FWD) return launch_col_n<T, OP_FWD>(h, a, nlines, s);
  if (op == OP_INV) return launch_col_n<T, OP_INV>(h, a, nlines, s);
  if (op == OP_DIAG) {
    timer_start(1, s);
    int r = launch_col_n<T, OP_DIAG>(h, a, nlines, s);
    timer_stop(s);
    return r;
  }
  return fail(DTMMX_E_ARG, "bad column pass");
}

template <typename T>
static int kspace(const dtmmx_plan* P, const cpx<T>* src, cpx<T>* dst, int nplanes, KArgs<T> a,
                  hipStream_t s) {
  a.src = src;
  a.dst = dst;
  a.H = P->d.height;
  a.W = P->d.width;
  a.nplanes = nplanes;
  a.amax = P->amax;
  a.bm = P->bm;
  const int HW = a.H * a.W;
  dim3 grid((HW + NT - 1) / NT, nplanes);
  hipLaunchKernelGGL((kspace_kernel<T>), grid, dim3(NT), 0, s, a);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

#define TRY(x)                 \
  do {                         \
    int r_ = (x);              \
    if (r_ != DTMMX_OK) return r_; \
  } while (0)

template <typename T>
static int transfer_impl(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout,
                         void* ws, int* status, hipStream_t s) {
  const int H = P->d.height, W = P->d.width;
  const size_t HW = size_t(H) * W;
  const int nplanes = P->d.n_rays * P->d.n_wavelengths;
  const int ppg = P->d.rays_per_beam * P->d.n_wavelengths;
  const int L = S->n_layers;
  const double scale = 1.0 / double(HW);
  auto eff = [&](int l) {
    return mk<double>(S->eps_eff_re[l], S->eps_eff_im ? S->eps_eff_im[l] : 0.);
  };
  cpx<T>* out = (cpx<T>*)fout;
  HIP_TRY(hipMemcpyAsync(out, fin, size_t(nplanes) * 4 * HW * sizeof(cpx<T>),
                         hipMemcpyDeviceToDevice, s));
  const size_t cs = size_t(L) * HW;
  auto rop_proto = [&](int l, int m4) {
    RowArgs<T> a{};
    a.eps_re = S->eps_re + size_t(l) * HW;
    a.eps_im = S->eps_im ? S->eps_im + size_t(l) * HW : nullptr;
    a.ang = S->angles + size_t(l) * HW;
    a.comp_stride = cs;
    a.d = S->thickness[l];
    a.method4 = m4;
    a.status = status;
    return a;
  };
  if (P->d.method == DTMMX_METHOD_4X4) {
    if (L == 0) return DTMMX_OK;  // ifft2(fft2(field_in))
    TRY(row_pass<T>(P, out, nplanes, 4, OP_FWD, s));
    TRY(col_pass<T>(P, out, nplanes * 4, 4, OP_FWD, s));
    KArgs<T> k{};
    k.kop = KOP_HEAD4;
    k.d0 = S->thickness[0];
    k.e0 = eff(0);
    k.scale = scale;
    TRY(kspace<T>(P, out, out, nplanes, k, s));
    TRY(col_pass<T>(P, out, nplanes * 4, 4, OP_INV, s));
    for (int l = 0; l < L; ++l) {
      RowArgs<T> ra = rop_proto(l, 1);
      TRY(row_pass<T>(P, out, nplanes, 4, OP_ROP, s, &ra, ppg));
      if (l + 1 < L) {
        cd e0 = eff(l), e1 = eff(l + 1);
        if (eq(e0, e1)) {
          ColArgs<T> ca{};
          ca.da = S->thickness[l];
          ca.db = S->thickness[l + 1];
          ca.ea = e0;
          ca.eb = e1;
          ca.scale = scale;
          ca.dense_sign = 1;
          TRY(col_pass<T>(P, out, nplanes * 4, 4, OP_DIAG, s, &ca));
        } else {
          TRY(col_pass<T>(P, out, nplanes * 4, 4, OP_FWD, s));
          KArgs<T> b{};
          b.kop = KOP_BOUND4;
          b.d0 = S->thickness[l];
          b.d1 = S->thickness[l + 1];
          b.e0 = e0;
          b.e1 = e1;
          b.scale = scale;
          TRY(kspace<T>(P, out, out, nplanes, b, s));
          TRY(col_pass<T>(P, out, nplanes * 4, 4, OP_INV, s));
        }
      }
    }
    TRY(col_pass<T>(P, out, nplanes * 4, 4, OP_FWD, s));
    KArgs<T> t{};
    t.kop = KOP_TAIL4;
    t.d0 = S->thickness[L - 1];
    t.e0 = eff(L - 1);
    t.scale = scale;
    TRY(kspace<T>(P, out, out, nplanes, t, s));
    TRY(col_pass<T>(P, out, nplanes * 4, 4, OP_INV, s));
    TRY(row_pass<T>(P, out, nplanes, 4, OP_INV, s));
    return DTMMX_OK;
  }
  // 2x2: out doubles as the 4-component scratch of the input projection.
  cpx<T>* e2 = (cpx<T>*)ws;
  TRY(row_pass<T>(P, out, nplanes, 4, OP_FWD, s));
  TRY(col_pass<T>(P, out, nplanes * 4, 4, OP_FWD, s));
  KArgs<T> h{};
  h.kop = KOP_PROJ2;
  h.d0 = L > 0 ? S->thickness[0] : 0.;
  h.e0 = L > 0 ? eff(0) : mk<double>(S->n_out * S->n_out);
  h.ein = mk<double>(S->n_in * S->n_in);
  h.scale = scale;
  TRY(kspace<T>(P, out, e2, nplanes, h, s));
  TRY(col_pass<T>(P, e2, nplanes * 2, 2, OP_INV, s));
  for (int l = 0; l < L; ++l) {
    RowArgs<T> ra = rop_proto(l, 0);
    TRY(row_pass<T>(P, e2, nplanes, 2, OP_ROP, s, &ra, ppg));
    ColArgs<T> ca{};
    ca.da = S->thickness[l];
    ca.ea = eff(l);
    if (l + 1 < L) {
      ca.db = S->thickness[l + 1];
      ca.eb = eff(l + 1);
    } else {
      ca.db = 0.;  // output medium: zero thickness, only the betamax mask
      ca.eb = mk<double>(S->n_out * S->n_out);
    }
    ca.scale = scale;
    ca.dense_sign = 0;
    TRY(col_pass<T>(P, e2, nplanes * 2, 2, OP_DIAG, s, &ca));
  }
  if (L == 0) {
    // no layer: the mask of the output medium is already in the projection;
    // undo the scale there by applying the inverse transform only.
  }
  TRY(row_pass<T>(P, e2, nplanes, 2, OP_INV, s));
  E2HArgs<T> ea{};
  ea.src = e2;
  ea.dst = out;
  ea.H = H;
  ea.W = W;
  ea.eout = mk<double>(S->n_out * S->n_out);
  ea.bm = P->bm;
  dim3 grid((unsigned)((HW + NT - 1) / NT), nplanes);
  hipLaunchKernelGGL((e2h_kernel<T>), grid, dim3(NT), 0, s, ea);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

template <typename T>
static int fft2_impl(void* data, int batch, int H, int W, int inverse, hipStream_t s) {
  dtmmx_plan P{};
  P.d.height = H;
  P.d.width = W;
  auto tw = twiddles<T>(W), th = twiddles<T>(H);
  HIP_TRY(hipMalloc(&P.tw_w, tw.size() * sizeof(cpx<T>)));
  HIP_TRY(hipMalloc(&P.tw_h, th.size() * sizeof(cpx<T>)));
  HIP_TRY(hipMemcpy(P.tw_w, tw.data(), tw.size() * sizeof(cpx<T>), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(P.tw_h, th.data(), th.size() * sizeof(cpx<T>), hipMemcpyHostToDevice));
  P.bm.nlam = 1;
  P.bm.rays_per_beam = 1;
  int r = row_pass<T>(&P, (cpx<T>*)data, batch, 1, inverse ? OP_INV : OP_FWD, s);
  if (r == DTMMX_OK) r = col_pass<T>(&P, (cpx<T>*)data, batch, 1, inverse ? OP_INV : OP_FWD, s);
  hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(P.tw_w);
  (void)hipFree(P.tw_h);
  if (r != DTMMX_OK) return r;
  if (e != hipSuccess) return fail(DTMMX_E_HIP, std::string("fft2: ") + hipGetErrorString(e));
  return DTMMX_OK;
}

}  // namespace dtmmx

extern "C" {

int dtmmx_abi_version(void) { return DTMMX_ABI_VERSION; }

const char* dtmmx_last_error(void) { return g_err.c_str(); }

int dtmmx_plan_create(dtmmx_plan** out, const dtmmx_plan_desc* d) {
  if (!out || !d) return fail(DTMMX_E_ARG, "null argument");
  *out = nullptr;
  if (d->height <= 0 || d->width <= 0) return fail(DTMMX_E_ARG, "empty grid");
  if (!size_supported(d->height) || !size_supported(d->width))
    return fail(DTMMX_E_UNSUPPORTED, "grid " + std::to_string(d->height) + "x" +
                                         std::to_string(d->width) + " not built");
  if (d->n_wavelengths <= 0 || d->n_wavelengths > MAXL)
    return fail(DTMMX_E_ARG, "n_wavelengths out of range");
  if (d->n_rays <= 0 || d->rays_per_beam <= 0 || d->n_rays % d->rays_per_beam != 0)
    return fail(DTMMX_E_ARG, "n_rays must be a positive multiple of rays_per_beam");
  const int nbeam = d->n_rays / d->rays_per_beam;
  if (nbeam > MAXB) return fail(DTMMX_E_ARG, "too many beams");
  if (d->method != DTMMX_METHOD_2X2 && d->method != DTMMX_METHOD_4X4)
    return fail(DTMMX_E_ARG, "method must be 2 or 4");
  if (d->precision != DTMMX_C64 && d->precision != DTMMX_C128)
    return fail(DTMMX_E_ARG, "precision must be DTMMX_C64 or DTMMX_C128");
  if (!d->k0 || !d->beta || !d->phi) return fail(DTMMX_E_ARG, "null k0/beta/phi");
  dtmmx_plan* P = new dtmmx_plan();
  P->d = *d;
  P->k0.assign(d->k0, d->k0 + d->n_wavelengths);
  P->beta.assign(d->beta, d->beta + nbeam);
  P->phi.assign(d->phi, d->phi + nbeam);
  P->d.k0 = P->k0.data();
  P->d.beta = P->beta.data();
  P->d.phi = P->phi.data();
  std::memset(&P->bm, 0, sizeof(P->bm));
  P->bm.nlam = d->n_wavelengths;
  P->bm.rays_per_beam = d->rays_per_beam;
  for (int i = 0; i < d->n_wavelengths; ++i) P->bm.k0[i] = d->k0[i];
  for (int i = 0; i < nbeam; ++i) {
    P->bm.beta[i] = d->beta[i];
    P->bm.phi[i] = d->phi[i];
  }
  (void)hipGetDevice(&P->device);
  // betamax mask per (wavelength, ky): beta >= betamax is cut
  // (dtmm/diffract.py:31, :55, :86) with beta computed as in
  // dtmm/wave.py:77-78.
  const int H = d->height, W = d->width;
  std::vector<int> amax(size_t(d->n_wavelengths) * H, -1);
  for (int l = 0; l < d->n_wavelengths; ++l) {
    const double k0 = std::fabs(d->k0[l]);
    const double f = 2 * M_PI / k0;
    for (int y = 0; y < H; ++y) {
      const double fy = fftfreq(y, H);
      int best = -1;
      for (int m = 0; m <= W / 2; ++m) {
#pragma clang fp contract(off)
        const double fx = fftfreq(m, W);
        const double s2 = fx * fx + fy * fy;
        const double b = f * std::sqrt(s2);
        if (b >= d->betamax) break;
        best = m;
      }
      amax[size_t(l) * H + y] = best;
    }
  }
  size_t cb = d->precision == DTMMX_C64 ? sizeof(cpx<float>) : sizeof(cpx<double>);
  hipError_t e = hipMalloc(&P->tw_w, W * cb);
  if (e == hipSuccess) e = hipMalloc(&P->tw_h, H * cb);
  if (e == hipSuccess) e = hipMalloc((void**)&P->amax, amax.size() * sizeof(int));
  if (e == hipSuccess) {
    if (d->precision == DTMMX_C64) {
      auto a = twiddles<float>(W), b = twiddles<float>(H);
      e = hipMemcpy(P->tw_w, a.data(), W * cb, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(P->tw_h, b.data(), H * cb, hipMemcpyHostToDevice);
    } else {
      auto a = twiddles<double>(W), b = twiddles<double>(H);
      e = hipMemcpy(P->tw_w, a.data(), W * cb, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(P->tw_h, b.data(), H * cb, hipMemcpyHostToDevice);
    }
  }
  if (e == hipSuccess)
    e = hipMemcpy(P->amax, amax.data(), amax.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    dtmmx_plan_destroy(P);
    return fail(e == hipErrorOutOfMemory ? DTMMX_E_NOMEM : DTMMX_E_HIP,
                std::string("plan tables: ") + hipGetErrorString(e));
  }
  *out = P;
  return DTMMX_OK;
}

int dtmmx_plan_destroy(dtmmx_plan* P) {
  if (!P) return DTMMX_OK;
  if (P->tw_w) (void)hipFree(P->tw_w);
  if (P->tw_h) (void)hipFree(P->tw_h);
  if (P->amax) (void)hipFree(P->amax);
  delete P;
  return DTMMX_OK;
}

size_t dtmmx_plan_workspace_size(const dtmmx_plan* P) {
  if (!P) return 0;
  size_t cb = P->d.precision == DTMMX_C64 ? sizeof(cpx<float>) : sizeof(cpx<double>);
  size_t planes = size_t(P->d.n_rays) * P->d.n_wavelengths;
  if (P->d.method == DTMMX_METHOD_2X2) return planes * 2 * P->d.height * P->d.width * cb;
  return 0;
}

int dtmmx_transfer(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout,
                   void* ws, size_t ws_bytes, int* status, void* stream) {
  if (!P || !S || !fin || !fout) return fail(DTMMX_E_ARG, "null argument");
  if (S->n_layers < 0) return fail(DTMMX_E_ARG, "negative layer count");
  if (S->n_layers > 0 && (!S->thickness || !S->eps_eff_re || !S->eps_re || !S->angles))
    return fail(DTMMX_E_ARG, "null stack array");
  if (ws_bytes < dtmmx_plan_workspace_size(P) || (dtmmx_plan_workspace_size(P) && !ws))
    return fail(DTMMX_E_ARG, "workspace too small");
  if (fin == fout) return fail(DTMMX_E_ARG, "field_in and field_out must not alias");
  hipStream_t s = (hipStream_t)stream;
  if (P->d.precision == DTMMX_C64) return transfer_impl<float>(P, S, fin, fout, ws, status, s);
  return transfer_impl<double>(P, S, fin, fout, ws, status, s);
}

int dtmmx_timing_begin(int capacity) {
  if (capacity <= 0) return fail(DTMMX_E_ARG, "capacity must be positive");
  for (auto e : g_timer.ev) (void)hipEventDestroy(e);
  g_timer.ev.assign(2 * size_t(capacity), nullptr);
  g_timer.cls.assign(capacity, -1);
  for (auto& e : g_timer.ev) HIP_TRY(hipEventCreate(&e));
  g_timer.cap = capacity;
  g_timer.used = 0;
  g_timer.on = true;
  return DTMMX_OK;
}

int dtmmx_timing_end(double* ms, int* count, int nclass) {
  if (!g_timer.on) return fail(DTMMX_E_ARG, "timing not started");
  g_timer.on = false;
  for (int c = 0; c < nclass; ++c) {
    ms[c] = 0.;
    count[c] = 0;
  }
  int rc = DTMMX_OK;
  for (int i = 0; i < g_timer.used; ++i) {
    float t = 0.f;
    hipError_t e = hipEventSynchronize(g_timer.ev[2 * i + 1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&t, g_timer.ev[2 * i], g_timer.ev[2 * i + 1]);
    if (e != hipSuccess) {
      rc = fail(DTMMX_E_HIP, std::string("timing: ") + hipGetErrorString(e));
      break;
    }
    const int c = g_timer.cls[i];
    if (c >= 0 && c < nclass) {
      ms[c] += t;
      count[c] += 1;
    }
  }
  for (auto e : g_timer.ev) (void)hipEventDestroy(e);
  g_timer.ev.clear();
  g_timer.cls.clear();
  g_timer.cap = g_timer.used = 0;
  return rc;
}

int dtmmx_fft2(void* data, int batch, int height, int width, int inverse, int precision,
               void* stream) {
  if (!data || batch <= 0) return fail(DTMMX_E_ARG, "bad fft2 arguments");
  if (!size_supported(height) || !size_supported(width))
    return fail(DTMMX_E_UNSUPPORTED, "fft size not built");
  hipStream_t s = (hipStream_t)stream;
  if (precision == DTMMX_C64) return fft2_impl<float>(data, batch, height, width, inverse, s);
  if (precision == DTMMX_C128) return fft2_impl<double>(data, batch, height, width, inverse, s);
  return fail(DTMMX_E_ARG, "bad precision");
}

int dtmmx_host_alphaf(int n, const double* beta, const double* phi, const double* eps_re,
                      const double* eps_im, const double* angles, double* alpha, double* fmat) {
  if (n < 0 || !beta || !phi || !eps_re || !angles || !alpha || !fmat)
    return fail(DTMMX_E_ARG, "null argument");
  int rc = DTMMX_OK;
  for (int i = 0; i < n; ++i) {
    cd e[3];
    for (int c = 0; c < 3; ++c) e[c] = mk<double>(eps_re[3 * i + c], eps_im ? eps_im[3 * i + c] : 0.);
    Modes m;
    int kind;
    layer_modes(beta[i], std::cos(phi[i]), std::sin(phi[i]), phi[i], e[0], e[1], e[2],
                angles[3 * i + 1], angles[3 * i + 2], m, kind);
    if (kind == KIND_BIAXIAL) rc = fail(DTMMX_E_UNSUPPORTED, "biaxial voxel");
    for (int j = 0; j < 4; ++j) {
      alpha[8 * i + 2 * j] = m.a[j].x;
      alpha[8 * i + 2 * j + 1] = m.a[j].y;
      for (int k = 0; k < 4; ++k) {
        fmat[32 * i + 8 * j + 2 * k] = m.F[j][k].x;
        fmat[32 * i + 8 * j + 2 * k + 1] = m.F[j][k].y;
      }
    }
  }
  return rc;
}

}  // extern "C"
