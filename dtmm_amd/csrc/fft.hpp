// LDS-staged mixed-radix Stockham FFT for one work-group.
//
// A work-group transforms NL independent lines of length N that sit in LDS
// (line l at buf + l * lstride).  Each stage reads R inputs per butterfly
// from LDS into registers, barriers, applies twiddles and the radix-R DFT,
// writes the outputs to their self-sorted positions and barriers again, so a
// single buffer suffices (no ping-pong copy).  Lengths are 2^a 3^b 5^c; the
// radix plan is fixed at compile time (8s first, then one 4 or 2, then 3s,
// then 5s).  Transforms are unnormalised; DIR = -1 is the forward transform
// exp(-2 pi i jk/N), DIR = +1 the inverse.
#pragma once
#include "cpx.hpp"

namespace dtmmx {

struct RadixPlan {
  int n;
  int ns;
  int r[16];
};

__host__ __device__ constexpr RadixPlan make_plan(int n) {
  RadixPlan p{n, 0, {0}};
  int m = n;
  int e2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++e2;
  }
  for (int i = 0; i < e2 / 3; ++i) p.r[p.ns++] = 8;
  if (e2 % 3 == 1) p.r[p.ns++] = 2;
  if (e2 % 3 == 2) p.r[p.ns++] = 4;
  while (m % 3 == 0) {
    m /= 3;
    p.r[p.ns++] = 3;
  }
  while (m % 5 == 0) {
    m /= 5;
    p.r[p.ns++] = 5;
  }
  if (m != 1) p.ns = -1;  // unsupported length
  return p;
}

__host__ __device__ constexpr int plan_prod(RadixPlan p, int s) {
  int v = 1;
  for (int i = 0; i < s; ++i) v *= p.r[i];
  return v;
}

template <int DIR, typename T>
__device__ __forceinline__ cpx<T> mul_i(cpx<T> a) {  // a * (DIR * i)
  return DIR > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

template <int DIR, typename T>
__device__ __forceinline__ void bfly2(cpx<T>& a, cpx<T>& b) {
  cpx<T> t = a - b;
  a = a + b;
  b = t;
}

template <int DIR, typename T>
__device__ __forceinline__ void bfly4(cpx<T>& x0, cpx<T>& x1, cpx<T>& x2, cpx<T>& x3) {
  cpx<T> t0 = x0 + x2, t1 = x0 - x2, t2 = x1 + x3, t3 = mul_i<DIR>(x1 - x3);
  x0 = t0 + t2;
  x2 = t0 - t2;
  x1 = t1 + t3;
  x3 = t1 - t3;
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void butterfly(cpx<T>* v) {
  if constexpr (R == 2) {
    bfly2<DIR>(v[0], v[1]);
  } else if constexpr (R == 3) {
    const T h = T(0.86602540378443864676);  // sin(2 pi / 3)
    cpx<T> t = v[1] + v[2];
    cpx<T> s = mul_i<DIR>(v[1] - v[2]) * h;
    cpx<T> m = v[0] - t * T(0.5);
    v[0] = v[0] + t;
    v[1] = m + s;
    v[2] = m - s;
  } else if constexpr (R == 4) {
    bfly4<DIR>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 5) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
    cpx<T> t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
    cpx<T> a1 = v[0] + t1 * c1 + t2 * c2;
    cpx<T> a2 = v[0] + t1 * c2 + t2 * c1;
    cpx<T> b1 = mul_i<DIR>(t3 * s1 + t4 * s2);
    cpx<T> b2 = mul_i<DIR>(t3 * s2 - t4 * s1);
    v[0] = v[0] + t1 + t2;
    v[1] = a1 + b1;
    v[4] = a1 - b1;
    v[2] = a2 + b2;
    v[3] = a2 - b2;
  } else if constexpr (R == 8) {
    const T r = T(0.70710678118654752440);
    bfly4<DIR>(v[0], v[2], v[4], v[6]);
    bfly4<DIR>(v[1], v[3], v[5], v[7]);
    // twiddle the odd half by w8^k, w8 = exp(DIR 2 pi i / 8)
    cpx<T> o1 = mk<T>((v[3].x - T(DIR) * v[3].y) * r, (v[3].y + T(DIR) * v[3].x) * r);
    cpx<T> o2 = mul_i<DIR>(v[5]);
    cpx<T> o3 = mk<T>((-v[7].x - T(DIR) * v[7].y) * r, (-v[7].y + T(DIR) * v[7].x) * r);
    cpx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
    v[0] = e0 + o0;
    v[4] = e0 - o0;
    v[1] = e1 + o1;
    v[5] = e1 - o1;
    v[2] = e2 + o2;
    v[6] = e2 - o2;
    v[3] = e3 + o3;
    v[7] = e3 - o3;
  }
}

// One Stockham stage of radix R with PP = product of the earlier radices.
template <typename T, int N, int R, int PP, int NL, int NT, int DIR>
__device__ __forceinline__ void stockham_stage(cpx<T>* buf, int lstride,
                                               const cpx<T>* __restrict__ tw, int tid) {
  constexpr int NB = N / R;
  constexpr int JOBS = NL * NB;
  constexpr int JPT = (JOBS + NT - 1) / NT;
  cpx<T> v[JPT][R];
#pragma unroll
  for (int jj = 0; jj < JPT; ++jj) {
    const int job = tid + jj * NT;
    if ((JOBS % NT) == 0 || job < JOBS) {
      const int line = job / NB;
      const int i = job - line * NB;
      const cpx<T>* b = buf + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) v[jj][r] = b[i + r * NB];
    }
  }
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < JPT; ++jj) {
    const int job = tid + jj * NT;
    if ((JOBS % NT) == 0 || job < JOBS) {
      const int line = job / NB;
      const int i = job - line * NB;
      const int k = i % PP;
      if constexpr (PP > 1) {
        constexpr int STEP = N / (PP * R);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cpx<T> w = tw[r * k * STEP];
          if (DIR > 0) w.y = -w.y;
          v[jj][r] = v[jj][r] * w;
        }
      }
      butterfly<R, DIR>(v[jj]);
      cpx<T>* b = buf + line * lstride;
      const int base = (i / PP) * PP * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) b[base + r * PP] = v[jj][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, int NL, int NT, int DIR, int S>
__device__ __forceinline__ void stockham_from(cpx<T>* buf, int lstride, const cpx<T>* tw, int tid) {
  constexpr RadixPlan P = make_plan(N);
  if constexpr (S < P.ns) {
    constexpr int R = P.r[S];
    constexpr int PP = plan_prod(P, S);
    stockham_stage<T, N, R, PP, NL, NT, DIR>(buf, lstride, tw, tid);
    stockham_from<T, N, NL, NT, DIR, S + 1>(buf, lstride, tw, tid);
  }
}

// Full in-LDS transform of NL lines.  The caller must have synchronised the
// work-group after filling `buf`; the routine ends with a barrier.
// `tw[t] = exp(-2 pi i t / N)`, t < N.
template <typename T, int N, int NL, int NT, int DIR>
__device__ __forceinline__ void lds_fft(cpx<T>* buf, int lstride, const cpx<T>* __restrict__ tw,
                                        int tid) {
  static_assert(make_plan(N).ns > 0, "FFT length must be 2^a 3^b 5^c");
  stockham_from<T, N, NL, NT, DIR, 0>(buf, lstride, tw, tid);
}

}  // namespace dtmmx
