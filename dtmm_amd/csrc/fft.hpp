// Mixed-radix Stockham FFT building blocks (radix 2, 3, 4, 5, 8).
//
// Lengths are 2^a 3^b 5^c; the radix plan is fixed at compile time (8s
// first, then one 4 or 2, then 3s, then 5s).  Transforms are unnormalised;
// DIR = -1 is the forward transform exp(-2 pi i jk/N), DIR = +1 the inverse.
// Stage s with radix R and PP = product of the earlier radices maps
// butterfly i (of N/R) from inputs x[i + r N/R] to outputs
// y[(i/PP) PP R + i%PP + r PP] after twiddling input r by w_N^(r (i%PP) N/(PP R)).
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

// ---- register-resident transforms ----

// ---------------------------------------------------------------------------
// Register-resident Stockham transforms.
//
// A group of NL lanes transforms one line of length N.  Stage s uses radix
// R_s of the plan (REV selects the reversed radix order), NB_s = N / R_s
// butterflies, and lane l owns butterflies i = l + NL*m.  Before a stage the
// lane holds the stage inputs v[m*R + r] = x[i + r*NB]; after the last stage
// it holds the outputs in natural order, v[m*R + r] = y[i + r*NB_last].
// Stages exchange data through a per-line LDS buffer (addressed through a
// stride-and-pad map chosen by the caller); the first stage inputs and the
// last stage outputs therefore never touch LDS.  Using the forward plan for
// one transform and the reversed plan for the next one makes the last radix
// of the first equal the first radix of the second, so a pointwise operator
// between them applies to registers without any data movement.
// ---------------------------------------------------------------------------

template <int N, bool REV>
__host__ __device__ constexpr int radix_at(int s) {
  return REV ? make_plan(N).r[make_plan(N).ns - 1 - s] : make_plan(N).r[s];
}

template <int N, bool REV>
__host__ __device__ constexpr int pp_at(int s) {
  int v = 1;
  for (int i = 0; i < s; ++i) v *= radix_at<N, REV>(i);
  return v;
}

template <int N, bool REV, int NL>
__host__ __device__ constexpr int max_vals() {
  int m = 0;
  for (int s = 0; s < make_plan(N).ns; ++s) {
    int R = radix_at<N, REV>(s);
    int J = (N / R + NL - 1) / NL;
    if (J * R > m) m = J * R;
  }
  return m;
}

// LDS address of line position x: x * STRIDE + (x >> DTMMX_LDS_PAD_SHIFT) *
// PADW; PADW padding slots after every 2^shift positions (8 by default)
// break up the power-of-two strides of the stage exchanges across the 64
// LDS banks (0 disables padding).
#ifndef DTMMX_LDS_PAD_SHIFT
#define DTMMX_LDS_PAD_SHIFT 3
#endif
template <int STRIDE, int PADW>
__device__ __forceinline__ int lds_at(int x) {
  return x * STRIDE + (PADW ? (x >> DTMMX_LDS_PAD_SHIFT) * PADW : 0);
}

// Twiddle tables in LDS, one per radix plan (forward order and, unless the
// plan reads the same backwards, reversed order).  Stage s >= 1 of a plan
// with radix R, PP = product of the earlier radices and STEP = N / (PP R)
// reads w^(r k STEP) for r = 1..R-1 and k = i % PP; it is stored at
// r PP + k - R0 (R0 = first radix: the stage blocks (R - 1) PP telescope to
// PP - R0).  A lane reads k + a compile-time offset, so the address is one
// register plus immediate offsets, and lanes with consecutive k read
// consecutive slots (no bank conflicts).
template <int N>
__host__ __device__ constexpr bool plan_palindrome() {
  for (int s = 0; s < make_plan(N).ns; ++s)
    if (radix_at<N, false>(s) != radix_at<N, true>(s)) return false;
  return true;
}
template <int N, bool REV>
__host__ __device__ constexpr int tw_table_off() {
  return (REV && !plan_palindrome<N>()) ? N - radix_at<N, false>(0) : 0;
}
template <int N>
__host__ __device__ constexpr int tw_table_slots() {
  const int f = N - radix_at<N, false>(0);
  const int n = plan_palindrome<N>() ? f : f + N - radix_at<N, true>(0);
  return n > 0 ? n : 1;
}

template <typename T, int N, bool REV, int S>
__device__ __forceinline__ void fill_tw_stage(cpx<T>* tws, const cpx<T>* tw, int tid, int nth) {
  if constexpr (S < make_plan(N).ns) {
    constexpr int R = radix_at<N, REV>(S), PP = pp_at<N, REV>(S);
    constexpr int STEP = N / (PP * R);
    constexpr int BASE = tw_table_off<N, REV>() + PP - radix_at<N, REV>(0);
    for (int e = tid; e < (R - 1) * PP; e += nth) {
      const int r = 1 + e / PP, k = e % PP;
      tws[BASE + e] = tw[r * k * STEP];
    }
    fill_tw_stage<T, N, REV, S + 1>(tws, tw, tid, nth);
  }
}

// Two-phase fill of the twiddle tables by NTH threads: load() reads this
// thread's entries into registers, store() writes them to LDS.  Vector loads
// complete in issue order, so a kernel issues these small (L2-resident)
// loads first, then its large field loads, and stores the table after its
// setup arithmetic without waiting for the field.
template <int N, bool REV, int S, int NTH>
__host__ __device__ constexpr int tw_regs_from() {
  int n = 0;
  for (int s = S; s < make_plan(N).ns; ++s)
    n += ((radix_at<N, REV>(s) - 1) * pp_at<N, REV>(s) + NTH - 1) / NTH;
  return n;
}
template <int N, int NTH>
__host__ __device__ constexpr int tw_regs() {
  return tw_regs_from<N, false, 1, NTH>() + (plan_palindrome<N>() ? 0 : tw_regs_from<N, true, 1, NTH>());
}

template <typename T, int N, bool REV, int S, int NTH, int Q, int MAXE>
__device__ __forceinline__ void tw_load_stage(cpx<T> (&r)[MAXE], const cpx<T>* tw, int tid) {
  if constexpr (S < make_plan(N).ns) {
    constexpr int R = radix_at<N, REV>(S), PP = pp_at<N, REV>(S);
    constexpr int STEP = N / (PP * R), CNT = (R - 1) * PP, IT = (CNT + NTH - 1) / NTH;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int e = tid + i * NTH;
      if (e < CNT) r[Q + i] = tw[(1 + e / PP) * (e % PP) * STEP];
    }
    tw_load_stage<T, N, REV, S + 1, NTH, Q + IT>(r, tw, tid);
  }
}

template <typename T, int N, bool REV, int S, int NTH, int Q, int MAXE>
__device__ __forceinline__ void tw_store_stage(const cpx<T> (&r)[MAXE], cpx<T>* tws, int tid) {
  if constexpr (S < make_plan(N).ns) {
    constexpr int R = radix_at<N, REV>(S), PP = pp_at<N, REV>(S);
    constexpr int CNT = (R - 1) * PP, IT = (CNT + NTH - 1) / NTH;
    constexpr int BASE = tw_table_off<N, REV>() + PP - radix_at<N, REV>(0);
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int e = tid + i * NTH;
      if (e < CNT) tws[BASE + e] = r[Q + i];
    }
    tw_store_stage<T, N, REV, S + 1, NTH, Q + IT>(r, tws, tid);
  }
}

template <typename T, int N, int NTH>
struct TwFill {
  static constexpr int MAXE = tw_regs<N, NTH>() > 0 ? tw_regs<N, NTH>() : 1;
  cpx<T> r[MAXE];
  __device__ __forceinline__ void load(const cpx<T>* tw, int tid) {
    tw_load_stage<T, N, false, 1, NTH, 0>(r, tw, tid);
    if constexpr (!plan_palindrome<N>())
      tw_load_stage<T, N, true, 1, NTH, tw_regs_from<N, false, 1, NTH>()>(r, tw, tid);
  }
  __device__ __forceinline__ void store(cpx<T>* tws, int tid) const {
    tw_store_stage<T, N, false, 1, NTH, 0>(r, tws, tid);
    if constexpr (!plan_palindrome<N>())
      tw_store_stage<T, N, true, 1, NTH, tw_regs_from<N, false, 1, NTH>()>(r, tws, tid);
  }
};

// Fills the twiddle tables of length-N plans from the plain table
// tw[j] = exp(-2 pi i j / N) (nth threads, thread tid; the caller
// synchronises before the first transform).
template <typename T, int N>
__device__ __forceinline__ void fill_tw_tables(cpx<T>* tws, const cpx<T>* tw, int tid, int nth) {
  fill_tw_stage<T, N, false, 1>(tws, tw, tid, nth);
  if constexpr (!plan_palindrome<N>()) fill_tw_stage<T, N, true, 1>(tws, tw, tid, nth);
}

// Barrier between LDS write and read phases: the whole work-group, or only
// the wave when a line is owned by a single wave (LDS operations of one wave
// execute in order, so a compiler fence at wavefront scope is enough).
template <bool WAVE>
__device__ __forceinline__ void exchange_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <typename T, int N, bool REV, int DIR, int NL, int STRIDE, int PADW, int S, int MAXV,
          bool WAVE = false>
__device__ __forceinline__ void reg_stages(cpx<T> (&v)[MAXV], cpx<T>* lds,
                                           const cpx<T>* __restrict__ tw, int l) {
  constexpr int NS = make_plan(N).ns;
  constexpr int R = radix_at<N, REV>(S);
  constexpr int PP = pp_at<N, REV>(S);
  constexpr int NB = N / R;
  constexpr int J = (NB + NL - 1) / NL;
#pragma unroll
  for (int m = 0; m < J; ++m) {
    const int i = l + NL * m;
    if ((NB % NL) == 0 || i < NB) {
      if constexpr (PP > 1) {
        constexpr int STEP = N / (PP * R);
        const int k = i % PP;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cpx<T> w = tw[tw_table_off<N, REV>() + r * PP + k - radix_at<N, REV>(0)];
          if (DIR > 0) w.y = -w.y;
          v[m * R + r] = v[m * R + r] * w;
        }
      }
      butterfly<R, DIR>(&v[m * R]);
    }
  }
  if constexpr (S + 1 < NS) {
#pragma unroll
    for (int m = 0; m < J; ++m) {
      const int i = l + NL * m;
      if ((NB % NL) == 0 || i < NB) {
        const int base = (i / PP) * PP * R + (i % PP);
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lds_at<STRIDE, PADW>(base + r * PP)] = v[m * R + r];
      }
    }
    exchange_sync<WAVE>();
    constexpr int R2 = radix_at<N, REV>(S + 1);
    constexpr int NB2 = N / R2;
    constexpr int J2 = (NB2 + NL - 1) / NL;
#pragma unroll
    for (int m = 0; m < J2; ++m) {
      const int i = l + NL * m;
      if ((NB2 % NL) == 0 || i < NB2) {
#pragma unroll
        for (int r = 0; r < R2; ++r) v[m * R2 + r] = lds[lds_at<STRIDE, PADW>(i + r * NB2)];
      }
    }
    exchange_sync<WAVE>();
    reg_stages<T, N, REV, DIR, NL, STRIDE, PADW, S + 1, MAXV, WAVE>(v, lds, tw, l);
  }
}

template <typename T, int N, bool REV, int DIR, int NL, int STRIDE, int PADW, int MAXV,
          bool WAVE = false>
__device__ __forceinline__ void reg_fft(cpx<T> (&v)[MAXV], cpx<T>* lds,
                                        const cpx<T>* __restrict__ tw, int l) {
  static_assert(make_plan(N).ns > 0, "FFT length must be 2^a 3^b 5^c");
  reg_stages<T, N, REV, DIR, NL, STRIDE, PADW, 0, MAXV, WAVE>(v, lds, tw, l);
}

// Position of value slot j (= m*R + r) of the first stage (inputs) or of the
// last stage (natural-order outputs): i + r*NB with i = l + NL*m.
template <int N, bool REV, int NL, bool LAST>
__device__ __forceinline__ int slot_pos(int j, int l) {
  constexpr int S = LAST ? make_plan(N).ns - 1 : 0;
  constexpr int R = radix_at<N, REV>(S);
  constexpr int NB = N / R;
  const int m = j / R, r = j - (j / R) * R;
  return l + NL * m + r * NB;
}

template <int N, bool REV, int NL, bool LAST>
__device__ __forceinline__ bool slot_valid(int j, int l) {
  constexpr int S = LAST ? make_plan(N).ns - 1 : 0;
  constexpr int R = radix_at<N, REV>(S);
  constexpr int NB = N / R;
  const int m = j / R;
  return (NB % NL) == 0 || (l + NL * m) < NB;
}

template <int N, bool REV, int NL, bool LAST>
__host__ __device__ constexpr int slot_count() {
  constexpr int S = LAST ? make_plan(N).ns - 1 : 0;
  constexpr int R = radix_at<N, REV>(S);
  return ((N / R + NL - 1) / NL) * R;
}

}  // namespace dtmmx
