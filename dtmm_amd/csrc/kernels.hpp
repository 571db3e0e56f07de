// Device kernels of the MI355X (gfx950) diffractive transfer-matrix engine.
// Included by dtmmx.hip (host driver and C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "berreman.hpp"
#include "cpx.hpp"
#include "rop_grad.hpp"
#include "dtmmx.h"
#include "fft.hpp"

namespace dtmmx {


constexpr int NT = 256;  // threads per work-group for every kernel
constexpr int MAXL = DTMMX_MAX_WAVELENGTHS;
constexpr int MAXB = DTMMX_MAX_BEAMS;

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
// Launch geometry shared by the per-layer kernels.
// ---------------------------------------------------------------------------

enum { OP_FWD = 0, OP_INV = 1, OP_ROP = 2, OP_DIAG = 3, OP_WIN = 4, OP_SPLIT = 5, OP_ROPW = 6,
       OP_FIELD = 7, OP_SPEC = 8, OP_COMB = 9 };

// Work-item index for block b of n such that consecutive work items land on
// the same XCD (blocks are dispatched round-robin over the 8 XCDs): work
// that shares data (the planes of one row) then shares one L2.
// Marks the prefetched values v[0..n) as needed here: the compiler waits for
// their loads at this point (called on every line, also where nothing was
// prefetched, so that no path reaches the stores with loads pending).  The per-line kernels call it before storing a
// line, because the counter of outstanding vector memory operations also
// counts stores and drains in issue order: a wait for the prefetch placed
// after the stores (where the compiler would put it, at the loop's end)
// would also wait for the stores to be acknowledged, exposing their latency
// on every line.  The memory clobber keeps the stores after the wait.
template <int NV, typename V>
__device__ __forceinline__ void complete_loads(const V (&v)[NV], int n) {
#pragma unroll
  for (int j = 0; j < NV; ++j)
    if (j < n) asm volatile("" ::"v"(v[j].x), "v"(v[j].y) : "memory");
}

// Wave vote in a scalar register: true when every active lane's p holds.
// The result of __all is uniform, but the compiler does not know it; read
// from the first lane it is, so a branch on it stays a scalar branch
// instead of being flattened into both paths with per-lane selects.
__device__ __forceinline__ bool wave_all(bool p) {
  return __builtin_amdgcn_readfirstlane(__all(p)) != 0;
}

__device__ __forceinline__ int xcd_swizzle(int b, int n) {
  const int x = b & 7, k = b >> 3, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + k;
}

// Planes of the current batch: ray r (0..nrays-1) and wavelength lam0 + j
// (j < nlam_b); plane index q = r * nlam + lam0 + j in the field layout
// [ray][wavelength][component][H][W].
struct Batch {
  int nrays;
  int nlam;
  int lam0, nlam_b;
};

constexpr int EIG_COMPS = 13;  // per-pixel operator data: alpha0..3, P0, P1, ok

// ---------------------------------------------------------------------------
// Row kernel: a work-group of NR*NC waves owns one row y of NR rays sharing
// one beam and loops over the wavelengths of the batch; wave li transforms
// component line li with its values in registers (first-stage inputs read
// from HBM, last-stage outputs written to HBM).  OP_ROP: inverse
// x-transform, per-pixel layer operator (through LDS, eigen data of the
// row held in registers across wavelengths), forward x-transform.
// ---------------------------------------------------------------------------

// LDS address maps of the row kernel's layer-operator exchange (padding
// slots every 2^DTMMX_LDS_PAD_SHIFT positions: 1, or none: 0) and of the
// column kernel's transform exchanges (0: x * TX + c; 1: padded).
#ifndef DTMMX_ROP_PADW
#define DTMMX_ROP_PADW 1
#endif
#ifndef DTMMX_COL_PADW
#define DTMMX_COL_PADW 0
#endif

// Poynting z of one (Ex, Hy, Ey, Hx) sample (dtmm/field.py:771-799).
__host__ __device__ inline double intensity4(const cd v[4]) {
  return (v[0].x * v[1].x + v[0].y * v[1].y) - (v[2].x * v[3].x + v[2].y * v[3].y);
}

__device__ __forceinline__ void mat4v(const cd* m, const cd v[4], cd w[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    cd t = mk<double>(0.);
#pragma unroll
    for (int j = 0; j < 4; ++j) t = t + m[i * 4 + j] * v[j];
    w[i] = t;
  }
}

template <typename T>
struct RowArgs {
  cpx<T>* field;
  int H, W;
  Batch b;
  int rays_per_group;   // NR
  const cpx<T>* tw;     // twiddles of length W
  // layer operator (OP_ROP): material of this layer, [3][L*H*W] fp64
  const double* eps_re;
  const double* eps_im;
  const double* ang;
  size_t comp_stride;   // L*H*W
  double d;             // layer thickness
  int method4;          // 1: 4x4 (E invertible), 0: 2x2 (E+ invertible)
  int bwd2;             // 2x2 (NC == 2): 1 propagates the backward modes (mode -1)
  int* status;
  // OP_ROP output: acc_mode 0 writes the result back to `field`; 1 writes
  // it to `acc`; 2 adds it to `acc` (sum over k-space windows)
  cpx<T>* acc;
  int acc_mode;
  // OP_ROPW (OP_ROP with an input window): the input line (indexed by kx)
  // is multiplied by the tent factor cx [beam][lam][wn][W] of window column
  // wi and by the window's mean-beam correction
  // exp(-+i k0 wd sqrt(we - beta^2)) on forward / backward components
  const double* wcx;
  int wn, wi;
  cd we;
  double wd;
  // OP_ROP: when set, the input lines are also stored here (field layout),
  // the adjoint's record of the row-pass input
  cpx<T>* save;
  // OP_ROP / OP_ROPW: per-pixel operator data of this layer precomputed by
  // rop_table_kernel ([EIG_COMPS][etab_nbeam][H*W], layers with biaxial
  // voxels); NULL: solved in the kernel.  etab_ws (host side): table space
  // to fill before the launch.
  const cpx<T>* etab;
  int etab_nbeam;
  cpx<T>* etab_ws;
  // OP_SPEC (the polarizing-microscope specter after the inverse x
  // transforms): specter [H*W][nlam] (the sum over the work-group's rays of
  // the Poynting z of the analysed field), the analyser: none (npm 0), the
  // shared matrix pm0 (npm 1) or per wavelength pmat [nlam][4][4] (npm > 1)
  T* spec;
  int spec_npm;
  const cd* spec_pmat;
  cd spec_pm0[16];
  // OP_COMB (forward x transforms of the Jones combination of two polarised
  // fields): source [ray][2][nlam][NC][H][W] and the weights of x and y
  const cpx<T>* csrc;
  cd cjc, cjs;
  Beams bm;
  // work-groups per (row, ray group): each takes a contiguous chunk of the
  // batch's wavelengths (values below 2: one work-group takes them all).
  // More chunks mean more work-groups in flight for grids that would
  // otherwise leave SIMDs idle, at the price of one more per-pixel solve per
  // chunk in the layer operator kernels.
  int lam_split;
};

template <int N>
__host__ __device__ constexpr int line_stride() {  // per-line LDS footprint with padding
  return N + N / 8 + 1;
}

// Minimum waves per SIMD the register allocator must allow: 4 for complex64
// (at most 128 VGPRs), 2 for complex128.
template <typename T>
__host__ __device__ constexpr int min_waves() {
  return sizeof(T) == 4 ? 4 : 2;
}

// Lanes per row line: one wave up to 512 points; at 1024 points two waves
// share a line (8 values per lane instead of 16, and half the pixels of the
// layer operator per lane), exchanging through LDS with work-group barriers.
template <int N>
__host__ __device__ constexpr int row_lanes() {
  return N >= 1024 ? 128 : 64;
}

// Row kernel of 1024 points: the line takes 8 values per lane but the eigen
// data of 2 pixels per lane takes 52 VGPRs, so the budget is 256 VGPRs
// (2 waves per SIMD; complex128: 1).
template <typename T, int N>
__host__ __device__ constexpr int row_min_waves() {
  return N >= 1024 ? (sizeof(T) == 4 ? 2 : 1) : min_waves<T>();
}

// Column kernel of 1024 points: 168 VGPRs (3 waves per SIMD) hold the
// diagonal factor, the line and the prefetched line without spilling.
template <typename T, int N, int OP>
__host__ __device__ constexpr int col_min_waves() {
  return (OP == OP_SPLIT || OP == OP_FIELD) ? (sizeof(T) == 4 ? 2 : 1)
                        : ((N >= 1024 && sizeof(T) == 4) ? 3 : min_waves<T>());
}

// The 2x2 blocks p I + (p' - p) P of one pixel's forward (f) and backward
// (b, 4x4 only) modes for the phase kd: f = {f00, f01, f10, f11}.
// FAST: every pixel of the wave has real mode pairs a1 = -a0, a3 = -a2, real
// projectors and ok = 1 (lossless uniaxial or isotropic media at normal
// incidence): one sine and cosine per pair and real-by-complex products.
// The general path holds an empty volatile asm, which cannot be executed
// speculatively: without it the compiler merges the two paths into per-lane
// selects and every wave evaluates both.
template <typename T, int NC>
__device__ __forceinline__ void layer_blocks(const cpx<T>* e, T kd, bool fast, cpx<T> f[4],
                                             cpx<T> b[4]) {
  if (fast) {
    const cpx<T> p0 = cexpi_layer(mk<T>(e[0].x * kd, T(0)));
    const cpx<T> p2 = cexpi_layer(mk<T>(e[2].x * kd, T(0)));
    const cpx<T> dp = p0 - p2;
    f[0] = p2 + dp * e[4].x;
    f[1] = dp * e[5].x;
    f[2] = dp * e[6].x;
    f[3] = p2 + dp * e[7].x;
    if constexpr (NC == 4) {
      // p1 = conj(p0), p3 = conj(p2)
      const cpx<T> p3 = conj(p2), dn = conj(dp);
      b[0] = p3 + dn * e[8].x;
      b[1] = dn * e[9].x;
      b[2] = dn * e[10].x;
      b[3] = p3 + dn * e[11].x;
    }
  } else {
    asm volatile("");
    const T ok = e[12].x;
    cpx<T> p0 = cexpi_layer(e[0] * kd);
    cpx<T> p2 = cexpi_layer(e[2] * kd);
    const cpx<T> dp = (p0 - p2) * ok;
    p2 = p2 * ok;
    f[0] = p2 + dp * e[4];
    f[1] = dp * e[5];
    f[2] = dp * e[6];
    f[3] = p2 + dp * e[7];
    if constexpr (NC == 4) {
      const cpx<T> p1 = cexpi_layer(e[1] * kd);
      cpx<T> p3 = cexpi_layer(e[3] * kd);
      const cpx<T> dn = (p1 - p3) * ok;
      p3 = p3 * ok;
      b[0] = p3 + dn * e[8];
      b[1] = dn * e[9];
      b[2] = dn * e[10];
      b[3] = p3 + dn * e[11];
    }
  }
}

// Whether one pixel's operator data allows layer_blocks' fast path.
template <typename T, int NC>
__device__ __forceinline__ bool fast_blocks_ok(const cpx<T>* e) {
  bool f = e[0].y == T(0) && e[2].y == T(0) && e[12].x == T(1) && e[4].y == T(0) &&
           e[5].y == T(0) && e[6].y == T(0) && e[7].y == T(0);
  if constexpr (NC == 4)
    f = f && e[1].x == -e[0].x && e[1].y == T(0) && e[3].x == -e[2].x && e[3].y == T(0) &&
        e[8].y == T(0) && e[9].y == T(0) && e[10].y == T(0) && e[11].y == T(0);
  return f;
}

// NW > 1 (layer operator kernels with one wave per line): the work-group
// transforms NW wavelengths at once in separate waves, so that one fp64
// solve per pixel serves twice the waves (the two-component lines of the
// 2x2 method otherwise give work-groups of only 4 waves).
template <typename T, int N, int NR, int NC, int OP, int NW = 1>
__global__ __launch_bounds__(row_lanes<N>() * NR * NC * NW, (row_min_waves<T, N>())) void row_kernel(
    RowArgs<T> A) {
  constexpr int NL = row_lanes<N>();
  constexpr bool WV = (NL == 64);  // one wave per line: wave-local exchanges
  static_assert(NW == 1 || (WV && OP == OP_ROP), "wavelength-parallel waves: layer operator, one wave per line");
  constexpr int LINES = NR * NC * NW;
  constexpr int NTW = NL * LINES;
  constexpr int LS = line_stride<N>();
  constexpr bool FWDX = (OP == OP_FWD || OP == OP_COMB);  // forward x transforms
  constexpr bool REV1 = !FWDX;
  constexpr int DIR1 = FWDX ? -1 : +1;
  constexpr int MV1 = max_vals<N, REV1, NL>();
  constexpr int MV2 = max_vals<N, false, NL>();
  constexpr int MAXV = MV1 > MV2 ? MV1 : MV2;
  constexpr int NIN = slot_count<N, REV1, NL, false>();
  constexpr int NMID = slot_count<N, REV1, NL, true>();
  constexpr bool ROPX = (OP == OP_ROP || OP == OP_ROPW);  // layer operator kernels
  constexpr bool REVL = ROPX ? false : REV1;
  constexpr int NOUT = slot_count<N, REVL, NL, true>();
  constexpr int PPX = (N + NTW - 1) / NTW;  // pixels per thread in the layer operator
  __shared__ cpx<T> lds[LINES * LS];
  __shared__ cpx<T> tws[tw_table_slots<N>()];  // twiddles in LDS: their reads never wait on global loads
  const int tid = threadIdx.x;
  const int li = tid / NL, lane = tid % NL;
  // Prologue order where registers allow (EARLY): twiddles and this
  // thread's material first, then the first field line, then the fp64 solve
  // (loads complete in issue order, so the solve does not wait for the
  // line).  Otherwise the twiddle table is filled at once and the material
  // read inside the solve.
  constexpr bool EARLY = tw_regs<N, NTW>() <= 2 || row_min_waves<T, N>() <= 2;
  TwFill<T, N, NTW> twf;
  if constexpr (EARLY)
    twf.load(A.tw, tid);
  else
    fill_tw_tables<T, N>(tws, A.tw, tid, NTW);
  const int wl = li / (NR * NC);  // wavelength slot of this wave (NW > 1)
  const int rr = (li - wl * (NR * NC)) / NC, comp = li - wl * (NR * NC) - rr * NC;
  // consecutive work items (the wavelength chunks of one row first) land on
  // one XCD, so the chunks of a row share its material in that XCD's L2
  const int lsplit = A.lam_split > 1 ? A.lam_split : 1;
  const int wsw = xcd_swizzle(blockIdx.x, gridDim.x);
  const int chunk = wsw % lsplit;
  const int w = wsw / lsplit;
  const int ngroups = A.b.nrays / NR;
  const int rg = w % ngroups;
  const int y = w / ngroups;
  // this work-group's wavelengths jb <= jl < je of the batch
  const int lper = (A.b.nlam_b + lsplit - 1) / lsplit;
  const int jb = min(chunk * lper, A.b.nlam_b);
  const int je = min(jb + lper, A.b.nlam_b);
  const size_t HW = size_t(A.H) * N;
  cpx<T>* line_lds = lds + li * LS;

  constexpr int PPM = (ROPX && EARLY) ? PPX : 1;
  double me[PPM][6], mth[PPM], mph[PPM];  // eps (re, im) x 3, theta, phi
  auto load_material = [&](int k) {
    const size_t o = size_t(y) * N + tid + k * NTW;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      me[k][2 * q] = A.eps_re[q * A.comp_stride + o];
      me[k][2 * q + 1] = A.eps_im ? A.eps_im[q * A.comp_stride + o] : 0.;
    }
    mth[k] = A.ang[A.comp_stride + o];
    mph[k] = A.ang[2 * A.comp_stride + o];
  };
  if constexpr (ROPX && EARLY) {
    if (!A.etab) {
#pragma unroll
      for (int k = 0; k < PPX; ++k)
        if (tid + k * NTW < N) load_material(k);
    }
  }
  // line of wavelength lam0 + jl for this wave
  auto line_ptr = [&](int jl) {
    const size_t q = size_t(rg * NR + rr) * A.b.nlam + A.b.lam0 + jl;
    return A.field + (q * NC + comp) * HW + size_t(y) * N;
  };
  constexpr bool PF = true;
  // prefetched inputs of the next wavelength; the first one is requested
  // before the eigen data so that its latency overlaps the fp64 solve
  cpx<T> nxt[PF ? NIN : 1];
  // OP_COMB: the lines of the x and y polarised fields [ray][2][nlam][NC][H][W]
  // are read and combined (jc x + js y, fp64) on their way in
  cpx<T> nxt2[OP == OP_COMB ? NIN : 1];
  auto pol_ptr = [&](int jl) {
    const size_t q = size_t(rg * NR + rr) * 2 * A.b.nlam + A.b.lam0 + jl;
    return A.csrc + (q * NC + comp) * HW + size_t(y) * N;
  };
  const size_t ypol = size_t(A.b.nlam) * NC * HW;
  if constexpr (PF) {
    const cpx<T>* g0 = (OP == OP_COMB) ? pol_ptr(jb + wl) : line_ptr(jb + wl);
#pragma unroll
    for (int j = 0; j < NIN; ++j)
      if (jb + wl < je && slot_valid<N, REV1, NL, false>(j, lane)) {
        nxt[j] = g0[slot_pos<N, REV1, NL, false>(j, lane)];
        if constexpr (OP == OP_COMB) nxt2[j] = g0[ypol + slot_pos<N, REV1, NL, false>(j, lane)];
      }
  }
  // per-pixel operator data of this row, solved in fp64 once and reused for
  // every wavelength and ray of the work-group
  cpx<T> eg[ROPX ? PPX : 1][EIG_COMPS];
  if constexpr (ROPX) {
   if (A.etab) {
    const int beam = (rg * NR) / A.bm.rays_per_beam;
    const size_t cs = size_t(A.etab_nbeam) * HW;
#pragma unroll
    for (int k = 0; k < PPX; ++k) {
      const int x = tid + k * NTW;
      if (x < N) {
        const cpx<T>* t = A.etab + size_t(beam) * HW + size_t(y) * N + x;
        for (int m = 0; m < EIG_COMPS; ++m) eg[k][m] = t[m * cs];
      } else {
        for (int m = 0; m < EIG_COMPS; ++m) eg[k][m] = mk<T>(T(0), T(0));
      }
    }
   } else {
    const int beam = (rg * NR) / A.bm.rays_per_beam;
    const double beta = A.bm.beta[beam], phi0 = A.bm.phi[beam];
    double s0, c0;
    sincos(phi0, &s0, &c0);
#pragma unroll
    for (int k = 0; k < PPX; ++k) {
      const int x = tid + k * NTW;
      if (x < N) {
        RopData r;
        if constexpr (EARLY) {
          rop_lean(beta, c0, s0, phi0, mk<double>(me[k][0], me[k][1]), mk<double>(me[k][2], me[k][3]),
                   mk<double>(me[k][4], me[k][5]), mth[k], mph[k], r);
        } else {
          const size_t o = size_t(y) * N + x;
          const cd e0 = mk<double>(A.eps_re[o], A.eps_im ? A.eps_im[o] : 0.);
          const cd e1 = mk<double>(A.eps_re[A.comp_stride + o], A.eps_im ? A.eps_im[A.comp_stride + o] : 0.);
          const cd e2 = mk<double>(A.eps_re[2 * A.comp_stride + o],
                                   A.eps_im ? A.eps_im[2 * A.comp_stride + o] : 0.);
          rop_lean(beta, c0, s0, phi0, e0, e1, e2, A.ang[A.comp_stride + o],
                   A.ang[2 * A.comp_stride + o], r);
        }
        if (r.kind == KIND_BIAXIAL && A.status) atomicOr(A.status, 1);
        if constexpr (NC == 2) {
          // the 2x2 operator reads a[0], a[2] and P0: the backward modes'
          // a[1], a[3] and P1 take their place for mode -1
          if (A.bwd2) {
            const cd t0 = r.a[0], t2 = r.a[2];
            r.a[0] = r.a[1];
            r.a[2] = r.a[3];
            r.a[1] = t0;
            r.a[3] = t2;
            for (int i = 0; i < 2; ++i)
              for (int j = 0; j < 2; ++j) {
                const cd t = r.P0[i][j];
                r.P0[i][j] = r.P1[i][j];
                r.P1[i][j] = t;
              }
            r.ok2 = r.okb;
          }
        }
        for (int i = 0; i < 4; ++i) eg[k][i] = cast<T>(r.a[i]);
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) {
            eg[k][4 + 2 * i + j] = cast<T>(r.P0[i][j]);
            eg[k][8 + 2 * i + j] = cast<T>(r.P1[i][j]);
          }
        T ok = (A.method4 ? r.ok4 : r.ok2) ? T(1) : T(0);
        if (r.kind == KIND_BIAXIAL) ok = T(__builtin_nan(""));
        eg[k][12] = mk<T>(ok, T(0));
      } else {
        for (int m = 0; m < EIG_COMPS; ++m) eg[k][m] = mk<T>(T(0), T(0));
      }
    }
   }
  }
  if constexpr (EARLY) twf.store(tws, tid);  // read after the barrier below
  // Every pixel of this thread's wave allows layer_blocks' fast path
  // (lossless media at normal incidence).  A wave vote in a scalar register
  // keeps the choice uniform, so no lane runs both paths.
  bool fast_ph = false;
  if constexpr (ROPX) {
    bool f = true;
#pragma unroll
    for (int k = 0; k < PPX; ++k)
      if (tid + k * NTW < N) f = f && fast_blocks_ok<T, NC>(eg[k]);
    fast_ph = wave_all(f);
  }
  __syncthreads();

  const int acc_mode = ROPX ? A.acc_mode : 0;
  for (int j0 = jb; j0 < je; j0 += NW) {
    const int jl = j0 + wl;
    const bool act = jl < je;  // false only for the last slots when NW > 1
    const int lam = A.b.lam0 + (act ? jl : j0);
    cpx<T>* gl = line_ptr(jl);
    cpx<T> v[MAXV];
    if constexpr (PF) {
#pragma unroll
      for (int j = 0; j < NIN; ++j) {
        if constexpr (OP == OP_COMB)
          v[j] = cast<T>(A.cjc * cast<double>(nxt[j]) + A.cjs * cast<double>(nxt2[j]));
        else
          v[j] = nxt[j];
      }
      if (jl + NW < je) {
        const cpx<T>* gn = (OP == OP_COMB) ? pol_ptr(jl + NW) : line_ptr(jl + NW);
#pragma unroll
        for (int j = 0; j < NIN; ++j)
          if (slot_valid<N, REV1, NL, false>(j, lane)) {
            nxt[j] = gn[slot_pos<N, REV1, NL, false>(j, lane)];
            if constexpr (OP == OP_COMB) nxt2[j] = gn[ypol + slot_pos<N, REV1, NL, false>(j, lane)];
          }
      }
    } else {
#pragma unroll
      for (int j = 0; j < NIN; ++j)
        if (slot_valid<N, REV1, NL, false>(j, lane)) v[j] = gl[slot_pos<N, REV1, NL, false>(j, lane)];
    }
    if constexpr (OP == OP_ROP) {
      if (A.save && act) {
        cpx<T>* sl = A.save + (gl - A.field);
#pragma unroll
        for (int j = 0; j < NIN; ++j)
          if (slot_valid<N, REV1, NL, false>(j, lane)) sl[slot_pos<N, REV1, NL, false>(j, lane)] = v[j];
      }
    }
    if constexpr (OP == OP_ROPW) {
      // window column factor and mean-beam correction of this line
      const int beam = (rg * NR + rr) / A.bm.rays_per_beam;
      const T bw = T(A.bm.beta[beam]);
      const cpx<T> ac = csqrt(cast<T>(A.we) - mk<T>(bw * bw));
      const cpx<T> cr = cexpi_layer(ac * T(((comp & 1) ? 1. : -1.) * A.bm.k0[lam] * A.wd));
      const double* cxl = A.wcx + ((size_t(beam) * A.b.nlam + lam) * A.wn + A.wi) * N;
#pragma unroll
      for (int j = 0; j < NIN; ++j)
        if (slot_valid<N, REV1, NL, false>(j, lane))
          v[j] = v[j] * (cr * T(cxl[slot_pos<N, REV1, NL, false>(j, lane)]));
    }
    reg_fft<T, N, REV1, DIR1, NL, 1, 1, MAXV, WV>(v, line_lds, tws, lane);

    if constexpr (OP == OP_SPEC) {
      // the real-space lines of the rays into LDS, then per pixel the sum
      // over rays of the Poynting z of the analysed field; no field store
#pragma unroll
      for (int j = 0; j < NMID; ++j)
        if (slot_valid<N, REV1, NL, true>(j, lane))
          line_lds[lds_at<1, DTMMX_ROP_PADW>(slot_pos<N, REV1, NL, true>(j, lane))] = v[j];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < PPX; ++k) {
        const int x = tid + k * NTW;
        if (x < N) {
          const int ax = lds_at<1, DTMMX_ROP_PADW>(x);
          double acc = 0.;
          for (int r = 0; r < NR; ++r) {
            const cpx<T>* L = lds + (r * 4) * LS + ax;
            const cd u[4] = {cast<double>(L[0]), cast<double>(L[LS]), cast<double>(L[2 * LS]),
                             cast<double>(L[3 * LS])};
            if (A.spec_npm == 0) {
              acc += intensity4(u);
            } else if (A.spec_npm == 1) {
              cd w[4];
              mat4v(A.spec_pm0, u, w);
              acc += intensity4(w);
            } else {
              cd w[4];
              mat4v(A.spec_pmat + size_t(lam) * 16, u, w);
              acc += intensity4(w);
            }
          }
          A.spec[(size_t(y) * N + x) * A.b.nlam + lam] = T(acc);
        }
      }
      __syncthreads();
      continue;
    }

    if constexpr (ROPX) {
#pragma unroll
      for (int j = 0; j < NMID; ++j)
        if (slot_valid<N, REV1, NL, true>(j, lane))
          line_lds[lds_at<1, DTMMX_ROP_PADW>(slot_pos<N, REV1, NL, true>(j, lane))] = v[j];
      __syncthreads();
#pragma unroll
      for (int ws = 0; ws < NW; ++ws) {
       if (j0 + ws < je) {
        const T kd = T(A.bm.k0[A.b.lam0 + j0 + ws] * A.d);
        cpx<T>* lw = lds + (ws * NR * NC) * LS;  // the lines of wavelength slot ws
#pragma unroll
      for (int k = 0; k < PPX; ++k) {
        const int x = tid + k * NTW;
        if (x < N) {
          const int ax = lds_at<1, DTMMX_ROP_PADW>(x);
          cpx<T> fb[4], bb[4];
          layer_blocks<T, NC>(eg[k], kd, fast_ph, fb, bb);
          const cpx<T> f00 = fb[0], f01 = fb[1], f10 = fb[2], f11 = fb[3];
          if constexpr (NC == 4) {
            const cpx<T> b00 = bb[0], b01 = bb[1], b10 = bb[2], b11 = bb[3];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              cpx<T>* L = lw + (r * 4) * LS + ax;
              const cpx<T> u0 = L[0], u1 = L[LS], u2 = L[2 * LS], u3 = L[3 * LS];
              L[0] = f00 * u0 + f01 * u2;
              L[2 * LS] = f10 * u0 + f11 * u2;
              L[LS] = b00 * u1 + b01 * u3;
              L[3 * LS] = b10 * u1 + b11 * u3;
            }
          } else {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              cpx<T>* L = lw + (r * 2) * LS + ax;
              const cpx<T> u0 = L[0], u1 = L[LS];
              L[0] = f00 * u0 + f01 * u1;
              L[LS] = f10 * u0 + f11 * u1;
            }
          }
        }
      }
       }
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < NMID; ++j)
        if (slot_valid<N, REV1, NL, true>(j, lane))
          v[j] = line_lds[lds_at<1, DTMMX_ROP_PADW>(slot_pos<N, REV1, NL, true>(j, lane))];
      exchange_sync<WV>();
      reg_fft<T, N, false, -1, NL, 1, 1, MAXV, WV>(v, line_lds, tws, lane);
    }
    if constexpr (PF) complete_loads(nxt, NIN);
    if (!act) {
    } else if (acc_mode == 0) {
#pragma unroll
      for (int j = 0; j < NOUT; ++j)
        if (slot_valid<N, REVL, NL, true>(j, lane)) gl[slot_pos<N, REVL, NL, true>(j, lane)] = v[j];
    } else {
      cpx<T>* al = A.acc + (gl - A.field);
      if (acc_mode == 2) {
#pragma unroll
        for (int j = 0; j < NOUT; ++j)
          if (slot_valid<N, REVL, NL, true>(j, lane)) {
            const int p = slot_pos<N, REVL, NL, true>(j, lane);
            al[p] = al[p] + v[j];
          }
      } else {
#pragma unroll
        for (int j = 0; j < NOUT; ++j)
          if (slot_valid<N, REVL, NL, true>(j, lane)) al[slot_pos<N, REVL, NL, true>(j, lane)] = v[j];
      }
    }
    exchange_sync<WV>();
  }
}

// ---------------------------------------------------------------------------
// Fused window row kernel (diffraction n > 1, single-step layers): every
// tent window of the layer in one launch.  The column split pass has written
// the n window rows t_j = IFFT_y(cy_j Q FFT_y(field)); window (i, j) needs
//   FFT_x( R_w IFFT_x( cx_i c_w t_j ) )
// with R_w the per-pixel layer operator of the window's mean beam and c_w
// its mean-beam correction, and the layer's output is the sum over windows
// (dtmm/propagate_4x4.py:390-419).  The forward transform is linear, so the
// sum is formed in real space and transformed once per line:
//   out = FFT_x( sum_w R_w IFFT_x( cx_i c_w t_j ) ).
// A work-group owns one row y, NR rays of one beam and a chunk of
// wavelengths; per window it solves the operator of its pixels once (fp64,
// reused for the chunk's wavelengths and rays) and adds R_w u to an
// accumulator kept in LDS, one slot per thread, pixel, line and wavelength.
// Each window row t_j is read from HBM once per layer (windows come in
// tent-row order, so the n windows of a row read the same lines back to
// back) and the output is written once.
// ---------------------------------------------------------------------------

template <typename T>
struct WinRowArgs {
  const cpx<T>* tmp;  // window rows [wn][ray][lam][comp][H][W]
  cpx<T>* out;        // layer output [ray][lam][comp][H][W]
  size_t fsz;         // elements of one window-row field
  int H, W;
  Batch b;
  int nchunk;         // wavelengths per work-group (C)
  const cpx<T>* tw;
  const double* eps_re;
  const double* eps_im;
  const double* ang;
  size_t comp_stride;
  double d;           // layer thickness (single step)
  int method4;
  int* status;
  const double* wcx;  // [beam][lam][wn][W]
  int wn;
  int nwin;
  const int* win_i;   // [nwin], tent-row order
  const int* win_j;
  const double* win_beta;  // [nwin][nbeam]
  const double* win_phi;
  int nbeam;
  cd we;              // window correction: effective eps and thickness
  double wd;
  Beams bm;
};

// The window sums fill the LDS, so one work-group runs per CU: the register
// budget is that of two waves per SIMD (complex128: one).
template <typename T, int N, int NR, int NC>
__global__ __launch_bounds__(row_lanes<N>() * NR * NC, (sizeof(T) == 4 ? 2 : 1)) void winrow_kernel(
    WinRowArgs<T> A) {
  constexpr int NL = row_lanes<N>();
  constexpr bool WV = (NL == 64);
  constexpr int LINES = NR * NC;
  constexpr int NTW = NL * LINES;
  constexpr int LS = line_stride<N>();
  constexpr int MV1 = max_vals<N, true, NL>();
  constexpr int MV2 = max_vals<N, false, NL>();
  constexpr int MAXV = MV1 > MV2 ? MV1 : MV2;
  constexpr int NIN = slot_count<N, true, NL, false>();
  constexpr int NMID = slot_count<N, true, NL, true>();
  constexpr int NOUT = slot_count<N, false, NL, true>();
  constexpr int PPX = (N + NTW - 1) / NTW;
  __shared__ cpx<T> lds[LINES * LS];
  __shared__ cpx<T> tws[tw_table_slots<N>()];
  extern __shared__ __align__(16) unsigned char winrow_dyn[];
  cpx<T>* accs = reinterpret_cast<cpx<T>*>(winrow_dyn);  // [C][PPX][LINES][NTW]
  const int tid = threadIdx.x;
  const int li = tid / NL, lane = tid % NL;
  const int rr = li / NC, comp = li - rr * NC;
  const int w = xcd_swizzle(blockIdx.x, gridDim.x);
  const int ngroups = A.b.nrays / NR;
  const int nchunks = (A.b.nlam_b + A.nchunk - 1) / A.nchunk;
  const int rg = w % ngroups;
  const int t = w / ngroups;
  const int ch = t % nchunks;
  const int y = t / nchunks;
  const int lam_lo = A.b.lam0 + ch * A.nchunk;
  const int nl = min(A.nchunk, A.b.nlam_b - ch * A.nchunk);
  const size_t HW = size_t(A.H) * N;
  const int beam = (rg * NR) / A.bm.rays_per_beam;
  cpx<T>* line_lds = lds + li * LS;
  fill_tw_tables<T, N>(tws, A.tw, tid, NTW);

  // line of (window row j, wavelength lam) for this wave
  auto line_off = [&](int lam) {
    const size_t q = size_t(rg * NR + rr) * A.b.nlam + lam;
    return (q * NC + comp) * HW + size_t(y) * N;
  };
  // first window's input requested before the material loads and the solve
  cpx<T> nxt[NIN];
  {
    const cpx<T>* g0 = A.tmp + size_t(A.win_j[0]) * A.fsz + line_off(lam_lo);
#pragma unroll
    for (int j = 0; j < NIN; ++j)
      if (slot_valid<N, true, NL, false>(j, lane)) nxt[j] = g0[slot_pos<N, true, NL, false>(j, lane)];
  }
  const size_t acc_stride = size_t(PPX) * LINES * NTW;  // elements per wavelength
  for (int jl = 0; jl < nl; ++jl)
#pragma unroll
    for (int k = 0; k < PPX; ++k)
#pragma unroll
      for (int l = 0; l < LINES; ++l) accs[jl * acc_stride + (k * LINES + l) * NTW + tid] = mk<T>(T(0), T(0));
  __syncthreads();

  cpx<T> eg[PPX][EIG_COMPS];
  bool fast_ph = false;
  const int nit = A.nwin * nl;
  for (int it = 0; it < nit; ++it) {
    const int wi = it / nl, jl = it - wi * nl;
    const int lam = lam_lo + jl;
    const double bw = A.win_beta[size_t(wi) * A.nbeam + beam];
    if (jl == 0) {
      // per-pixel operator of this window's mean beam, solved while this
      // window's first line is in flight (the material is re-read from the
      // cache: registers are kept for the lines)
      const double phi0 = A.win_phi[size_t(wi) * A.nbeam + beam];
      double s0, c0;
      sincos(phi0, &s0, &c0);
      bool f = true;
#pragma unroll
      for (int k = 0; k < PPX; ++k) {
        const int x = tid + k * NTW;
        const size_t o = size_t(y) * N + (x < N ? x : 0);
        const cd e0 = mk<double>(A.eps_re[o], A.eps_im ? A.eps_im[o] : 0.);
        const cd e1 = mk<double>(A.eps_re[A.comp_stride + o], A.eps_im ? A.eps_im[A.comp_stride + o] : 0.);
        const cd e2 = mk<double>(A.eps_re[2 * A.comp_stride + o],
                                 A.eps_im ? A.eps_im[2 * A.comp_stride + o] : 0.);
        RopData r;
        rop_lean(bw, c0, s0, phi0, e0, e1, e2, A.ang[A.comp_stride + o], A.ang[2 * A.comp_stride + o], r);
        if (r.kind == KIND_BIAXIAL && x < N && A.status) atomicOr(A.status, 1);
        for (int i = 0; i < 4; ++i) eg[k][i] = cast<T>(r.a[i]);
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) {
            eg[k][4 + 2 * i + j] = cast<T>(r.P0[i][j]);
            eg[k][8 + 2 * i + j] = cast<T>(r.P1[i][j]);
          }
        T ok = (A.method4 ? r.ok4 : r.ok2) ? T(1) : T(0);
        if (r.kind == KIND_BIAXIAL) ok = T(__builtin_nan(""));
        eg[k][12] = mk<T>(ok, T(0));
        if (x < N) f = f && fast_blocks_ok<T, NC>(eg[k]);
      }
      fast_ph = wave_all(f);
    }
    cpx<T> v[MAXV];
#pragma unroll
    for (int j = 0; j < NIN; ++j) v[j] = nxt[j];
    if (it + 1 < nit) {
      const int wn1 = (it + 1) / nl, jn = (it + 1) - wn1 * nl;
      const cpx<T>* gn = A.tmp + size_t(A.win_j[wn1]) * A.fsz + line_off(lam_lo + jn);
#pragma unroll
      for (int j = 0; j < NIN; ++j)
        if (slot_valid<N, true, NL, false>(j, lane)) nxt[j] = gn[slot_pos<N, true, NL, false>(j, lane)];
    }
    {
      // tent column factor and mean-beam correction of this window
      const int i = A.win_i[wi];
      const T bwt = T(bw);
      const cpx<T> ac = csqrt(cast<T>(A.we) - mk<T>(bwt * bwt));
      // forward components (4x4: even, 2x2: both) exp(-i k0 d a_c), backward exp(+i ...)
      const double sg = (NC == 4 && (comp & 1)) ? 1. : -1.;
      const cpx<T> cr = cexpi_layer(ac * T(sg * A.bm.k0[lam] * A.wd));
      const double* cxl = A.wcx + ((size_t(beam) * A.b.nlam + lam) * A.wn + i) * N;
#pragma unroll
      for (int j = 0; j < NIN; ++j)
        if (slot_valid<N, true, NL, false>(j, lane))
          v[j] = v[j] * (cr * T(cxl[slot_pos<N, true, NL, false>(j, lane)]));
    }
    reg_fft<T, N, true, +1, NL, 1, 1, MAXV, WV>(v, line_lds, tws, lane);
#pragma unroll
    for (int j = 0; j < NMID; ++j)
      if (slot_valid<N, true, NL, true>(j, lane))
        line_lds[lds_at<1, DTMMX_ROP_PADW>(slot_pos<N, true, NL, true>(j, lane))] = v[j];
    __syncthreads();
    const T kd = T(A.bm.k0[lam] * A.d);
    cpx<T>* acl = accs + jl * acc_stride;
#pragma unroll
    for (int k = 0; k < PPX; ++k) {
      const int x = tid + k * NTW;
      if (x < N) {
        const int ax = lds_at<1, DTMMX_ROP_PADW>(x);
        cpx<T> fb[4], bb[4];
        layer_blocks<T, NC>(eg[k], kd, fast_ph, fb, bb);
        const cpx<T> f00 = fb[0], f01 = fb[1], f10 = fb[2], f11 = fb[3];
        cpx<T>* a = acl + size_t(k) * LINES * NTW + tid;
        if constexpr (NC == 4) {
          const cpx<T> b00 = bb[0], b01 = bb[1], b10 = bb[2], b11 = bb[3];
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const cpx<T>* L = lds + (r * 4) * LS + ax;
            const cpx<T> u0 = L[0], u1 = L[LS], u2 = L[2 * LS], u3 = L[3 * LS];
            cpx<T>* ar = a + (r * 4) * NTW;
            ar[0] = ar[0] + (f00 * u0 + f01 * u2);
            ar[2 * NTW] = ar[2 * NTW] + (f10 * u0 + f11 * u2);
            ar[NTW] = ar[NTW] + (b00 * u1 + b01 * u3);
            ar[3 * NTW] = ar[3 * NTW] + (b10 * u1 + b11 * u3);
          }
        } else {
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const cpx<T>* L = lds + (r * 2) * LS + ax;
            const cpx<T> u0 = L[0], u1 = L[LS];
            cpx<T>* ar = a + (r * 2) * NTW;
            ar[0] = ar[0] + (f00 * u0 + f01 * u1);
            ar[NTW] = ar[NTW] + (f10 * u0 + f11 * u1);
          }
        }
      }
    }
    __syncthreads();
  }

  // one forward transform per line and wavelength of the summed windows
  for (int jl = 0; jl < nl; ++jl) {
    const cpx<T>* acl = accs + jl * acc_stride;
#pragma unroll
    for (int k = 0; k < PPX; ++k) {
      const int x = tid + k * NTW;
      if (x < N) {
        const int ax = lds_at<1, DTMMX_ROP_PADW>(x);
#pragma unroll
        for (int l = 0; l < LINES; ++l) lds[l * LS + ax] = acl[(size_t(k) * LINES + l) * NTW + tid];
      }
    }
    __syncthreads();
    cpx<T> v[MAXV];
#pragma unroll
    for (int j = 0; j < NMID; ++j)
      if (slot_valid<N, true, NL, true>(j, lane))
        v[j] = line_lds[lds_at<1, DTMMX_ROP_PADW>(slot_pos<N, true, NL, true>(j, lane))];
    exchange_sync<WV>();
    reg_fft<T, N, false, -1, NL, 1, 1, MAXV, WV>(v, line_lds, tws, lane);
    cpx<T>* gl = A.out + line_off(lam_lo + jl);
#pragma unroll
    for (int j = 0; j < NOUT; ++j)
      if (slot_valid<N, false, NL, true>(j, lane)) gl[slot_pos<N, false, NL, true>(j, lane)] = v[j];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Adjoint row kernel (4x4): for one sub-step of a layer, the transpose of
// the row pass F_x R B_x applied to the adjoint field (B_x, R^H, F_x, in
// place) and the director-angle gradient
//   dL/dtheta, dL/dphi += 2 Re( vhat^H (dR/dangle) u )
// with u = B_x g the real-space field the forward row pass saw (g recorded
// on the tape) and vhat = B_x of the incoming adjoint.  One work-group per
// (row y, ray); the per-pixel operator data and derivatives come from the
// layer's eig_grad_kernel table.
// ---------------------------------------------------------------------------

// Per-pixel operator data and its theta / phi derivatives for one layer and
// every beam, solved once per layer (fp64 forward-mode, rop_grad) and read by
// the adjoint row kernel for every wavelength and ray: table
// [GRAD_COMPS][nbeam][H*W] of a[4], P0[2][2], P1[2][2], da[2][4],
// dP0[2][2][2], dP1[2][2][2] and the invertibility flag.
constexpr int GRAD_COMPS = 37;

template <typename T>
struct EigArgs {
  const double* eps_re;
  const double* eps_im;
  const double* ang;
  size_t comp_stride;
  size_t HW;
  int nbeam;
  cpx<T>* table;
  int* status;
  int method4;  // rop_table_kernel: 1 = 4x4 (E invertible), 0 = 2x2 (E+ invertible)
  int bwd2;     // rop_table_kernel, 2x2: backward modes (mode -1) in the forward slots
  Beams bm;
};

// Per-pixel operator data of one layer for every beam, in the row kernel's
// order (a[4], P0, P1, invertibility factor): table [EIG_COMPS][nbeam][H*W].
// For layers with biaxial voxels, whose numeric eigensolve (rop_full) runs
// here once per layer instead of inside the row kernel.
template <typename T>
__global__ __launch_bounds__(NT) void rop_table_kernel(EigArgs<T> A) {
  const size_t o = size_t(blockIdx.x) * NT + threadIdx.x;
  const int beam = blockIdx.y;
  if (o >= A.HW) return;
  const double beta = A.bm.beta[beam], phi0 = A.bm.phi[beam];
  double s0, c0;
  sincos(phi0, &s0, &c0);
  const size_t cs = A.comp_stride;
  const cd e0 = mk<double>(A.eps_re[o], A.eps_im ? A.eps_im[o] : 0.);
  const cd e1 = mk<double>(A.eps_re[cs + o], A.eps_im ? A.eps_im[cs + o] : 0.);
  const cd e2 = mk<double>(A.eps_re[2 * cs + o], A.eps_im ? A.eps_im[2 * cs + o] : 0.);
  RopData r;
  rop_full(beta, c0, s0, phi0, e0, e1, e2, A.ang[o], A.ang[cs + o], A.ang[2 * cs + o], r);
  cpx<T>* t = A.table + size_t(beam) * A.HW + o;
  const size_t ts = size_t(A.nbeam) * A.HW;
  const bool bw = !A.method4 && A.bwd2;
  const int f0 = bw ? 1 : 0, f2 = bw ? 3 : 2, b1 = bw ? 0 : 1, b3 = bw ? 2 : 3;
  t[0] = cast<T>(r.a[f0]);
  t[ts] = cast<T>(r.a[b1]);
  t[2 * ts] = cast<T>(r.a[f2]);
  t[3 * ts] = cast<T>(r.a[b3]);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      t[(4 + 2 * i + j) * ts] = cast<T>(bw ? r.P1[i][j] : r.P0[i][j]);
      t[(8 + 2 * i + j) * ts] = cast<T>(bw ? r.P0[i][j] : r.P1[i][j]);
    }
  t[12 * ts] = mk<T>((A.method4 ? r.ok4 : (bw ? r.okb : r.ok2)) ? T(1) : T(0), T(0));
}

template <typename T>
__global__ __launch_bounds__(NT) void eig_grad_kernel(EigArgs<T> A) {
  const size_t o = size_t(blockIdx.x) * NT + threadIdx.x;
  const int beam = blockIdx.y;
  if (o >= A.HW) return;
  const double beta = A.bm.beta[beam], phi0 = A.bm.phi[beam];
  double s0, c0;
  sincos(phi0, &s0, &c0);
  const cd e0 = mk<double>(A.eps_re[o], A.eps_im ? A.eps_im[o] : 0.);
  const cd e1 = mk<double>(A.eps_re[A.comp_stride + o], A.eps_im ? A.eps_im[A.comp_stride + o] : 0.);
  const cd e2 = mk<double>(A.eps_re[2 * A.comp_stride + o],
                           A.eps_im ? A.eps_im[2 * A.comp_stride + o] : 0.);
  RopGrad g;
  rop_grad(beta, c0, s0, phi0, e0, e1, e2, A.ang[A.comp_stride + o], A.ang[2 * A.comp_stride + o], g);
  if (g.kind == KIND_BIAXIAL && A.status) atomicOr(A.status, 1);
  cpx<T>* t = A.table + size_t(beam) * A.HW + o;
  const size_t cs = size_t(A.nbeam) * A.HW;
  int c = 0;
  auto put = [&](cd z) { t[(c++) * cs] = cast<T>(z); };
  for (int i = 0; i < 4; ++i) put(g.a[i]);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) put(g.P0[i][j]);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) put(g.P1[i][j]);
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < 4; ++i) put(g.da[k][i]);
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) put(g.dP0[k][i][j]);
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) put(g.dP1[k][i][j]);
  put(mk<double>(g.ok4 ? 1. : 0.));
}

template <typename T>
struct AdjArgs {
  cpx<T>* adj;          // adjoint field (mixed domain), in place
  const cpx<T>* tape;   // recorded row-pass input of this sub-step
  int H, W;
  Batch b;
  const cpx<T>* tw;
  const double* eps_re;
  const double* eps_im;
  const double* ang;
  size_t comp_stride;
  double d;
  int* status;
  double* gpart;        // [nrays][nlam_b][2][H][W]: this sub-step's gradient per ray, wavelength
  const cpx<T>* table;  // eig_grad_kernel output of this layer
  int nbeam;
  Beams bm;
};

template <typename T, int N>
__global__ __launch_bounds__(row_lanes<N>() * 4, 2) void row_adj_kernel(AdjArgs<T> A) {
  constexpr int NL = row_lanes<N>();
  constexpr bool WV = (NL == 64);
  constexpr int LINES = 4;
  constexpr int NTW = NL * LINES;
  constexpr int LS = line_stride<N>();
  constexpr int MV1 = max_vals<N, true, NL>();
  constexpr int MV2 = max_vals<N, false, NL>();
  constexpr int MAXV = MV1 > MV2 ? MV1 : MV2;
  constexpr int NIN = slot_count<N, true, NL, false>();
  constexpr int NMID = slot_count<N, true, NL, true>();
  constexpr int NOUT = slot_count<N, false, NL, true>();
  constexpr int PPX = (N + NTW - 1) / NTW;
  __shared__ cpx<T> la[LINES * LS];  // adjoint lines
  __shared__ cpx<T> lu[LINES * LS];  // forward lines
  __shared__ cpx<T> tws[tw_table_slots<N>()];
  const int tid = threadIdx.x;
  const int comp = tid / NL, lane = tid % NL;
  fill_tw_tables<T, N>(tws, A.tw, tid, NTW);
  const int w = xcd_swizzle(blockIdx.x, gridDim.x);
  const int rl = w % (A.b.nrays * A.b.nlam_b);  // (ray, wavelength) of this work-group
  const int ray = rl / A.b.nlam_b, jl = rl % A.b.nlam_b;
  const int y = w / (A.b.nrays * A.b.nlam_b);
  const size_t HW = size_t(A.H) * N;
  const int beam = ray / A.bm.rays_per_beam;
  double gt[PPX], gp[PPX];
  for (int k = 0; k < PPX; ++k) gt[k] = gp[k] = 0.;
  __syncthreads();
  {
    const int lam = A.b.lam0 + jl;
    const size_t off = ((size_t(ray) * A.b.nlam + lam) * 4 + comp) * HW + size_t(y) * N;
    cpx<T>* ga = A.adj + off;
    const cpx<T>* gu = A.tape + off;
    cpx<T> v[MAXV], u[MAXV];
#pragma unroll
    for (int j = 0; j < NIN; ++j)
      if (slot_valid<N, true, NL, false>(j, lane)) {
        v[j] = ga[slot_pos<N, true, NL, false>(j, lane)];
        u[j] = gu[slot_pos<N, true, NL, false>(j, lane)];
      }
    cpx<T>* lva = la + comp * LS;
    cpx<T>* lvu = lu + comp * LS;
    reg_fft<T, N, true, +1, NL, 1, 1, MAXV, WV>(v, lva, tws, lane);
    reg_fft<T, N, true, +1, NL, 1, 1, MAXV, WV>(u, lvu, tws, lane);
#pragma unroll
    for (int j = 0; j < NMID; ++j)
      if (slot_valid<N, true, NL, true>(j, lane)) {
        const int p = lds_at<1, 1>(slot_pos<N, true, NL, true>(j, lane));
        lva[p] = v[j];
        lvu[p] = u[j];
      }
    __syncthreads();
    const double kd = A.bm.k0[lam] * A.d;
#pragma unroll
    for (int k = 0; k < PPX; ++k) {
      const int x = tid + k * NTW;
      if (x >= N) continue;
      const size_t o = size_t(y) * N + x;
      // table entries of this pixel: a 0-3, P0 4-7, P1 8-11, da 12-19,
      // dP0 20-27, dP1 28-35, ok 36
      const cpx<T>* tp = A.table + size_t(beam) * HW + o;
      const size_t cs = size_t(A.nbeam) * HW;
      auto tb = [&](int c) { return cast<double>(tp[c * cs]); };
      const double ok = tb(GRAD_COMPS - 1).x;
      const int ax = lds_at<1, 1>(x);
      // blocks (E+x, E+y) = comps (0, 2) with P0, (E-x, E-y) = comps (1, 3) with P1
      for (int blk = 0; blk < 2; ++blk) {
        const int ca = blk, cb = blk + 2;
        const int ia = blk, ib = blk + 2;  // alpha indices: a0, a2 (+), a1, a3 (-)
        const cd aa = tb(ia), ab = tb(ib);
        const cd pa = cexpi(aa * kd), pb = cexpi(ab * kd);
        cd P[2][2];
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) P[i][j] = tb(4 + 4 * blk + 2 * i + j);
        const cd uh[2] = {cast<double>(lu[ca * LS + ax]), cast<double>(lu[cb * LS + ax])};
        const cd vh[2] = {cast<double>(la[ca * LS + ax]), cast<double>(la[cb * LS + ax])};
        const cd dp = (pa - pb) * ok, p2 = pb * ok;
        // gradient: vhat^H (dp2 I + (dpa - dpb) P + (pa - pb) dP) u
        for (int dir = 0; dir < 2; ++dir) {
          cd dP[2][2];
          for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) dP[i][j] = tb(20 + 8 * blk + 4 * dir + 2 * i + j);
          const cd dpa = mk<double>(0., kd) * tb(12 + 4 * dir + ia) * pa;
          const cd dpb = mk<double>(0., kd) * tb(12 + 4 * dir + ib) * pb;
          cd s = mk<double>(0.);
          for (int i = 0; i < 2; ++i) {
            cd t = dpb * uh[i];
            for (int j = 0; j < 2; ++j) t = t + ((dpa - dpb) * P[i][j] + (pa - pb) * dP[i][j]) * uh[j];
            s = s + conj(vh[i]) * t;
          }
          const double gv = 2. * s.x * ok;
          if (dir == 0) gt[k] += gv;
          else gp[k] += gv;
        }
        // adjoint: uhat = (p2 I + dp P)^H vhat
        cd r[2];
        for (int i = 0; i < 2; ++i) {
          cd t = conj(p2) * vh[i];
          for (int j = 0; j < 2; ++j) t = t + conj(dp * P[j][i]) * vh[j];
          r[i] = t;
        }
        la[ca * LS + ax] = cast<T>(r[0]);
        la[cb * LS + ax] = cast<T>(r[1]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NMID; ++j)
      if (slot_valid<N, true, NL, true>(j, lane)) v[j] = lva[lds_at<1, 1>(slot_pos<N, true, NL, true>(j, lane))];
    exchange_sync<WV>();
    reg_fft<T, N, false, -1, NL, 1, 1, MAXV, WV>(v, lva, tws, lane);
#pragma unroll
    for (int j = 0; j < NOUT; ++j)
      if (slot_valid<N, false, NL, true>(j, lane)) ga[slot_pos<N, false, NL, true>(j, lane)] = v[j];
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < PPX; ++k) {
    const int x = tid + k * NTW;
    if (x >= N) continue;
    const size_t o = size_t(y) * N + x;
    A.gpart[(size_t(rl) * 2) * HW + o] = gt[k];
    A.gpart[(size_t(rl) * 2 + 1) * HW + o] = gp[k];
  }
}

// grad[0|1][pixel] += sum over (ray, wavelength) r of gpart[r][0|1][pixel].
static __global__ __launch_bounds__(NT) void grad_sum_kernel(const double* gpart, int nrays, size_t HW,
                                                       double* gtheta, double* gphi) {
  const size_t i = size_t(blockIdx.x) * NT + threadIdx.x;
  if (i >= HW) return;
  double a = 0., b = 0.;
  for (int r = 0; r < nrays; ++r) {
    a += gpart[(size_t(r) * 2) * HW + i];
    b += gpart[(size_t(r) * 2 + 1) * HW + i];
  }
  gtheta[i] += a;
  gphi[i] += b;
}

// ---------------------------------------------------------------------------
// Column kernel: a work-group owns TX adjacent columns of one wavelength
// (TX lanes read 8*TX contiguous bytes of a row; NLC lanes share a column,
// one radix-8 butterfly each) and transforms the lines of ncl components in
// turn.  OP_DIAG: forward y-transform, diagonal k-space operator (computed
// once per work-group and reused for every line), inverse y-transform.
// ---------------------------------------------------------------------------

template <typename T>
struct ColArgs {
  cpx<T>* field;
  int H, W;
  Batch b;
  int NC;             // components per plane
  int ncl;            // lines per work-group (divides nrays * NC)
  const cpx<T>* tw;   // twiddles of length H
  const int* amax;    // [nlam][H]: largest passing |kx| index per ky row, -1 if none
  double da, db;      // thicknesses of the two half steps
  cd ea, eb;          // isotropic effective eps of the two half steps
  double kca, kcb;    // weights of the mean-beam correction of each half step
  double scale;
  int dense_sign;     // 1: backward components (odd c) get exp(-i phase)
  int conj_k;         // 1: the adjoint (complex conjugate) of the diagonal operator
  // OP_WIN: destination (the source `field` is left unchanged) and the tent
  // window (wi, wj) of wn x wn: separable factors cx [beam][lam][wn][W] and
  // cy [beam][lam][wn][H]
  cpx<T>* out;
  const double* wcx;
  const double* wcy;
  int wn, wi, wj;
  // OP_FIELD: the field diffraction operator of the isotropic medium ea over
  // the thickness da (0 both directions, +1 forward modes, -1 backward)
  int fmode;
  Beams bm;
  // 1: every mode passing the betamax mask propagates in the half-step
  // medium ea (real, eps > betamax^2 with a margin over the float rounding
  // of beta^2), so the diagonal factor of a real medium is a pure phase
  int prop;
};

// F P(kd) F^-1 of an isotropic medium e0 in closed form at in-plane wave
// vector beta with direction (c, sn) (dtmm/diffract.py:107-119
// field_diffraction_matrix, phase_matrix :96-105): the mode frame g = R^T v,
// amplitudes m = F^-1 g, phases of the kept directions (fmode 0 both, +1
// forward only, -1 backward only), back to the field; zero where the mode
// does not exist.
__device__ inline void field4_apply(cd e0, double kd, int fmode, double beta, double c, double sn,
                                    const cd v[4], cd r[4]) {
  for (int i = 0; i < 4; ++i) r[i] = mk<double>(0.);
  const cd a = csqrt(e0 - mk<double>(beta * beta));
  if (is_zero(a)) return;
  const cd igp = cdiv(a, e0), igs = cdiv(mk<double>(-1.), a);
  const cd g0 = v[0] * c + v[2] * sn, g1 = v[1] * c - v[3] * sn;
  const cd g2 = v[2] * c - v[0] * sn, g3 = v[1] * sn + v[3] * c;
  const cd zero = mk<double>(0.);
  const cd p = fmode < 0 ? zero : cexpi(a * kd);
  const cd pi = fmode > 0 ? zero : cexpi(a * -kd);
  const cd m0 = (g0 + g1 * igp) * p, m1 = (g0 - g1 * igp) * pi;
  const cd m2 = (g2 + g3 * igs) * p, m3 = (g2 - g3 * igs) * pi;
  const cd gp = cdiv(e0, a), gs = -a;
  const cd h0 = (m0 + m1) * 0.5, h1 = gp * (m0 - m1) * 0.5;
  const cd h2 = (m2 + m3) * 0.5, h3 = gs * (m2 - m3) * 0.5;
  r[0] = h0 * c - h2 * sn;
  r[1] = h1 * c + h3 * sn;
  r[2] = h0 * sn + h2 * c;
  r[3] = h3 * c - h1 * sn;
}

// field4_apply for a real medium eps (eps > beta^2 wherever the betamax mask
// passes, so a = sqrt(eps - beta^2) and the mode-frame factors are real and
// exp(-i a kd) = conj(exp(i a kd))), in the field's precision and scaled:
// the column pass OP_FIELD.
template <typename T>
struct FieldConst {
  T e, ie, kd, l2, scale;  // eps, 1 / eps, k0 d, (2 pi / k0)^2, scale
  bool fwd, bwd;
};

template <typename T>
__device__ __forceinline__ FieldConst<T> field_const(double eps, double k0, double d, int fmode,
                                                     double scale) {
  FieldConst<T> c;
  c.e = T(eps);
  c.ie = T(1. / eps);
  c.kd = T(k0 * d);
  const double l = 2 * M_PI / k0;
  c.l2 = T(l * l);
  c.scale = T(scale);
  c.fwd = fmode >= 0;
  c.bwd = fmode <= 0;
  return c;
}

template <typename T>
__device__ __forceinline__ void field4_real(const FieldConst<T>& K, bool ok, T fx, T fy,
                                            cpx<T> v0, cpx<T> v1, cpx<T> v2, cpx<T> v3,
                                            cpx<T> r[4]) {
  const T f2 = fx * fx + fy * fy;
  const T a2 = K.e - K.l2 * f2;
  if (!ok || !(a2 > T(0))) {
    r[0] = r[1] = r[2] = r[3] = mk<T>(T(0), T(0));
    return;
  }
  const T fr = sqrt(f2);
  const T c = fr > T(0) ? fx / fr : T(1), sn = fr > T(0) ? fy / fr : T(0);
  const T a = sqrt(a2), ia = T(1) / a;
  const T igp = a * K.ie, gp = K.e * ia;
  const cpx<T> g0 = v0 * c + v2 * sn, g1 = v1 * c - v3 * sn;
  const cpx<T> g2 = v2 * c - v0 * sn, g3 = v1 * sn + v3 * c;
  const cpx<T> ph = cexpi_layer(mk<T>(a * K.kd, T(0)));
  const cpx<T> zero = mk<T>(T(0), T(0));
  const cpx<T> p = K.fwd ? ph : zero, pi = K.bwd ? conj(ph) : zero;
  const cpx<T> m0 = (g0 + g1 * igp) * p, m1 = (g0 - g1 * igp) * pi;
  const cpx<T> m2 = (g2 - g3 * ia) * p, m3 = (g2 + g3 * ia) * pi;  // igs = -1 / a
  const T hs = T(0.5) * K.scale;
  const cpx<T> h0 = (m0 + m1) * hs, h1 = (m0 - m1) * (gp * hs);
  const cpx<T> h2 = (m2 + m3) * hs, h3 = (m2 - m3) * (-a * hs);  // gs = -a
  r[0] = h0 * c - h2 * sn;
  r[1] = h1 * c + h3 * sn;
  r[2] = h0 * sn + h2 * c;
  r[3] = h3 * c - h1 * sn;
}

// A value known to be equal in every lane, moved to a scalar register.
__device__ __forceinline__ float uniform(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double uniform(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readfirstlane(int(b & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane(int(b >> 32));
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Constants of the diagonal k-space operator shared by every k of a
// work-group (its beam and wavelength are fixed): with
// b2(k) = cx mx^2 + cy my^2 (= beta(k)^2) the phase is
//   k0 (da/2 sqrt(ea - b2) + db/2 sqrt(eb - b2)) - pc,
//   pc = k0 (da/2 kca sqrt(ea - bc^2) + db/2 kcb sqrt(eb - bc^2)).
// `fast`: ea == eb and real, so one real square root per k.
template <typename T>
struct DiagConst {
  T k0, cx, cy, ha, hb, er;
  cpx<T> ea, eb, pc;
  bool fast;
};

template <typename T>
__device__ __forceinline__ DiagConst<T> diag_const(const ColArgs<T>& A, int lam, int beam) {
  DiagConst<T> c;
  const double k0 = A.bm.k0[lam];
  const double l = 2 * M_PI / k0;
  c.k0 = T(k0);
  c.cx = T((l / A.W) * (l / A.W));
  c.cy = T((l / A.H) * (l / A.H));
  c.ea = cast<T>(A.ea);
  c.eb = cast<T>(A.eb);
  c.ha = T(0.5 * A.da * k0);
  c.hb = T(0.5 * A.db * k0);
  const T bc = T(A.bm.beta[beam]);
  c.pc = mk<T>(T(0), T(0));
  if (A.da != 0. && A.kca != 0.) c.pc = c.pc + csqrt(c.ea - mk<T>(bc * bc)) * T(0.5 * A.da * A.kca * k0);
  if (A.db != 0. && A.kcb != 0.) c.pc = c.pc + csqrt(c.eb - mk<T>(bc * bc)) * T(0.5 * A.db * A.kcb * k0);
  c.fast = (A.ea.x == A.eb.x && A.ea.y == A.eb.y && A.ea.y == 0.) || (A.db == 0. && A.ea.y == 0.);
  c.er = T(A.ea.x);
  if (A.db == 0.) c.hb = T(0);
  if (A.da == 0.) c.ha = T(0);
  // every field is uniform over the work-group: keep them in scalar registers
  c.k0 = uniform(c.k0);
  c.cx = uniform(c.cx);
  c.cy = uniform(c.cy);
  c.ha = uniform(c.ha);
  c.hb = uniform(c.hb);
  c.er = uniform(c.er);
  c.ea = mk<T>(uniform(c.ea.x), uniform(c.ea.y));
  c.eb = mk<T>(uniform(c.eb.x), uniform(c.eb.y));
  c.pc = mk<T>(uniform(c.pc.x), uniform(c.pc.y));
  return c;
}

template <typename T, bool FAST, bool PROP = false>
__device__ __forceinline__ cpx<T> diag_factor(const ColArgs<T>& A, const DiagConst<T>& c, int amax,
                                              int ky, int kx) {
  // amax: largest |kx| index passing betamax in row ky (-1: none)
  const int mx = mode_index(kx, A.W), my = mode_index(ky, A.H);
  if (mx > amax) return mk<T>(T(0), T(0));
  const T b2 = c.cx * T(mx * mx) + c.cy * T(my * my);
  if constexpr (PROP) {
    // a real medium (FAST) whose passing modes all propagate and a real
    // mean-beam correction: a phase only, no attenuation
    const T q = sqrt(fmax(c.er - b2, T(0)));
    return cexpi_layer_real(q * (c.ha + c.hb) - c.pc.x) * T(A.scale);
  }
  cpx<T> ph;
  if constexpr (FAST) {
    // sqrt(er - b2) of a real medium: real, or positive imaginary past er
    const T x = c.er - b2;
    const T q = sqrt(fabs(x));
    ph = (x >= T(0) ? mk<T>(q, T(0)) : mk<T>(T(0), q)) * (c.ha + c.hb) - c.pc;
  } else {
    ph = mk<T>(T(0), T(0)) - c.pc;
    if (A.da != 0.) ph = ph + csqrt(c.ea - mk<T>(b2)) * c.ha;
    if (A.db != 0.) ph = ph + csqrt(c.eb - mk<T>(b2)) * c.hb;
  }
  return cexpi_layer(ph) * T(A.scale);
}

// s^2 / k for the backward components (k = s exp(i ph) -> s exp(-i ph)),
// zero where k is.
__device__ __forceinline__ cpx<float> back_factor(cpx<float> k, float s2) {
  const float n2 = norm2(k);
  return n2 > 0.f ? conj(k) * (s2 * __builtin_amdgcn_rcpf(n2)) : mk<float>(0.f, 0.f);
}
__device__ __forceinline__ cpx<double> back_factor(cpx<double> k, double s2) {
  const double n2 = norm2(k);
  return n2 > 0. ? conj(k) * (s2 / n2) : mk<double>(0., 0.);
}

// Lanes per column: one radix-8 butterfly per lane, 16..128 lanes.
template <int N>
__host__ __device__ constexpr int col_lanes() {
  return N / 8 < 16 ? 16 : (N / 8 > 128 ? 128 : N / 8);
}

template <typename T, int N, int TX, int OP>
__global__ __launch_bounds__(TX * col_lanes<N>(), (col_min_waves<T, N, OP>())) void col_kernel(
    ColArgs<T> A) {
  constexpr int NL = col_lanes<N>();
  constexpr bool REV1 = (OP == OP_INV);
  constexpr int DIR1 = (OP == OP_INV) ? +1 : -1;
  constexpr int MV1 = max_vals<N, REV1, NL>();
  constexpr int MV2 = max_vals<N, true, NL>();
  constexpr int MAXV = MV1 > MV2 ? MV1 : MV2;
  constexpr int NIN = slot_count<N, REV1, NL, false>();
  constexpr int NMID = slot_count<N, REV1, NL, true>();
  __shared__ cpx<T> lds[(N + (DTMMX_COL_PADW ? (N >> DTMMX_LDS_PAD_SHIFT) : 0)) * TX];
  __shared__ cpx<T> tws[tw_table_slots<N>()];  // twiddles in LDS: their reads never wait on global loads
  const int tid = threadIdx.x;
  const int c = tid % TX, g = tid / TX;
  // Prologue order: the small L2-resident loads (twiddles, betamax mask)
  // first, then the first field line, then the setup arithmetic; loads
  // complete in issue order, so the setup waits only for the small ones.
  TwFill<T, N, TX * NL> twf;
  twf.load(A.tw, tid);
  const int ntiles = (A.W + TX - 1) / TX;
  const int w = xcd_swizzle(blockIdx.x, gridDim.x);
  const int tile = w % ntiles;
  const int t = w / ntiles;
  const int lam = A.b.lam0 + t % A.b.nlam_b;
  const int lg = t / A.b.nlam_b;
  const int x = tile * TX + c;
  const bool xin = x < A.W;
  const size_t HW = size_t(N) * A.W;
  constexpr bool KOP = (OP == OP_DIAG || OP == OP_WIN || OP == OP_SPLIT);
  int am[(KOP || OP == OP_FIELD) ? NMID : 1];  // largest passing |kx| of each slot's ky row
  if constexpr (KOP || OP == OP_FIELD) {
#pragma unroll
    for (int j = 0; j < NMID; ++j)
      am[j] = slot_valid<N, REV1, NL, true>(j, g) ? A.amax[lam * A.H + slot_pos<N, REV1, NL, true>(j, g)] : -1;
  }

  auto line_off = [&](int li) {
    const int line = lg * A.ncl + li;
    const int ray = line / A.NC, comp = line % A.NC;
    const size_t q = size_t(ray) * A.b.nlam + lam;
    return (q * A.NC + comp) * HW + x;
  };
  auto line_base = [&](int li) { return A.field + line_off(li); };
  // prefetched inputs of the next line; the first line is requested before
  // the k-space setup so that its latency overlaps the setup arithmetic
  cpx<T> nxt[NIN];
  {
    const cpx<T>* b0 = line_base(0);
#pragma unroll
    for (int j = 0; j < NIN; ++j) {
      nxt[j] = mk<T>(T(0), T(0));
      if (xin && slot_valid<N, REV1, NL, false>(j, g))
        nxt[j] = b0[size_t(slot_pos<N, REV1, NL, false>(j, g)) * A.W];
    }
  }
  cpx<T> K[KOP ? NMID : 1];
  T Wv[(OP == OP_WIN) ? NMID : 1];  // window factor cx(kx) * cy(ky)
  const double* cys = nullptr;        // OP_SPLIT: cy [wn][H] of this beam and wavelength
  if constexpr (KOP) {
    const int ray0 = (lg * A.ncl) / A.NC;  // all lines of a work-group share one beam
    const int beam = ray0 / A.bm.rays_per_beam;
    double cx = 0.;
    const double* cyt = nullptr;
    if constexpr (OP == OP_WIN) {
      const size_t bl = size_t(beam) * A.b.nlam + lam;
      if (xin) cx = A.wcx[(bl * A.wn + A.wi) * A.W + x];
      cyt = A.wcy + (bl * A.wn + A.wj) * N;
    }
    if constexpr (OP == OP_SPLIT) cys = A.wcy + (size_t(beam) * A.b.nlam + lam) * A.wn * N;
    const DiagConst<T> dc = diag_const(A, lam, beam);
#pragma unroll
    for (int j = 0; j < NMID; ++j) K[j] = mk<T>(T(0), T(0));
    if (dc.fast && A.prop && dc.pc.y == T(0)) {
#pragma unroll
      for (int j = 0; j < NMID; ++j)
        if (xin && slot_valid<N, REV1, NL, true>(j, g))
          K[j] = diag_factor<T, true, true>(A, dc, am[j], slot_pos<N, REV1, NL, true>(j, g), x);
    } else if (dc.fast) {
      asm volatile("");  // a branch of its own, not merged with the one above
#pragma unroll
      for (int j = 0; j < NMID; ++j)
        if (xin && slot_valid<N, REV1, NL, true>(j, g))
          K[j] = diag_factor<T, true>(A, dc, am[j], slot_pos<N, REV1, NL, true>(j, g), x);
    } else {
#pragma unroll
      for (int j = 0; j < NMID; ++j)
        if (xin && slot_valid<N, REV1, NL, true>(j, g))
          K[j] = diag_factor<T, false>(A, dc, am[j], slot_pos<N, REV1, NL, true>(j, g), x);
    }
#pragma unroll
    for (int j = 0; j < NMID; ++j) {
      if (A.conj_k) K[j] = conj(K[j]);
      if constexpr (OP == OP_WIN) {
        Wv[j] = T(0);
        if (xin && slot_valid<N, REV1, NL, true>(j, g)) Wv[j] = T(cx * cyt[slot_pos<N, REV1, NL, true>(j, g)]);
      }
    }
  }
  // visible to the other waves after the first exchange barrier of the
  // first transform, before the first twiddle read (stage 1)
  twf.store(tws, tid);
  if constexpr (OP == OP_FIELD) {
    // the four components of a plane (lines li .. li + 3): forward y
    // transforms, the field diffraction operator per k, inverse transforms
    const FieldConst<T> fc = field_const<T>(A.ea.x, A.bm.k0[lam], A.da, A.fmode, A.scale);
    const T fxt = T(fftfreq(x, A.W));
    const int mx = mode_index(x, A.W);
    for (int l4 = 0; l4 < A.ncl; l4 += 4) {
      cpx<T> vv[4][MAXV];
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const cpx<T>* bl = line_base(l4 + cc);
#pragma unroll
        for (int j = 0; j < MAXV; ++j) vv[cc][j] = mk<T>(T(0), T(0));
#pragma unroll
        for (int j = 0; j < NIN; ++j)
          if (xin && slot_valid<N, false, NL, false>(j, g))
            vv[cc][j] = bl[size_t(slot_pos<N, false, NL, false>(j, g)) * A.W];
        reg_fft<T, N, false, -1, NL, TX, DTMMX_COL_PADW * TX>(vv[cc], lds + c, tws, g);
      }
#pragma unroll
      for (int j = 0; j < NMID; ++j) {
        const bool ok = xin && slot_valid<N, false, NL, true>(j, g) && mx <= am[j];
        const T fy = T(fftfreq(slot_pos<N, false, NL, true>(j, g), A.H));
        cpx<T> r[4];
        field4_real<T>(fc, ok, fxt, fy, vv[0][j], vv[1][j], vv[2][j], vv[3][j], r);
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) vv[cc][j] = r[cc];
      }
      constexpr int NOUT = slot_count<N, true, NL, true>();
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        reg_fft<T, N, true, +1, NL, TX, DTMMX_COL_PADW * TX>(vv[cc], lds + c, tws, g);
        cpx<T>* ob = line_base(l4 + cc);
#pragma unroll
        for (int j = 0; j < NOUT; ++j)
          if (xin && slot_valid<N, true, NL, true>(j, g))
            ob[size_t(slot_pos<N, true, NL, true>(j, g)) * A.W] = vv[cc][j];
      }
    }
    return;
  }
  for (int li = 0; li < A.ncl; ++li) {
    const int comp = (lg * A.ncl + li) % A.NC;
    cpx<T>* base = (OP == OP_WIN || OP == OP_SPLIT) ? A.out + line_off(li) : line_base(li);
    cpx<T> v[MAXV];
#pragma unroll
    for (int j = 0; j < NIN; ++j) v[j] = nxt[j];
    if (li + 1 < A.ncl) {
      const cpx<T>* bn = line_base(li + 1);
#pragma unroll
      for (int j = 0; j < NIN; ++j)
        if (xin && slot_valid<N, REV1, NL, false>(j, g))
          nxt[j] = bn[size_t(slot_pos<N, REV1, NL, false>(j, g)) * A.W];
    }
    reg_fft<T, N, REV1, DIR1, NL, TX, DTMMX_COL_PADW * TX>(v, lds + c, tws, g);
    if constexpr (KOP) {
      const bool back = A.dense_sign && (comp & 1);
      const T s2 = T(A.scale * A.scale);
#pragma unroll
      for (int j = 0; j < NMID; ++j) {
        cpx<T> k = K[j];
        if (back) k = back_factor(k, s2);
        if constexpr (OP == OP_WIN) k = k * Wv[j];
        v[j] = v[j] * k;
      }
      if constexpr (OP == OP_SPLIT) {
        complete_loads(nxt, NIN);
        // one inverse transform per window row j: cy_j(ky) times the line
        constexpr int NOUT = slot_count<N, true, NL, true>();
        const size_t fstride = size_t(A.b.nrays) * A.b.nlam * A.NC * HW;
        for (int wj = 0; wj < A.wn; ++wj) {
          cpx<T> u[MAXV];
#pragma unroll
          for (int j = 0; j < NMID; ++j) {
            T cy = T(0);
            if (slot_valid<N, REV1, NL, true>(j, g)) cy = T(cys[wj * N + slot_pos<N, REV1, NL, true>(j, g)]);
            u[j] = v[j] * cy;
          }
          __syncthreads();
          reg_fft<T, N, true, +1, NL, TX, DTMMX_COL_PADW * TX>(u, lds + c, tws, g);
          cpx<T>* ob = base + wj * fstride;
#pragma unroll
          for (int j = 0; j < NOUT; ++j)
            if (xin && slot_valid<N, true, NL, true>(j, g))
              ob[size_t(slot_pos<N, true, NL, true>(j, g)) * A.W] = u[j];
        }
        __syncthreads();
        continue;
      }
      reg_fft<T, N, true, +1, NL, TX, DTMMX_COL_PADW * TX>(v, lds + c, tws, g);
    }
    constexpr bool REVL = KOP ? true : REV1;
    constexpr int NOUT = slot_count<N, REVL, NL, true>();
    complete_loads(nxt, NIN);
#pragma unroll
    for (int j = 0; j < NOUT; ++j)
      if (xin && slot_valid<N, REVL, NL, true>(j, g))
        base[size_t(slot_pos<N, REVL, NL, true>(j, g)) * A.W] = v[j];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Dense per-k operators (fp64), one thread per (k, plane).
// ---------------------------------------------------------------------------

enum { KOP_HEAD4 = 0, KOP_TAIL4 = 1, KOP_BOUND4 = 2, KOP_PROJ2 = 3, KOP_TR2 = 4, KOP_FIELD4 = 5,
       KOP_E2H4 = 6, KOP_IF2W = 7, KOP_OUT2W = 8 };

template <typename T>
struct KArgs {
  const cpx<T>* src;
  cpx<T>* dst;
  int H, W;
  Batch b;
  const int* amax;
  int kop;
  double d0, d1;  // thickness of the layer(s) involved
  cd e0, e1;      // effective eps of the layer(s) involved
  cd ein;         // input medium eps (PROJ2)
  double scale;
  int nocorr;     // 1: no mean-beam correction (HEAD4, TAIL4, BOUND4)
  int fresnel;    // PROJ2: 1 applies the interface transmission ein -> e0
  int adjoint;    // 1: apply the conjugate transpose (4-component operators)
  int fmode;      // FIELD4: 0 both directions, +1 forward modes only, -1 backward only
  // OUT2W: the tent windows (separable factors [beam][lam][wn][W or H], the
  // window list and each window's mean beam [nwin][nbeam])
  const double* wcx;
  const double* wcy;
  const int* win_i;
  const int* win_j;
  const double* win_beta;
  const double* win_phi;
  int wn, nwin, nbeam;
  Beams bm;
};

// Fresnel transmission of the tangential E field from isotropic medium e0
// into e1 at in-plane wave vector beta: p (in the plane of incidence)
// 2 g0 / (g0 + g1) with g = eps / a, s 2 a0 / (a0 + a1), a = sqrt(eps - beta^2);
// zero where a mode does not exist (a == 0).
__device__ inline void fresnel_t(double beta, cd e0, cd e1, cd& tp, cd& ts) {
  const cd a0 = csqrt(e0 - mk<double>(beta * beta)), a1 = csqrt(e1 - mk<double>(beta * beta));
  if (is_zero(a0) || is_zero(a1)) {
    tp = ts = mk<double>(0.);
    return;
  }
  const cd g0 = cdiv(e0, a0), g1 = cdiv(e1, a1);
  tp = cdiv(g0 * 2., g0 + g1);
  ts = cdiv(a0 * 2., a0 + a1);
}

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

// One dense k-space operator r = M(k) v (fp64) at wave vector (beta, phi)
// with in-plane direction (c, sn), for plane wavenumber k0 and beam (bb, pb).
template <typename T, int KOP>
__device__ __forceinline__ void kop_apply(const KArgs<T>& A, bool pass, double beta, double phi,
                                          double c, double sn, double k0, double bb, double pb,
                                          const cd v[4], cd r[4]) {
  for (int i = 0; i < 4; ++i) r[i] = mk<double>(0.);
  if constexpr (KOP == KOP_HEAD4) {
    if (!pass) return;
    // C . E P(kd) F^-1 in closed form for an isotropic medium: field ->
    // frame of the mode (g = R^T f) -> mode amplitudes m = F^-1 g -> half-step
    // phases -> (E+x, E-x, E+y, E-y) -> mean-beam correction.
    const cd a = csqrt(A.e0 - mk<double>(beta * beta));
    if (!is_zero(a)) {
      const double kd = k0 * A.d0 * 0.5;
      const cd igp = cdiv(a, A.e0), igs = cdiv(mk<double>(-1.), a);
      const cd g0 = v[0] * c + v[2] * sn, g1 = v[1] * c - v[3] * sn;
      const cd g2 = v[2] * c - v[0] * sn, g3 = v[1] * sn + v[3] * c;
      const cd p = cexpi(a * kd), pi = cexpi(a * -kd);
      const cd m0 = (g0 + g1 * igp) * p, m1 = (g0 - g1 * igp) * pi;
      const cd m2 = (g2 + g3 * igs) * p, m3 = (g2 - g3 * igs) * pi;
      const cd ac = csqrt(A.e0 - mk<double>(bb * bb));
      const cd cf = A.nocorr ? mk<double>(1.) : cexpi(ac * -kd);
      const cd cb = A.nocorr ? mk<double>(1.) : cexpi(ac * kd);
      r[0] = (m0 * c - m2 * sn) * cf * 0.5;
      r[1] = (m1 * c - m3 * sn) * cb * 0.5;
      r[2] = (m0 * sn + m2 * c) * cf * 0.5;
      r[3] = (m1 * sn + m3 * c) * cb * 0.5;
    }
  } else if constexpr (KOP == KOP_TAIL4) {
    if (!pass) return;
    // F P(kd) E^-1 . C in closed form (inverse path of KOP_HEAD4).
    const cd a = csqrt(A.e0 - mk<double>(beta * beta));
    if (!is_zero(a)) {
      const double kd = k0 * A.d0 * 0.5;
      const cd ac = csqrt(A.e0 - mk<double>(bb * bb));
      const cd cf = A.nocorr ? mk<double>(1.) : cexpi(ac * -kd);
      const cd cb = A.nocorr ? mk<double>(1.) : cexpi(ac * kd);
      const cd epx = v[0] * cf, enx = v[1] * cb, epy = v[2] * cf, eny = v[3] * cb;
      const cd p = cexpi(a * kd), pi = cexpi(a * -kd);
      const cd m0 = (epx * c + epy * sn) * p * 2., m2 = (epy * c - epx * sn) * p * 2.;
      const cd m1 = (enx * c + eny * sn) * pi * 2., m3 = (eny * c - enx * sn) * pi * 2.;
      const cd gp = cdiv(A.e0, a), gs = -a;
      const cd g0 = (m0 + m1) * 0.5, g1 = gp * (m0 - m1) * 0.5;
      const cd g2 = (m2 + m3) * 0.5, g3 = gs * (m2 - m3) * 0.5;
      r[0] = g0 * c - g2 * sn;
      r[1] = g1 * c + g3 * sn;
      r[2] = g0 * sn + g2 * c;
      r[3] = g3 * c - g1 * sn;
    }
  } else if constexpr (KOP == KOP_PROJ2) {
    if (!pass) return;
    // rows (Ex, Ey) of the forward projection F diag(1,0,1,0) F^-1 in the
    // input medium (dtmm/diffract.py:242-251) times the first 2x2 half step
    // exp(i k0 d/2 (a(k) - a_c)) of the first layer.
    const cd ain = csqrt(A.ein - mk<double>(beta * beta));
    if (!is_zero(ain)) {
      const cd igp = cdiv(ain, A.ein), igs = cdiv(mk<double>(-1.), ain);
      const cd g0 = v[0] * c + v[2] * sn, g1 = v[1] * c - v[3] * sn;
      const cd g2 = v[2] * c - v[0] * sn, g3 = v[1] * sn + v[3] * c;
      cd h0 = (g0 + g1 * igp) * 0.5, h2 = (g2 + g3 * igs) * 0.5;
      if (A.fresnel) {
        cd tp, ts;
        fresnel_t(beta, A.ein, A.e0, tp, ts);
        h0 = h0 * tp;
        h2 = h2 * ts;
      }
      cd ph = mk<double>(0.);
      if (A.d0 != 0.) {
        cd ak = csqrt(A.e0 - mk<double>(beta * beta));
        cd ac = csqrt(A.e0 - mk<double>(bb * bb));
        ph = (ak - ac) * (0.5 * A.d0 * k0);
      }
      const cd K = cexpi(ph);
      r[0] = K * (h0 * c - h2 * sn);
      r[1] = K * (h0 * sn + h2 * c);
    }
  } else if constexpr (KOP == KOP_TR2) {
    if (!pass) return;
    // K_b . T(e0 -> e1) . K_a on (Ex, Ey): half step of the layer before
    // the interface, Fresnel transmission of the tangential E field
    // (dtmm/tmm.py:925-950 for isotropic media), half step after it.
    cd ph = mk<double>(0.);
    if (A.d0 != 0.)
      ph = ph + (csqrt(A.e0 - mk<double>(beta * beta)) - csqrt(A.e0 - mk<double>(bb * bb))) *
                    (0.5 * A.d0 * k0);
    if (A.d1 != 0.)
      ph = ph + (csqrt(A.e1 - mk<double>(beta * beta)) - csqrt(A.e1 - mk<double>(bb * bb))) *
                    (0.5 * A.d1 * k0);
    const cd K = cexpi(ph);
    cd tp, ts;
    fresnel_t(beta, A.e0, A.e1, tp, ts);
    const cd ep = (v[0] * c + v[1] * sn) * tp, es = (v[1] * c - v[0] * sn) * ts;
    r[0] = K * (ep * c - es * sn);
    r[1] = K * (ep * sn + es * c);
  } else if constexpr (KOP == KOP_IF2W) {
    if (!pass) return;
    // 2x2 with tent windows, interface between layers of different
    // effective media: the common half step exp(i k0 d0/2 a0(k)) of the
    // layer before it (each window applies its own mean-beam correction with
    // its part), then the Fresnel transmission e0 -> e1 in reflection mode 1
    cd K = mk<double>(1.);
    if (A.d0 != 0.) K = cexpi(csqrt(A.e0 - mk<double>(beta * beta)) * (0.5 * A.d0 * k0));
    cd tp = mk<double>(1.), ts = mk<double>(1.);
    if (A.fresnel) fresnel_t(beta, A.e0, A.e1, tp, ts);
    const cd ep = (v[0] * c + v[1] * sn) * tp, es = (v[1] * c - v[0] * sn) * ts;
    r[0] = K * (ep * c - es * sn);
    r[1] = K * (ep * sn + es * c);
  } else if constexpr (KOP == KOP_FIELD4) {
    if (!pass) return;
    field4_apply(A.e0, k0 * A.d0, A.fmode, beta, c, sn, v, r);
  } else if constexpr (KOP == KOP_E2H4) {
    // E components kept as they are (at every k), H of the forward modes of
    // the isotropic medium e0 from them, zero above betamax
    // (dtmm/transfer.py:255-268 jones2H): p (E along k) with eps / a, s with
    // -a in the frame of the mode
    r[0] = v[0];
    r[2] = v[2];
    if (!pass) return;
    const cd a = csqrt(A.e0 - mk<double>(beta * beta));
    if (is_zero(a)) return;
    const cd g0 = v[0] * c + v[2] * sn, g2 = v[2] * c - v[0] * sn;
    const cd g1 = cdiv(A.e0, a) * g0, g3 = -(a * g2);
    r[1] = g1 * c + g3 * sn;
    r[3] = g3 * c - g1 * sn;
  } else if constexpr (KOP == KOP_BOUND4) {
    if (!pass) return;
    cd C0[4][4], D0[4][4], C1[4][4], D1[4][4], X[4][4], Y[4][4], M[4][4];
    if (A.nocorr) {
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) C0[i][j] = C1[i][j] = mk<double>(i == j ? 1. : 0.);
    } else {
      correction(bb, pb, A.e0, k0 * A.d0 * 0.5, C0);
      correction(bb, pb, A.e1, k0 * A.d1 * 0.5, C1);
    }
    second_half(beta, phi, A.e0, k0 * A.d0 * 0.5, D0);
    first_half(beta, phi, A.e1, k0 * A.d1 * 0.5, D1);
    matmul4(D0, C0, X);  // second corrected of layer 0
    matmul4(C1, D1, Y);  // first corrected of layer 1
    matmul4(Y, X, M);
    for (int i = 0; i < 4; ++i) {
      cd t = mk<double>(0.);
      for (int j = 0; j < 4; ++j) t = t + M[i][j] * v[j];
      r[i] = t;
    }
  }
}

template <typename T, int KOP, bool ADJ>
__global__ __launch_bounds__(NT) void kspace_kernel(KArgs<T> A) {
  const int k = blockIdx.x * NT + threadIdx.x;
  const int ray = blockIdx.y / A.b.nlam_b;
  const int lam = A.b.lam0 + blockIdx.y % A.b.nlam_b;
  const int q = ray * A.b.nlam + lam;
  const size_t HW = size_t(A.H) * A.W;
  if (k >= int(HW)) return;
  const int ky = k / A.W, kx = k - ky * A.W;
  const int beam = ray / A.bm.rays_per_beam;
  const double k0 = A.bm.k0[lam];
  const double bb = A.bm.beta[beam], pb = A.bm.phi[beam];
  const bool pass = mode_index(kx, A.W) <= A.amax[lam * A.H + ky];
  const double fx = fftfreq(kx, A.W), fy = fftfreq(ky, A.H);
  const double beta = (2 * M_PI / k0) * sqrt(fx * fx + fy * fy);
  const double phi = atan2(fy, fx);
  constexpr int nin = (KOP == KOP_TR2 || KOP == KOP_IF2W || KOP == KOP_OUT2W) ? 2 : 4;
  constexpr int nout = (KOP == KOP_PROJ2 || KOP == KOP_TR2 || KOP == KOP_IF2W) ? 2 : 4;
  const cpx<T>* s = A.src + size_t(q) * nin * HW + k;
  cpx<T>* o = A.dst + size_t(q) * nout * HW + k;
  cd v[4];
  for (int c = 0; c < nin; ++c) v[c] = cast<double>(s[c * HW]);
  // in-plane direction of the mode: (cos, sin) of atan2(fy, fx), (1, 0) at k = 0
  const double fr = sqrt(fx * fx + fy * fy);
  const double c = fr > 0. ? fx / fr : 1., sn = fr > 0. ? fy / fr : 0.;
  cd r[4] = {mk<double>(0.), mk<double>(0.), mk<double>(0.), mk<double>(0.)};
  if constexpr (KOP == KOP_OUT2W) {
    // 2x2 with tent windows, output medium e1 (dtmm/propagate_2x2.py:
    // 212-257 with the zero-thickness output layer): the last layer's common
    // half step exp(i k0 d0/2 a0(k)), the Fresnel transmission e0 -> e1 in
    // reflection mode 1, then E = sum_w w(k) E' and H = sum_w w(k) G_w E'
    // with G_w the E -> H matrix of the forward modes of e1 at window w's
    // mean beam (each window part builds its H with its own beam)
    if (pass) {
      kop_apply<T, KOP_IF2W>(A, pass, beta, phi, c, sn, k0, bb, pb, v, r);
      const size_t bl = size_t(beam) * A.b.nlam + lam;
      cd sw = mk<double>(0.), G00 = sw, G01 = sw, G10 = sw, G11 = sw;
      for (int w = 0; w < A.nwin; ++w) {
        const double f = A.wcx[(bl * A.wn + A.win_i[w]) * A.W + kx] *
                         A.wcy[(bl * A.wn + A.win_j[w]) * A.H + ky];
        if (f == 0.) continue;
        const double bw = A.win_beta[size_t(w) * A.nbeam + beam];
        const double pw = A.win_phi[size_t(w) * A.nbeam + beam];
        Modes m;
        iso_modes(bw, A.e1, m);
        rotate_rows(m, cos(pw), sin(pw));
        cd Am[2][2] = {{m.F[0][0], m.F[0][2]}, {m.F[2][0], m.F[2][2]}};
        cd Bm[2][2] = {{m.F[1][0], m.F[1][2]}, {m.F[3][0], m.F[3][2]}};
        cd Ai[2][2];
        inv2(Am, Ai);
        sw = sw + mk<double>(f);
        G00 = G00 + (Bm[0][0] * Ai[0][0] + Bm[0][1] * Ai[1][0]) * f;
        G01 = G01 + (Bm[0][0] * Ai[0][1] + Bm[0][1] * Ai[1][1]) * f;
        G10 = G10 + (Bm[1][0] * Ai[0][0] + Bm[1][1] * Ai[1][0]) * f;
        G11 = G11 + (Bm[1][0] * Ai[0][1] + Bm[1][1] * Ai[1][1]) * f;
      }
      const cd ex = r[0], ey = r[1];
      r[0] = sw * ex;
      r[1] = G00 * ex + G01 * ey;
      r[2] = sw * ey;
      r[3] = G10 * ex + G11 * ey;
    }
  } else if constexpr (!ADJ) {
    kop_apply<T, KOP>(A, pass, beta, phi, c, sn, k0, bb, pb, v, r);
  } else if (pass) {
    // adjoint: r = M^H v with the columns of M from the unit vectors
    for (int j = 0; j < 4; ++j) {
      cd e[4] = {mk<double>(0.), mk<double>(0.), mk<double>(0.), mk<double>(0.)}, m[4];
      e[j] = mk<double>(1.);
      kop_apply<T, KOP>(A, pass, beta, phi, c, sn, k0, bb, pb, e, m);
      cd t = mk<double>(0.);
      for (int i = 0; i < 4; ++i) t = t + conj(m[i]) * v[i];
      r[j] = t;
    }
  }
  for (int i = 0; i < nout; ++i) o[i * HW] = cast<T>(r[i] * A.scale);
}

// ---------------------------------------------------------------------------
// Anisotropic effective media (eff_data 1, 2, per-layer symmetries or a
// custom effective block; dtmm/transfer.py:926-976, dtmm/data.py:1364-1419)
// for the 4x4 method in reflection mode 1.  The diffraction step of a layer
// (dtmm/propagate_4x4.py:376-388, dtmm/matrix.py:109-174) is
//   first  = C E P(kd/2) F^-1,   second = F P(kd/2) E^-1 C,
//   C = E_c P_c(-kd/2) E_c^-1 at the mean beam,
// with F, E, alpha the modes of the layer's effective medium at each k (zero
// above betamax) and kd = k0 d_eff.  For an anisotropic medium these are
// dense per k, so every layer boundary applies second(l) then first(l + 1)
// on the full spectrum, as the boundary between different isotropic media
// does.  One thread per (k, wavelength, beam) builds the 4x4 operator in
// fp64 and applies it to every ray of the beam.
// ---------------------------------------------------------------------------

struct GMed {
  cd e[3];        // eigenvalues of eps
  double ang[3];  // Euler angles psi, theta, phi
  double d;       // thickness of the effective step (negative: backward pass)
};

enum { GOP_HEAD = 0, GOP_TAIL = 1, GOP_BOUND = 2 };

template <typename T>
struct GArgs {
  const cpx<T>* src;
  cpx<T>* dst;
  int H, W;
  Batch b;
  const int* amax;
  GMed m0, m1;  // effective media before and after the boundary
  double scale;
  Beams bm;
};

__device__ inline void g_modes(double beta, double phi, const GMed& g, Modes& m) {
  int kind;
  layer_modes(beta, cos(phi), sin(phi), phi, g.e[0], g.e[1], g.e[2], g.ang[0], g.ang[1],
              g.ang[2], m, kind);
}

// E_c P_c(-kd) E_c^-1 at the mean beam (dtmm/matrix.py:132-139)
__device__ inline void g_correction(double bb, double pb, const GMed& g, double kd, cd C[4][4]) {
  Modes m;
  g_modes(bb, pb, g, m);
  cd E[4][4], Ei[4][4];
  emat(m, E);
  inv4(E, Ei);
  xpy(E, m.a, -kd, Ei, C);
}

// C E P(kd) F^-1 (dtmm/matrix.py:141-157 with :109-118), kd = k0 d / 2
__device__ inline void g_first(double beta, double phi, double bb, double pb, const GMed& g,
                               double k0, cd M[4][4]) {
  const double kd = 0.5 * k0 * g.d;
  Modes m;
  g_modes(beta, phi, g, m);
  cd E[4][4], Fi[4][4], D[4][4], C[4][4];
  emat(m, E);
  inv4(m.F, Fi);
  xpy(E, m.a, kd, Fi, D);
  g_correction(bb, pb, g, kd, C);
  matmul4(C, D, M);
}

// F P(kd) E^-1 C (dtmm/matrix.py:159-174 with :120-130), kd = k0 d / 2
__device__ inline void g_second(double beta, double phi, double bb, double pb, const GMed& g,
                                double k0, cd M[4][4]) {
  const double kd = 0.5 * k0 * g.d;
  Modes m;
  g_modes(beta, phi, g, m);
  cd E[4][4], Ei[4][4], D[4][4], C[4][4];
  emat(m, E);
  inv4(E, Ei);
  xpy(m.F, m.a, kd, Ei, D);
  g_correction(bb, pb, g, kd, C);
  matmul4(D, C, M);
}

template <typename T, int GOP>
__global__ __launch_bounds__(NT) void gkspace_kernel(GArgs<T> A) {
  const int k = blockIdx.x * NT + threadIdx.x;
  const int nbeam = A.b.nrays / A.bm.rays_per_beam;
  const int beam = blockIdx.y / A.b.nlam_b;
  const int lam = A.b.lam0 + blockIdx.y % A.b.nlam_b;
  const size_t HW = size_t(A.H) * A.W;
  if (k >= int(HW) || beam >= nbeam) return;
  const int ky = k / A.W, kx = k - ky * A.W;
  const double k0 = A.bm.k0[lam];
  const double bb = A.bm.beta[beam], pb = A.bm.phi[beam];
  const bool pass = mode_index(kx, A.W) <= A.amax[lam * A.H + ky];
  const double fx = fftfreq(kx, A.W), fy = fftfreq(ky, A.H);
  const double beta = (2 * M_PI / k0) * sqrt(fx * fx + fy * fy);
  const double phi = atan2(fy, fx);
  cd M[4][4];
  if (pass) {
    if constexpr (GOP == GOP_HEAD) {
      g_first(beta, phi, bb, pb, A.m1, k0, M);
    } else if constexpr (GOP == GOP_TAIL) {
      g_second(beta, phi, bb, pb, A.m0, k0, M);
    } else {
      cd X[4][4], Y[4][4];
      g_second(beta, phi, bb, pb, A.m0, k0, X);
      g_first(beta, phi, bb, pb, A.m1, k0, Y);
      matmul4(Y, X, M);
    }
  }
  for (int r = 0; r < A.bm.rays_per_beam; ++r) {
    const size_t q = size_t(beam * A.bm.rays_per_beam + r) * A.b.nlam + lam;
    const cpx<T>* s = A.src + q * 4 * HW + k;
    cpx<T>* o = A.dst + q * 4 * HW + k;
    cd v[4];
    for (int c = 0; c < 4; ++c) v[c] = cast<double>(s[c * HW]);
    for (int i = 0; i < 4; ++i) {
      cd t = mk<double>(0.);
      if (pass)
        for (int j = 0; j < 4; ++j) t = t + M[i][j] * v[j];
      o[i * HW] = cast<T>(t * A.scale);
    }
  }
}

// ---------------------------------------------------------------------------
// Interface of the 2x2 method in reflection mode 1 with stored reflections
// (npass > 1, dtmm/propagate_2x2.py:45-58 with E_tr_matrix of
// dtmm/diffract.py:215-225), one thread per (k, plane) on the full spectrum
// of the E field:
//   u = K_a v scale   (second half step of the layer before the interface)
//   g = T u + refl    (Fresnel transmission between the isotropic effective
//                      media, plus the wave reflected there by the previous
//                      pass)
//   refl <- (T - 1) u (the reflected tangential E: t = 1 + r)
//   dst = K_b g       (first half step of the layer after it)
// K = exp(i k0 d/2 (a(k) - a(beta_beam))), zero above betamax.  The
// backward direction (mode -1) has the same scalars with the thickness |d|
// and the same tangential Fresnel coefficients (a mirror in z keeps the
// tangential E field).  Without fresnel (equal media) T = 1 and refl is not
// touched.  With g set, g also receives T u + refl (the field entering the
// layer, for the bulk of a backward pass).
// ---------------------------------------------------------------------------

template <typename T>
struct TrrArgs {
  const cpx<T>* src;
  cpx<T>* dst;
  cpx<T>* refl;
  cpx<T>* g;
  int H, W;
  Batch b;
  const int* amax;
  double da, db;  // sub-step thickness (>= 0) of the layers before / after
  cd ea, eb;      // their effective eps
  int fresnel;    // 1: transmission from ea into eb
  double scale;
  Beams bm;
};

template <typename T>
__global__ __launch_bounds__(NT) void trr2_kernel(TrrArgs<T> A) {
  const int k = blockIdx.x * NT + threadIdx.x;
  const int ray = blockIdx.y / A.b.nlam_b;
  const int lam = A.b.lam0 + blockIdx.y % A.b.nlam_b;
  const int q = ray * A.b.nlam + lam;
  const size_t HW = size_t(A.H) * A.W;
  if (k >= int(HW)) return;
  const int ky = k / A.W, kx = k - ky * A.W;
  const int beam = ray / A.bm.rays_per_beam;
  const double k0 = A.bm.k0[lam];
  const double bb = A.bm.beta[beam];
  const bool pass = mode_index(kx, A.W) <= A.amax[lam * A.H + ky];
  const double fx = fftfreq(kx, A.W), fy = fftfreq(ky, A.H);
  const double beta = (2 * M_PI / k0) * sqrt(fx * fx + fy * fy);
  const double fr = sqrt(fx * fx + fy * fy);
  const double c = fr > 0. ? fx / fr : 1., sn = fr > 0. ? fy / fr : 0.;
  const size_t o = size_t(q) * 2 * HW + k;
  cd u0 = mk<double>(0.), u1 = mk<double>(0.), r0 = mk<double>(0.), r1 = mk<double>(0.);
  cd g0 = mk<double>(0.), g1 = mk<double>(0.), d0 = mk<double>(0.), d1 = mk<double>(0.);
  if (pass) {
    cd pa = mk<double>(0.), pb = mk<double>(0.);
    if (A.da != 0.)
      pa = (csqrt(A.ea - mk<double>(beta * beta)) - csqrt(A.ea - mk<double>(bb * bb))) *
           (0.5 * A.da * k0);
    if (A.db != 0.)
      pb = (csqrt(A.eb - mk<double>(beta * beta)) - csqrt(A.eb - mk<double>(bb * bb))) *
           (0.5 * A.db * k0);
    const cd Ka = cexpi(pa) * A.scale, Kb = cexpi(pb);
    u0 = cast<double>(A.src[o]) * Ka;
    u1 = cast<double>(A.src[o + HW]) * Ka;
    cd t0 = u0, t1 = u1;
    if (A.fresnel) {
      cd tp, ts;
      fresnel_t(beta, A.ea, A.eb, tp, ts);
      const cd ep = (u0 * c + u1 * sn) * tp, es = (u1 * c - u0 * sn) * ts;
      t0 = ep * c - es * sn;
      t1 = ep * sn + es * c;
      r0 = cast<double>(A.refl[o]);
      r1 = cast<double>(A.refl[o + HW]);
    }
    g0 = t0 + r0;
    g1 = t1 + r1;
    d0 = g0 * Kb;
    d1 = g1 * Kb;
    u0 = t0 - u0;  // the new stored reflection
    u1 = t1 - u1;
  }
  if (A.fresnel) {
    A.refl[o] = cast<T>(u0);
    A.refl[o + HW] = cast<T>(u1);
  }
  if (A.g) {
    A.g[o] = cast<T>(g0);
    A.g[o + HW] = cast<T>(g1);
  }
  A.dst[o] = cast<T>(d0);
  A.dst[o + HW] = cast<T>(d1);
}

// ---------------------------------------------------------------------------
// Polarizing-microscope specter (dtmm/field_viewer.py:70-143, :1262-1289;
// dtmm/field.py:800-845) and its colour conversion (dtmm/color.py:52-200).
// ---------------------------------------------------------------------------


template <typename T>
struct PomArgs {
  const cpx<T>* src;  // fields [nf][npol][nlam][4][HW]
  cpx<T>* dst;        // combined fields [nf][nlam][4][HW]
  T* spec;            // specter [HW][nlam]
  long long HW;
  int nlam, nf, npol;
  cd jc, js;          // weights of the x and y polarised fields (npol == 2)
  const cd* pmat;     // [npm][4][4] analyser matrices (npm > 1), or null
  int npm;            // 0: no analyser, 1: pm0, > 1: pmat per wavelength
  cd pm0[16];         // the shared analyser matrix (npm == 1)
};


// One thread per pixel: for every wavelength, the sum over fields of the
// Poynting z of (analyser matrix) . (field), written with the wavelength
// last.
template <typename T>
__global__ __launch_bounds__(NT) void pom_specter_kernel(PomArgs<T> A) {
  const long long p = (long long)blockIdx.x * NT + threadIdx.x;
  if (p >= A.HW) return;
  const cpx<T>* __restrict__ src = A.src;
  T* __restrict__ spec = A.spec;
  const size_t HW = size_t(A.HW);
  // four wavelengths per round: all their loads are issued before any math
  constexpr int LB = 4;
  for (int lam0 = 0; lam0 < A.nlam; lam0 += LB) {
    const int nb = min(LB, A.nlam - lam0);
    double acc[LB] = {0., 0., 0., 0.};
    for (int f = 0; f < A.nf; ++f) {
      cd v[LB][4];
#pragma unroll
      for (int k = 0; k < LB; ++k) {
        if (k >= nb) break;
        const int lam = lam0 + k;
        if (A.npol == 2) {
          const cpx<T>* s = src + (size_t(f) * 2 * A.nlam + lam) * 4 * HW + p;
          const size_t py = size_t(A.nlam) * 4 * HW;
#pragma unroll
          for (int c = 0; c < 4; ++c)
            v[k][c] = A.jc * cast<double>(s[c * HW]) + A.js * cast<double>(s[c * HW + py]);
        } else {
          const cpx<T>* s = src + (size_t(f) * A.nlam + lam) * 4 * HW + p;
#pragma unroll
          for (int c = 0; c < 4; ++c) v[k][c] = cast<double>(s[c * HW]);
        }
      }
#pragma unroll
      for (int k = 0; k < LB; ++k) {
        if (k >= nb) break;
        // separate branches keep each matrix in its own address space
        // (kernel arguments / global), so its uniform loads stay scalar
        if (A.npm == 0) {
          acc[k] += intensity4(v[k]);
        } else if (A.npm == 1) {
          cd w[4];
          mat4v(A.pm0, v[k], w);
          acc[k] += intensity4(w);
        } else {
          cd w[4];
          mat4v(A.pmat + (lam0 + k) * 16, v[k], w);
          acc[k] += intensity4(w);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < LB; ++k)
      if (k < nb) spec[size_t(p) * A.nlam + lam0 + k] = T(acc[k]);
  }
}

constexpr int COLOR_MAXL = 128;  // wavelengths whose cmf travels in the kernel arguments

template <typename T>
struct ColorArgs {
  const T* spec;      // [n][nlam]
  T* out;             // [n][3]
  long long n;
  int nlam;
  const double* cmf;  // device [nlam][3] (nlam > COLOR_MAXL), else null
  double cmf_v[COLOR_MAXL * 3];
  double* part;       // per-block maxima of the XYZ values (stage 0)
  int nblocks;
  double norm;        // divide XYZ by this (stage 2 reads part[0] when auto)
  int norm_auto;
  int gray;
  int gamma_mode;     // 0 none, 1 sRGB curve, 2 power 1/gamma
  double gamma;
};

// sRGB of D65 white: the IEC 61966-2-1 XYZ -> linear RGB matrix and transfer
// curve constants (dtmm/color.py:41-50).
__device__ inline double srgb_curve(double v) {
  if (v < 0.0031308) return v < 0. ? 0. : 12.92 * v;
  return v > 1. ? 1. : (1. + 0.055) * pow(v, 1. / 2.4) - 0.055;
}

__device__ inline double power_curve(double v, double g) {
  if (v > 1.) return 1.;
  if (v < 0.) return 0.;
  return pow(v, 1. / g);
}

template <typename T>
__device__ inline void spec_to_xyz(const T* spec, const double* cmf, int nlam, double xyz[3]) {
  xyz[0] = xyz[1] = xyz[2] = 0.;
  for (int i = 0; i < nlam; ++i) {
    const double s = double(spec[i]);
    xyz[0] = xyz[0] + cmf[3 * i] * s;
    xyz[1] = xyz[1] + cmf[3 * i + 1] * s;
    xyz[2] = xyz[2] + cmf[3 * i + 2] * s;
  }
}

template <typename T>
__device__ inline void xyz_to_out(const ColorArgs<T>& A, double xyz[3], double norm, long long p) {
  if (norm != 0.)
    for (int k = 0; k < 3; ++k) xyz[k] = xyz[k] / norm;
  double o[3];
  if (A.gray) {
    o[0] = o[1] = o[2] = xyz[1];
  } else {
    o[0] = 3.2404542 * xyz[0] + -1.5371385 * xyz[1] + -0.4985314 * xyz[2];
    o[1] = -0.9692660 * xyz[0] + 1.8760108 * xyz[1] + 0.0415560 * xyz[2];
    o[2] = 0.0556434 * xyz[0] + -0.2040259 * xyz[1] + 1.0572252 * xyz[2];
  }
  for (int k = 0; k < 3; ++k) {
    double v = o[k];
    if (A.gamma_mode == 1) v = srgb_curve(v);
    else if (A.gamma_mode == 2) v = power_curve(v, A.gamma);
    A.out[size_t(p) * 3 + k] = T(v);
  }
}

// STAGE 0: XYZ into out, per-block maximum into part (automatic norm).
// STAGE 1: one work-group reduces part to part[0].
// STAGE 2: XYZ (from out when automatic, else from the specter) -> colour.
template <typename T, int STAGE>
__global__ __launch_bounds__(NT) void color_kernel(ColorArgs<T> A) {
  __shared__ double red[NT];
  if constexpr (STAGE == 1) {
    double m = -INFINITY;
    for (int i = threadIdx.x; i < A.nblocks; i += NT) m = fmax(m, A.part[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) A.part[0] = red[0];
    return;
  } else {
    const long long p = (long long)blockIdx.x * NT + threadIdx.x;
    double xyz[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (p < A.n) {
      if (STAGE == 2 && A.norm_auto) {
        for (int k = 0; k < 3; ++k) xyz[k] = double(A.out[size_t(p) * 3 + k]);
      } else {
        spec_to_xyz<T>(A.spec + size_t(p) * A.nlam, A.cmf ? A.cmf : A.cmf_v, A.nlam, xyz);
      }
    }
    if constexpr (STAGE == 0) {
      if (p < A.n)
        for (int k = 0; k < 3; ++k) A.out[size_t(p) * 3 + k] = T(xyz[k]);
      red[threadIdx.x] = fmax(xyz[0], fmax(xyz[1], xyz[2]));
      __syncthreads();
      for (int w = NT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
      }
      if (threadIdx.x == 0) A.part[blockIdx.x] = red[0];
    } else {
      if (p >= A.n) return;
      xyz_to_out<T>(A, xyz, A.norm_auto ? A.part[0] : A.norm, p);
    }
  }
}

// ---------------------------------------------------------------------------
// Optical data of a director field (dtmm/data.py:277-320 director2data with
// director2order :962-972, uniaxial_order :1215-1274, director2angles
// :975-989): one thread per voxel, grid (blocks per layer, layers).
// ---------------------------------------------------------------------------

struct DirArgs {
  const double* director;     // [n][3]
  const unsigned char* mask;  // [n] or null
  long long layer_size, n;    // voxels per layer, all voxels
  double eo, ee, eh;          // no^2, ne^2, nhost^2
  double scale;               // scale_factor (the order is divided by it)
  double* eps_re;             // [3][n]
  double* angles;             // [3][n]
  double* sums;               // [layers][gridDim.x]: sums of (e0 + e1 + e2) / 3
};

static __global__ __launch_bounds__(NT) void director2data_kernel(DirArgs A) {
#pragma clang fp contract(off)
  __shared__ double red[NT];
  const long long x = (long long)blockIdx.x * NT + threadIdx.x;
  const long long v = (long long)blockIdx.y * A.layer_size + x;
  double iso = 0.;
  if (x < A.layer_size) {
    const double dx = A.director[3 * v], dy = A.director[3 * v + 1], dz = A.director[3 * v + 2];
    double e0, e2;
    if (A.mask && !A.mask[v]) {
      e0 = e2 = A.eh;
    } else {
      const double o = sqrt(sqrt(dx * dx + dy * dy + dz * dz)) / A.scale;
      const double m = (A.eo + A.eo + A.ee) / 3.;
      const double delta = A.ee - (A.eo + A.eo) / 2.;
      e0 = o == 0. ? m : m - 1. / 3. * o * delta;
      e2 = o == 0. ? m : m + 2. / 3. * o * delta;
    }
    A.eps_re[v] = e0;
    A.eps_re[A.n + v] = e0;
    A.eps_re[2 * A.n + v] = e2;
    A.angles[v] = 0.;
    A.angles[A.n + v] = atan2(sqrt(dx * dx + dy * dy), dz);
    A.angles[2 * A.n + v] = atan2(dy, dx);
    iso = (e0 + e0 + e2) / 3.;
  }
  red[threadIdx.x] = iso;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) A.sums[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// Optical data of a Q-tensor field (dtmm/data.py:322-373 Q2data): per voxel
// eps = (ne^2 - no^2) / scale_factor * Q + (2 no^2 + ne^2) / 3 (Q2eps,
// data.py:432-487), its eigenvalues and the Euler angles of its eigenframe
// (eps2epsva, data.py:489-513).  The eigenvalues are sorted as tensor_eig
// sorts them (dtmm/linalg.py:172-222): ascending, then the most distinct one
// is moved to index 2 by a cyclic permutation, which keeps the frame
// right-handed.  The eigensolve is cyclic Jacobi (a different method from the
// reference's closed form: the eigenvalues agree to rounding and each
// eigenvector up to its sign, so the rebuilt tensor R diag(eps) R^T is the
// same).
struct QArgs {
  const double* q;            // [n][6]: Q00, Q11, Q22, Q01, Q02, Q12
  const unsigned char* mask;  // [n] or null
  long long layer_size, n;
  double epsa, epsm, eh;      // (ne^2 - no^2) / scale_factor, (2 no^2 + ne^2) / 3, nhost^2
  int biaxial;                // 0: uniaxial_order(1, eig) (e0 = e1 = their mean)
  double* eps_re;             // [3][n]
  double* angles;             // [3][n]
  double* sums;               // [layers][gridDim.x]: sums of (e0 + e1 + e2) / 3
};

__device__ inline void jacobi_rotate(double a[3][3], double v[3][3], int p, int q) {
  if (a[p][q] == 0.) return;
  const double th = (a[q][q] - a[p][p]) / (2. * a[p][q]);
  const double t = fabs(th) > 1e150 ? 0.5 / th
                                    : copysign(1., th) / (fabs(th) + sqrt(th * th + 1.));
  const double c = 1. / sqrt(t * t + 1.), s = t * c;
  const double apq = a[p][q];
  a[p][p] -= t * apq;
  a[q][q] += t * apq;
  a[p][q] = a[q][p] = 0.;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    if (r != p && r != q) {
      const double arp = a[r][p], arq = a[r][q];
      a[r][p] = a[p][r] = c * arp - s * arq;
      a[r][q] = a[q][r] = s * arp + c * arq;
    }
    const double vrp = v[r][p], vrq = v[r][q];
    v[r][p] = c * vrp - s * vrq;
    v[r][q] = s * vrp + c * vrq;
  }
}

__device__ inline void swap_eig(double w[3], double v[3][3], int i, int j) {
  const double t = w[i];
  w[i] = w[j];
  w[j] = t;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double u = v[r][i];
    v[r][i] = v[r][j];
    v[r][j] = u;
  }
}

static __global__ __launch_bounds__(NT) void q2data_kernel(QArgs A) {
  __shared__ double red[NT];
  const long long x = (long long)blockIdx.x * NT + threadIdx.x;
  const long long v = (long long)blockIdx.y * A.layer_size + x;
  double iso = 0.;
  if (x < A.layer_size) {
    const double* q = A.q + 6 * v;
    double a[3][3], R[3][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
    a[0][0] = A.epsa * q[0] + A.epsm;
    a[1][1] = A.epsa * q[1] + A.epsm;
    a[2][2] = A.epsa * q[2] + A.epsm;
    a[0][1] = a[1][0] = A.epsa * q[3];
    a[0][2] = a[2][0] = A.epsa * q[4];
    a[1][2] = a[2][1] = A.epsa * q[5];
    // A tensor whose off-diagonal part, relative to its largest element, has
    // a squared norm of at most 1e-12 counts as diagonal: its eigenvalues
    // keep the input order with the identity frame (dtmm/linalg.py:251-318).
    const double scale = fmax(fmax(fmax(fabs(a[0][0]), fabs(a[1][1])), fmax(fabs(a[2][2]),
                              fabs(a[0][1]))), fmax(fabs(a[0][2]), fabs(a[1][2])));
    const double sm = scale > 0. ? 1. / scale : 0.;
    const double s01 = a[0][1] * sm, s02 = a[0][2] * sm, s12 = a[1][2] * sm;
    const bool diagonal = !(s01 * s01 + s02 * s02 + s12 * s12 > 1e-12);
    double w[3];
    if (diagonal) {
      w[0] = a[0][0] * sm * scale;
      w[1] = a[1][1] * sm * scale;
      w[2] = a[2][2] * sm * scale;
    } else {
      for (int sweep = 0; sweep < 16; ++sweep) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (!(off > 1e-36 * dia)) break;
        jacobi_rotate(a, R, 0, 1);
        jacobi_rotate(a, R, 0, 2);
        jacobi_rotate(a, R, 1, 2);
      }
      w[0] = a[0][0];
      w[1] = a[1][1];
      w[2] = a[2][2];
      if (w[1] < w[0]) swap_eig(w, R, 0, 1);
      if (w[2] < w[1]) swap_eig(w, R, 1, 2);
      if (w[1] < w[0]) swap_eig(w, R, 0, 1);
    }
    // right-handed frame: column 2 = column 0 x column 1
    R[0][2] = R[1][0] * R[2][1] - R[2][0] * R[1][1];
    R[1][2] = R[2][0] * R[0][1] - R[0][0] * R[2][1];
    R[2][2] = R[0][0] * R[1][1] - R[1][0] * R[0][1];
    // the most distinct eigenvalue to index 2 (dtmm/linalg.py:179-193)
    const double m2 = fabs(w[1] - w[0]), m1 = fabs(w[2] - w[0]), m0 = fabs(w[2] - w[1]);
    int i0 = 0, i1 = 1, i2 = 2;
    if (m2 > m1 || m2 > m0) {
      if (m1 > m0) {
        i0 = 1; i1 = 2; i2 = 0;
      } else {
        i0 = 2; i1 = 0; i2 = 1;
      }
    }
    double e0 = w[i0], e1 = w[i1], e2 = w[i2];
    double S[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      S[r][0] = R[r][i0];
      S[r][1] = R[r][i1];
      S[r][2] = R[r][i2];
    }
    // Euler angles (dtmm/rotation.py:154-171)
    const double r22 = fmax(-1., fmin(1., S[2][2]));
    const double theta = acos(r22);
    double psi, phi;
    if ((S[1][2] == 0. && S[0][2] == 0.) || theta == 0.) {
      psi = 0.;
      phi = acos(fmax(-1., fmin(1., S[1][1])));
    } else {
      phi = atan2(S[1][2], S[0][2]);
      psi = atan2(S[2][1], -S[2][0]);
    }
    if (!A.biaxial) {  // uniaxial_order(1, eig) (dtmm/data.py:1215-1233)
      const double m = (e0 + e1 + e2) / 3.;
      const double delta = e2 - (e0 + e1) / 2.;
      e0 = e1 = m - 1. / 3. * delta;
      e2 = m + 2. / 3. * delta;
    }
    if (A.mask && !A.mask[v]) e0 = e1 = e2 = A.eh;
    A.eps_re[v] = e0;
    A.eps_re[A.n + v] = e1;
    A.eps_re[2 * A.n + v] = e2;
    A.angles[v] = psi;
    A.angles[A.n + v] = theta;
    A.angles[2 * A.n + v] = phi;
    iso = (e0 + e1 + e2) / 3.;
  }
  red[threadIdx.x] = iso;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) A.sums[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// Non-iterative 1D transfer matrix method (dtmm/tmm.py:1112-1673, method
// "4x4"): one thread per independent stack, fp64 throughout.
// ---------------------------------------------------------------------------

struct TmmArgs {
  long long n;          // stacks
  int L;                // layers
  const double* kd;     // [L][n]
  const double* eps_re; // [3][L][n]
  const double* eps_im; // [3][L][n] or null
  const double* ang;    // [3][L][n]
  double beta, phi;     // beam of the layers
  double cfact;         // 4x4_2, 4x4_r: backward-mode damping
  cd fin[16], fini[16], fout[16], fouti[16];  // isotropic input/output media
  cd* fvec_in;          // [n][4]: incident field, replaced by incident + reflected
                        // (4x4 methods; read only by the 2x2 methods)
  cd* fvec_out;         // [n][4]: transmitted field (incoming backward waves read
                        // first when has_out, 4x4 methods)
  int has_out;
  cd* cmat;             // characteristic matrix [n][4][4] (4x4_r: [n][8][8],
                        // 2x2 methods: [n][2][2]), or null
  // 2x2 methods: Fresnel transmission at the first / last interface, with the
  // first / last layer's modes at the media beam (mbeta, mphi)
  int reflect_in, reflect_out;
  double mbeta, mphi;
};

// Columns scaled to unit |Poynting z| (dtmm/tmm.py:699-724); zero columns stay zero.
__host__ __device__ inline void normalize_cols(cd F[4][4]) {
  for (int j = 0; j < 4; ++j) {
    const double p = (F[0][j].x * F[1][j].x + F[0][j].y * F[1][j].y) -
                     (F[2][j].x * F[3][j].x + F[2][j].y * F[3][j].y);
    const double n = sqrt(fabs(p));
    const double sc = n == 0. ? 0. : 1. / n;
    for (int i = 0; i < 4; ++i) F[i][j] = F[i][j] * sc;
  }
}

__device__ inline void mat16(const cd* a, cd o[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) o[i][j] = a[i * 4 + j];
}

__device__ inline void matv4(const cd m[4][4], const cd v[4], cd o[4]) {
  for (int i = 0; i < 4; ++i) {
    cd t = mk<double>(0.);
    for (int j = 0; j < 4; ++j) t = t + m[i][j] * v[j];
    o[i] = t;
  }
}

// Biaxial voxels (numeric eigensolve) out of line: the closed-form layers
// keep the kernel's registers.
__device__ __noinline__ void tmm_biaxial_modes(double beta, double c0, double s0, double phi0,
                                               cd e0, cd e1, cd e2, double psi, double theta,
                                               double phi, Modes& m) {
  biaxial_modes(beta, e0, e1, e2, psi, theta, phi - phi0, m);
  rotate_rows(m, c0, s0);
}

// Modes (alphaf, dtmm/tmm.py:435-464) of layer l of stack e at the beam
// (beta, phi0), rotated to the lab frame, not normalised.
__device__ inline void tmm_layer_modes(const TmmArgs& A, int l, long long e, double beta,
                                       double phi0, double c0, double s0, Modes& m) {
  const size_t cs = size_t(A.L) * size_t(A.n);
  const size_t o = size_t(l) * size_t(A.n) + size_t(e);
  const cd e0 = mk<double>(A.eps_re[o], A.eps_im ? A.eps_im[o] : 0.);
  const cd e1 = mk<double>(A.eps_re[cs + o], A.eps_im ? A.eps_im[cs + o] : 0.);
  const cd e2 = mk<double>(A.eps_re[2 * cs + o], A.eps_im ? A.eps_im[2 * cs + o] : 0.);
  const int kind = material_kind(e0, e1, e2);
  if (kind == KIND_BIAXIAL) {
    tmm_biaxial_modes(beta, c0, s0, phi0, e0, e1, e2, A.ang[o], A.ang[cs + o], A.ang[2 * cs + o], m);
  } else {
    if (kind == KIND_ISO)
      iso_modes(beta, e0, m);
    else
      uniaxial_modes(beta, e0, e2, A.ang[cs + o], A.ang[2 * cs + o] - phi0, m);
    rotate_rows(m, c0, s0);
  }
}

// C <- C . F . diag(p) . F^-1, one row at a time in place
__device__ inline void mul_layer(cd C[4][4], const cd F[4][4], const cd p[4], const cd Fi[4][4]) {
  for (int i = 0; i < 4; ++i) {
    cd r[4];
    for (int jj = 0; jj < 4; ++jj) {
      cd t = mk<double>(0.);
      for (int k = 0; k < 4; ++k) t = t + C[i][k] * F[k][jj];
      r[jj] = t * p[jj];
    }
    for (int jj = 0; jj < 4; ++jj) {
      cd t = mk<double>(0.);
      for (int k = 0; k < 4; ++k) t = t + r[k] * Fi[k][jj];
      C[i][jj] = t;
    }
  }
}

// The 4x4 methods of dtmm/tmm.py:1112-1213 (layer_mat), chained as in
// stack_mat (:1215-1296) and solved as in transmit4x4 (:1491-1563):
//   M = 0 "4x4":   F P(-kd) F^-1, F normalised;
//   M = 1 "4x4_1": backward phases zeroed;
//   M = 2 "4x4_2": backward modes damped, alpha += 2i cfact Re(alpha);
//   M = 3 "4x4_r": the 8x8 matrix [[F P2 F^-1, 0], [F P F^-1, 0]] (P2 the
//                  "4x4_2" phases, P the plain ones) of layer_mat; its
//                  chained product is [[A, 0], [B, 0]] with A the "4x4_2"
//                  stack and B = F1 P F1^-1 . A_2 ... A_n, and system_mat
//                  keeps rows 0, 5, 2, 7 of its first four columns after
//                  transforming only the diagonal blocks: S = (Fin^-1 A
//                  Fout rows 0, 2; B rows 1, 3).
template <int M>
static __global__ __launch_bounds__(NT) void tmm_kernel(TmmArgs A) {
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= A.n) return;
  double s0, c0;
  sincos(A.phi, &s0, &c0);
  cd C[4][4], B[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) C[i][j] = B[i][j] = mk<double>(i == j ? 1. : 0.);
  for (int l = 0; l < A.L; ++l) {
    Modes m;
    tmm_layer_modes(A, l, e, A.beta, A.phi, c0, s0, m);
    normalize_cols(m.F);
    cd Fi[4][4];
    inv4(m.F, Fi);
    cd p[4];
    const double kd = -A.kd[size_t(l) * size_t(A.n) + size_t(e)];
    for (int q = 0; q < 4; ++q) {
      cd a = m.a[q];
      if ((M == 2 || M == 3) && (q & 1)) a = a + mk<double>(0., 2. * A.cfact * a.x);
      p[q] = cexpi(a * kd);
      if (M == 1 && (q & 1)) p[q] = mk<double>(0.);
    }
    if constexpr (M == 3) {
      if (l == 0) {
        cd pp[4];
        for (int q = 0; q < 4; ++q) pp[q] = cexpi(m.a[q] * kd);
        mul_layer(B, m.F, pp, Fi);
      } else {
        mul_layer(B, m.F, p, Fi);
      }
    }
    mul_layer(C, m.F, p, Fi);
  }
  if (A.cmat) {
    if constexpr (M == 3) {
      cd* o = A.cmat + size_t(e) * 64;
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j)
          o[i * 8 + j] = (j < 4) ? (i < 4 ? C[i][j] : B[i - 4][j]) : mk<double>(0.);
    } else {
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) A.cmat[size_t(e) * 16 + i * 4 + j] = C[i][j];
    }
  }
  if (!A.fvec_in) return;
  // transmit4x4 (tmm.py:1491-1563) with system_mat (:1298-1336) and
  // reflection_mat (:1338-1357); the media matrices are read from the
  // kernel arguments
  cd S[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      cd t = mk<double>(0.);
      for (int k = 0; k < 4; ++k) t = t + A.fini[i * 4 + k] * C[k][j];
      S[i][j] = t;
    }
  for (int i = 0; i < 4; ++i) {
    cd r[4];
    for (int j = 0; j < 4; ++j) {
      cd t = mk<double>(0.);
      for (int k = 0; k < 4; ++k) t = t + S[i][k] * A.fout[k * 4 + j];
      r[j] = t;
    }
    for (int j = 0; j < 4; ++j) S[i][j] = r[j];
  }
  if constexpr (M == 3) {
    for (int j = 0; j < 4; ++j) {
      S[1][j] = B[1][j];
      S[3][j] = B[3][j];
    }
  }
  // reflection_mat: m1 = [-S[:, 0], e1, -S[:, 2], e3], m2 = [-e0, S[:, 1], -e2, S[:, 3]]
  cd m1[4][4];
  for (int i = 0; i < 4; ++i) {
    m1[i][0] = -S[i][0];
    m1[i][1] = mk<double>(i == 1 ? 1. : 0.);
    m1[i][2] = -S[i][2];
    m1[i][3] = mk<double>(i == 3 ? 1. : 0.);
  }
  cd m1i[4][4];
  inv4(m1, m1i);
  cd fv[4], av[4], a[4], bv[4], out[4];
  for (int i = 0; i < 4; ++i) fv[i] = A.fvec_in[size_t(e) * 4 + i];
  for (int i = 0; i < 4; ++i) {
    cd t = mk<double>(0.);
    for (int k = 0; k < 4; ++k) t = t + A.fini[i * 4 + k] * fv[k];
    av[i] = t;
  }
  av[1] = av[3] = mk<double>(0.);
  for (int i = 0; i < 4; ++i) {
    a[i] = av[i];
    bv[i] = mk<double>(0.);
  }
  if (A.has_out) {
    cd fo[4];
    for (int i = 0; i < 4; ++i) fo[i] = A.fvec_out[size_t(e) * 4 + i];
    for (int q = 1; q < 4; q += 2) {
      cd t = mk<double>(0.);
      for (int k = 0; k < 4; ++k) t = t + A.fouti[q * 4 + k] * fo[k];
      bv[q] = a[q] = t;
    }
  }
  // out = m1^-1 . (m2 . a), with m2 . a = (-a0, S a odd, ...) written out
  cd m2a[4];
  for (int i = 0; i < 4; ++i) {
    cd t = S[i][1] * a[1] + S[i][3] * a[3];
    if (i == 0) t = t - a[0];
    if (i == 2) t = t - a[2];
    m2a[i] = t;
  }
  for (int i = 0; i < 4; ++i) {
    cd t = mk<double>(0.);
    for (int k = 0; k < 4; ++k) t = t + m1i[i][k] * m2a[k];
    out[i] = t;
  }
  av[1] = out[1];
  av[3] = out[3];
  bv[0] = out[0];
  bv[2] = out[2];
  for (int i = 0; i < 4; ++i) {
    cd t = mk<double>(0.), u = mk<double>(0.);
    for (int k = 0; k < 4; ++k) {
      t = t + A.fin[i * 4 + k] * av[k];
      u = u + A.fout[i * 4 + k] * bv[k];
    }
    A.fvec_in[size_t(e) * 4 + i] = t;
    A.fvec_out[size_t(e) * 4 + i] = u;
  }
}

__device__ inline void mat2mul(const cd a[2][2], const cd b[2][2], cd o[2][2]) {
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) o[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j];
}

// E rows / forward-mode columns of F (E_mat mode +1, dtmm/tmm.py:974-997)
__device__ inline void e_fwd(const cd F[4][4], cd E[2][2]) {
  E[0][0] = F[0][0];
  E[0][1] = F[0][2];
  E[1][0] = F[2][0];
  E[1][1] = F[2][2];
}

// Eti_mat (dtmm/tmm.py:999-1022): inv(St) . inv(A), A = E_mat(Fin, +1),
// St = (Fin^-1 Fout)[::2, ::2]
__device__ inline void eti_mat(const cd Fin[4][4], const cd Fout[4][4], cd O[2][2]) {
  cd Fii[4][4];
  inv4(Fin, Fii);
  cd St[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      cd t = mk<double>(0.);
      for (int k = 0; k < 4; ++k) t = t + Fii[2 * i][k] * Fout[k][2 * j];
      St[i][j] = t;
    }
  cd Ea[2][2], Sti[2][2], Eai[2][2];
  e_fwd(Fin, Ea);
  inv2(St, Sti);
  inv2(Ea, Eai);
  mat2mul(Sti, Eai, O);
}

// t_mat (dtmm/tmm.py:952-972): E_mat(Fout, +1) . Eti_mat(Fin, Fout)
__device__ inline void t_mat2(const cd Fin[4][4], const cd Fout[4][4], cd O[2][2]) {
  cd Et[2][2], Eti[2][2];
  e_fwd(Fout, Et);
  eti_mat(Fin, Fout, Eti);
  mat2mul(Et, Eti, O);
}

// The 2x2 methods (dtmm/tmm.py:1112-1213, :1215-1296, :1675-1748):
// M = E P(kd) E^-1 of the forward modes (F not normalised), chained as
// M_n ... M_1; ETI ("2x2_1") enters every layer after the first through
// Eti_mat from the previous layer's modes.  transfer: E = the forward part
// of fvec_in in the input medium (fvec2E), optional Fresnel t_mat into the
// first layer and out of the last one (their modes at the media beam),
// then E2fvec in the output medium.
template <bool ETI>
static __global__ __launch_bounds__(NT) void tmm2_kernel(TmmArgs A) {
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= A.n) return;
  double s0, c0;
  sincos(A.phi, &s0, &c0);
  cd C[2][2] = {{mk<double>(1.), mk<double>(0.)}, {mk<double>(0.), mk<double>(1.)}};
  cd Fp[4][4];
  for (int l = 0; l < A.L; ++l) {
    Modes m;
    tmm_layer_modes(A, l, e, A.beta, A.phi, c0, s0, m);
    cd E[2][2], Ei[2][2];
    e_fwd(m.F, E);
    if (ETI && l > 0)
      eti_mat(Fp, m.F, Ei);
    else
      inv2(E, Ei);
    const double kd = A.kd[size_t(l) * size_t(A.n) + size_t(e)];
    const cd p0 = cexpi(m.a[0] * kd), p2 = cexpi(m.a[2] * kd);
    cd Ml[2][2], T[2][2];
    for (int i = 0; i < 2; ++i) {
      Ml[i][0] = E[i][0] * p0 * Ei[0][0] + E[i][1] * p2 * Ei[1][0];
      Ml[i][1] = E[i][0] * p0 * Ei[0][1] + E[i][1] * p2 * Ei[1][1];
    }
    mat2mul(Ml, C, T);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) C[i][j] = T[i][j];
    if (ETI)
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Fp[i][j] = m.F[i][j];
  }
  if (A.cmat)
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) A.cmat[size_t(e) * 4 + i * 2 + j] = C[i][j];
  if (!A.fvec_in) return;
  cd fin[4][4], fini[4][4], fout[4][4];
  mat16(A.fin, fin);
  mat16(A.fini, fini);
  mat16(A.fout, fout);
  // fvec2E: rows Ex, Ey of Fin diag(1, 0, 1, 0) Fin^-1 applied to fvec_in
  cd fv[4], am[4];
  for (int i = 0; i < 4; ++i) fv[i] = A.fvec_in[size_t(e) * 4 + i];
  matv4(fini, fv, am);
  cd ev[2];
  for (int i = 0; i < 2; ++i) ev[i] = fin[2 * i][0] * am[0] + fin[2 * i][2] * am[2];
  double ms, mc;
  sincos(A.mphi, &ms, &mc);
  auto apply = [&](const cd M2[2][2]) {
    const cd x = M2[0][0] * ev[0] + M2[0][1] * ev[1], y = M2[1][0] * ev[0] + M2[1][1] * ev[1];
    ev[0] = x;
    ev[1] = y;
  };
  if (A.reflect_in) {
    Modes m;
    tmm_layer_modes(A, 0, e, A.mbeta, A.mphi, mc, ms, m);
    cd T[2][2];
    t_mat2(fin, m.F, T);
    apply(T);
  }
  apply(C);
  if (A.reflect_out) {
    Modes m;
    tmm_layer_modes(A, A.L - 1, e, A.mbeta, A.mphi, mc, ms, m);
    cd T[2][2];
    t_mat2(m.F, fout, T);
    apply(T);
  }
  // E2fvec in the output medium: H = B A^-1 E with the forward modes of Fout
  cd Ea[2][2], Eai[2][2];
  e_fwd(fout, Ea);
  inv2(Ea, Eai);
  cd* o = A.fvec_out + size_t(e) * 4;
  for (int i = 0; i < 2; ++i) {
    const cd b0 = fout[2 * i + 1][0] * Eai[0][0] + fout[2 * i + 1][2] * Eai[1][0];
    const cd b1 = fout[2 * i + 1][0] * Eai[0][1] + fout[2 * i + 1][2] * Eai[1][1];
    o[2 * i] = ev[i];
    o[2 * i + 1] = b0 * ev[0] + b1 * ev[1];
  }
}

// ---------------------------------------------------------------------------
// Illumination field on the device (dtmm/field.py:586-713): plane waves of
// each beam (wave.py:376-409; times a window, or snapped to the strongest FFT
// eigen-wave, wave.py:432-476, when there is none) carried by the normalised
// field vector of the isotropic medium with the ray's Jones vector
// (field.py:566-582, tmm.py:1750-1843).  One thread per (pixel, plane); per
// block the partial sums of the Poynting z of the planes (for the power
// normalisation of field.py:703-711).
// ---------------------------------------------------------------------------

template <typename T>
struct IllumArgs {
  cpx<T>* field;        // [ray][lam][4][H][W]
  int H, W;
  int nlam, nrays, rays_per_beam;
  double n;             // refractive index of the medium
  int mode;             // +1 forward, -1 backward
  int polarized;        // 1: jones for every ray; 0: x (pol 0) and y (pol 1)
  cd jones[2];
  const double* window; // [H][W] or null (eigen-wave snapping)
  double* part;         // [planes][gridDim.x] partial Poynting sums, or null
  const double* scale;  // FSCALE: [planes] factors, applied in place
  Beams bm;
};

enum { ILL_BUILD = 0, ILL_POWER = 1, ILL_SCALE = 2 };

// The strongest FFT mode of a sampled plane wave exp(i k x), x = 0..n-1: the
// integer frequency nearest to k n / (2 pi), ties to the lower index (the
// first maximum of |fft|^2, as numpy's argmax reads it).
__device__ inline int eigen_index(double k, int n) {
  const double u = k * n / (2 * M_PI);
  long long m = (long long)ceil(u - 0.5);
  m %= n;
  if (m < 0) m += n;
  return int(m);
}

template <typename T, int OP>
__global__ __launch_bounds__(NT) void illum_kernel(IllumArgs<T> A) {
  __shared__ double red[NT];
  const int tid = threadIdx.x;
  const size_t HW = size_t(A.H) * A.W;
  const size_t k = size_t(blockIdx.x) * NT + tid;
  const int q = blockIdx.y;  // plane = ray * nlam + lam
  const int ray = q / A.nlam, lam = q - ray * A.nlam;
  cpx<T>* f = A.field + size_t(q) * 4 * HW;
  double pz = 0.;
  if (k < HW) {
    if constexpr (OP == ILL_BUILD) {
      const int y = int(k / A.W), x = int(k - size_t(y) * A.W);
      const int beam = ray / A.rays_per_beam, pol = ray - beam * A.rays_per_beam;
      const double beta = A.bm.beta[beam], phi = A.bm.phi[beam], k0 = A.bm.k0[lam];
      double sp, cp;
      sincos(phi, &sp, &cp);
      const double kx = k0 * beta * cp, ky = k0 * beta * sp;
      cd wave;
      if (A.window) {
        double s, c;
        sincos(kx * x + ky * y, &s, &c);
        wave = mk<double>(c, s) * A.window[k];
      } else {
        const int jj = eigen_index(kx, A.W), ii = eigen_index(ky, A.H);
        const double a = 2 * M_PI * (double((long long)jj * x % A.W) / A.W +
                                     double((long long)ii * y % A.H) / A.H);
        double s, c;
        sincos(a, &s, &c);
        wave = mk<double>(c, s);
      }
      // Jones vector of the ray, normalised, in the frame of the beam
      cd j0 = A.polarized ? A.jones[0] : mk<double>(pol == 0 ? 1. : 0.);
      cd j1 = A.polarized ? A.jones[1] : mk<double>(pol == 0 ? 0. : 1.);
      const double nj = sqrt(norm2(j0) + norm2(j1));
      j0 = j0 * (1. / nj);
      j1 = j1 * (1. / nj);
      const cd jc = j0 * cp + j1 * sp, js = j1 * cp - j0 * sp;  // rotation by -phi
      Modes m;
      iso_modes(beta, mk<double>(A.n * A.n), m);
      rotate_rows(m, cp, sp);
      normalize_cols(m.F);
      const int c0 = A.mode > 0 ? 0 : 1, c1 = c0 + 2;
      for (int i = 0; i < 4; ++i) {
        const cd v = (m.F[i][c0] * jc + m.F[i][c1] * js) * wave;
        f[i * HW + k] = cast<T>(v);
      }
    } else if constexpr (OP == ILL_SCALE) {
      const T sc = T(A.scale[q]);
      for (int i = 0; i < 4; ++i) f[i * HW + k] = f[i * HW + k] * sc;
    }
    if (A.part) {
      const cd ex = cast<double>(f[k]), hy = cast<double>(f[HW + k]);
      const cd ey = cast<double>(f[2 * HW + k]), hx = cast<double>(f[3 * HW + k]);
      pz = (ex.x * hy.x + ex.y * hy.y) - (ey.x * hx.x + ey.y * hx.y);
    }
  }
  if (!A.part) return;
  red[tid] = pz;
  __syncthreads();
  for (int m = NT / 2; m > 0; m >>= 1) {
    if (tid < m) red[tid] += red[tid + m];
    __syncthreads();
  }
  if (tid == 0) A.part[size_t(q) * gridDim.x + blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// The 2D non-iterative method (dtmm/tmm2d.py) for one wavelength, in the
// space of the n masked 1D FFT modes of the x axis.  Block matrices are kept
// flat: element (4 i + a, 4 j + b) of a 4n x 4n complex128 matrix is
// component (a, b) of block (i, j).  fp64 throughout.
// ---------------------------------------------------------------------------

constexpr int T2D_MAXL = 256;  // layers of one chunk (their k0 d travel in the arguments)

struct T2dArgs {
  int W, n;
  const int* idx;        // [n] FFT bin of every mode
  double k0, betay;
  double kd[T2D_MAXL];   // k0 * thickness of each layer of the chunk
  int nl;                // layers in the chunk
  const double* eps_re;  // first layer of the chunk: [3][stride], layer l at l * W
  const double* eps_im;
  const double* ang;
  size_t stride;
  int method;            // 0: 4x4, 1: 4x4_1 (backward phases zeroed)
  cd* M;                 // [nl][n][W][16]: F P(-kd) F^-1 of pixel x at mode j's beam
  const cd* tw;          // [W]: exp(2 pi i t / W)
  cd* C;                 // [nl][4n][4n]
};

// M[l][j][x] = F P(-k0 d_l) F^-1 of pixel x of layer l for the beam of mode
// j (beta_x = 2 pi / k0 fftfreq(idx_j), beta_y; dtmm/tmm2d.py:67-79).
static __global__ __launch_bounds__(NT) void t2d_mode_kernel(T2dArgs A) {
  const int x = blockIdx.x * NT + threadIdx.x;
  const int j = blockIdx.y, l = blockIdx.z;
  if (x >= A.W) return;
  const double bx = (2 * M_PI / A.k0) * fftfreq(A.idx[j], A.W);
  const double beta = sqrt(bx * bx + A.betay * A.betay), phi = atan2(A.betay, bx);
  double s0, c0;
  sincos(phi, &s0, &c0);
  const size_t o = size_t(l) * A.W + x;
  const cd e0 = mk<double>(A.eps_re[o], A.eps_im ? A.eps_im[o] : 0.);
  const cd e1 = mk<double>(A.eps_re[A.stride + o], A.eps_im ? A.eps_im[A.stride + o] : 0.);
  const cd e2 = mk<double>(A.eps_re[2 * A.stride + o], A.eps_im ? A.eps_im[2 * A.stride + o] : 0.);
  Modes m;
  int kind;
  layer_modes(beta, c0, s0, phi, e0, e1, e2, A.ang[o], A.ang[A.stride + o],
              A.ang[2 * A.stride + o], m, kind);
  cd Fi[4][4];
  inv4(m.F, Fi);
  cd p[4];
  for (int q = 0; q < 4; ++q) {
    p[q] = cexpi(m.a[q] * (-A.kd[l]));
    if (A.method == 1 && (q & 1)) p[q] = mk<double>(0.);
  }
  cd* out = A.M + ((size_t(l) * A.n + j) * A.W + x) * 16;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      cd t = mk<double>(0.);
      for (int q = 0; q < 4; ++q) t = t + m.F[a][q] * p[q] * Fi[q][b];
      out[a * 4 + b] = t;
    }
}

// C_l(4i + a, 4j + b) = 1/W sum_x M[l][j][x][a][b] exp(2 pi i (idx_j - idx_i) x / W):
// the masked FFT of M_j times mode j's eigen-wave (dtmm/tmm2d.py:81-87).
// One wave per block (l, i, j): lanes split x and read 256 contiguous bytes
// each (the 16 entries of one pixel), then the 16 sums are reduced in LDS.
static __global__ __launch_bounds__(64) void t2d_dft_kernel(T2dArgs A) {
  const int n = A.n, N = 4 * n;
  const size_t blk = blockIdx.x;
  const int l = int(blk / (size_t(n) * n));
  const int i = int((blk / n) % n), j = int(blk % n);
  const int lane = threadIdx.x;
  __shared__ cd red[64][17];
  int dk = (A.idx[j] - A.idx[i]) % A.W;
  if (dk < 0) dk += A.W;
  const cd* mj = A.M + (size_t(l) * n + j) * A.W * 16;
  cd acc[16];
  for (int e = 0; e < 16; ++e) acc[e] = mk<double>(0.);
  for (int x = lane; x < A.W; x += 64) {
    const cd w = A.tw[int((long long)dk * x % A.W)];
    const cd* m = mj + size_t(x) * 16;
    for (int e = 0; e < 16; ++e) acc[e] = acc[e] + m[e] * w;
  }
  for (int e = 0; e < 16; ++e) red[lane][e] = acc[e];
  __syncthreads();
  if (lane < 16) {
    cd t = mk<double>(0.);
    for (int r = 0; r < 64; ++r) t = t + red[r][lane];
    const int a = lane >> 2, b = lane & 3;
    A.C[size_t(l) * N * N + size_t(4 * i + a) * N + 4 * j + b] = t * (1. / A.W);
  }
}

// out_z = a_z . b_z for N x N complex128 matrices (row-major), 16 x 16
// tiles; batch z at offsets z * sa, z * sb, z * so.
static __global__ __launch_bounds__(256) void zgemm_kernel(const cd* a, size_t sa, const cd* b,
                                                         size_t sb, cd* out, size_t so, int N) {
  __shared__ cd ta[16][17], tb[16][17];
  a += blockIdx.z * sa;
  b += blockIdx.z * sb;
  out += blockIdx.z * so;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
  cd acc = mk<double>(0.);
  for (int k0 = 0; k0 < N; k0 += 16) {
    ta[ty][tx] = (row < N && k0 + tx < N) ? a[size_t(row) * N + k0 + tx] : mk<double>(0.);
    tb[ty][tx] = (k0 + ty < N && col < N) ? b[size_t(k0 + ty) * N + col] : mk<double>(0.);
    __syncthreads();
    for (int k = 0; k < 16; ++k) acc = acc + ta[ty][k] * tb[k][tx];
    __syncthreads();
  }
  if (row < N && col < N) out[size_t(row) * N + col] = acc;
}

// S = diag(Fin_i^-1) . C . diag(Fout_j) block-wise (dtmm/tmm2d.py:144-148),
// with fini and fout [n][4][4].
static __global__ __launch_bounds__(NT) void t2d_system_kernel(const cd* fini, const cd* C,
                                                             const cd* fout, cd* S, int n) {
  const int N = 4 * n;
  const size_t e = size_t(blockIdx.x) * NT + threadIdx.x;
  if (e >= size_t(N) * N) return;
  const int r = int(e / N), c = int(e - size_t(r) * N);
  const int i = r >> 2, a = r & 3, j = c >> 2, b = c & 3;
  cd t = mk<double>(0.);
  for (int p = 0; p < 4; ++p) {
    cd u = mk<double>(0.);
    for (int q = 0; q < 4; ++q) u = u + C[size_t(4 * i + p) * N + 4 * j + q] * fout[(j * 4 + q) * 4 + b];
    t = t + fini[(i * 4 + a) * 4 + p] * u;
  }
  S[e] = t;
}

// reflection_mat (dtmm/tmm.py:1338-1357) of an N x N system matrix: the
// augmented [m1 | m2] with m1 = -S on even columns and e_c on odd ones,
// m2 = -e_c on even columns and S on odd ones; Gauss-Jordan elimination
// then leaves m1^-1 m2 in the right half.
static __global__ __launch_bounds__(NT) void gj_init_kernel(const cd* S, cd* G, int N) {
  const size_t e = size_t(blockIdx.x) * NT + threadIdx.x;
  if (e >= size_t(N) * 2 * N) return;
  const int r = int(e / (2 * N)), c = int(e - size_t(r) * 2 * N);
  cd v;
  if (c < N) {
    v = (c & 1) ? mk<double>(r == c ? 1. : 0.) : -S[size_t(r) * N + c];
  } else {
    const int cc = c - N;
    v = (cc & 1) ? S[size_t(r) * N + cc] : mk<double>(r == cc ? -1. : 0.);
  }
  G[e] = v;
}

// Pivot of column k: the row r >= k with the largest |re| + |im| (the
// measure of LAPACK's izamax), the first one on ties; rows k and p are
// swapped and row k is scaled by its pivot.  One work-group.
static __global__ __launch_bounds__(NT) void gj_pivot_kernel(cd* G, int N, int k, cd* fac) {
  __shared__ double bv[NT];
  __shared__ int bi[NT];
  const int tid = threadIdx.x;
  const int W2 = 2 * N;
  double best = -1.;
  int bidx = k;
  for (int r = k + tid; r < N; r += NT) {
    const cd v = G[size_t(r) * W2 + k];
    const double m = fabs(v.x) + fabs(v.y);
    if (m > best) {
      best = m;
      bidx = r;
    }
  }
  bv[tid] = best;
  bi[tid] = bidx;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const double o = bv[tid + s];
      const int oi = bi[tid + s];
      if (o > bv[tid] || (o == bv[tid] && oi < bi[tid])) {
        bv[tid] = o;
        bi[tid] = oi;
      }
    }
    __syncthreads();
  }
  const int p = bi[0];
  if (p != k)
    for (int c = tid; c < W2; c += NT) {
      const cd t = G[size_t(k) * W2 + c];
      G[size_t(k) * W2 + c] = G[size_t(p) * W2 + c];
      G[size_t(p) * W2 + c] = t;
    }
  __syncthreads();
  const cd piv = G[size_t(k) * W2 + k];
  const double n2 = norm2(piv);
  const cd ip = n2 > 0. ? conj(piv) * (1. / n2) : mk<double>(0.);
  __syncthreads();
  for (int c = tid; c < W2; c += NT) G[size_t(k) * W2 + c] = G[size_t(k) * W2 + c] * ip;
  __syncthreads();
  // the elimination factors of column k (read before any row changes)
  for (int r = tid; r < N; r += NT) fac[r] = (r == k) ? mk<double>(0.) : G[size_t(r) * W2 + k];
}

static __global__ __launch_bounds__(NT) void gj_elim_kernel(cd* G, int N, int k, const cd* fac) {
  const size_t e = size_t(blockIdx.x) * NT + threadIdx.x;
  const int W2 = 2 * N;
  if (e >= size_t(N) * W2) return;
  const int r = int(e / W2), c = int(e - size_t(r) * W2);
  if (r == k) return;
  const cd f = fac[r];
  if (f.x == 0. && f.y == 0.) return;
  G[e] = G[e] - f * G[size_t(k) * W2 + c];
}

static __global__ __launch_bounds__(NT) void gj_extract_kernel(const cd* G, cd* R, int N) {
  const size_t e = size_t(blockIdx.x) * NT + threadIdx.x;
  if (e >= size_t(N) * N) return;
  const int r = int(e / N), c = int(e - size_t(r) * N);
  R[e] = G[size_t(r) * 2 * N + N + c];
}

// reflection_mat for small systems (N <= GJ_SMALL) in one work-group: the
// same augmented matrix, pivot rule and elimination steps as gj_init /
// gj_pivot / gj_elim / gj_extract, held in LDS.
constexpr int GJ_SMALL = 40;

static __global__ __launch_bounds__(256) void gj_small_kernel(const cd* S, cd* R, int N) {
  __shared__ cd G[GJ_SMALL][2 * GJ_SMALL];
  __shared__ cd fac[GJ_SMALL];
  __shared__ int piv_row;
  const int tid = threadIdx.x, W2 = 2 * N;
  for (int e = tid; e < N * W2; e += 256) {
    const int r = e / W2, c = e - r * W2;
    cd v;
    if (c < N) {
      v = (c & 1) ? mk<double>(r == c ? 1. : 0.) : -S[size_t(r) * N + c];
    } else {
      const int cc = c - N;
      v = (cc & 1) ? S[size_t(r) * N + cc] : mk<double>(r == cc ? -1. : 0.);
    }
    G[r][c] = v;
  }
  __syncthreads();
  for (int k = 0; k < N; ++k) {
    if (tid == 0) {
      double best = -1.;
      int p = k;
      for (int r = k; r < N; ++r) {
        const double m = fabs(G[r][k].x) + fabs(G[r][k].y);
        if (m > best) {
          best = m;
          p = r;
        }
      }
      piv_row = p;
    }
    __syncthreads();
    const int p = piv_row;
    if (p != k)
      for (int c = tid; c < W2; c += 256) {
        const cd t = G[k][c];
        G[k][c] = G[p][c];
        G[p][c] = t;
      }
    __syncthreads();
    const cd pv = G[k][k];
    const double n2 = norm2(pv);
    const cd ip = n2 > 0. ? conj(pv) * (1. / n2) : mk<double>(0.);
    __syncthreads();
    for (int c = tid; c < W2; c += 256) G[k][c] = G[k][c] * ip;
    __syncthreads();
    for (int r = tid; r < N; r += 256) fac[r] = (r == k) ? mk<double>(0.) : G[r][k];
    __syncthreads();
    for (int e = tid; e < N * W2; e += 256) {
      const int r = e / W2, c = e - r * W2;
      if (r == k) continue;
      const cd f = fac[r];
      if (f.x == 0. && f.y == 0.) continue;
      G[r][c] = G[r][c] - f * G[k][c];
    }
    __syncthreads();
  }
  for (int e = tid; e < N * N; e += 256) {
    const int r = e / N, c = e - r * N;
    R[e] = G[r][N + c];
  }
}

// reflect2d (dtmm/tmm2d.py:184-213) for nvec sets of n mode vectors
// [nvec][n][4]: a = (Fin^-1 fin on even, Fout^-1 fout on odd components),
// out = R a; fin <- Fin (a_even, out_odd), fout <- Fout (out_even, 0 odd).
static __global__ __launch_bounds__(NT) void t2d_reflect_kernel(cd* fvin, const cd* fin,
                                                              const cd* fini, const cd* R,
                                                              const cd* fout, const cd* fouti,
                                                              cd* fvout, int has_out, int n,
                                                              long long nvec, cd* tmp) {
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  const int N = 4 * n;
  if (e >= nvec * N) return;
  const long long v = e / N;
  const int r = int(e - v * N), i = r >> 2, a = r & 3;
  // a vector of this set, assembled by every thread of the set (read only)
  auto avec = [&](int rr) {
    const int ii = rr >> 2, aa = rr & 3;
    const cd* fv = (aa & 1) ? (has_out ? fvout : nullptr) : fvin;
    if (!fv) return mk<double>(0.);
    const cd* M = (aa & 1) ? fouti : fini;
    cd t = mk<double>(0.);
    for (int q = 0; q < 4; ++q) t = t + M[(ii * 4 + aa) * 4 + q] * fv[(v * n + ii) * 4 + q];
    return t;
  };
  cd o = mk<double>(0.);
  for (int c = 0; c < N; ++c) o = o + R[size_t(r) * N + c] * avec(c);
  // keep: (a_even of the input, out_odd) for the input side, (out_even, 0) out
  tmp[e] = (a & 1) ? o : avec(r);
  tmp[nvec * N + e] = (a & 1) ? avec(r) : o;
  (void)i;
}

static __global__ __launch_bounds__(NT) void t2d_reflect_out_kernel(cd* fvin, const cd* fin,
                                                                  const cd* fout, cd* fvout,
                                                                  int n, long long nvec,
                                                                  const cd* tmp) {
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  const int N = 4 * n;
  if (e >= nvec * N) return;
  const long long v = e / N;
  const int r = int(e - v * N), i = r >> 2, a = r & 3;
  const cd* ai = tmp + v * N + 4 * i;
  const cd* bo = tmp + nvec * N + v * N + 4 * i;
  cd s = mk<double>(0.), t = mk<double>(0.);
  for (int q = 0; q < 4; ++q) {
    s = s + fin[(i * 4 + a) * 4 + q] * ai[q];
    t = t + fout[(i * 4 + a) * 4 + q] * bo[q];
  }
  fvin[e] = s;
  fvout[e] = t;
}

// exp(2 pi i t / W), t = 0 .. W - 1
static __global__ __launch_bounds__(NT) void t2d_tw_kernel(cd* tw, int W) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= W) return;
  double s, c;
  sincospi(2.0 * t / W, &s, &c);
  tw[t] = mk<double>(c, s);
}

// inverses of n 4x4 field matrices
static __global__ __launch_bounds__(NT) void t2d_inv_kernel(const cd* f, cd* fi, int n) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  cd m[4][4], o[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) m[a][b] = f[(i * 4 + a) * 4 + b];
  inv4(m, o);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) fi[(i * 4 + a) * 4 + b] = o[a][b];
}

// Masked 1D DFT of the field (dtmm/field.py:216-250) and its inverse
// (:307-340): modes[v][i][c] = sum_x f[v][c][x] exp(-2 pi i idx_i x / W),
// f[v][c][x] = 1/W sum_i modes[v][i][c] exp(2 pi i idx_i x / W).
static __global__ __launch_bounds__(NT) void t2d_field2modes_kernel(const cd* f, cd* modes,
                                                                  const int* idx, const cd* tw,
                                                                  int W, int n, long long nvec) {
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= nvec * n * 4) return;
  const long long v = e / (4 * n);
  const int r = int(e - v * 4 * n), i = r >> 2, c = r & 3;
  const cd* fl = f + (v * 4 + c) * W;
  cd t = mk<double>(0.);
  int ph = 0;
  const int k = idx[i] % W;
  for (int x = 0; x < W; ++x) {
    t = t + fl[x] * conj(tw[ph]);
    ph += k;
    if (ph >= W) ph -= W;
  }
  modes[e] = t;
}

static __global__ __launch_bounds__(NT) void t2d_modes2field_kernel(const cd* modes, cd* f,
                                                                  const int* idx, const cd* tw,
                                                                  int W, int n, long long nvec) {
  const long long e = (long long)blockIdx.x * NT + threadIdx.x;
  if (e >= nvec * 4 * W) return;
  const long long v = e / (4 * W);
  const int r = int(e - v * 4 * W), c = r / W, x = r - c * W;
  cd t = mk<double>(0.);
  for (int i = 0; i < n; ++i) {
    const int ph = int((long long)(idx[i] % W) * x % W);
    t = t + modes[(v * n + i) * 4 + c] * tw[ph];
  }
  f[e] = t * (1. / W);
}

// ---------------------------------------------------------------------------
// Pass normalisation kernels (npass > 1, dtmm/transfer.py:815-872): one
// thread per (k, plane) on the unnormalised 2D spectrum of a whole field.
// ---------------------------------------------------------------------------

enum { POP_PROJ = 0, POP_NORMFFT = 1, POP_TOTALSUM = 2, POP_TOTALAPPLY = 3, POP_NOTCH = 4 };

template <typename T>
struct PArgs {
  const cpx<T>* src;
  cpx<T>* dst;
  int H, W;
  int nlam;
  const int* amax;
  int pop;
  cd eps;             // isotropic medium of the forward projection
  int notch;          // 1: multiply by the notch filter of dtmm/denoise.py
  double x0, sigma;   // notch centre (beta) and width
  double scale;       // applied to the written values
  double* part;       // per-block partial sums [2][nplanes][gridDim.x]
  const double* sums; // per-plane sums [2][nplanes] (POP_TOTALAPPLY)
  const cpx<T>* ref;  // POP_NORMFFT: spectrum of the reference field whose
                      // intensities the modes are rescaled to (NULL: the
                      // field's own, dtmm/transfer.py:78-91)
  double ref_scale;   // amplitude factor of the reference
  Beams bm;
};

// Forward-mode projection F diag(1, 0, 1, 0) F^-1 of an isotropic medium at
// in-plane wave vector beta (direction (c, sn)) in closed form
// (dtmm/diffract.py:242-251): rotate into the mode frame, keep the forward
// p and s amplitudes, rebuild.  Zero where a = sqrt(eps - beta^2) == 0.
__host__ __device__ inline void iso_forward_projection(double beta, double c, double sn, cd eps,
                                                       const cd v[4], cd r[4]) {
  const cd a = csqrt(eps - mk<double>(beta * beta));
  if (is_zero(a)) {
    for (int i = 0; i < 4; ++i) r[i] = mk<double>(0.);
    return;
  }
  const cd g0 = v[0] * c + v[2] * sn, g1 = v[1] * c - v[3] * sn;
  const cd g2 = v[2] * c - v[0] * sn, g3 = v[1] * sn + v[3] * c;
  const cd mp = (g0 + g1 * cdiv(a, eps)) * 0.5;  // forward p amplitude / 2
  const cd ms = (g2 - cdiv(g3, a)) * 0.5;        // forward s amplitude / 2
  const cd h0 = mp, h1 = cdiv(eps, a) * mp, h2 = ms, h3 = -(a * ms);
  r[0] = h0 * c - h2 * sn;
  r[1] = h1 * c + h3 * sn;
  r[2] = h0 * sn + h2 * c;
  r[3] = h3 * c - h1 * sn;
}

// Notch filter around beta = x0 (dtmm/denoise.py:21-22).
__device__ inline double notch_filter(double b, double x0, double sigma) {
  double d = fabs(b - x0);
  d = d < 0. ? 0. : (d > x0 ? x0 : d);
  return (1. - exp(-d / sigma)) / (1. - exp(-x0 / sigma));
}

template <typename T>
__global__ __launch_bounds__(NT) void pass_kernel(PArgs<T> A) {
  __shared__ double red[2][NT];
  const int tid = threadIdx.x;
  const int k = blockIdx.x * NT + tid;
  const int q = blockIdx.y;
  const int nplanes = gridDim.y;
  const int lam = q % A.nlam;
  const size_t HW = size_t(A.H) * A.W;
  double s0 = 0., s1 = 0.;
  if (k < int(HW)) {
    const int ky = k / A.W, kx = k - ky * A.W;
    const double k0 = A.bm.k0[lam];
    const bool pass = mode_index(kx, A.W) <= A.amax[lam * A.H + ky];
    const double fx = fftfreq(kx, A.W), fy = fftfreq(ky, A.H);
    const double fr = sqrt(fx * fx + fy * fy);
    const double beta = (2 * M_PI / fabs(k0)) * fr;
    const double c = fr > 0. ? fx / fr : 1., sn = fr > 0. ? fy / fr : 0.;
    const cpx<T>* src = A.src + size_t(q) * 4 * HW + k;
    cd v[4], p[4];
    for (int i = 0; i < 4; ++i) v[i] = cast<double>(src[i * HW]);
    if (A.pop == POP_NOTCH) {  // the notch filter alone, at every k
      for (int i = 0; i < 4; ++i) p[i] = v[i];
    } else if (pass) {
      iso_forward_projection(beta, c, sn, A.eps, v, p);
    } else {
      for (int i = 0; i < 4; ++i) p[i] = mk<double>(0.);
    }
    double f = 1.;
    if (A.notch) f = notch_filter(beta, A.x0, A.sigma);
    if (A.pop == POP_TOTALSUM) {
      s0 = intensity4(v);
      s1 = intensity4(p);
    } else {
      double g = 1.;
      if (A.pop == POP_NORMFFT) {
        double i1 = intensity4(v), i2 = intensity4(p);
        if (A.ref) {
          cd w[4];
          const cpx<T>* rs = A.ref + size_t(q) * 4 * HW + k;
          for (int i = 0; i < 4; ++i) w[i] = cast<double>(rs[i * HW]);
          i1 = intensity4(w) * (A.ref_scale * A.ref_scale);
        }
        if (i2 == 0.) {
          i2 = 1.;
          i1 = 0.;
        }
        g = fabs(fmin(fmax(i1 / i2, -1.), 1.));
      } else if (A.pop == POP_TOTALAPPLY) {
        double i1 = A.sums[q], i2 = A.sums[nplanes + q];
        if (i2 == 0.) {
          i2 = 1.;
          i1 = 0.;
        }
        g = fabs(fmin(fmax(i1 / i2, -1.), 1.));
      }
      for (int i = 0; i < 4; ++i) p[i] = p[i] * (g * f);
      s0 = intensity4(p);
      cpx<T>* dst = A.dst + size_t(q) * 4 * HW + k;
      for (int i = 0; i < 4; ++i) dst[i * HW] = cast<T>(p[i] * A.scale);
    }
  }
  if (A.part) {
    red[0][tid] = s0;
    red[1][tid] = s1;
    __syncthreads();
    for (int m = NT / 2; m > 0; m >>= 1) {
      if (tid < m) {
        red[0][tid] += red[0][tid + m];
        red[1][tid] += red[1][tid + m];
      }
      __syncthreads();
    }
    if (tid == 0) {
      const size_t nb = gridDim.x;
      A.part[size_t(q) * nb + blockIdx.x] = red[0][0];
      A.part[(size_t(nplanes) + q) * nb + blockIdx.x] = red[1][0];
    }
  }
}

// sums[j] = scale * sum_b part[j][b] for 2 * nplanes rows (one block each).
static __global__ __launch_bounds__(NT) void sum_rows_kernel(const double* part, int nb, double scale,
                                                       double* sums) {
  __shared__ double red[NT];
  const int r = blockIdx.x, tid = threadIdx.x;
  double a = 0.;
  for (int b = tid; b < nb; b += NT) a += part[size_t(r) * nb + b];
  red[tid] = a;
  __syncthreads();
  for (int m = NT / 2; m > 0; m >>= 1) {
    if (tid < m) red[tid] += red[tid + m];
    __syncthreads();
  }
  if (tid == 0) sums[r] = red[0] * scale;
}

// Local normalisation with a reference field (dtmm/transfer.py:103-138),
// one thread per (pixel, plane) on real-space fields: f is the forward part
// of the field in the output medium; its normal-incidence forward amplitudes
// (unit-flux modes of that medium, dtmm/tmm.py:701-725, :1851-1873) give the
// pixel's Jones vector j, P1 = F as4x4(j j^H) F^-1 (dtmm/jones4.py:75-103,
// jones.py:371-398), and
//   f <- f |clip(I(P1 ref) / I(f), -1, 1)| + (1 - P1) ref
// with the factor 0 where I(f) = 0 (transfer.py:66-75).  A pixel whose
// forward amplitudes are both zero gets NaN, as the reference's
// normalisation of a zero Jones vector does.
template <typename T>
struct LocalArgs {
  cpx<T>* f;
  const cpx<T>* ref;
  double ref_scale;  // amplitude factor of the reference
  size_t HW;
  cd eps;  // output medium
};

template <typename T>
__global__ __launch_bounds__(NT) void local_norm_kernel(LocalArgs<T> A) {
  const size_t px = size_t(blockIdx.x) * NT + threadIdx.x;
  if (px >= A.HW) return;
  const size_t o = size_t(blockIdx.y) * 4 * A.HW + px;
  Modes m;
  iso_modes(0., A.eps, m);
  // unit-flux columns
  for (int j = 0; j < 4; ++j) {
    const double pz = (m.F[0][j].x * m.F[1][j].x + m.F[0][j].y * m.F[1][j].y) -
                      (m.F[2][j].x * m.F[3][j].x + m.F[2][j].y * m.F[3][j].y);
    const double n = sqrt(fabs(pz));
    const double sc = n == 0. ? 0. : 1. / n;
    for (int i = 0; i < 4; ++i) m.F[i][j] = m.F[i][j] * sc;
  }
  cd Fi[4][4];
  inv4(m.F, Fi);
  cd fv[4], rv[4];
  for (int i = 0; i < 4; ++i) {
    fv[i] = cast<double>(A.f[o + i * A.HW]);
    rv[i] = cast<double>(A.ref[o + i * A.HW]) * A.ref_scale;
  }
  cd a[4];
  for (int i = 0; i < 4; ++i) {
    cd t = mk<double>(0.);
    for (int k = 0; k < 4; ++k) t = t + Fi[i][k] * fv[k];
    a[i] = t;
  }
  const double nrm = sqrt(a[0].x * a[0].x + a[0].y * a[0].y + a[2].x * a[2].x + a[2].y * a[2].y);
  const cd j0 = a[0] * (1. / nrm), j1 = a[2] * (1. / nrm);
  // J = j j^H on the forward modes (0, 2), conj(J) on the backward (1, 3)
  const cd J[2][2] = {{j0 * conj(j0), j0 * conj(j1)}, {j1 * conj(j0), j1 * conj(j1)}};
  cd b[4];  // as4x4(J) Fi rv
  for (int i = 0; i < 4; ++i) {
    cd t = mk<double>(0.);
    for (int k = 0; k < 4; ++k) t = t + Fi[i][k] * rv[k];
    b[i] = t;
  }
  const cd c0 = J[0][0] * b[0] + J[0][1] * b[2], c2 = J[1][0] * b[0] + J[1][1] * b[2];
  const cd c1 = conj(J[0][0]) * b[1] + conj(J[0][1]) * b[3];
  const cd c3 = conj(J[1][0]) * b[1] + conj(J[1][1]) * b[3];
  cd p1r[4];
  for (int i = 0; i < 4; ++i) p1r[i] = m.F[i][0] * c0 + m.F[i][1] * c1 + m.F[i][2] * c2 + m.F[i][3] * c3;
  double i1 = intensity4(p1r), i2 = intensity4(fv);
  if (i2 == 0.) {
    i2 = 1.;
    i1 = 0.;
  }
  const double g = fabs(fmin(fmax(i1 / i2, -1.), 1.));
  for (int i = 0; i < 4; ++i) A.f[o + i * A.HW] = cast<T>(fv[i] * g + (rv[i] - p1r[i]));
}

// Odd-pass update (dtmm/transfer.py:855-866) with fact = i0 / i0f per plane:
// out *= fact; fin = fact * fin + (f0 - fact * g); g = f0 - fact * g.
template <typename T>
__global__ __launch_bounds__(NT) void reflect_update_kernel(cpx<T>* g, cpx<T>* fout, cpx<T>* fin,
                                                             const cpx<T>* f0, const double* i0,
                                                             const double* i0f, size_t per_plane,
                                                             size_t n) {
  const size_t i = size_t(blockIdx.x) * NT + threadIdx.x;
  if (i >= n) return;
  const size_t q = i / per_plane;
  const double fact = i0[q] / i0f[q];
  const cd gi = cast<double>(g[i]) * fact;
  fout[i] = cast<T>(cast<double>(fout[i]) * fact);
  const cd d = cast<double>(f0[i]) - gi;
  fin[i] = cast<T>(cast<double>(fin[i]) * fact + d);
  g[i] = cast<T>(d);
}

// Even-pass accumulation (dtmm/transfer.py:838-840): out += g; g = out
// (first: out = g).
// One tent window's part of a 2D spectrum (reflection mode 2 with
// diffraction > 1, dtmm/propagate_4x4.py:176-204): dst = src * cx_i(kx) *
// cy_j(ky) * scale where the mode passes the window mask, else 0; one thread
// per (k, plane), planes [ray][wavelength][component].
template <typename T>
__global__ __launch_bounds__(NT) void win_part_kernel(const cpx<T>* src, cpx<T>* dst, int H,
                                                      int W, int nlam, int rays_per_beam,
                                                      int wn, int wi, int wj, const double* wcx,
                                                      const double* wcy, const int* wmask,
                                                      double scale) {
  const size_t HW = size_t(H) * W;
  const size_t k = size_t(blockIdx.x) * NT + threadIdx.x;
  if (k >= HW) return;
  const int p = blockIdx.y;
  const int ray = p / (nlam * 4), lam = (p / 4) % nlam;
  const int beam = ray / rays_per_beam;
  const int ky = int(k / W), kx = int(k - size_t(ky) * W);
  const size_t bl = size_t(beam) * nlam + lam;
  const size_t o = size_t(p) * HW + k;
  cpx<T> v = mk<T>(T(0), T(0));
  if (mode_index(kx, W) <= wmask[lam * H + ky]) {
    const double f = wcx[(bl * wn + wi) * W + kx] * wcy[(bl * wn + wj) * H + ky] * scale;
    v = src[o] * T(f);
  }
  dst[o] = v;
}

template <typename T>
__global__ __launch_bounds__(NT) void accumulate_kernel(cpx<T>* g, cpx<T>* fout, int first,
                                                         size_t n) {
  const size_t i = size_t(blockIdx.x) * NT + threadIdx.x;
  if (i >= n) return;
  const cpx<T> v = first ? g[i] : fout[i] + g[i];
  fout[i] = v;
  g[i] = v;
}

// Sets status bit 2 when any of the n values is NaN or infinite (the
// output check of a transfer; one atomic per wave that finds one).
template <typename T>
__global__ __launch_bounds__(NT) void finite_kernel(const T* p, size_t n, int* status) {
  bool bad = false;
  for (size_t i = size_t(blockIdx.x) * NT + threadIdx.x; i < n; i += size_t(gridDim.x) * NT)
    bad = bad || !isfinite(p[i]);
  if (__any(bad) && (threadIdx.x % warpSize) == 0) atomicOr(status, 4);
}

// Zero n values of type T.
template <typename T>
__global__ __launch_bounds__(NT) void zero_kernel(T* p, size_t n) {
  const size_t i = size_t(blockIdx.x) * NT + threadIdx.x;
  if (i < n) p[i] = T(0);
}

// (Ex, Ey) -> (Ex, Hy, Ey, Hx) with the mean-beam H/E relation of the output
// medium (dtmm/propagate_2x2.py:81-85, dtmm/tmm.py:1051-1072).
template <typename T>
struct E2HArgs {
  const cpx<T>* src;
  cpx<T>* dst;
  int H, W;
  Batch b;
  cd eout;
  Beams bm;
};

template <typename T>
__global__ __launch_bounds__(NT) void e2h_kernel(E2HArgs<T> A) {
  const int k = blockIdx.x * NT + threadIdx.x;
  const int ray = blockIdx.y / A.b.nlam_b;
  const int q = ray * A.b.nlam + A.b.lam0 + blockIdx.y % A.b.nlam_b;
  const size_t HW = size_t(A.H) * A.W;
  if (k >= int(HW)) return;
  const int beam = ray / A.bm.rays_per_beam;
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
// Modes without diffraction (diffraction = 0): the field stays in real space
// and every pixel is an independent 1D stack.  One thread per (pixel, beam)
// solves the layer in fp64 once and applies it to every ray of the beam and
// every wavelength, reading and writing the field in place.
//   PIX_LAYER4: 4x4 reflection mode 2, F diag(exp(i k0 d alpha)) F^-1 on
//               (Ex, Hy, Ey, Hx) (dtmm/propagate_4x4.py:18-51, dmatpn None);
//   PIX_LAYER2: 2x2 reflection mode 0, E diag(p) E^-1 of the forward modes
//               on (Ex, Ey) (dtmm/propagate_2x2.py:91-190, dmat None);
//   PIX_IN2:    forward part of the 4-component input in the isotropic input
//               medium, its E components to the 2-component field
//               (dtmm/transfer.py:94-101);
//   PIX_OUT2:   H of the forward modes of the isotropic output medium from E
//               (dtmm/tmm.py:1051-1072), the 4-component output;
//   PIX_A4:     4x4 reflection mode 2 with diffraction, first half of a
//               sub-step: e diag(exp(i k0 d alpha)) F^-1 (d = half the
//               sub-step) maps (Ex, Hy, Ey, Hx) to the E parts of the
//               forward and backward modes (E+x, E-x, E+y, E-y)
//               (dtmm/propagate_4x4.py:45);
//   PIX_B4:     the second half, F diag(exp(i k0 d alpha)) e^-1 back to
//               (Ex, Hy, Ey, Hx) (dtmm/propagate_4x4.py:48).
// A layer's sub-steps are one application with the full thickness:
// (F P F^-1)^n = F P^n F^-1.
// ---------------------------------------------------------------------------

//   PIX_E2H:    as PIX_OUT2 with the forward modes of the layer's own voxel
//               (the 2x2 field of every layer, dtmm/propagate_2x2.py:80-85).
//   PIX_T2:     2x2 reflection mode 2, first half step of a layer entered
//               through the interface from the previous layer's voxel (or
//               the isotropic input medium): e diag(p) Eti with
//               Eti = St^-1 A^-1 (dtmm/tmm.py:999-1022,
//               dtmm/propagate_2x2.py:172-177); the current layer may be the
//               isotropic output medium (emed, cur_iso).
//   PIX_TRR:    2x2 reflection mode 2 with stored reflections (npass > 1,
//               dtmm/propagate_2x2.py:120-151): at the interface from the
//               previous layer's voxel (or isotropic medium) into this one,
//               t = e_out St^-1 A^-1 of tr_mat (dtmm/tmm.py:925-950) for the
//               modes of `mode`; refl <- field - t field, field <- t field +
//               (old refl); with dst (mode -1) that field and its H are added
//               to dst; then the half step e diag(p) e^-1 of this layer.
// PIX_LAYER2 and PIX_TRR use the backward modes for mode -1; PIX_E2H and
// PIX_OUT2 then build H of the backward modes, and with accum add the
// 4-component result to dst instead of writing it.  PIX_IN2 with dst also
// writes the whole 4-component forward part there (dst may be src).
enum { PIX_LAYER4 = 0, PIX_LAYER2 = 1, PIX_IN2 = 2, PIX_OUT2 = 3, PIX_A4 = 4, PIX_B4 = 5,
       PIX_E2H = 6, PIX_T2 = 7, PIX_TRR = 8 };

template <typename T>
struct PixArgs {
  cpx<T>* field;        // LAYER: [ray][lam][NC][H*W] in place; IN2: output E field
  const cpx<T>* src;    // IN2: 4-component input; OUT2: E field
  cpx<T>* dst;          // OUT2: 4-component output
  size_t HW;
  int nlam;
  // layer material: [3][L*H*W] with this layer's offset applied
  const double* eps_re;
  const double* eps_im;
  const double* ang;
  size_t comp_stride;
  double d;             // full layer thickness
  cd emed;              // IN2 / OUT2: isotropic medium eps
  int lam0, nlam_b;     // OUT2 / E2H: wavelengths lam0 .. lam0 + nlam_b (nlam_b 0: all)
  // T2: the previous layer's material (same layout) or, with prev_iso, the
  // isotropic medium pemed; cur_iso: the current layer is emed
  const double* peps_re;
  const double* peps_im;
  const double* pang;
  int prev_iso, cur_iso;
  cd pemed;
  int mode;             // +1 (or 0): forward modes; -1: backward modes (npass > 1)
  int accum;            // OUT2 / E2H: add to dst
  cpx<T>* refl;         // TRR: stored reflection of this interface, field layout
  Beams bm;
};

__device__ inline void pixel_modes(const double* er, const double* ei, const double* ang,
                                   size_t cs, size_t o, double beta, double c0, double s0,
                                   double phi0, Modes& m) {
  const cd e0 = mk<double>(er[o], ei ? ei[o] : 0.);
  const cd e1 = mk<double>(er[cs + o], ei ? ei[cs + o] : 0.);
  const cd e2 = mk<double>(er[2 * cs + o], ei ? ei[2 * cs + o] : 0.);
  int kind;
  layer_modes(beta, c0, s0, phi0, e0, e1, e2, ang[o], ang[cs + o], ang[2 * cs + o], m, kind);
}

template <typename T, int PK>
__global__ __launch_bounds__(NT) void pixel_kernel(PixArgs<T> A) {
  const size_t px = size_t(blockIdx.x) * NT + threadIdx.x;
  const int beam = blockIdx.y;
  if (px >= A.HW) return;
  const double beta = A.bm.beta[beam], phi0 = A.bm.phi[beam];
  double s0, c0;
  sincos(phi0, &s0, &c0);
  const int r0 = beam * A.bm.rays_per_beam, r1 = r0 + A.bm.rays_per_beam;
  auto plane = [&](cpx<T>* base, int ray, int lam, int nc) {
    return base + (size_t(ray) * A.nlam + lam) * nc * A.HW + px;
  };
  if constexpr (PK == PIX_A4 || PK == PIX_B4) {
    const size_t o = px;
    const cd e0 = mk<double>(A.eps_re[o], A.eps_im ? A.eps_im[o] : 0.);
    const cd e1 = mk<double>(A.eps_re[A.comp_stride + o], A.eps_im ? A.eps_im[A.comp_stride + o] : 0.);
    const cd e2 = mk<double>(A.eps_re[2 * A.comp_stride + o],
                             A.eps_im ? A.eps_im[2 * A.comp_stride + o] : 0.);
    Modes m;
    int kind;
    layer_modes(beta, c0, s0, phi0, e0, e1, e2, A.ang[o], A.ang[A.comp_stride + o],
                A.ang[2 * A.comp_stride + o], m, kind);
    cpx<T> F[4][4], a[4];
    for (int i = 0; i < 4; ++i) {
      a[i] = cast<T>(m.a[i]);
      for (int j = 0; j < 4; ++j) F[i][j] = cast<T>(m.F[i][j]);
    }
    // PIX_A4: F^-1.  PIX_B4: inverses of the forward and backward E blocks
    // (the 4x4 inverse of e, all zero when singular as dtmm/linalg.py:510-513)
    cpx<T> G[4][4];
    if constexpr (PK == PIX_A4) {
      cd fi[4][4];
      inv4(m.F, fi);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) G[i][j] = cast<T>(fi[i][j]);
    } else {
      const cd ef[2][2] = {{m.F[0][0], m.F[0][2]}, {m.F[2][0], m.F[2][2]}};
      const cd eb[2][2] = {{m.F[0][1], m.F[0][3]}, {m.F[2][1], m.F[2][3]}};
      cd fi[2][2], bi[2][2];
      const bool ok = inv2(ef, fi) & inv2(eb, bi);
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
          G[i][j] = ok ? cast<T>(fi[i][j]) : mk<T>(T(0), T(0));
          G[2 + i][j] = ok ? cast<T>(bi[i][j]) : mk<T>(T(0), T(0));
        }
    }
    for (int lam = 0; lam < A.nlam; ++lam) {
      const T kd = T(A.bm.k0[lam] * A.d);
      cpx<T> p[4];
      for (int i = 0; i < 4; ++i) p[i] = cexpi_layer(a[i] * kd);
      for (int ray = r0; ray < r1; ++ray) {
        cpx<T>* f = plane(A.field, ray, lam, 4);
        cpx<T> v[4], c[4];
        for (int q = 0; q < 4; ++q) v[q] = f[q * A.HW];
        if constexpr (PK == PIX_A4) {
          for (int i = 0; i < 4; ++i) {
            cpx<T> t = G[i][0] * v[0];
            for (int q = 1; q < 4; ++q) t = t + G[i][q] * v[q];
            c[i] = t * p[i];
          }
          f[0] = F[0][0] * c[0] + F[0][2] * c[2];       // E+x
          f[A.HW] = F[0][1] * c[1] + F[0][3] * c[3];    // E-x
          f[2 * A.HW] = F[2][0] * c[0] + F[2][2] * c[2];  // E+y
          f[3 * A.HW] = F[2][1] * c[1] + F[2][3] * c[3];  // E-y
        } else {
          c[0] = (G[0][0] * v[0] + G[0][1] * v[2]) * p[0];
          c[2] = (G[1][0] * v[0] + G[1][1] * v[2]) * p[2];
          c[1] = (G[2][0] * v[1] + G[2][1] * v[3]) * p[1];
          c[3] = (G[3][0] * v[1] + G[3][1] * v[3]) * p[3];
          for (int i = 0; i < 4; ++i) {
            cpx<T> t = F[i][0] * c[0];
            for (int q = 1; q < 4; ++q) t = t + F[i][q] * c[q];
            f[i * A.HW] = t;
          }
        }
      }
    }
  } else if constexpr (PK == PIX_LAYER4 || PK == PIX_LAYER2) {
    const size_t o = px;
    const cd e0 = mk<double>(A.eps_re[o], A.eps_im ? A.eps_im[o] : 0.);
    const cd e1 = mk<double>(A.eps_re[A.comp_stride + o], A.eps_im ? A.eps_im[A.comp_stride + o] : 0.);
    const cd e2 = mk<double>(A.eps_re[2 * A.comp_stride + o],
                             A.eps_im ? A.eps_im[2 * A.comp_stride + o] : 0.);
    const double psi = A.ang[o], theta = A.ang[A.comp_stride + o], phi = A.ang[2 * A.comp_stride + o];
    if constexpr (PK == PIX_LAYER4) {
      Modes m;
      int kind;
      layer_modes(beta, c0, s0, phi0, e0, e1, e2, psi, theta, phi, m, kind);
      cd fi[4][4];
      inv4(m.F, fi);  // singular: zeros, as dtmm/linalg.py:510-513
      cpx<T> F[4][4], Fi[4][4], a[4];
      for (int i = 0; i < 4; ++i) {
        a[i] = cast<T>(m.a[i]);
        for (int j = 0; j < 4; ++j) {
          F[i][j] = cast<T>(m.F[i][j]);
          Fi[i][j] = cast<T>(fi[i][j]);
        }
      }
      for (int lam = 0; lam < A.nlam; ++lam) {
        const T kd = T(A.bm.k0[lam] * A.d);
        cpx<T> p[4];
        for (int i = 0; i < 4; ++i) p[i] = cexpi_layer(a[i] * kd);
        for (int ray = r0; ray < r1; ++ray) {
          cpx<T>* f = plane(A.field, ray, lam, 4);
          cpx<T> v[4], u[4];
          for (int c = 0; c < 4; ++c) v[c] = f[c * A.HW];
          for (int i = 0; i < 4; ++i) {
            cpx<T> t = Fi[i][0] * v[0];
            for (int c = 1; c < 4; ++c) t = t + Fi[i][c] * v[c];
            u[i] = t * p[i];
          }
          for (int i = 0; i < 4; ++i) {
            cpx<T> t = F[i][0] * u[0];
            for (int c = 1; c < 4; ++c) t = t + F[i][c] * u[c];
            f[i * A.HW] = t;
          }
        }
      }
    } else {
      RopData r;
      rop_full(beta, c0, s0, phi0, e0, e1, e2, psi, theta, phi, r);
      const bool bw = A.mode < 0;
      const T ok = (bw ? r.okb : r.ok2) ? T(1) : T(0);
      const cpx<T> a0 = cast<T>(r.a[bw ? 1 : 0]), a2 = cast<T>(r.a[bw ? 3 : 2]);
      cpx<T> P[2][2];
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) P[i][j] = cast<T>(bw ? r.P1[i][j] : r.P0[i][j]);
      for (int lam = 0; lam < A.nlam; ++lam) {
        const T kd = T(A.bm.k0[lam] * A.d);
        const cpx<T> p0 = cexpi_layer(a0 * kd);
        cpx<T> p2 = cexpi_layer(a2 * kd);
        const cpx<T> dp = (p0 - p2) * ok;
        p2 = p2 * ok;
        const cpx<T> m00 = p2 + dp * P[0][0], m01 = dp * P[0][1], m10 = dp * P[1][0],
                     m11 = p2 + dp * P[1][1];
        for (int ray = r0; ray < r1; ++ray) {
          cpx<T>* f = plane(A.field, ray, lam, 2);
          const cpx<T> ex = f[0], ey = f[A.HW];
          f[0] = m00 * ex + m01 * ey;
          f[A.HW] = m10 * ex + m11 * ey;
        }
      }
    }
  } else if constexpr (PK == PIX_T2) {
    Modes mi, mo;
    int kind;
    if (A.prev_iso)
      layer_modes(beta, c0, s0, phi0, A.pemed, A.pemed, A.pemed, 0., 0., 0., mi, kind);
    else
      pixel_modes(A.peps_re, A.peps_im, A.pang, A.comp_stride, px, beta, c0, s0, phi0, mi);
    if (A.cur_iso)
      layer_modes(beta, c0, s0, phi0, A.emed, A.emed, A.emed, 0., 0., 0., mo, kind);
    else
      pixel_modes(A.eps_re, A.eps_im, A.ang, A.comp_stride, px, beta, c0, s0, phi0, mo);
    // Eti = St^-1 A^-1: A the E block of the incoming forward modes, St the
    // forward-forward block of F_in^-1 F_out
    cd fi[4][4];
    inv4(mi.F, fi);
    cd St[2][2], Am[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        cd t = mk<double>(0.);
        for (int k = 0; k < 4; ++k) t = t + fi[2 * i][k] * mo.F[k][2 * j];
        St[i][j] = t;
        Am[i][j] = mi.F[2 * i][2 * j];
      }
    cd Sti[2][2], Ai[2][2], Eti[2][2];
    inv2(St, Sti);
    inv2(Am, Ai);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) Eti[i][j] = Sti[i][0] * Ai[0][j] + Sti[i][1] * Ai[1][j];
    const cd e[2][2] = {{mo.F[0][0], mo.F[0][2]}, {mo.F[2][0], mo.F[2][2]}};
    for (int lam = 0; lam < A.nlam; ++lam) {
      const double kd = A.bm.k0[lam] * A.d;
      const cd p0 = cexpi(mo.a[0] * kd), p2 = cexpi(mo.a[2] * kd);
      cpx<T> M[2][2];
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          M[i][j] = cast<T>(e[i][0] * p0 * Eti[0][j] + e[i][1] * p2 * Eti[1][j]);
      for (int ray = r0; ray < r1; ++ray) {
        cpx<T>* f = plane(A.field, ray, lam, 2);
        const cpx<T> ex = f[0], ey = f[A.HW];
        f[0] = M[0][0] * ex + M[0][1] * ey;
        f[A.HW] = M[1][0] * ex + M[1][1] * ey;
      }
    }
  } else if constexpr (PK == PIX_TRR) {
    Modes mi, mo;
    int kind;
    if (A.prev_iso)
      layer_modes(beta, c0, s0, phi0, A.pemed, A.pemed, A.pemed, 0., 0., 0., mi, kind);
    else
      pixel_modes(A.peps_re, A.peps_im, A.pang, A.comp_stride, px, beta, c0, s0, phi0, mi);
    if (A.cur_iso)
      layer_modes(beta, c0, s0, phi0, A.emed, A.emed, A.emed, 0., 0., 0., mo, kind);
    else
      pixel_modes(A.eps_re, A.eps_im, A.ang, A.comp_stride, px, beta, c0, s0, phi0, mo);
    // columns of the modes of this direction (dtmm/tmm.py:850-878 S_mat,
    // :974-997 E_mat)
    const int q0 = A.mode < 0 ? 1 : 0, q1 = A.mode < 0 ? 3 : 2;
    const int qs[2] = {q0, q1};
    cd fi[4][4];
    inv4(mi.F, fi);
    cd St[2][2], Am[2][2], e[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        cd t = mk<double>(0.);
        for (int k = 0; k < 4; ++k) t = t + fi[qs[i]][k] * mo.F[k][qs[j]];
        St[i][j] = t;
        Am[i][j] = mi.F[2 * i][qs[j]];
        e[i][j] = mo.F[2 * i][qs[j]];
      }
    cd Sti[2][2], Ai[2][2], ei[2][2];
    inv2(St, Sti);
    inv2(Am, Ai);
    inv2(e, ei);
    cd Eti[2][2], Tm[2][2], G[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) Eti[i][j] = Sti[i][0] * Ai[0][j] + Sti[i][1] * Ai[1][j];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        Tm[i][j] = e[i][0] * Eti[0][j] + e[i][1] * Eti[1][j];
        // E2H = B e^-1, B the H rows of the same columns (dtmm/tmm.py:1051-1072)
        G[i][j] = mo.F[2 * i + 1][q0] * ei[0][j] + mo.F[2 * i + 1][q1] * ei[1][j];
      }
    cpx<T> Tt[2][2], Gt[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        Tt[i][j] = cast<T>(Tm[i][j]);
        Gt[i][j] = cast<T>(G[i][j]);
      }
    for (int lam = 0; lam < A.nlam; ++lam) {
      const double kd = A.bm.k0[lam] * A.d;
      const cd p0 = A.d != 0. ? cexpi(mo.a[q0] * kd) : mk<double>(1.);
      const cd p1 = A.d != 0. ? cexpi(mo.a[q1] * kd) : mk<double>(1.);
      cpx<T> L[2][2];
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) L[i][j] = cast<T>(e[i][0] * p0 * ei[0][j] + e[i][1] * p1 * ei[1][j]);
      for (int ray = r0; ray < r1; ++ray) {
        cpx<T>* f = plane(A.field, ray, lam, 2);
        cpx<T>* rf = plane(A.refl, ray, lam, 2);
        const cpx<T> ex = f[0], ey = f[A.HW];
        const cpx<T> tx = Tt[0][0] * ex + Tt[0][1] * ey, ty = Tt[1][0] * ex + Tt[1][1] * ey;
        const cpx<T> gx = tx + rf[0], gy = ty + rf[A.HW];
        rf[0] = ex - tx;
        rf[A.HW] = ey - ty;
        if (A.dst) {
          cpx<T>* o = plane(A.dst, ray, lam, 4);
          o[0] = o[0] + gx;
          o[A.HW] = o[A.HW] + (Gt[0][0] * gx + Gt[0][1] * gy);
          o[2 * A.HW] = o[2 * A.HW] + gy;
          o[3 * A.HW] = o[3 * A.HW] + (Gt[1][0] * gx + Gt[1][1] * gy);
        }
        f[0] = L[0][0] * gx + L[0][1] * gy;
        f[A.HW] = L[1][0] * gx + L[1][1] * gy;
      }
    }
  } else {
    // isotropic medium of the beam (PIX_E2H: the layer's voxel): F and its
    // inverse (fp64)
    Modes m;
    int kind;
    if constexpr (PK == PIX_E2H) {
      const size_t o = px;
      const cd e0 = mk<double>(A.eps_re[o], A.eps_im ? A.eps_im[o] : 0.);
      const cd e1 = mk<double>(A.eps_re[A.comp_stride + o],
                               A.eps_im ? A.eps_im[A.comp_stride + o] : 0.);
      const cd e2 = mk<double>(A.eps_re[2 * A.comp_stride + o],
                               A.eps_im ? A.eps_im[2 * A.comp_stride + o] : 0.);
      layer_modes(beta, c0, s0, phi0, e0, e1, e2, A.ang[o], A.ang[A.comp_stride + o],
                  A.ang[2 * A.comp_stride + o], m, kind);
    } else {
      layer_modes(beta, c0, s0, phi0, A.emed, A.emed, A.emed, 0., 0., 0., m, kind);
    }
    if constexpr (PK == PIX_IN2) {
      cd fi[4][4];
      inv4(m.F, fi);
      // F diag(1, 0, 1, 0) F^-1 (rows Ex, Ey without dst)
      cpx<T> Pm[4][4];
      for (int i = 0; i < 4; ++i)
        for (int c = 0; c < 4; ++c)
          Pm[i][c] = cast<T>(m.F[i][0] * fi[0][c] + m.F[i][2] * fi[2][c]);
      for (int lam = 0; lam < A.nlam; ++lam)
        for (int ray = r0; ray < r1; ++ray) {
          const cpx<T>* s = plane(const_cast<cpx<T>*>(A.src), ray, lam, 4);
          cpx<T> v[4], w[4];
          for (int c = 0; c < 4; ++c) v[c] = s[c * A.HW];
          for (int i = 0; i < 4; ++i) {
            cpx<T> t = Pm[i][0] * v[0];
            for (int c = 1; c < 4; ++c) t = t + Pm[i][c] * v[c];
            w[i] = t;
          }
          cpx<T>* f = plane(A.field, ray, lam, 2);
          f[0] = w[0];
          f[A.HW] = w[2];
          if (A.dst) {
            cpx<T>* o = plane(A.dst, ray, lam, 4);
            for (int i = 0; i < 4; ++i) o[i * A.HW] = w[i];
          }
        }
    } else {
      // E2H = B A^-1 with A, B the E and H rows of the forward (mode -1:
      // backward) columns
      const int q0 = A.mode < 0 ? 1 : 0, q1 = A.mode < 0 ? 3 : 2;
      const cd Am[2][2] = {{m.F[0][q0], m.F[0][q1]}, {m.F[2][q0], m.F[2][q1]}};
      cd Ai[2][2];
      inv2(Am, Ai);
      cpx<T> G[2][2];
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          G[i][j] = cast<T>(m.F[2 * i + 1][q0] * Ai[0][j] + m.F[2 * i + 1][q1] * Ai[1][j]);
      const int la0 = A.nlam_b ? A.lam0 : 0, la1 = A.nlam_b ? A.lam0 + A.nlam_b : A.nlam;
      for (int lam = la0; lam < la1; ++lam)
        for (int ray = r0; ray < r1; ++ray) {
          const cpx<T>* e = plane(const_cast<cpx<T>*>(A.src), ray, lam, 2);
          const cpx<T> ex = e[0], ey = e[A.HW];
          cpx<T>* o = plane(A.dst, ray, lam, 4);
          const cpx<T> hy = G[0][0] * ex + G[0][1] * ey, hx = G[1][0] * ex + G[1][1] * ey;
          if (A.accum) {
            o[0] = o[0] + ex;
            o[A.HW] = o[A.HW] + hy;
            o[2 * A.HW] = o[2 * A.HW] + ey;
            o[3 * A.HW] = o[3 * A.HW] + hx;
          } else {
            o[0] = ex;
            o[A.HW] = hy;
            o[2 * A.HW] = ey;
            o[3 * A.HW] = hx;
          }
        }
    }
  }
}


// ---------------------------------------------------------------------------
// Exact per-mode propagation (diffraction < 0 or inf;
// dtmm/propagate_4x4.py:515-564): every FFT mode of the field is a plane
// wave with its own (beta, phi), propagated at each pixel by that pixel's
// F P(kd) F^-1 at the mode's (beta, phi), and the results are summed.
// One thread per pixel, a chunk of the wavelength's mode list per grid z.
// ---------------------------------------------------------------------------

constexpr int FULL_RAYS = 2;  // rays per thread (they share the per-mode solve)

template <typename T>
struct FullArgs {
  const cpx<T>* spec;   // unnormalised FFT of the field, [ray][lam][4][H*W]
  cpx<T>* part;         // partial sums [chunk][ray][lam][4][H*W]
  int H, W;
  size_t HW;
  int nrays, nlam;
  const double* eps_re;  // this layer's material, [3][L*H*W] with the offset applied
  const double* eps_im;
  const double* ang;
  size_t comp_stride;
  double d;              // kd = k0 * d for one step
  const int* modes;      // [lam][HW]: k indices (ky * W + kx) with beta < betamax, ascending beta
  const int* nmodes;     // [lam]
  int nchunk;
  Beams bm;
};

template <typename T>
__global__ __launch_bounds__(NT) void full_kernel(FullArgs<T> A) {
  const size_t p = size_t(blockIdx.x) * NT + threadIdx.x;
  const int ngroup = (A.nrays + FULL_RAYS - 1) / FULL_RAYS;
  const int lam = blockIdx.y / ngroup;
  const int r0 = (blockIdx.y % ngroup) * FULL_RAYS;
  const int chunk = blockIdx.z;
  if (p >= A.HW) return;
  const int y = int(p / A.W), x = int(p - size_t(y) * A.W);
  const size_t cs = A.comp_stride;
  const cd e0 = mk<double>(A.eps_re[p], A.eps_im ? A.eps_im[p] : 0.);
  const cd e1 = mk<double>(A.eps_re[cs + p], A.eps_im ? A.eps_im[cs + p] : 0.);
  const cd e2 = mk<double>(A.eps_re[2 * cs + p], A.eps_im ? A.eps_im[2 * cs + p] : 0.);
  const double psi = A.ang[p], theta = A.ang[cs + p], phi = A.ang[2 * cs + p];
  const double k0 = A.bm.k0[lam];
  const double kd = k0 * A.d;
  const int nm = A.nmodes[lam];
  const int m0 = int((long long)nm * chunk / A.nchunk);
  const int m1 = int((long long)nm * (chunk + 1) / A.nchunk);
  const double inv_hw = 1. / double(A.HW);
  cd acc[FULL_RAYS][4];
  for (int r = 0; r < FULL_RAYS; ++r)
    for (int c = 0; c < 4; ++c) acc[r][c] = mk<double>(0.);
  for (int m = m0; m < m1; ++m) {
    const int k = A.modes[size_t(lam) * A.HW + m];
    const int ky = k / A.W, kx = k - ky * A.W;
    const double fx = fftfreq(kx, A.W), fy = fftfreq(ky, A.H);
    const double beta = (2 * M_PI / fabs(k0)) * sqrt(fx * fx + fy * fy);
    const double phim = atan2(fy, fx);
    Modes md;
    int kind;
    layer_modes(beta, cos(phim), sin(phim), phim, e0, e1, e2, psi, theta, phi, md, kind);
    cd Fi[4][4];
    if (!inv4(md.F, Fi))
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Fi[i][j] = mk<double>(0.);
    cd pm[4];
    for (int i = 0; i < 4; ++i) pm[i] = cexpi(md.a[i] * kd);
    // the mode's plane wave at this pixel, exp(2 pi i (ky y / H + kx x / W)) / HW
    const long long ny = (long long)ky * y % A.H, nx = (long long)kx * x % A.W;
    const double arg = 2 * M_PI * (double(ny) / A.H + double(nx) / A.W);
    double sw, cw;
    sincos(arg, &sw, &cw);
    const cd wave = mk<double>(cw * inv_hw, sw * inv_hw);
    for (int r = 0; r < FULL_RAYS; ++r) {
      if (r0 + r >= A.nrays) break;
      const cpx<T>* sp = A.spec + (size_t(r0 + r) * A.nlam + lam) * 4 * A.HW + k;
      cd v[4], u[4];
      for (int c = 0; c < 4; ++c) v[c] = cast<double>(sp[c * A.HW]) * wave;
      for (int i = 0; i < 4; ++i) {  // u = P F^-1 v
        cd t = mk<double>(0.);
        for (int j = 0; j < 4; ++j) t = t + Fi[i][j] * v[j];
        u[i] = t * pm[i];
      }
      for (int i = 0; i < 4; ++i) {
        cd t = mk<double>(0.);
        for (int j = 0; j < 4; ++j) t = t + md.F[i][j] * u[j];
        acc[r][i] = acc[r][i] + t;
      }
    }
  }
  for (int r = 0; r < FULL_RAYS; ++r) {
    if (r0 + r >= A.nrays) break;
    cpx<T>* o = A.part + ((size_t(chunk) * A.nrays + r0 + r) * A.nlam + lam) * 4 * A.HW + p;
    for (int c = 0; c < 4; ++c) o[c * A.HW] = cast<T>(acc[r][c]);
  }
}

// dst[i] = sum over chunks of part[chunk][i], chunks in order (deterministic).
template <typename T>
__global__ __launch_bounds__(NT) void chunk_sum_kernel(const cpx<T>* part, cpx<T>* dst, size_t n,
                                                       int nchunk) {
  const size_t i = size_t(blockIdx.x) * NT + threadIdx.x;
  if (i >= n) return;
  cpx<T> s = part[i];
  for (int c = 1; c < nchunk; ++c) s = s + part[size_t(c) * n + i];
  dst[i] = s;
}

}  // namespace dtmmx
