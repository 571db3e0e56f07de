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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"

// The library is compiled as one translation unit (DTMMX_PART 0) or, to
// build in parallel, as three: 1 = host code and the C ABI, 2 = the
// complex64 kernels and driver, 3 = the complex128 ones.
#ifndef DTMMX_PART
#define DTMMX_PART 0
#endif
#define DTMMX_HOST_PART (DTMMX_PART == 0 || DTMMX_PART == 1)
#define DTMMX_C64_PART (DTMMX_PART == 0 || DTMMX_PART == 2)
#define DTMMX_C128_PART (DTMMX_PART == 0 || DTMMX_PART == 3)

namespace dtmmx {
#if DTMMX_HOST_PART
thread_local std::string g_err;
#else
extern thread_local std::string g_err;
#endif

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Optional per-launch timing of the per-layer kernels with HIP events on the
// launch stream (class 0: row layer kernel, 1: column k-space kernel,
// 3: dense k-space head/tail kernel).
struct KTimer {
  bool on = false;
  bool armed = false;  // the current launch is being timed
  int cap = 0, used = 0, stride = 1;
  int seen[4] = {0, 0, 0, 0};
  std::vector<hipEvent_t> ev;
  std::vector<int> cls;
};
#if DTMMX_HOST_PART
KTimer g_timer;
#else
extern KTimer g_timer;
#endif

// A timed launch passes its event pair to hipExtLaunchKernel, which stamps
// them with the kernel's own start and end (the dispatch's timestamps, as a
// kernel trace reads them), so the measured time holds no queue or event
// overhead.  timer_start arms the pair for the next DTMMX_LAUNCH on the
// stream; timer_stop keeps it.
static void timer_start(int cls, hipStream_t) {
  g_timer.armed = false;
  if (!g_timer.on || g_timer.used >= g_timer.cap) return;
  if (g_timer.seen[cls & 3]++ % g_timer.stride != 0) return;
  g_timer.armed = true;
  g_timer.cls[g_timer.used] = cls;
}

static void timer_stop(hipStream_t) {
  if (!g_timer.armed) return;
  g_timer.armed = false;
  ++g_timer.used;
}

#define DTMMX_LAUNCH(K, G, B, S, A)                                                     \
  do {                                                                                  \
    if (g_timer.armed)                                                                  \
      hipExtLaunchKernelGGL(K, G, B, 0, S, g_timer.ev[2 * g_timer.used],                \
                            g_timer.ev[2 * g_timer.used + 1], 0, A);                    \
    else                                                                                \
      hipLaunchKernelGGL(K, G, B, 0, S, A);                                             \
  } while (0)

#define DTMMX_LAUNCH_SHM(K, G, B, SHM, S, A)                                            \
  do {                                                                                  \
    if (g_timer.armed)                                                                  \
      hipExtLaunchKernelGGL(K, G, B, SHM, S, g_timer.ev[2 * g_timer.used],              \
                            g_timer.ev[2 * g_timer.used + 1], 0, A);                    \
    else                                                                                \
      hipLaunchKernelGGL(K, G, B, SHM, S, A);                                           \
  } while (0)

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(DTMMX_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_));            \
  } while (0)

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

#define DTMMX_SIZES(X)                                                                        \
  X(8) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) \
  X(320) X(384) X(512) X(640) X(768) X(1024)

static bool size_supported(int n) {
  switch (n) {
#define X(N) case N:
    DTMMX_SIZES(X)
#undef X
    return true;
    default:
      return false;
  }
}

// Rays per row work item: two rays share the eigen data and phases when the
// line registers fit (complex64 up to 512 points).
template <typename T, int N>
constexpr int row_nr_max() {
  return (sizeof(T) == 4 && N <= 512) ? 2 : 1;
}

template <typename T, int NC, int OP>
static int launch_row(int n, int nr, const RowArgs<T>& a, int nwork, hipStream_t s, int nw = 1) {
  switch (n) {
#define X(N)                                                                                \
  case N:                                                                                   \
    if constexpr (OP == OP_ROP && NC == 2 && row_nr_max<T, N>() >= 2 && row_lanes<N>() == 64) { \
      if (nw == 2 && nr == 2) {                                                             \
        DTMMX_LAUNCH((row_kernel<T, N, 2, 2, OP_ROP, 2>), dim3(nwork), dim3(64 * 8), s, a);  \
        break;                                                                              \
      }                                                                                     \
    }                                                                                       \
    if (nr == 2 && row_nr_max<T, N>() >= 2)                                                 \
      DTMMX_LAUNCH((row_kernel<T, N, (row_nr_max<T, N>() >= 2 ? 2 : 1), NC, OP>),           \
                   dim3(nwork), dim3(row_lanes<N>() * 2 * NC), s, a);                       \
    else                                                                                    \
      DTMMX_LAUNCH((row_kernel<T, N, 1, NC, OP>), dim3(nwork), dim3(row_lanes<N>() * NC), s, \
                   a);                                                                      \
    break;
    DTMMX_SIZES(X)
#undef X
    default:
      return fail(DTMMX_E_UNSUPPORTED, "unsupported width " + std::to_string(n));
  }
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

// Compute units of the device current at the first call (256 if unknown),
// kept for the process; the initialisation of a function-local static runs
// once even when several threads plan at the same time.
static int device_cus() {
  static const int v = [] {
    int dev = 0, x = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&x, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || x <= 0)
      x = 256;
    return x;
  }();
  return v;
}

// Fused window row pass: NR rays of one beam per work-group (two for
// complex64 lines up to 512 points) and a chunk of wavelengths whose
// per-window sums fill the LDS beside the kernel's own buffers.  Returns
// DTMMX_E_UNSUPPORTED (nothing launched) when not even one wavelength fits.
static int max_lds_per_block() {
  static const int v = [] {  // initialised once, also under concurrent plans
    int dev = 0, x = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&x, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      x = 65536;
    return x;
  }();
  return v;
}

template <typename T, int NC>
static int launch_winrow(int n, int nr, WinRowArgs<T> a, int nlam, hipStream_t s) {
  switch (n) {
#define X(N)                                                                                   \
  case N: {                                                                                    \
    constexpr int NRM = row_nr_max<T, N>() >= 2 ? 2 : 1;                                       \
    auto go = [&](auto nrc) -> int {                                                           \
      constexpr int NRX = decltype(nrc)::value;                                                \
      constexpr int NTW = row_lanes<N>() * NRX * NC;                                           \
      constexpr int PPX = (N + NTW - 1) / NTW;                                                 \
      const size_t stat = (size_t(NRX * NC) * line_stride<N>() + tw_table_slots<N>()) * sizeof(cpx<T>); \
      const size_t per = size_t(PPX) * NRX * NC * NTW * sizeof(cpx<T>);                        \
      const size_t cap = size_t(max_lds_per_block());                                          \
      int c = cap > stat ? int((cap - stat) / per) : 0;                                        \
      if (const char* e = std::getenv("DTMMX_WIN_C")) c = std::min(c, std::atoi(e));          \
      c = std::min(c, nlam);                                                                   \
      if (c < 1) return fail(DTMMX_E_UNSUPPORTED, "window sums do not fit in LDS");           \
      a.nchunk = c;                                                                            \
      const int nwork = a.H * (a.b.nrays / NRX) * ((nlam + c - 1) / c);                        \
      const size_t shm = size_t(c) * per;                                                      \
      (void)hipFuncSetAttribute((const void*)winrow_kernel<T, N, NRX, NC>,                     \
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(shm));         \
      timer_start(0, s);                                                                       \
      DTMMX_LAUNCH_SHM((winrow_kernel<T, N, NRX, NC>), dim3(nwork), dim3(NTW), shm, s, a);     \
      timer_stop(s);                                                                           \
      return DTMMX_OK;                                                                         \
    };                                                                                         \
    int r = (nr == 2 && NRM == 2) ? go(std::integral_constant<int, NRM>{})                     \
                                  : go(std::integral_constant<int, 1>{});                      \
    if (r != DTMMX_OK) return r;                                                               \
    break;                                                                                     \
  }
    DTMMX_SIZES(X)
#undef X
    default:
      return fail(DTMMX_E_UNSUPPORTED, "unsupported width " + std::to_string(n));
  }
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

// Columns per column work-group: the widest of 16, 8, 4, 2, 1 (16 gives
// 128-byte row segments for complex64) whose LDS stays within 64 KiB and
// whose work-group stays within 1024 threads.
template <typename T, int N>
#ifndef DTMMX_COL_LDS_CAP
#define DTMMX_COL_LDS_CAP 65536
#endif
constexpr int col_tx() {
  constexpr int cb = int(sizeof(cpx<T>));
  constexpr int nl = col_lanes<N>();
#ifdef DTMMX_COL_TX
  int tx = DTMMX_COL_TX;
#else
  int tx = 16;
#endif
  while (tx > 1 && (N * (tx + 1) * cb > DTMMX_COL_LDS_CAP || tx * nl > 1024)) tx /= 2;
  return tx;
}

// Lines per work-group: halved (down to two) while the launch has fewer
// than COL_WORK_TARGET work-groups, about 8 per CU, so that short
// work-groups overlap their setup with other work-groups' streaming (the
// 512^2 x 18-plane pass: 2 lines 16.7 ms, 4 lines 17.1 ms, 1 line 17.9 ms);
// then, for small grids, down to one until there are COL_MIN_WORK, two per
// CU.
constexpr int COL_WORK_TARGET = 2048;
constexpr int COL_MIN_WORK = 512;

template <typename T, int OP>
static int launch_col(int n, const ColArgs<T>& a0, int nlines_per_lam, bool tune, hipStream_t s) {
  ColArgs<T> a = a0;
  switch (n) {
#define X(N)                                                                                \
  case N: {                                                                                 \
    constexpr int TX = col_tx<T, N>();                                                      \
    const int ntile = (a.W + TX - 1) / TX;                                                  \
    while (tune && a.ncl % 2 == 0 && a.ncl > 2 &&                                           \
           ntile * a.b.nlam_b * (nlines_per_lam / a.ncl) < COL_WORK_TARGET)                   \
      a.ncl /= 2;                                                                           \
    while (tune && a.ncl % 2 == 0 && ntile * a.b.nlam_b * (nlines_per_lam / a.ncl) < COL_MIN_WORK) \
      a.ncl /= 2;                                                                           \
    const int nwork = ntile * a.b.nlam_b * (nlines_per_lam / a.ncl);                        \
    DTMMX_LAUNCH((col_kernel<T, N, TX, OP>), dim3(nwork), dim3(TX * col_lanes<N>()), s, a);  \
    break;                                                                                  \
  }
    DTMMX_SIZES(X)
#undef X
    default:
      return fail(DTMMX_E_UNSUPPORTED, "unsupported height " + std::to_string(n));
  }
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

template <typename T>
static int launch_row_adj(int n, const AdjArgs<T>& a, int nwork, hipStream_t s) {
  switch (n) {
#define X(N)                                                                                \
  case N:                                                                                   \
    if constexpr (sizeof(T) == 8 && N > 512) {                                              \
      return fail(DTMMX_E_UNSUPPORTED, "complex128 adjoint is built up to 512 points");      \
    } else {                                                                                \
      DTMMX_LAUNCH((row_adj_kernel<T, N>), dim3(nwork), dim3(row_lanes<N>() * 4), s, a);    \
    }                                                                                       \
    break;
    DTMMX_SIZES(X)
#undef X
    default:
      return fail(DTMMX_E_UNSUPPORTED, "unsupported width " + std::to_string(n));
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
  // diffraction > 1: non-zero tent windows (i, j) of wn x wn in the
  // reference's order i * wn + j, their mean beams [window][beam], the
  // separable tent factors on the device and the mask of the window grid
  int wn = 0;
  std::vector<int> win_i, win_j;
  std::vector<double> win_beta, win_phi;
  double* wcx = nullptr;  // [beam][nlam][wn][W]
  double* wcy = nullptr;  // [beam][nlam][wn][H]
  int* wmask = nullptr;   // [nlam][H]: min(amax, largest |kx| the windows pass)
  // the windows in tent-row order (j, then i) for the fused window row
  // kernel: win_ord [2][nwin] (i, j), win_bp [2][nwin][nbeam] (beta, phi)
  int* win_ord = nullptr;
  double* win_bp = nullptr;
  // diffraction < 0 (exact per-mode propagation): per wavelength the k
  // indices with beta < betamax in ascending beta, their counts, and the
  // number of mode chunks one launch splits them into
  int* full_modes = nullptr;   // [nlam][H*W]
  int* full_nmodes = nullptr;  // [nlam]
  int full_chunks = 0;
  // launch shapes (0: chosen from the grid): column lines per work-group,
  // row rays per work-group (dtmmx_plan_set_tiles, the autotuner's knobs)
  int col_lines = 0;
  int row_rays = 0;
  // row work-groups per (row, ray group), each taking a chunk of the
  // wavelengths (dtmmx_plan_set_row_split; 0: chosen from the grid)
  int row_split = 0;
};

namespace dtmmx {



template <typename T>
static int row_pass(const dtmmx_plan* P, cpx<T>* field, Batch b, int nc, int op, hipStream_t s,
                    const RowArgs<T>* proto = nullptr, const Beams* bm = nullptr) {
  RowArgs<T> a{};
  if (proto) a = *proto;
  a.field = field;
  a.H = P->d.height;
  a.W = P->d.width;
  a.b = b;
  a.tw = (const cpx<T>*)P->tw_w;
  a.bm = bm ? *bm : P->bm;
  int nr = 1;
  if (sizeof(T) == 4 && a.W <= 512 && b.nrays % 2 == 0 &&
      ((op != OP_ROP && op != OP_ROPW) || P->bm.rays_per_beam % 2 == 0) && P->row_rays != 1)
    nr = 2;
  a.rays_per_group = nr;
  const int w = a.W;
  // two-component layer passes of two rays (complex64, one wave per line):
  // two wavelengths per work-group in separate waves, 8 waves sharing one
  // per-pixel solve
  const int nw = (op == OP_ROP && nc == 2 && nr == 2 && sizeof(T) == 4 && w <= 512) ? 2 : 1;
  // wavelength chunks per (row, ray group): the plan's setting, else enough
  // work-groups to give every SIMD of the device the row kernel's occupancy
  // (4 waves of complex64, 2 of complex128), at most 8 and at most the
  // batch's wavelengths
  int split = P->row_split;
  if (split <= 0) {
    const long long waves = (long long)a.H * (b.nrays / nr) * nr * nc * nw * (w >= 1024 ? 2 : 1);
    const long long want = (long long)device_cus() * 4 * (sizeof(T) == 4 ? 4 : 2);
    split = int((want + waves - 1) / (waves > 0 ? waves : 1));
  }
  split = std::max(1, std::min({split, 8, std::max(1, b.nlam_b)}));
  a.lam_split = split;
  const int nwork = a.H * (b.nrays / nr) * split;
  if ((op == OP_ROP || op == OP_ROPW) && a.etab_ws) {
    // layer with biaxial voxels: per-pixel data for this pass's beams first
    const size_t HW = size_t(a.H) * a.W;
    EigArgs<T> e{};
    e.eps_re = a.eps_re;
    e.eps_im = a.eps_im;
    e.ang = a.ang;
    e.comp_stride = a.comp_stride;
    e.HW = HW;
    e.nbeam = b.nrays / a.bm.rays_per_beam;
    e.table = a.etab_ws;
    e.status = a.status;
    e.method4 = a.method4;
    e.bwd2 = a.bwd2;
    e.bm = a.bm;
    hipLaunchKernelGGL((rop_table_kernel<T>), dim3(unsigned((HW + NT - 1) / NT), e.nbeam), dim3(NT), 0,
                       s, e);
    HIP_TRY(hipGetLastError());
    a.etab = a.etab_ws;
    a.etab_nbeam = e.nbeam;
  }
  if (op == OP_ROP) {
    timer_start(0, s);
    int r = nc == 2 ? launch_row<T, 2, OP_ROP>(w, nr, a, nwork, s, nw)
                    : launch_row<T, 4, OP_ROP>(w, nr, a, nwork, s);
    timer_stop(s);
    return r;
  }
  if (op == OP_ROPW) {
    if (nc != 4 || !a.wcx) return fail(DTMMX_E_ARG, "window row pass needs 4 components");
    timer_start(0, s);
    int r = launch_row<T, 4, OP_ROPW>(w, nr, a, nwork, s);
    timer_stop(s);
    return r;
  }
  if (nc == 1 && op == OP_FWD) return launch_row<T, 1, OP_FWD>(w, nr, a, nwork, s);
  if (nc == 1 && op == OP_INV) return launch_row<T, 1, OP_INV>(w, nr, a, nwork, s);
  if (nc == 2 && op == OP_FWD) return launch_row<T, 2, OP_FWD>(w, nr, a, nwork, s);
  if (nc == 2 && op == OP_INV) return launch_row<T, 2, OP_INV>(w, nr, a, nwork, s);
  if (nc == 4 && op == OP_FWD) return launch_row<T, 4, OP_FWD>(w, nr, a, nwork, s);
  if (nc == 4 && op == OP_INV) return launch_row<T, 4, OP_INV>(w, nr, a, nwork, s);
  if (nc == 4 && op == OP_COMB) {
    if (!a.csrc) return fail(DTMMX_E_ARG, "combination row pass without a source");
    return launch_row<T, 4, OP_COMB>(w, nr, a, nwork, s);
  }
  if (nc == 4 && op == OP_SPEC) {
    // every ray of the batch in one work-group: the specter sums over them
    if (nr != b.nrays) return fail(DTMMX_E_ARG, "specter row pass needs all rays per work-group");
    return launch_row<T, 4, OP_SPEC>(w, nr, a, nwork, s);
  }
  return fail(DTMMX_E_ARG, "bad row pass");
}

template <typename T>
static int col_pass(const dtmmx_plan* P, cpx<T>* field, Batch b, int nc, int op, hipStream_t s,
                    const ColArgs<T>* proto = nullptr, const Beams* bm = nullptr) {
  ColArgs<T> a{};
  if (proto) a = *proto;
  a.field = field;
  a.H = P->d.height;
  a.W = P->d.width;
  a.b = b;
  a.NC = nc;
  {
    // the passing modes all propagate in ea: real, eps above betamax^2 by a
    // margin over the float rounding of beta^2 in the kernel
    const double bm2 = P->d.betamax * P->d.betamax;
    a.prop = (a.ea.y == 0. && a.ea.x > bm2 * (1. + 1e-5) + 1e-12) ? 1 : 0;
  }
  // lines per work-group: the components of one ray, or of two rays sharing a beam
  a.ncl = nc;
  if (b.nrays % 2 == 0 && P->bm.rays_per_beam % 2 == 0 && nc <= 2) a.ncl = 2 * nc;
  bool tune = true;
  {
    int v = P->col_lines;
    if (const char* e = std::getenv("DTMMX_COL_NCL")) v = std::atoi(e);
    const bool pair = b.nrays % 2 == 0 && P->bm.rays_per_beam % 2 == 0;
    if (v > 0 && (nc % v == 0 || (pair && v == 2 * nc))) {
      a.ncl = v;
      tune = false;
    }
  }
  a.tw = (const cpx<T>*)P->tw_h;
  a.amax = P->amax;
  a.bm = bm ? *bm : P->bm;
  const int nlines = b.nrays * nc;
  const int h = a.H;
  if (op == OP_FWD) return launch_col<T, OP_FWD>(h, a, nlines, tune, s);
  if (op == OP_INV) return launch_col<T, OP_INV>(h, a, nlines, tune, s);
  if (op == OP_FIELD) {
    // a work-group takes the four components of a plane
    if (nc != 4) return fail(DTMMX_E_ARG, "the field operator pass takes 4 components");
    a.ncl = 4;
    timer_start(1, s);
    int r = launch_col<T, OP_FIELD>(h, a, nlines, false, s);
    timer_stop(s);
    return r;
  }
  if (op == OP_DIAG) {
    timer_start(1, s);
    int r = launch_col<T, OP_DIAG>(h, a, nlines, tune, s);
    timer_stop(s);
    return r;
  }
  if (op == OP_WIN || op == OP_SPLIT) {
    if (!P->wcx || !a.out) return fail(DTMMX_E_ARG, "window pass without windows");
    a.amax = P->wmask;
    a.wcx = P->wcx;
    a.wcy = P->wcy;
    a.wn = P->wn;
    timer_start(1, s);
    int r = op == OP_WIN ? launch_col<T, OP_WIN>(h, a, nlines, tune, s)
                         : launch_col<T, OP_SPLIT>(h, a, nlines, tune, s);
    timer_stop(s);
    return r;
  }
  return fail(DTMMX_E_ARG, "bad column pass");
}

// Boundary operator of anisotropic effective media on the full spectrum
// (in place): GOP_HEAD first(m1), GOP_TAIL second(m0), GOP_BOUND both.
template <typename T>
static int gkspace(const dtmmx_plan* P, cpx<T>* f, Batch b, int gop, const GMed* m0,
                   const GMed* m1, hipStream_t s) {
  GArgs<T> a{};
  a.src = f;
  a.dst = f;
  a.H = P->d.height;
  a.W = P->d.width;
  a.b = b;
  a.amax = P->amax;
  a.bm = P->bm;
  a.scale = 1.0 / (double(a.H) * a.W);
  if (m0) a.m0 = *m0;
  if (m1) a.m1 = *m1;
  const int HW = a.H * a.W;
  const int nbeam = b.nrays / P->bm.rays_per_beam;
  dim3 grid((HW + NT - 1) / NT, nbeam * b.nlam_b);
  timer_start(3, s);
  if (gop == GOP_HEAD)
    DTMMX_LAUNCH((gkspace_kernel<T, GOP_HEAD>), grid, dim3(NT), s, a);
  else if (gop == GOP_TAIL)
    DTMMX_LAUNCH((gkspace_kernel<T, GOP_TAIL>), grid, dim3(NT), s, a);
  else
    DTMMX_LAUNCH((gkspace_kernel<T, GOP_BOUND>), grid, dim3(NT), s, a);
  timer_stop(s);
  return DTMMX_OK;
}

template <typename T>
static int kspace(const dtmmx_plan* P, const cpx<T>* src, cpx<T>* dst, Batch b, KArgs<T> a,
                  hipStream_t s) {
  a.src = src;
  a.dst = dst;
  a.H = P->d.height;
  a.W = P->d.width;
  a.b = b;
  a.amax = P->amax;
  a.bm = P->bm;
  const int HW = a.H * a.W;
  dim3 grid((HW + NT - 1) / NT, b.nrays * b.nlam_b);
  timer_start(3, s);
  switch (a.kop) {
    case KOP_HEAD4:
      DTMMX_LAUNCH((kspace_kernel<T, KOP_HEAD4, false>), grid, dim3(NT), s, a);
      break;
    case KOP_TAIL4:
      if (a.adjoint)
        DTMMX_LAUNCH((kspace_kernel<T, KOP_TAIL4, true>), grid, dim3(NT), s, a);
      else
        DTMMX_LAUNCH((kspace_kernel<T, KOP_TAIL4, false>), grid, dim3(NT), s, a);
      break;
    case KOP_BOUND4:
      if (a.adjoint)
        DTMMX_LAUNCH((kspace_kernel<T, KOP_BOUND4, true>), grid, dim3(NT), s, a);
      else
        DTMMX_LAUNCH((kspace_kernel<T, KOP_BOUND4, false>), grid, dim3(NT), s, a);
      break;
    case KOP_PROJ2:
      DTMMX_LAUNCH((kspace_kernel<T, KOP_PROJ2, false>), grid, dim3(NT), s, a);
      break;
    case KOP_TR2:
      DTMMX_LAUNCH((kspace_kernel<T, KOP_TR2, false>), grid, dim3(NT), s, a);
      break;
    case KOP_FIELD4:
      DTMMX_LAUNCH((kspace_kernel<T, KOP_FIELD4, false>), grid, dim3(NT), s, a);
      break;
    case KOP_E2H4:
      DTMMX_LAUNCH((kspace_kernel<T, KOP_E2H4, false>), grid, dim3(NT), s, a);
      break;
    case KOP_IF2W:
      DTMMX_LAUNCH((kspace_kernel<T, KOP_IF2W, false>), grid, dim3(NT), s, a);
      break;
    case KOP_OUT2W:
      DTMMX_LAUNCH((kspace_kernel<T, KOP_OUT2W, false>), grid, dim3(NT), s, a);
      break;
    default:
      timer_stop(s);
      return fail(DTMMX_E_ARG, "bad k-space operator");
  }
  timer_stop(s);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

// Row-kernel arguments of the layer operator of layer l, sub-step thickness d.
template <typename T>
static RowArgs<T> layer_args(const dtmmx_stack* S, size_t HW, int l, double d, int method4,
                             int* status, void* etab_ws = nullptr) {
  RowArgs<T> a{};
  if (S->biaxial && S->biaxial[l]) a.etab_ws = (cpx<T>*)etab_ws;
  a.eps_re = S->eps_re + size_t(l) * HW;
  a.eps_im = S->eps_im ? S->eps_im + size_t(l) * HW : nullptr;
  a.ang = S->angles + size_t(l) * HW;
  a.comp_stride = size_t(S->n_layers) * HW;
  a.d = d;
  a.method4 = method4;
  a.status = status;
  return a;
}

// One propagation step: layer l of the stack with sub-step thickness d and
// isotropic effective eps e (transfer.py:933-934 splits layer l into
// n_steps[l] steps of thickness[l] / n_steps[l]).
struct Step {
  int l;
  double d;
  cd e;
  GMed g;  // anisotropic effective medium (stacks with eff_eps_re)
};

// dir = -1: the backward pass of npass > 1, layers in reverse order with
// negative thickness (transfer.py:744-760).
static std::vector<Step> expand_steps(const dtmmx_stack* S, int dir) {
  std::vector<Step> v;
  const int L = S->n_layers;
  for (int k = 0; k < L; ++k) {
    const int l = dir > 0 ? k : L - 1 - k;
    const int ns = S->n_steps ? S->n_steps[l] : 1;
    const cd e = mk<double>(S->eps_eff_re[l], S->eps_eff_im ? S->eps_eff_im[l] : 0.);
    GMed g{};
    if (S->eff_eps_re) {
      for (int c = 0; c < 3; ++c) {
        g.e[c] = mk<double>(S->eff_eps_re[3 * l + c], S->eff_eps_im ? S->eff_eps_im[3 * l + c] : 0.);
        g.ang[c] = S->eff_angles[3 * l + c];
      }
      g.d = ((S->eff_thickness ? S->eff_thickness[l] : S->thickness[l]) * dir) / double(ns);
    }
    for (int j = 0; j < ns; ++j) v.push_back(Step{l, (S->thickness[l] * dir) / double(ns), e, g});
  }
  return v;
}

// Workspace layout of a transfer call (byte offsets; 256-byte aligned).
struct WsLayout {
  size_t pass_field, pass_field0;  // npass > 1: propagated field, transmitted input
  size_t win_tmp, win_acc;         // diffraction > 1: window rows (wn fields; the
                                   // first is the window part of sub-stepped layers), sum
  size_t e2;                       // 2x2: the 2-component field
  size_t part, sums;               // npass > 1: reduction partials and sums
  size_t etab;                     // per-pixel data of a layer with biaxial voxels
  size_t full_spec, full_part;     // diffraction < 0: spectrum, per-chunk partial sums
  size_t tmp2;                     // 2x2 npass > 1: a 2-component field
  size_t refb;                     // 4x4 npass > 1 with DTMMX_NORM_REF: reference field
  size_t total;                    // the stored reflections of 2x2 npass > 1 follow
                                   // (stack_ws_bytes)
  int nblk;
};

static WsLayout ws_layout(const dtmmx_plan* P) {
  WsLayout w{};
  const size_t cb = P->d.precision == DTMMX_C64 ? sizeof(cpx<float>) : sizeof(cpx<double>);
  const size_t HW = size_t(P->d.height) * P->d.width;
  const size_t planes = size_t(P->d.n_rays) * P->d.n_wavelengths;
  const size_t f4 = planes * 4 * HW * cb;
  auto take = [&](size_t bytes) {
    const size_t o = w.total;
    w.total += (bytes + 255) / 256 * 256;
    return o;
  };
  w.nblk = int((HW + NT - 1) / NT);
  const bool m4 = P->d.method == DTMMX_METHOD_4X4;
  if (m4 && P->d.npass > 1) {
    w.pass_field = take(f4);
    w.pass_field0 = take(f4);
    w.part = take(2 * planes * size_t(w.nblk) * sizeof(double));
    w.sums = take(4 * planes * sizeof(double));
    if (P->d.norm & DTMMX_NORM_REF) {
      w.refb = take(f4);
      w.e2 = take(planes * 2 * HW * cb);
    }
  }
  if (m4 && P->wn > 1) {
    w.win_tmp = take(f4 * P->wn);
    w.win_acc = take(f4);
  }
  if (!m4) w.e2 = take(planes * 2 * HW * cb);
  if (!m4 && P->wn > 1) {
    w.win_tmp = take(planes * 2 * HW * cb * P->wn);
    w.win_acc = take(planes * 2 * HW * cb);
  }
  if (!m4 && P->d.npass > 1) w.tmp2 = take(planes * 2 * HW * cb);
  if (P->d.diffraction < 0) {
    w.full_spec = take(f4);
    w.full_part = take(f4 * size_t(P->full_chunks > 0 ? P->full_chunks : 1));
  }
  w.etab = take(size_t(EIG_COMPS) * (P->d.n_rays / P->d.rays_per_beam) * HW * cb);
  return w;
}

// Workspace that depends on the stack: the reflected wave of every interface
// (input medium, layers, output medium: n_layers + 1) kept between the
// passes of the 2x2 method with npass > 1 (dtmm/transfer.py:1081-1083).
static size_t stack_ws_bytes(const dtmmx_plan* P, const dtmmx_stack* S) {
  if (P->d.method != DTMMX_METHOD_2X2 || P->d.npass <= 1 || !S) return 0;
  const size_t cb = P->d.precision == DTMMX_C64 ? sizeof(cpx<float>) : sizeof(cpx<double>);
  const size_t HW = size_t(P->d.height) * P->d.width;
  const size_t planes = size_t(P->d.n_rays) * P->d.n_wavelengths;
  return size_t(S->n_layers + 1) * ((planes * 2 * HW * cb + 255) / 256 * 256);
}

#define TRY(x)                     \
  do {                             \
    int r_ = (x);                  \
    if (r_ != DTMMX_OK) return r_; \
  } while (0)

// Wavelengths per batch: the whole batch stays resident in the 256 MiB
// memory-side cache across the row and column passes of a layer when it is
// small enough.  DTMMX_BATCH_MB overrides the budget (0: no batching).
static int lam_batch(const dtmmx_plan* P, size_t bytes_per_lam) {
  long mb = 0;
  if (const char* e = std::getenv("DTMMX_BATCH_MB")) mb = std::atol(e);
  const int nlam = P->d.n_wavelengths;
  if (mb <= 0) return nlam;
  int nb = int(size_t(mb) * 1048576 / (bytes_per_lam ? bytes_per_lam : 1));
  if (nb < 1) nb = 1;
  return nb < nlam ? nb : nlam;
}

// 4x4 method with diffraction > 1 (dtmm/propagate_4x4.py:390-419): per layer
// and window w, the field (kept in the (E+x, E-x, E+y, E-y) basis of the
// effective medium, without the mean-beam corrections) is multiplied in
// k-space by the window, the common half-step phases and the window's
// correction exp(-+i k0 d a_c(w)) of the whole layer (scalars per direction
// commute with the per-pixel operator), propagated through the layer's
// sub-steps with the window's mean beam, and summed into `acc`.
// DTMMX_FUSED_WINDOWS=0 runs one row launch per window instead of the fused
// window row kernel (same arithmetic, summed in a different order).
static bool fused_windows_off() {
  const char* e = std::getenv("DTMMX_FUSED_WINDOWS");
  return e && e[0] == '0';
}

template <typename T>
static int transfer_windows(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* out, cpx<T>* tmp,
                            cpx<T>* acc, Batch b, int* status, hipStream_t s, int dir,
                            void* etab) {
  const int H = P->d.height, W = P->d.width;
  const size_t HW = size_t(H) * W;
  const int L = S->n_layers;
  const double scale = 1.0 / double(HW);
  const int nbeam = P->d.n_rays / P->d.rays_per_beam;
  const int nwin = int(P->win_i.size());
  auto eff = [&](int l) {
    return mk<double>(S->eps_eff_re[l], S->eps_eff_im ? S->eps_eff_im[l] : 0.);
  };
  // Layers with biaxial voxels get the per-beam operator table of the
  // window's mean beams (built by the row pass that receives etab_ws).
  auto win_layer_args = [&](int l, double ds) {
    return layer_args<T>(S, HW, l, ds, 1, status, etab);
  };
  cpx<T>* A = out;
  TRY(row_pass<T>(P, A, b, 4, OP_FWD, s));
  TRY(col_pass<T>(P, A, b, 4, OP_FWD, s));
  KArgs<T> k{};
  k.kop = KOP_HEAD4;  // F -> Epn basis change only
  k.d0 = 0.;
  k.e0 = eff(dir > 0 ? 0 : L - 1);
  k.scale = 1.0 / double(H);
  k.nocorr = 1;
  TRY(kspace<T>(P, A, A, b, k, s));
  TRY(col_pass<T>(P, A, b, 4, OP_INV, s));
  double prev_d = 0.;  // half step of the previous layer still to apply
  cd prev_e = eff(dir > 0 ? 0 : L - 1);
  for (int kl = 0; kl < L; ++kl) {
    const int l = dir > 0 ? kl : L - 1 - kl;
    const int ns = S->n_steps ? S->n_steps[l] : 1;
    const double ds = (S->thickness[l] * dir) / double(ns);
    const cd el = eff(l);
    if (!eq(el, prev_e)) {
      TRY(col_pass<T>(P, A, b, 4, OP_FWD, s));
      KArgs<T> bd{};
      bd.kop = KOP_BOUND4;
      bd.d0 = prev_d;
      bd.e0 = prev_e;
      bd.d1 = 0.;
      bd.e1 = el;
      bd.scale = 1.0 / double(H);
      bd.nocorr = 1;
      TRY(kspace<T>(P, A, A, b, bd, s));
      TRY(col_pass<T>(P, A, b, 4, OP_INV, s));
      prev_d = 0.;
    }
    if (ns == 1) {
      // single-step layer: the tent windows are separable, cx(kx) cy(ky),
      // so one column pass writes the wn window rows cy_j * (common
      // half-step phases) * field, and the row pass of window (i, j) applies
      // cx_i and the window's mean-beam correction to row j on its input
      ColArgs<T> ca{};
      ca.da = prev_d;
      ca.ea = prev_e;
      ca.db = ds;
      ca.eb = el;
      ca.scale = scale;
      ca.dense_sign = 1;
      ca.out = tmp;
      TRY(col_pass<T>(P, A, b, 4, OP_SPLIT, s, &ca));
      const size_t fsz = size_t(b.nrays) * b.nlam * 4 * HW;
      const bool biax = S->biaxial && S->biaxial[l];
      if (!biax && !fused_windows_off()) {
        // every window in one launch, summed on chip
        WinRowArgs<T> wa{};
        wa.tmp = tmp;
        wa.out = acc;
        wa.fsz = fsz;
        wa.H = H;
        wa.W = W;
        wa.b = b;
        wa.tw = (const cpx<T>*)P->tw_w;
        const RowArgs<T> la = layer_args<T>(S, HW, l, ds, 1, status);
        wa.eps_re = la.eps_re;
        wa.eps_im = la.eps_im;
        wa.ang = la.ang;
        wa.comp_stride = la.comp_stride;
        wa.d = ds;
        wa.method4 = 1;
        wa.status = status;
        wa.wcx = P->wcx;
        wa.wn = P->wn;
        wa.nwin = nwin;
        wa.win_i = P->win_ord;
        wa.win_j = P->win_ord + nwin;
        wa.win_beta = P->win_bp;
        wa.win_phi = P->win_bp + size_t(nwin) * nbeam;
        wa.nbeam = nbeam;
        wa.we = el;
        wa.wd = ds;
        wa.bm = P->bm;
        int nr = 1;
        if (sizeof(T) == 4 && W <= 512 && b.nrays % 2 == 0 && P->bm.rays_per_beam % 2 == 0 &&
            P->row_rays != 1)
          nr = 2;
        const int r = launch_winrow<T, 4>(W, nr, wa, b.nlam_b, s);
        if (r == DTMMX_OK) {
          std::swap(A, acc);
          prev_d = ds;
          prev_e = el;
          continue;
        }
        if (r != DTMMX_E_UNSUPPORTED) return r;
      }
      for (int w = 0; w < nwin; ++w) {
        Beams bw = P->bm;
        for (int q = 0; q < nbeam; ++q) {
          bw.beta[q] = P->win_beta[size_t(w) * nbeam + q];
          bw.phi[q] = P->win_phi[size_t(w) * nbeam + q];
        }
        RowArgs<T> la = win_layer_args(l, ds);
        la.acc = acc;
        la.acc_mode = (w == 0) ? 1 : 2;
        la.wcx = P->wcx;
        la.wn = P->wn;
        la.wi = P->win_i[w];
        la.we = el;
        la.wd = ds;
        TRY(row_pass<T>(P, tmp + size_t(P->win_j[w]) * fsz, b, 4, OP_ROPW, s, &la, &bw));
      }
      std::swap(A, acc);
      prev_d = ds;
      prev_e = el;
      continue;
    }
    for (int w = 0; w < nwin; ++w) {
      Beams bw = P->bm;
      for (int q = 0; q < nbeam; ++q) {
        bw.beta[q] = P->win_beta[size_t(w) * nbeam + q];
        bw.phi[q] = P->win_phi[size_t(w) * nbeam + q];
      }
      ColArgs<T> ca{};
      ca.da = prev_d;
      ca.ea = prev_e;
      ca.kca = 0.;
      ca.db = ds;
      ca.eb = el;
      ca.kcb = 2.0 * ns;
      ca.scale = scale;
      ca.dense_sign = 1;
      ca.out = tmp;
      ca.wi = P->win_i[w];
      ca.wj = P->win_j[w];
      TRY(col_pass<T>(P, A, b, 4, OP_WIN, s, &ca, &bw));
      for (int st = 0; st < ns; ++st) {
        RowArgs<T> la = win_layer_args(l, ds);
        if (st > 0 && la.etab_ws) {
          // the table of this window and layer is built by the first sub-step
          la.etab = la.etab_ws;
          la.etab_nbeam = b.nrays / P->bm.rays_per_beam;
          la.etab_ws = nullptr;
        }
        if (st + 1 == ns) {
          la.acc = acc;
          la.acc_mode = (w == 0) ? 1 : 2;
        }
        TRY(row_pass<T>(P, tmp, b, 4, OP_ROP, s, &la, &bw));
        if (st + 1 < ns) {
          ColArgs<T> cs{};
          cs.da = ds;
          cs.db = ds;
          cs.ea = el;
          cs.eb = el;
          cs.scale = scale;
          cs.dense_sign = 1;
          TRY(col_pass<T>(P, tmp, b, 4, OP_DIAG, s, &cs, &bw));
        }
      }
    }
    std::swap(A, acc);
    prev_d = ds;
    prev_e = el;
  }
  TRY(col_pass<T>(P, A, b, 4, OP_FWD, s));
  KArgs<T> t{};
  t.kop = KOP_TAIL4;  // last half step, then Epn -> F
  t.d0 = prev_d;
  t.e0 = prev_e;
  t.scale = scale;
  t.nocorr = 1;
  TRY(kspace<T>(P, A, A, b, t, s));
  TRY(col_pass<T>(P, A, b, 4, OP_INV, s));
  TRY(row_pass<T>(P, A, b, 4, OP_INV, s));
  if (A != out) {
    const size_t bytes = size_t(b.nrays) * b.nlam * 4 * HW * sizeof(cpx<T>);
    HIP_TRY(hipMemcpyAsync(out, A, bytes, hipMemcpyDeviceToDevice, s));
  }
  return DTMMX_OK;
}

// One pass through the stack in place on `out` (real-space field in and
// out); dir = -1 runs the layers backward (npass > 1).
// Copies the planes of a wavelength batch (all rays, 4 components).
template <typename T>
static int copy_planes(cpx<T>* dst, const cpx<T>* src, const Batch& b, size_t HW, hipStream_t s,
                       int nc = 4) {
  const size_t per_lam = size_t(nc) * HW;
  for (int r = 0; r < b.nrays; ++r) {
    const size_t o = (size_t(r) * b.nlam + b.lam0) * per_lam;
    HIP_TRY(hipMemcpyAsync(dst + o, src + o, size_t(b.nlam_b) * per_lam * sizeof(cpx<T>),
                           hipMemcpyDeviceToDevice, s));
  }
  return DTMMX_OK;
}

// 2x2 method with tent windows (diffraction n > 1, reflection mode 0 or 1,
// npass 1; dtmm/propagate_2x2.py:212-257).  As the 4x4 windows: the E field
// is projected onto the forward modes of the input medium (and through the
// Fresnel interface into the first layer in reflection mode 1); per layer the
// column pass writes the window rows with the common half-step phases and
// the window row pass propagates every window's part with its own mean beam
// and sums them; interfaces between different effective media get the
// Fresnel transmission (mode 1); the output medium builds H per window.
// `out` holds the 4-component input on entry and the output on return; `e2`
// is the 2-component field.
template <typename T>
static int transfer_windows_2x2(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* out,
                                cpx<T>* e2, cpx<T>* tmp, cpx<T>* acc, Batch b, int* status,
                                hipStream_t s, void* etab) {
  const int H = P->d.height, W = P->d.width;
  const size_t HW = size_t(H) * W;
  const int L = S->n_layers;
  const double scale = 1.0 / double(HW);
  const int nbeam = P->d.n_rays / P->d.rays_per_beam;
  const int nwin = int(P->win_i.size());
  const cd ein = mk<double>(S->n_in * S->n_in), eout = mk<double>(S->n_out * S->n_out);
  auto eff = [&](int l) {
    return mk<double>(S->eps_eff_re[l], S->eps_eff_im ? S->eps_eff_im[l] : 0.);
  };
  TRY(row_pass<T>(P, out, b, 4, OP_FWD, s));
  TRY(col_pass<T>(P, out, b, 4, OP_FWD, s));
  KArgs<T> h{};
  h.kop = KOP_PROJ2;  // forward projection (and Fresnel into the first layer) only
  h.d0 = 0.;
  h.e0 = L > 0 ? eff(0) : eout;
  h.ein = ein;
  h.scale = 1.0 / double(H);
  h.fresnel = P->d.reflection;
  TRY(kspace<T>(P, out, e2, b, h, s));
  TRY(col_pass<T>(P, e2, b, 2, OP_INV, s));
  cpx<T>* A = e2;
  double prev_d = 0.;  // half step of the previous layer still to apply
  cd prev_e = L > 0 ? eff(0) : eout;
  const size_t fsz = size_t(b.nrays) * b.nlam * 2 * HW;
  for (int l = 0; l < L; ++l) {
    const int ns = S->n_steps ? S->n_steps[l] : 1;
    const double ds = S->thickness[l] / double(ns);
    const cd el = eff(l);
    if (!eq(el, prev_e)) {
      TRY(col_pass<T>(P, A, b, 2, OP_FWD, s));
      KArgs<T> t{};
      t.kop = KOP_IF2W;
      t.d0 = prev_d;
      t.e0 = prev_e;
      t.e1 = el;
      t.fresnel = P->d.reflection;
      t.scale = 1.0 / double(H);
      TRY(kspace<T>(P, A, A, b, t, s));
      TRY(col_pass<T>(P, A, b, 2, OP_INV, s));
      prev_d = 0.;
      prev_e = el;
    }
    const bool biax = S->biaxial && S->biaxial[l];
    bool done = false;
    if (ns == 1) {
      ColArgs<T> ca{};
      ca.da = prev_d;
      ca.ea = prev_e;
      ca.db = ds;
      ca.eb = el;
      ca.scale = scale;
      ca.dense_sign = 0;
      ca.out = tmp;
      if (!biax && !fused_windows_off()) {
        TRY(col_pass<T>(P, A, b, 2, OP_SPLIT, s, &ca));
        WinRowArgs<T> wa{};
        wa.tmp = tmp;
        wa.out = acc;
        wa.fsz = fsz;
        wa.H = H;
        wa.W = W;
        wa.b = b;
        wa.tw = (const cpx<T>*)P->tw_w;
        const RowArgs<T> la = layer_args<T>(S, HW, l, ds, 0, status);
        wa.eps_re = la.eps_re;
        wa.eps_im = la.eps_im;
        wa.ang = la.ang;
        wa.comp_stride = la.comp_stride;
        wa.d = ds;
        wa.method4 = 0;
        wa.status = status;
        wa.wcx = P->wcx;
        wa.wn = P->wn;
        wa.nwin = nwin;
        wa.win_i = P->win_ord;
        wa.win_j = P->win_ord + nwin;
        wa.win_beta = P->win_bp;
        wa.win_phi = P->win_bp + size_t(nwin) * nbeam;
        wa.nbeam = nbeam;
        wa.we = el;
        wa.wd = ds;
        wa.bm = P->bm;
        int nr = 1;
        if (sizeof(T) == 4 && W <= 512 && b.nrays % 2 == 0 && P->bm.rays_per_beam % 2 == 0 &&
            P->row_rays != 1)
          nr = 2;
        const int r = launch_winrow<T, 2>(W, nr, wa, b.nlam_b, s);
        if (r != DTMMX_OK && r != DTMMX_E_UNSUPPORTED) return r;
        done = r == DTMMX_OK;
      }
    }
    if (!done) {
      // one window at a time: its part through the layer's sub-steps with its
      // mean beam, added to the window sum
      for (int w = 0; w < nwin; ++w) {
        Beams bw = P->bm;
        for (int q = 0; q < nbeam; ++q) {
          bw.beta[q] = P->win_beta[size_t(w) * nbeam + q];
          bw.phi[q] = P->win_phi[size_t(w) * nbeam + q];
        }
        ColArgs<T> ca{};
        ca.da = prev_d;
        ca.ea = prev_e;
        ca.kca = 0.;
        ca.db = ds;
        ca.eb = el;
        ca.kcb = 2.0 * ns;
        ca.scale = scale;
        ca.dense_sign = 0;
        ca.out = tmp;
        ca.wi = P->win_i[w];
        ca.wj = P->win_j[w];
        TRY(col_pass<T>(P, A, b, 2, OP_WIN, s, &ca, &bw));
        for (int st = 0; st < ns; ++st) {
          RowArgs<T> la = layer_args<T>(S, HW, l, ds, 0, status, etab);
          if (st > 0 && la.etab_ws) {
            la.etab = la.etab_ws;
            la.etab_nbeam = b.nrays / P->bm.rays_per_beam;
            la.etab_ws = nullptr;
          }
          if (st + 1 == ns) {
            la.acc = acc;
            la.acc_mode = (w == 0) ? 1 : 2;
          }
          TRY(row_pass<T>(P, tmp, b, 2, OP_ROP, s, &la, &bw));
          if (st + 1 < ns) {
            ColArgs<T> cs{};
            cs.da = ds;
            cs.db = ds;
            cs.ea = el;
            cs.eb = el;
            cs.scale = scale;
            cs.dense_sign = 0;
            TRY(col_pass<T>(P, tmp, b, 2, OP_DIAG, s, &cs, &bw));
          }
        }
      }
    }
    std::swap(A, acc);
    prev_d = ds;
    prev_e = el;
  }
  // output medium: last half step, Fresnel (mode 1), H per window
  TRY(col_pass<T>(P, A, b, 2, OP_FWD, s));
  KArgs<T> o{};
  o.kop = KOP_OUT2W;
  o.d0 = prev_d;
  o.e0 = prev_e;
  o.e1 = eout;
  o.fresnel = P->d.reflection;
  o.scale = scale;
  o.wcx = P->wcx;
  o.wcy = P->wcy;
  o.win_i = P->win_ord;
  o.win_j = P->win_ord + nwin;
  o.win_beta = P->win_bp;
  o.win_phi = P->win_bp + size_t(nwin) * nbeam;
  o.wn = P->wn;
  o.nwin = nwin;
  o.nbeam = nbeam;
  TRY(kspace<T>(P, A, out, b, o, s));
  TRY(col_pass<T>(P, out, b, 4, OP_INV, s));
  return row_pass<T>(P, out, b, 4, OP_INV, s);
}

template <typename T>
static int propagate(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* out, void* ws,
                     int* status, hipStream_t s, int dir, cpx<T>* bulk = nullptr) {
  const int H = P->d.height, W = P->d.width;
  const size_t HW = size_t(H) * W;
  const int nrays = P->d.n_rays, nlam = P->d.n_wavelengths;
  const int nplanes = nrays * nlam;
  const double scale = 1.0 / double(HW);
  const bool m4 = P->d.method == DTMMX_METHOD_4X4;
  const WsLayout wl = ws_layout(P);
  char* wsb = (char*)ws;
  if (!m4 && P->wn > 1) {
    if (P->win_i.empty()) {  // no window passes betamax: every part is zero
      const size_t n = size_t(nplanes) * 4 * HW * 2;
      hipLaunchKernelGGL((zero_kernel<T>), dim3(unsigned((n + NT - 1) / NT)), dim3(NT), 0, s,
                         (T*)out, n);
      HIP_TRY(hipGetLastError());
      return DTMMX_OK;
    }
    Batch b{nrays, nlam, 0, nlam};
    return transfer_windows_2x2<T>(P, S, out, (cpx<T>*)(wsb + wl.e2), (cpx<T>*)(wsb + wl.win_tmp),
                                   (cpx<T>*)(wsb + wl.win_acc), b, status, s, wsb + wl.etab);
  }
  if (m4 && P->wn > 1) {
    if (S->n_layers == 0) return DTMMX_OK;  // ifft2(fft2(field_in))
    if (P->win_i.empty()) {  // no window passes betamax: every part is zero
      const size_t n = size_t(nplanes) * 4 * HW * 2;
      hipLaunchKernelGGL((zero_kernel<T>), dim3(unsigned((n + NT - 1) / NT)), dim3(NT), 0, s,
                         (T*)out, n);
      HIP_TRY(hipGetLastError());
      return DTMMX_OK;
    }
    cpx<T>* tmp = (cpx<T>*)(wsb + wl.win_tmp);
    cpx<T>* acc = (cpx<T>*)(wsb + wl.win_acc);
    Batch b{nrays, nlam, 0, nlam};
    return transfer_windows<T>(P, S, out, tmp, acc, b, status, s, dir, wsb + wl.etab);
  }
  const std::vector<Step> st = expand_steps(S, dir);
  const int L = int(st.size());
  // workspace: 2x2 method: the 2-component field
  cpx<T>* e2 = (cpx<T>*)(wsb + wl.e2);
  // with bulk output the 2x2 method runs all wavelengths as one batch: its
  // bulk step uses `out` as scratch once the input projection of every plane
  // has read it
  const int nlb = (bulk && !m4) ? nlam : lam_batch(P, size_t(nrays) * 4 * HW * sizeof(cpx<T>));
  for (int lam0 = 0; lam0 < nlam; lam0 += nlb) {
    Batch b{nrays, nlam, lam0, (lam0 + nlb <= nlam) ? nlb : nlam - lam0};
    if (m4 && S->eff_eps_re) {
      // anisotropic effective media: dense boundary operators on the full
      // spectrum between the row passes (dtmm/propagate_4x4.py:376-388)
      if (L == 0) break;
      TRY(row_pass<T>(P, out, b, 4, OP_FWD, s));
      TRY(col_pass<T>(P, out, b, 4, OP_FWD, s));
      TRY(gkspace<T>(P, out, b, GOP_HEAD, nullptr, &st[0].g, s));
      TRY(col_pass<T>(P, out, b, 4, OP_INV, s));
      for (int l = 0; l < L; ++l) {
        const RowArgs<T> la = layer_args<T>(S, HW, st[l].l, st[l].d, 1, status, wsb + wl.etab);
        TRY(row_pass<T>(P, out, b, 4, OP_ROP, s, &la));
        if (bulk && (l + 1 == L || st[l + 1].l != st[l].l)) {
          cpx<T>* slot = bulk + size_t(st[l].l) * nplanes * 4 * HW;
          TRY(copy_planes<T>(slot, out, b, HW, s));
          TRY(col_pass<T>(P, slot, b, 4, OP_FWD, s));
          TRY(gkspace<T>(P, slot, b, GOP_TAIL, &st[l].g, nullptr, s));
          TRY(col_pass<T>(P, slot, b, 4, OP_INV, s));
          TRY(row_pass<T>(P, slot, b, 4, OP_INV, s));
        }
        TRY(col_pass<T>(P, out, b, 4, OP_FWD, s));
        if (l + 1 < L)
          TRY(gkspace<T>(P, out, b, GOP_BOUND, &st[l].g, &st[l + 1].g, s));
        else
          TRY(gkspace<T>(P, out, b, GOP_TAIL, &st[l].g, nullptr, s));
        TRY(col_pass<T>(P, out, b, 4, OP_INV, s));
      }
      TRY(row_pass<T>(P, out, b, 4, OP_INV, s));
      continue;
    }
    if (m4) {
      if (L == 0) break;  // ifft2(fft2(field_in))
      TRY(row_pass<T>(P, out, b, 4, OP_FWD, s));
      TRY(col_pass<T>(P, out, b, 4, OP_FWD, s));
      KArgs<T> k{};
      k.kop = KOP_HEAD4;
      k.d0 = st[0].d;
      k.e0 = st[0].e;
      k.scale = scale;
      TRY(kspace<T>(P, out, out, b, k, s));
      TRY(col_pass<T>(P, out, b, 4, OP_INV, s));
      for (int l = 0; l < L; ++l) {
        const RowArgs<T> la = layer_args<T>(S, HW, st[l].l, st[l].d, 1, status, wsb + wl.etab);
        TRY(row_pass<T>(P, out, b, 4, OP_ROP, s, &la));
        if (bulk && (l + 1 == L || st[l + 1].l != st[l].l)) {
          // field after layer st[l].l in the (x, y) field basis: the tail of
          // this step applied to a copy (dtmm/transfer.py:777-803 with
          // bulk_out, :876-877)
          cpx<T>* slot = bulk + size_t(st[l].l) * nplanes * 4 * HW;
          TRY(copy_planes<T>(slot, out, b, HW, s));
          TRY(col_pass<T>(P, slot, b, 4, OP_FWD, s));
          KArgs<T> t{};
          t.kop = KOP_TAIL4;
          t.d0 = st[l].d;
          t.e0 = st[l].e;
          t.scale = scale;
          TRY(kspace<T>(P, slot, slot, b, t, s));
          TRY(col_pass<T>(P, slot, b, 4, OP_INV, s));
          TRY(row_pass<T>(P, slot, b, 4, OP_INV, s));
        }
        if (l + 1 < L) {
          cd e0 = st[l].e, e1 = st[l + 1].e;
          if (eq(e0, e1)) {
            ColArgs<T> ca{};
            ca.da = st[l].d;
            ca.db = st[l + 1].d;
            ca.ea = e0;
            ca.eb = e1;
            ca.kca = 1.;
            ca.kcb = 1.;
            ca.scale = scale;
            ca.dense_sign = 1;
            TRY(col_pass<T>(P, out, b, 4, OP_DIAG, s, &ca));
          } else {
            TRY(col_pass<T>(P, out, b, 4, OP_FWD, s));
            KArgs<T> bd{};
            bd.kop = KOP_BOUND4;
            bd.d0 = st[l].d;
            bd.d1 = st[l + 1].d;
            bd.e0 = e0;
            bd.e1 = e1;
            bd.scale = scale;
            TRY(kspace<T>(P, out, out, b, bd, s));
            TRY(col_pass<T>(P, out, b, 4, OP_INV, s));
          }
        }
      }
      TRY(col_pass<T>(P, out, b, 4, OP_FWD, s));
      KArgs<T> t{};
      t.kop = KOP_TAIL4;
      t.d0 = st[L - 1].d;
      t.e0 = st[L - 1].e;
      t.scale = scale;
      TRY(kspace<T>(P, out, out, b, t, s));
      TRY(col_pass<T>(P, out, b, 4, OP_INV, s));
      TRY(row_pass<T>(P, out, b, 4, OP_INV, s));
      continue;
    }
    // 2x2: out doubles as the 4-component scratch of the input projection.
    TRY(row_pass<T>(P, out, b, 4, OP_FWD, s));
    TRY(col_pass<T>(P, out, b, 4, OP_FWD, s));
    KArgs<T> h{};
    h.kop = KOP_PROJ2;
    h.d0 = L > 0 ? st[0].d : 0.;
    h.e0 = L > 0 ? st[0].e : mk<double>(S->n_out * S->n_out);
    h.ein = mk<double>(S->n_in * S->n_in);
    h.scale = scale;
    h.fresnel = P->d.reflection;
    TRY(kspace<T>(P, out, e2, b, h, s));
    TRY(col_pass<T>(P, e2, b, 2, OP_INV, s));
    const cd eout = mk<double>(S->n_out * S->n_out);
    for (int l = 0; l < L; ++l) {
      const RowArgs<T> la = layer_args<T>(S, HW, st[l].l, st[l].d, 0, status, wsb + wl.etab);
      TRY(row_pass<T>(P, e2, b, 2, OP_ROP, s, &la));
      if (bulk && (l + 1 == L || st[l + 1].l != st[l].l)) {
        // field after layer st[l].l (dtmm/propagate_2x2.py:80-85): E after the
        // layer's second half step, in real space, and H of the forward modes
        // of each voxel; `out` is free during the layer loop
        cpx<T>* tmp = out;
        TRY(copy_planes<T>(tmp, e2, b, HW, s, 2));
        ColArgs<T> cb{};
        cb.da = st[l].d;
        cb.ea = st[l].e;
        cb.db = 0.;
        cb.eb = st[l].e;
        cb.kca = 1.;
        cb.kcb = 1.;
        cb.scale = scale;
        cb.dense_sign = 0;
        TRY(col_pass<T>(P, tmp, b, 2, OP_DIAG, s, &cb));
        TRY(row_pass<T>(P, tmp, b, 2, OP_INV, s));
        PixArgs<T> pa{};
        pa.src = tmp;
        pa.dst = bulk + size_t(st[l].l) * nplanes * 4 * HW;
        pa.HW = HW;
        pa.nlam = nlam;
        pa.lam0 = b.lam0;
        pa.nlam_b = b.nlam_b;
        pa.eps_re = S->eps_re + size_t(st[l].l) * HW;
        pa.eps_im = S->eps_im ? S->eps_im + size_t(st[l].l) * HW : nullptr;
        pa.ang = S->angles + size_t(st[l].l) * HW;
        pa.comp_stride = size_t(S->n_layers) * HW;
        pa.bm = P->bm;
        hipLaunchKernelGGL((pixel_kernel<T, PIX_E2H>),
                           dim3(unsigned((HW + NT - 1) / NT), unsigned(nrays / P->d.rays_per_beam)),
                           dim3(NT), 0, s, pa);
        HIP_TRY(hipGetLastError());
      }
      const cd enext = (l + 1 < L) ? st[l + 1].e : eout;
      if (P->d.reflection && !eq(st[l].e, enext)) {
        // reflection mode 1: Fresnel transmission at the interface between
        // different effective media (propagate_2x2.py:28-30, :62-65)
        TRY(col_pass<T>(P, e2, b, 2, OP_FWD, s));
        KArgs<T> tr{};
        tr.kop = KOP_TR2;
        tr.d0 = st[l].d;
        tr.e0 = st[l].e;
        tr.d1 = (l + 1 < L) ? st[l + 1].d : 0.;
        tr.e1 = enext;
        tr.scale = scale;
        TRY(kspace<T>(P, e2, e2, b, tr, s));
        TRY(col_pass<T>(P, e2, b, 2, OP_INV, s));
        continue;
      }
      ColArgs<T> ca{};
      ca.da = st[l].d;
      ca.ea = st[l].e;
      ca.kca = 1.;
      ca.kcb = 1.;
      if (l + 1 < L) {
        ca.db = st[l + 1].d;
        ca.eb = st[l + 1].e;
      } else {
        ca.db = 0.;  // output medium: zero thickness, only the betamax mask
        ca.eb = mk<double>(S->n_out * S->n_out);
      }
      ca.scale = scale;
      ca.dense_sign = 0;
      TRY(col_pass<T>(P, e2, b, 2, OP_DIAG, s, &ca));
    }
    TRY(row_pass<T>(P, e2, b, 2, OP_INV, s));
    E2HArgs<T> ea{};
    ea.src = e2;
    ea.dst = out;
    ea.H = H;
    ea.W = W;
    ea.b = b;
    ea.eout = mk<double>(S->n_out * S->n_out);
    ea.bm = P->bm;
    dim3 grid((unsigned)((HW + NT - 1) / NT), nrays * b.nlam_b);
    hipLaunchKernelGGL((e2h_kernel<T>), grid, dim3(NT), 0, s, ea);
    HIP_TRY(hipGetLastError());
  }
  return DTMMX_OK;
}

template <typename T>
static int fft2_planes(const dtmmx_plan* P, cpx<T>* f, int inverse, hipStream_t s) {
  Batch b{P->d.n_rays, P->d.n_wavelengths, 0, P->d.n_wavelengths};
  if (!inverse) {
    TRY(row_pass<T>(P, f, b, 4, OP_FWD, s));
    return col_pass<T>(P, f, b, 4, OP_FWD, s);
  }
  TRY(col_pass<T>(P, f, b, 4, OP_INV, s));
  return row_pass<T>(P, f, b, 4, OP_INV, s);
}

// k-space normalisation step on the spectrum of a whole field (in place).
template <typename T>
static int pass_op(const dtmmx_plan* P, cpx<T>* f, PArgs<T> a, hipStream_t s) {
  a.src = f;
  a.dst = f;
  a.H = P->d.height;
  a.W = P->d.width;
  a.nlam = P->d.n_wavelengths;
  a.amax = P->amax;
  a.bm = P->bm;
  const size_t HW = size_t(a.H) * a.W;
  dim3 grid(unsigned((HW + NT - 1) / NT), unsigned(P->d.n_rays * P->d.n_wavelengths));
  hipLaunchKernelGGL((pass_kernel<T>), grid, dim3(NT), 0, s, a);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

static int sum_rows(const double* part, int rows, int nb, double scale, double* sums,
                    hipStream_t s) {
  hipLaunchKernelGGL(sum_rows_kernel, dim3(rows), dim3(NT), 0, s, part, nb, scale, sums);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

// The reference field of DTMMX_NORM_REF through one forward pass
// (dtmm/transfer.py:769-775, :805-806): its E components propagated with
// the 2x2 method without reflections in real space (propagate_2x2.py
// :284-305 with :91-190: per sub-step a half step e P e^-1 of each voxel's
// forward modes, the corrected E diffraction of the layer's effective
// medium, a second half step), then H of the forward modes of the output
// medium from E on the FFT grid (jones2H, zero above betamax).
template <typename T>
static int ref_forward(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* refb, cpx<T>* e2,
                       hipStream_t s) {
  const size_t HW = size_t(P->d.height) * P->d.width;
  const int nlam = P->d.n_wavelengths;
  const size_t nplanes = size_t(P->d.n_rays) * nlam;
  const int nbeam = P->d.n_rays / P->d.rays_per_beam;
  const Batch b{P->d.n_rays, nlam, 0, nlam};
  const double scale = 1.0 / double(HW);
  const size_t cb = sizeof(cpx<T>);
  HIP_TRY(hipMemcpy2DAsync(e2, HW * cb, refb, 2 * HW * cb, HW * cb, nplanes * 2,
                           hipMemcpyDeviceToDevice, s));
  PixArgs<T> a{};
  a.field = e2;
  a.HW = HW;
  a.nlam = nlam;
  a.bm = P->bm;
  a.comp_stride = size_t(S->n_layers) * HW;
  a.mode = 1;
  const dim3 grid(unsigned((HW + NT - 1) / NT), unsigned(nbeam));
  for (int l = 0; l < S->n_layers; ++l) {
    const int ns = S->n_steps ? S->n_steps[l] : 1;
    const double ds = S->thickness[l] / double(ns);
    const cd ee = mk<double>(S->eps_eff_re[l], S->eps_eff_im ? S->eps_eff_im[l] : 0.);
    a.eps_re = S->eps_re + size_t(l) * HW;
    a.eps_im = S->eps_im ? S->eps_im + size_t(l) * HW : nullptr;
    a.ang = S->angles + size_t(l) * HW;
    a.d = 0.5 * ds;
    for (int q = 0; q < ns; ++q) {
      if (ds != 0.) {
        hipLaunchKernelGGL((pixel_kernel<T, PIX_LAYER2>), grid, dim3(NT), 0, s, a);
        HIP_TRY(hipGetLastError());
      }
      if (P->d.diffraction) {
        TRY(row_pass<T>(P, e2, b, 2, OP_FWD, s));
        ColArgs<T> ca{};
        ca.da = ds;
        ca.db = ds;
        ca.ea = ee;
        ca.eb = ee;
        ca.kca = 1.;
        ca.kcb = 1.;
        ca.scale = scale;
        ca.dense_sign = 0;
        TRY(col_pass<T>(P, e2, b, 2, OP_DIAG, s, &ca));
        TRY(row_pass<T>(P, e2, b, 2, OP_INV, s));
      }
      if (ds != 0.) {
        hipLaunchKernelGGL((pixel_kernel<T, PIX_LAYER2>), grid, dim3(NT), 0, s, a);
        HIP_TRY(hipGetLastError());
      }
    }
  }
  HIP_TRY(hipMemcpy2DAsync(refb, 2 * HW * cb, e2, HW * cb, HW * cb, nplanes * 2,
                           hipMemcpyDeviceToDevice, s));
  TRY(fft2_planes<T>(P, refb, 0, s));
  KArgs<T> k{};
  k.kop = KOP_E2H4;
  k.e0 = mk<double>(S->n_out * S->n_out);
  k.scale = scale;
  TRY(kspace<T>(P, refb, refb, b, k, s));
  return fft2_planes<T>(P, refb, 1, s);
}

// npass > 1 (dtmm/transfer.py:697-872): forward and backward passes with the output normalised onto the
// forward modes of the output medium and the input-side field rescaled to
// the incoming intensity between passes.
// One pass of 4x4 reflection mode 2 in place (defined below).
template <typename T>
static int refl2_pass(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* fout, hipStream_t s,
                      int dir, cpx<T>* bulk = nullptr, void* ws = nullptr);

template <typename T>
static int transfer_npass(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* fin, cpx<T>* fout,
                          void* ws, int* status, hipStream_t s) {
  const size_t HW = size_t(P->d.height) * P->d.width;
  const int nplanes = P->d.n_rays * P->d.n_wavelengths;
  const size_t n = size_t(nplanes) * 4 * HW;
  const size_t bytes = n * sizeof(cpx<T>);
  const double scale = 1.0 / double(HW);
  const WsLayout wl = ws_layout(P);
  char* wsb = (char*)ws;
  cpx<T>* g = (cpx<T>*)(wsb + wl.pass_field);
  cpx<T>* f0 = (cpx<T>*)(wsb + wl.pass_field0);
  double* part = (double*)(wsb + wl.part);
  double* sums = (double*)(wsb + wl.sums);  // [i0 | unused | i0f or i1 | i2]
  double* i0 = sums;
  double* i0f = sums + 2 * nplanes;
  const int npass = P->d.npass;
  const double smooth = P->d.smooth < 0. ? 0. : P->d.smooth;
  const cd ein = mk<double>(S->n_in * S->n_in), eout = mk<double>(S->n_out * S->n_out);
  const unsigned eb = unsigned((n + NT - 1) / NT);
  // DTMMX_NORM_REF: the reference starts as the input field
  // (transfer.py:710-713)
  const bool cref = (P->d.norm & DTMMX_NORM_REF) != 0;
  cpx<T>* refb = cref ? (cpx<T>*)(wsb + wl.refb) : nullptr;
  cpx<T>* e2r = cref ? (cpx<T>*)(wsb + wl.e2) : nullptr;
  if (cref) HIP_TRY(hipMemcpyAsync(refb, fin, bytes, hipMemcpyDeviceToDevice, s));
  // field0 = transmitted_field(field_in, nin); i0 = its total intensity
  HIP_TRY(hipMemcpyAsync(f0, fin, bytes, hipMemcpyDeviceToDevice, s));
  TRY(fft2_planes<T>(P, f0, 0, s));
  {
    PArgs<T> a{};
    a.pop = POP_PROJ;
    a.eps = ein;
    a.scale = scale;
    a.part = part;
    TRY(pass_op<T>(P, f0, a, s));
    TRY(sum_rows(part, nplanes, wl.nblk, scale, i0, s));
  }
  TRY(fft2_planes<T>(P, f0, 1, s));
  HIP_TRY(hipMemcpyAsync(g, fin, bytes, hipMemcpyDeviceToDevice, s));
  // reflection mode 2 propagates in real space (work_in_fft False,
  // dtmm/transfer.py:719-725); the normalisations between passes are the same
  const bool r2 = P->d.method == DTMMX_METHOD_4X4 && P->d.reflection == 2;
  // the reference's amplitude factor before each normalisation
  // (transfer.py:822-823)
  const double ref_scale = std::sqrt(S->n_in / S->n_out);
  for (int i = 0; i < npass; ++i) {
    const int dir = (i % 2 == 0) ? 1 : -1;
    // the reference is read only by the normalisation after this pass
    if (cref && dir > 0 && i != npass - 1) TRY(ref_forward<T>(P, S, refb, e2r, s));
    if (r2)
      TRY(refl2_pass<T>(P, S, g, s, dir, nullptr, ws));
    else
      TRY(propagate<T>(P, S, g, ws, status, s, dir));
    const double sigma = smooth * double(npass - i) / double(npass);
    if (i % 2 == 0) {
      if (i != npass - 1) {
        TRY(fft2_planes<T>(P, g, 0, s));
        if (P->d.norm & DTMMX_NORM_FFT) {
          PArgs<T> a{};
          a.pop = POP_NORMFFT;
          a.eps = eout;
          a.notch = 1;
          a.x0 = S->n_in;
          a.sigma = sigma;
          a.scale = scale;
          if (cref) {
            TRY(fft2_planes<T>(P, refb, 0, s));
            a.ref = refb;
            a.ref_scale = ref_scale;
          }
          TRY(pass_op<T>(P, g, a, s));
        } else if (cref) {
          // "local": forward part in the output medium, per-pixel
          // rescaling against the reference, then the notch filter
          // (transfer.py:103-138, denoise.py:24-28)
          PArgs<T> a{};
          a.pop = POP_PROJ;
          a.eps = eout;
          a.scale = scale;
          TRY(pass_op<T>(P, g, a, s));
          TRY(fft2_planes<T>(P, g, 1, s));
          LocalArgs<T> la{};
          la.f = g;
          la.ref = refb;
          la.ref_scale = ref_scale;
          la.HW = HW;
          la.eps = eout;
          hipLaunchKernelGGL((local_norm_kernel<T>), dim3(unsigned((HW + NT - 1) / NT), nplanes),
                             dim3(NT), 0, s, la);
          HIP_TRY(hipGetLastError());
          TRY(fft2_planes<T>(P, g, 0, s));
          PArgs<T> d{};
          d.pop = POP_NOTCH;
          d.notch = 1;
          d.x0 = S->n_in;
          d.sigma = sigma;
          d.scale = scale;
          TRY(pass_op<T>(P, g, d, s));
        } else {
          PArgs<T> a{};
          a.pop = POP_TOTALSUM;
          a.eps = eout;
          a.part = part;
          TRY(pass_op<T>(P, g, a, s));
          TRY(sum_rows(part, 2 * nplanes, wl.nblk, scale, i0f, s));
          a.pop = POP_TOTALAPPLY;
          a.part = nullptr;
          a.sums = i0f;
          a.notch = 1;
          a.x0 = S->n_in;
          a.sigma = sigma;
          a.scale = scale;
          TRY(pass_op<T>(P, g, a, s));
        }
        TRY(fft2_planes<T>(P, g, 1, s));
      }
      hipLaunchKernelGGL((accumulate_kernel<T>), dim3(eb), dim3(NT), 0, s, g, fout,
                         i == 0 ? 1 : 0, n);
      HIP_TRY(hipGetLastError());
    } else {
      HIP_TRY(hipMemcpyAsync(fin, g, bytes, hipMemcpyDeviceToDevice, s));
      if (i != npass - 1) {
        TRY(fft2_planes<T>(P, g, 0, s));
        PArgs<T> a{};
        a.pop = POP_PROJ;
        a.eps = ein;
        a.notch = 1;
        a.x0 = S->n_in;
        a.sigma = sigma;
        a.scale = scale;
        a.part = part;
        TRY(pass_op<T>(P, g, a, s));
        TRY(sum_rows(part, nplanes, wl.nblk, scale, i0f, s));
        TRY(fft2_planes<T>(P, g, 1, s));
        hipLaunchKernelGGL((reflect_update_kernel<T>), dim3(eb), dim3(NT), 0, s, g, fout, fin,
                           (const cpx<T>*)f0, (const double*)i0, (const double*)i0f, 4 * HW, n);
        HIP_TRY(hipGetLastError());
        // the next forward pass starts its reference from this field
        // (transfer.py:867-868)
        if (cref) HIP_TRY(hipMemcpyAsync(refb, g, bytes, hipMemcpyDeviceToDevice, s));
      }
    }
  }
  return DTMMX_OK;
}

// diffraction 0 (4x4 reflection mode 2, 2x2 reflection mode 0, one pass):
// every pixel is an independent stack, the field stays in real space
// (dtmm/transfer.py:694-700 and :1071-1099 with work_in_fft False).  The
// 4x4 method propagates the input as it is, skipping the input and output
// media; the 2x2 method keeps the forward part of the input in the input
// medium, propagates E and builds H in the output medium.
template <typename T>
static int transfer_nodiff(const dtmmx_plan* P, const dtmmx_stack* S, const cpx<T>* fin,
                           cpx<T>* fout, void* ws, hipStream_t s, cpx<T>* bulk = nullptr) {
  const size_t HW = size_t(P->d.height) * P->d.width;
  const bool m4 = P->d.method == DTMMX_METHOD_4X4;
  const int nbeam = P->d.n_rays / P->d.rays_per_beam;
  PixArgs<T> a{};
  a.HW = HW;
  a.nlam = P->d.n_wavelengths;
  a.bm = P->bm;
  const dim3 grid(unsigned((HW + NT - 1) / NT), unsigned(nbeam));
  cpx<T>* f = m4 ? fout : (cpx<T>*)((char*)ws + ws_layout(P).e2);
  if (m4) {
    const size_t n = size_t(P->d.n_rays) * P->d.n_wavelengths * 4 * HW;
    HIP_TRY(hipMemcpyAsync(fout, fin, n * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
  } else {
    a.src = fin;
    a.field = f;
    a.emed = mk<double>(S->n_in * S->n_in);
    hipLaunchKernelGGL((pixel_kernel<T, PIX_IN2>), grid, dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
  }
  a.field = f;
  a.comp_stride = size_t(S->n_layers) * HW;
  for (int l = 0; l < S->n_layers; ++l) {
    a.eps_re = S->eps_re + size_t(l) * HW;
    a.eps_im = S->eps_im ? S->eps_im + size_t(l) * HW : nullptr;
    a.ang = S->angles + size_t(l) * HW;
    a.d = S->thickness[l];
    if (m4)
      hipLaunchKernelGGL((pixel_kernel<T, PIX_LAYER4>), grid, dim3(NT), 0, s, a);
    else
      hipLaunchKernelGGL((pixel_kernel<T, PIX_LAYER2>), grid, dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
    if (bulk && m4) {
      const size_t n = size_t(P->d.n_rays) * P->d.n_wavelengths * 4 * HW;
      HIP_TRY(hipMemcpyAsync(bulk + size_t(l) * n, fout, n * sizeof(cpx<T>),
                             hipMemcpyDeviceToDevice, s));
    } else if (bulk) {
      // 2x2: E after the layer and H of each voxel's forward modes
      // (dtmm/propagate_2x2.py:185-188)
      PixArgs<T> pa = a;
      pa.src = f;
      pa.dst = bulk + size_t(l) * P->d.n_rays * P->d.n_wavelengths * 4 * HW;
      hipLaunchKernelGGL((pixel_kernel<T, PIX_E2H>), grid, dim3(NT), 0, s, pa);
      HIP_TRY(hipGetLastError());
    }
  }
  if (!m4) {
    a.src = f;
    a.dst = fout;
    a.emed = mk<double>(S->n_out * S->n_out);
    hipLaunchKernelGGL((pixel_kernel<T, PIX_OUT2>), grid, dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
  }
  return DTMMX_OK;
}

// 4x4 reflection mode 2 with diffraction 1, one pass (dtmm/transfer.py
// :694-700 and :782-795 with work_in_fft False; propagate_4x4.py:156-175 and
// :18-51): the field stays in real space; each sub-step is e P/2 F^-1 per
// pixel, a 2D transform, the diagonal k-space operator
// exp(+-i k0 d (a(k) - a_c)) of corrected_Epn_diffraction_matrix
// (matrix.py:38-43, :62-68) on the forward / backward E parts, the inverse
// transform, and F P/2 e^-1 per pixel.
// One pass of 4x4 reflection mode 2 in place on the real-space field f: the
// layers in the order of dir with thickness times dir (dtmm/transfer.py
// :752-760, :782-785).  Diffraction 0: F P F^-1 per pixel and layer;
// diffraction 1: the sub-steps described above.
template <typename T>
static int refl2_steps(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* fout, hipStream_t s,
                       const std::vector<Step>& st, const Beams& bm, cpx<T>* bulk);

// diffraction > 1: per layer, the spectrum of the field is split by the tent
// windows; each part goes back to real space, through all of the layer's
// sub-steps with its window's mean beams, and the parts are summed
// (dtmm/propagate_4x4.py:176-204).  ws: the window-row region of the
// workspace (two fields) and the window sum.
template <typename T>
static int refl2_windows(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* f, hipStream_t s,
                         int dir, void* ws) {
  const int H = P->d.height, W = P->d.width;
  const size_t HW = size_t(H) * W;
  const int nbeam = P->d.n_rays / P->d.rays_per_beam;
  const int nlam = P->d.n_wavelengths;
  const size_t n = size_t(P->d.n_rays) * nlam * 4 * HW;
  if (size_t(P->d.n_rays) * nlam * 4 > 65535) return fail(DTMMX_E_ARG, "too many planes");
  const WsLayout wl = ws_layout(P);
  cpx<T>* fk = (cpx<T>*)((char*)ws + wl.win_tmp);
  cpx<T>* part = fk + n;
  cpx<T>* acc = (cpx<T>*)((char*)ws + wl.win_acc);
  const std::vector<Step> all = expand_steps(S, dir);
  const unsigned eb = unsigned((n + NT - 1) / NT);
  const dim3 pg(unsigned((HW + NT - 1) / NT), unsigned(size_t(P->d.n_rays) * nlam * 4));
  for (size_t k0 = 0; k0 < all.size();) {
    size_t k1 = k0;
    while (k1 < all.size() && all[k1].l == all[k0].l) ++k1;
    const std::vector<Step> st(all.begin() + k0, all.begin() + k1);
    HIP_TRY(hipMemcpyAsync(fk, f, n * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
    TRY(fft2_planes<T>(P, fk, 0, s));
    const int nwin = int(P->win_i.size());
    for (int w = 0; w < nwin; ++w) {
      Beams bw = P->bm;
      for (int q = 0; q < nbeam; ++q) {
        bw.beta[q] = P->win_beta[size_t(w) * nbeam + q];
        bw.phi[q] = P->win_phi[size_t(w) * nbeam + q];
      }
      hipLaunchKernelGGL((win_part_kernel<T>), pg, dim3(NT), 0, s, (const cpx<T>*)fk, part, H, W,
                         nlam, P->d.rays_per_beam, P->wn, P->win_i[w], P->win_j[w],
                         (const double*)P->wcx, (const double*)P->wcy, (const int*)P->wmask,
                         1.0 / double(HW));
      HIP_TRY(hipGetLastError());
      TRY(fft2_planes<T>(P, part, 1, s));
      TRY(refl2_steps<T>(P, S, part, s, st, bw, nullptr));
      hipLaunchKernelGGL((accumulate_kernel<T>), dim3(eb), dim3(NT), 0, s, part, acc,
                         w == 0 ? 1 : 0, n);
      HIP_TRY(hipGetLastError());
    }
    if (nwin == 0) {  // no window passes betamax: every part is zero
      hipLaunchKernelGGL((zero_kernel<T>), dim3(unsigned((2 * n + NT - 1) / NT)), dim3(NT), 0, s,
                         (T*)f, 2 * n);
      HIP_TRY(hipGetLastError());
    } else {
      HIP_TRY(hipMemcpyAsync(f, acc, n * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
    }
    k0 = k1;
  }
  return DTMMX_OK;
}

template <typename T>
static int refl2_pass(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* fout, hipStream_t s,
                      int dir, cpx<T>* bulk, void* ws) {
  if (P->d.diffraction > 1) {
    if (!ws) return fail(DTMMX_E_ARG, "reflection mode 2 with diffraction > 1 needs a workspace");
    return refl2_windows<T>(P, S, fout, s, dir, ws);
  }
  const size_t HW = size_t(P->d.height) * P->d.width;
  const int nbeam = P->d.n_rays / P->d.rays_per_beam;
  const size_t n = size_t(P->d.n_rays) * P->d.n_wavelengths * 4 * HW;
  PixArgs<T> a{};
  a.field = fout;
  a.HW = HW;
  a.nlam = P->d.n_wavelengths;
  a.bm = P->bm;
  a.comp_stride = size_t(S->n_layers) * HW;
  const dim3 grid(unsigned((HW + NT - 1) / NT), unsigned(nbeam));
  if (P->d.diffraction == 0) {
    const int L = S->n_layers;
    for (int k = 0; k < L; ++k) {
      const int l = dir > 0 ? k : L - 1 - k;
      a.eps_re = S->eps_re + size_t(l) * HW;
      a.eps_im = S->eps_im ? S->eps_im + size_t(l) * HW : nullptr;
      a.ang = S->angles + size_t(l) * HW;
      a.d = S->thickness[l] * dir;
      hipLaunchKernelGGL((pixel_kernel<T, PIX_LAYER4>), grid, dim3(NT), 0, s, a);
      HIP_TRY(hipGetLastError());
    }
    return DTMMX_OK;
  }
  return refl2_steps<T>(P, S, fout, s, expand_steps(S, dir), P->bm, bulk);
}

// The diffraction-1 sub-steps of reflection mode 2 in place on the
// real-space field, for the beams bm (the plan's, or a window's mean beams).
template <typename T>
static int refl2_steps(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* fout, hipStream_t s,
                       const std::vector<Step>& st, const Beams& bm, cpx<T>* bulk) {
  const size_t HW = size_t(P->d.height) * P->d.width;
  const int nbeam = P->d.n_rays / P->d.rays_per_beam;
  const size_t n = size_t(P->d.n_rays) * P->d.n_wavelengths * 4 * HW;
  PixArgs<T> a{};
  a.field = fout;
  a.HW = HW;
  a.nlam = P->d.n_wavelengths;
  a.bm = bm;
  a.comp_stride = size_t(S->n_layers) * HW;
  const dim3 grid(unsigned((HW + NT - 1) / NT), unsigned(nbeam));
  const Batch b{P->d.n_rays, P->d.n_wavelengths, 0, P->d.n_wavelengths};
  for (size_t k = 0; k < st.size(); ++k) {
    const Step& q = st[k];
    a.eps_re = S->eps_re + size_t(q.l) * HW;
    a.eps_im = S->eps_im ? S->eps_im + size_t(q.l) * HW : nullptr;
    a.ang = S->angles + size_t(q.l) * HW;
    a.d = 0.5 * q.d;
    hipLaunchKernelGGL((pixel_kernel<T, PIX_A4>), grid, dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
    TRY(row_pass<T>(P, fout, b, 4, OP_FWD, s));
    ColArgs<T> ca{};
    ca.da = q.d;
    ca.db = q.d;
    ca.ea = q.e;
    ca.eb = q.e;
    ca.kca = 1.;
    ca.kcb = 1.;
    ca.scale = 1.0 / double(HW);
    ca.dense_sign = 1;
    TRY(col_pass<T>(P, fout, b, 4, OP_DIAG, s, &ca, &bm));
    TRY(row_pass<T>(P, fout, b, 4, OP_INV, s));
    hipLaunchKernelGGL((pixel_kernel<T, PIX_B4>), grid, dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
    if (bulk && (k + 1 == st.size() || st[k + 1].l != q.l))
      HIP_TRY(hipMemcpyAsync(bulk + size_t(q.l) * n, fout, n * sizeof(cpx<T>),
                             hipMemcpyDeviceToDevice, s));
  }
  return DTMMX_OK;
}

template <typename T>
static int transfer_refl2(const dtmmx_plan* P, const dtmmx_stack* S, const cpx<T>* fin,
                          cpx<T>* fout, void* ws, hipStream_t s, cpx<T>* bulk = nullptr) {
  const size_t n = size_t(P->d.n_rays) * P->d.n_wavelengths * 4 * size_t(P->d.height) *
                   P->d.width;
  HIP_TRY(hipMemcpyAsync(fout, fin, n * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
  return refl2_pass<T>(P, S, fout, s, 1, bulk, ws);
}

// 2x2 reflection mode 2, one pass, diffraction 0 or 1 (dtmm/transfer.py
// :1071-1099 with work_in_fft False, propagate_2x2.py:91-190 with refl
// None): E in real space; each layer is entered through the per-pixel
// interface transmission from the previous one (PIX_T2, first half step),
// then the corrected E diffraction of the layer's effective medium, then a
// plain half step (PIX_LAYER2 at half thickness); the output medium is
// entered the same way and builds H.
template <typename T>
static int transfer_2x2_refl2(const dtmmx_plan* P, const dtmmx_stack* S, const cpx<T>* fin,
                              cpx<T>* fout, void* ws, hipStream_t s, cpx<T>* bulk) {
  const size_t HW = size_t(P->d.height) * P->d.width;
  const int nbeam = P->d.n_rays / P->d.rays_per_beam;
  const int nlam = P->d.n_wavelengths;
  const size_t nplanes = size_t(P->d.n_rays) * nlam;
  const double scale = 1.0 / double(HW);
  const bool diff = P->d.diffraction != 0;
  const Batch b{P->d.n_rays, nlam, 0, nlam};
  cpx<T>* e2 = (cpx<T>*)((char*)ws + ws_layout(P).e2);
  const dim3 grid(unsigned((HW + NT - 1) / NT), unsigned(nbeam));
  const cd ein = mk<double>(S->n_in * S->n_in), eout = mk<double>(S->n_out * S->n_out);
  PixArgs<T> a{};
  a.HW = HW;
  a.nlam = nlam;
  a.bm = P->bm;
  a.comp_stride = size_t(S->n_layers) * HW;
  // forward part of the input in the input medium (transfer.py:1071-1075)
  if (diff) {
    HIP_TRY(hipMemcpyAsync(fout, fin, nplanes * 4 * HW * sizeof(cpx<T>), hipMemcpyDeviceToDevice,
                           s));
    TRY(row_pass<T>(P, fout, b, 4, OP_FWD, s));
    TRY(col_pass<T>(P, fout, b, 4, OP_FWD, s));
    KArgs<T> h{};
    h.kop = KOP_PROJ2;
    h.d0 = 0.;
    h.e0 = ein;
    h.ein = ein;
    h.scale = scale;
    h.fresnel = 0;
    TRY(kspace<T>(P, fout, e2, b, h, s));
    TRY(col_pass<T>(P, e2, b, 2, OP_INV, s));
    TRY(row_pass<T>(P, e2, b, 2, OP_INV, s));
  } else {
    a.src = fin;
    a.field = e2;
    a.emed = ein;
    hipLaunchKernelGGL((pixel_kernel<T, PIX_IN2>), grid, dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
  }
  a.field = e2;
  // E diffraction of one layer's effective medium (corrected, full thickness d)
  auto diffract2 = [&](double d, cd e) -> int {
    TRY(row_pass<T>(P, e2, b, 2, OP_FWD, s));
    ColArgs<T> ca{};
    ca.da = d;
    ca.db = d;
    ca.ea = e;
    ca.eb = e;
    ca.kca = 1.;
    ca.kcb = 1.;
    ca.scale = scale;
    ca.dense_sign = 0;
    TRY(col_pass<T>(P, e2, b, 2, OP_DIAG, s, &ca));
    return row_pass<T>(P, e2, b, 2, OP_INV, s);
  };
  const int L = S->n_layers;
  for (int l = 0; l <= L; ++l) {
    const bool out_medium = l == L;
    a.prev_iso = l == 0;
    a.pemed = ein;
    if (l > 0) {
      a.peps_re = S->eps_re + size_t(l - 1) * HW;
      a.peps_im = S->eps_im ? S->eps_im + size_t(l - 1) * HW : nullptr;
      a.pang = S->angles + size_t(l - 1) * HW;
    }
    a.cur_iso = out_medium;
    a.emed = eout;
    if (!out_medium) {
      a.eps_re = S->eps_re + size_t(l) * HW;
      a.eps_im = S->eps_im ? S->eps_im + size_t(l) * HW : nullptr;
      a.ang = S->angles + size_t(l) * HW;
    }
    const int ns = out_medium ? 1 : (S->n_steps ? S->n_steps[l] : 1);
    const double ds = out_medium ? 0. : S->thickness[l] / double(ns);
    const cd ee = out_medium ? eout
                             : mk<double>(S->eps_eff_re[l], S->eps_eff_im ? S->eps_eff_im[l] : 0.);
    for (int j = 0; j < ns; ++j) {
      a.d = 0.5 * ds;
      if (j == 0) {
        hipLaunchKernelGGL((pixel_kernel<T, PIX_T2>), grid, dim3(NT), 0, s, a);
        HIP_TRY(hipGetLastError());
      } else if (ds != 0.) {
        hipLaunchKernelGGL((pixel_kernel<T, PIX_LAYER2>), grid, dim3(NT), 0, s, a);
        HIP_TRY(hipGetLastError());
      }
      if (diff) TRY(diffract2(ds, ee));
      if (ds != 0.) {
        hipLaunchKernelGGL((pixel_kernel<T, PIX_LAYER2>), grid, dim3(NT), 0, s, a);
        HIP_TRY(hipGetLastError());
      }
    }
    if (bulk && !out_medium) {
      PixArgs<T> pa = a;
      pa.src = e2;
      pa.dst = bulk + size_t(l) * nplanes * 4 * HW;
      hipLaunchKernelGGL((pixel_kernel<T, PIX_E2H>), grid, dim3(NT), 0, s, pa);
      HIP_TRY(hipGetLastError());
    }
  }
  PixArgs<T> po = a;
  po.src = e2;
  po.dst = fout;
  po.emed = eout;
  hipLaunchKernelGGL((pixel_kernel<T, PIX_OUT2>), grid, dim3(NT), 0, s, po);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

// The 2x2 method with npass > 1 (dtmm/transfer.py:1076-1178), reflection
// mode 2 (real space, diffraction 0 or 1, dtmm/propagate_2x2.py:91-190) or
// reflection mode 1 (k space, diffraction 1, :19-87).  The layers are
// indexed as the reference's list: 0 the input medium, 1..L the stack, L + 1
// the output medium; interface j lies between j and j + 1 and keeps the wave
// reflected there (refl[j], workspace after the plan's part) for the next
// pass.  field_in is overwritten with its forward part and receives the
// backward field of every backward pass; field_out sums the forward passes;
// bulk (L slots) sums the field after each layer of the forward passes and
// the field entering it in the backward passes.
template <typename T>
static int transfer_2x2_npass(const dtmmx_plan* P, const dtmmx_stack* S, cpx<T>* fin,
                              cpx<T>* fout, void* ws, int* status, hipStream_t s, cpx<T>* bulk) {
  const size_t HW = size_t(P->d.height) * P->d.width;
  const int nbeam = P->d.n_rays / P->d.rays_per_beam;
  const int nlam = P->d.n_wavelengths;
  const size_t nplanes = size_t(P->d.n_rays) * nlam;
  const size_t n4 = nplanes * 4 * HW, n2 = nplanes * 2 * HW;
  const double scale = 1.0 / double(HW);
  const bool diff = P->d.diffraction != 0;
  const bool r1 = P->d.reflection == 1;
  const Batch b{P->d.n_rays, nlam, 0, nlam};
  const WsLayout wl = ws_layout(P);
  cpx<T>* e2 = (cpx<T>*)((char*)ws + wl.e2);
  cpx<T>* tmp = (cpx<T>*)((char*)ws + wl.tmp2);
  const size_t rstride = (n2 * sizeof(cpx<T>) + 255) / 256 * 256;
  char* reflb = (char*)ws + wl.total;
  auto refl = [&](int j) { return (cpx<T>*)(reflb + size_t(j) * rstride); };
  const int L = S->n_layers;
  const cd ein = mk<double>(S->n_in * S->n_in), eout = mk<double>(S->n_out * S->n_out);
  const dim3 grid(unsigned((HW + NT - 1) / NT), unsigned(nbeam));
  auto zero = [&](void* p, size_t nreal) -> int {
    hipLaunchKernelGGL((zero_kernel<T>), dim3(unsigned((nreal + NT - 1) / NT)), dim3(NT), 0, s,
                       (T*)p, nreal);
    HIP_TRY(hipGetLastError());
    return DTMMX_OK;
  };
  // layer r of the reference's list
  struct Lay {
    bool iso;
    int l, ns;
    double t;  // sub-step thickness
    cd e, eeff;
  };
  auto lay = [&](int r) {
    Lay y{};
    if (r == 0 || r == L + 1) {
      y.iso = true;
      y.l = -1;
      y.ns = 1;
      y.t = 0.;
      y.e = y.eeff = r == 0 ? ein : eout;
    } else {
      y.iso = false;
      y.l = r - 1;
      y.ns = S->n_steps ? S->n_steps[y.l] : 1;
      y.t = S->thickness[y.l] / double(y.ns);
      y.eeff = mk<double>(S->eps_eff_re[y.l], S->eps_eff_im ? S->eps_eff_im[y.l] : 0.);
    }
    return y;
  };
  auto set_cur = [&](PixArgs<T>& a, const Lay& y) {
    a.cur_iso = y.iso;
    a.emed = y.e;
    if (!y.iso) {
      a.eps_re = S->eps_re + size_t(y.l) * HW;
      a.eps_im = S->eps_im ? S->eps_im + size_t(y.l) * HW : nullptr;
      a.ang = S->angles + size_t(y.l) * HW;
    }
  };
  // H of the modes of direction dir from the 2-component real-space field
  // src, added to the 4-component target
  auto add_bulk = [&](const cpx<T>* src, cpx<T>* dst, const Lay& y, int dir) -> int {
    PixArgs<T> pa{};
    pa.HW = HW;
    pa.nlam = nlam;
    pa.bm = P->bm;
    pa.comp_stride = size_t(L) * HW;
    pa.src = src;
    pa.dst = dst;
    pa.mode = dir;
    pa.accum = 1;
    set_cur(pa, y);
    if (y.iso)
      hipLaunchKernelGGL((pixel_kernel<T, PIX_OUT2>), grid, dim3(NT), 0, s, pa);
    else
      hipLaunchKernelGGL((pixel_kernel<T, PIX_E2H>), grid, dim3(NT), 0, s, pa);
    HIP_TRY(hipGetLastError());
    return DTMMX_OK;
  };
  TRY(zero(fout, 2 * n4));
  if (L > 0 && bulk) TRY(zero(bulk, 2 * n4 * size_t(L)));
  for (int j = 0; j <= L; ++j) TRY(zero(refl(j), 2 * n2));
  // forward part of the input in the input medium, left in field_in
  // (transfer.py:1071-1080)
  if (diff) {
    TRY(fft2_planes<T>(P, fin, 0, s));
    PArgs<T> a{};
    a.pop = POP_PROJ;
    a.eps = ein;
    a.scale = scale;
    TRY(pass_op<T>(P, fin, a, s));
    TRY(fft2_planes<T>(P, fin, 1, s));
    // E components (0 and 2 of each plane) to the 2-component field
    HIP_TRY(hipMemcpy2DAsync(e2, HW * sizeof(cpx<T>), fin, 2 * HW * sizeof(cpx<T>),
                             HW * sizeof(cpx<T>), nplanes * 2, hipMemcpyDeviceToDevice, s));
  } else {
    PixArgs<T> a{};
    a.HW = HW;
    a.nlam = nlam;
    a.bm = P->bm;
    a.src = fin;
    a.dst = fin;
    a.field = e2;
    a.emed = ein;
    hipLaunchKernelGGL((pixel_kernel<T, PIX_IN2>), grid, dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
  }
  // target of the bulk terms of layer r: field_in, field_out or a bulk slot
  auto target = [&](int r) -> cpx<T>* {
    if (r == 0) return fin;
    if (r == L + 1) return fout;
    return bulk ? bulk + size_t(r - 1) * n4 : nullptr;
  };
  for (int i = 0; i < P->d.npass; ++i) {
    const int dir = (i % 2 == 0) ? 1 : -1;
    if (i > 0)
      TRY(zero(e2, 2 * n2));
    else if (r1)
      TRY(row_pass<T>(P, e2, b, 2, OP_FWD, s));  // real space -> the row-pass output form
    for (int k = 0; k <= L; ++k) {
      const int j = dir > 0 ? k : L - k;
      const int jin = dir > 0 ? j : j + 1, jout = dir > 0 ? j + 1 : j;
      const Lay yi = lay(jin), yo = lay(jout);
      cpx<T>* tg = target(jout);
      if (r1) {
        // k space: interface (D2 of the layer left, Fresnel, stored
        // reflection, D1 of the layer entered), then the sub-steps
        TRY(col_pass<T>(P, e2, b, 2, OP_FWD, s));
        TrrArgs<T> ta{};
        ta.src = e2;
        ta.dst = e2;
        ta.refl = refl(j);
        ta.g = (dir < 0 && tg) ? tmp : nullptr;
        ta.H = P->d.height;
        ta.W = P->d.width;
        ta.b = b;
        ta.amax = P->amax;
        ta.da = yi.t;
        ta.db = yo.t;
        ta.ea = yi.eeff;
        ta.eb = yo.eeff;
        ta.fresnel = !eq(yi.eeff, yo.eeff);
        ta.scale = scale;
        ta.bm = P->bm;
        hipLaunchKernelGGL((trr2_kernel<T>), dim3(unsigned((HW + NT - 1) / NT), unsigned(nplanes)),
                           dim3(NT), 0, s, ta);
        HIP_TRY(hipGetLastError());
        if (ta.g) {
          TRY(col_pass<T>(P, tmp, b, 2, OP_INV, s));
          TRY(row_pass<T>(P, tmp, b, 2, OP_INV, s));
          TRY(add_bulk(tmp, tg, yo, -1));
        }
        if (dir < 0 && jout == 0) break;  // the rest of this pass reaches no output
        TRY(col_pass<T>(P, e2, b, 2, OP_INV, s));
        if (yo.iso) {  // output medium: E and H of its forward modes
          TRY(row_pass<T>(P, e2, b, 2, OP_INV, s));
          TRY(add_bulk(e2, tg, yo, +1));
          continue;
        }
        for (int q = 0; q < yo.ns; ++q) {
          RowArgs<T> la = layer_args<T>(S, HW, yo.l, yo.t * dir, 0, status, (char*)ws + wl.etab);
          la.bwd2 = dir < 0;
          TRY(row_pass<T>(P, e2, b, 2, OP_ROP, s, &la));
          if (q + 1 < yo.ns) {
            ColArgs<T> ca{};
            ca.da = yo.t;
            ca.db = yo.t;
            ca.ea = yo.eeff;
            ca.eb = yo.eeff;
            ca.kca = 1.;
            ca.kcb = 1.;
            ca.scale = scale;
            ca.dense_sign = 0;
            TRY(col_pass<T>(P, e2, b, 2, OP_DIAG, s, &ca));
          }
        }
        if (dir > 0 && tg) {
          // field after the layer (its second half step) in real space
          TRY(copy_planes<T>(tmp, e2, b, HW, s, 2));
          ColArgs<T> cb{};
          cb.da = yo.t;
          cb.ea = yo.eeff;
          cb.db = 0.;
          cb.eb = yo.eeff;
          cb.kca = 1.;
          cb.kcb = 1.;
          cb.scale = scale;
          cb.dense_sign = 0;
          TRY(col_pass<T>(P, tmp, b, 2, OP_DIAG, s, &cb));
          TRY(row_pass<T>(P, tmp, b, 2, OP_INV, s));
          TRY(add_bulk(tmp, tg, yo, +1));
        }
        continue;
      }
      // reflection mode 2, real space
      PixArgs<T> a{};
      a.HW = HW;
      a.nlam = nlam;
      a.bm = P->bm;
      a.comp_stride = size_t(L) * HW;
      a.field = e2;
      a.mode = dir;
      a.prev_iso = yi.iso;
      a.pemed = yi.e;
      if (!yi.iso) {
        a.peps_re = S->eps_re + size_t(yi.l) * HW;
        a.peps_im = S->eps_im ? S->eps_im + size_t(yi.l) * HW : nullptr;
        a.pang = S->angles + size_t(yi.l) * HW;
      }
      set_cur(a, yo);
      const double ds = yo.t * dir;
      for (int q = 0; q < yo.ns; ++q) {
        a.d = 0.5 * ds;
        if (q == 0) {
          PixArgs<T> t = a;
          t.refl = refl(j);
          t.dst = dir < 0 ? tg : nullptr;
          hipLaunchKernelGGL((pixel_kernel<T, PIX_TRR>), grid, dim3(NT), 0, s, t);
          HIP_TRY(hipGetLastError());
          if (dir < 0 && jout == 0) break;
        } else if (ds != 0.) {
          hipLaunchKernelGGL((pixel_kernel<T, PIX_LAYER2>), grid, dim3(NT), 0, s, a);
          HIP_TRY(hipGetLastError());
        }
        if (diff) {
          // corrected E diffraction of the layer's effective medium; the
          // backward one at -d equals the forward one at d (isotropic medium)
          TRY(row_pass<T>(P, e2, b, 2, OP_FWD, s));
          ColArgs<T> ca{};
          ca.da = yo.t;
          ca.db = yo.t;
          ca.ea = yo.eeff;
          ca.eb = yo.eeff;
          ca.kca = 1.;
          ca.kcb = 1.;
          ca.scale = scale;
          ca.dense_sign = 0;
          TRY(col_pass<T>(P, e2, b, 2, OP_DIAG, s, &ca));
          TRY(row_pass<T>(P, e2, b, 2, OP_INV, s));
        }
        if (ds != 0.) {
          hipLaunchKernelGGL((pixel_kernel<T, PIX_LAYER2>), grid, dim3(NT), 0, s, a);
          HIP_TRY(hipGetLastError());
        }
      }
      if (dir < 0 && jout == 0) break;
      if (dir > 0 && tg) TRY(add_bulk(e2, tg, yo, +1));
    }
  }
  return DTMMX_OK;
}

// One layer step on a k-space field (the layer-level functions
// dtmm/propagate_4x4.py:376-389 propagate_4x4_effective_1 with :80-101 and
// dtmm/propagate_2x2.py:193-211 propagate_2x2_effective_1 with :19-76, mode
// +1, diffraction 1): every stack layer with its sub-steps, the field in
// and out in the reference's k-space convention (numpy fft2 of the
// real-space field).
//   4x4: D1 (closed form, first corrected Epn half step), then per sub-step
//        the inverse transform, per-pixel operator and forward transform,
//        the full diagonal step between sub-steps, and D2 at the end.
//   2x2: the Fresnel transmission of the tangential E field from the
//        isotropic medium e_in (reflection mode 1; e_in equal to the
//        layer's effective medium gives 1) with the first E half step,
//        then the same sub-steps, and the second E half step.
template <typename T>
static int propagate_layer_k(const dtmmx_plan* P, const dtmmx_stack* S, const cpx<T>* fin,
                             cpx<T>* fout, cd e_in, void* ws, int* status, hipStream_t s) {
  const size_t HW = size_t(P->d.height) * P->d.width;
  const bool m4 = P->d.method == DTMMX_METHOD_4X4;
  const int nc = m4 ? 4 : 2;
  const Batch b{P->d.n_rays, P->d.n_wavelengths, 0, P->d.n_wavelengths};
  const std::vector<Step> st = expand_steps(S, 1);
  const int L = int(st.size());
  const double scale = 1.0 / double(HW);
  const WsLayout wl = ws_layout(P);
  void* etab = ws ? (char*)ws + wl.etab : nullptr;
  KArgs<T> h{};
  if (m4) {
    h.kop = KOP_HEAD4;
    h.d0 = st[0].d;
    h.e0 = st[0].e;
  } else {
    h.kop = KOP_TR2;
    h.d0 = 0.;
    h.e0 = P->d.reflection ? e_in : st[0].e;
    h.d1 = st[0].d;
    h.e1 = st[0].e;
  }
  h.scale = scale;
  TRY(kspace<T>(P, fin, fout, b, h, s));
  TRY(col_pass<T>(P, fout, b, nc, OP_INV, s));
  for (int l = 0; l < L; ++l) {
    const RowArgs<T> la = layer_args<T>(S, HW, st[l].l, st[l].d, m4 ? 1 : 0, status, etab);
    TRY(row_pass<T>(P, fout, b, nc, OP_ROP, s, &la));
    if (l + 1 < L) {
      if (!eq(st[l].e, st[l + 1].e))
        return fail(DTMMX_E_UNSUPPORTED, "the layer step takes layers of one effective medium");
      ColArgs<T> ca{};
      ca.da = st[l].d;
      ca.db = st[l + 1].d;
      ca.ea = st[l].e;
      ca.eb = st[l + 1].e;
      ca.kca = 1.;
      ca.kcb = 1.;
      ca.scale = scale;
      ca.dense_sign = m4 ? 1 : 0;
      TRY(col_pass<T>(P, fout, b, nc, OP_DIAG, s, &ca));
    }
  }
  TRY(col_pass<T>(P, fout, b, nc, OP_FWD, s));
  KArgs<T> t{};
  if (m4) {
    t.kop = KOP_TAIL4;
    t.d0 = st[L - 1].d;
    t.e0 = st[L - 1].e;
  } else {
    t.kop = KOP_TR2;
    t.d0 = st[L - 1].d;
    t.e0 = st[L - 1].e;
    t.d1 = 0.;
    t.e1 = st[L - 1].e;
  }
  t.scale = 1.;
  return kspace<T>(P, fout, fout, b, t, s);
}

// 4x4 method with diffraction < 0, one pass (dtmm/transfer.py:719-725 in
// real space, :800-803; propagate_4x4.py:515-564): every sub-step of every
// stack layer (input and output media skipped) transforms the field,
// propagates each FFT mode exactly at every pixel and sums the modes.  The
// reference divides the sub-step thickness t / n by n again (:522), so a
// sub-step propagates t / n^2.
template <typename T>
static int transfer_full(const dtmmx_plan* P, const dtmmx_stack* S, const cpx<T>* fin,
                         cpx<T>* fout, void* ws, hipStream_t s) {
  const size_t HW = size_t(P->d.height) * P->d.width;
  const int nrays = P->d.n_rays, nlam = P->d.n_wavelengths;
  const size_t n = size_t(nrays) * nlam * 4 * HW;
  HIP_TRY(hipMemcpyAsync(fout, fin, n * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
  if (!ws) return fail(DTMMX_E_ARG, "diffraction < 0 needs a workspace");
  const WsLayout wl = ws_layout(P);
  cpx<T>* spec = (cpx<T>*)((char*)ws + wl.full_spec);
  cpx<T>* part = (cpx<T>*)((char*)ws + wl.full_part);
  const Batch b{nrays, nlam, 0, nlam};
  const int ngroup = (nrays + FULL_RAYS - 1) / FULL_RAYS;
  for (int l = 0; l < S->n_layers; ++l) {
    const int ns = S->n_steps ? S->n_steps[l] : 1;
    FullArgs<T> a{};
    a.spec = spec;
    a.part = part;
    a.H = P->d.height;
    a.W = P->d.width;
    a.HW = HW;
    a.nrays = nrays;
    a.nlam = nlam;
    const size_t off = size_t(l) * HW;
    a.eps_re = S->eps_re + off;
    a.eps_im = S->eps_im ? S->eps_im + off : nullptr;
    a.ang = S->angles + off;
    a.comp_stride = size_t(S->n_layers) * HW;
    a.d = S->thickness[l] / double(ns) / double(ns);
    a.modes = P->full_modes;
    a.nmodes = P->full_nmodes;
    a.nchunk = P->full_chunks;
    a.bm = P->bm;
    for (int st = 0; st < ns; ++st) {
      HIP_TRY(hipMemcpyAsync(spec, fout, n * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
      TRY(row_pass<T>(P, spec, b, 4, OP_FWD, s));
      TRY(col_pass<T>(P, spec, b, 4, OP_FWD, s));
      hipLaunchKernelGGL((full_kernel<T>),
                         dim3(unsigned((HW + NT - 1) / NT), unsigned(nlam * ngroup),
                              unsigned(P->full_chunks)),
                         dim3(NT), 0, s, a);
      HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL((chunk_sum_kernel<T>), dim3(unsigned((n + NT - 1) / NT)), dim3(NT), 0, s,
                         (const cpx<T>*)part, fout, n, P->full_chunks);
      HIP_TRY(hipGetLastError());
    }
  }
  return DTMMX_OK;
}

template <typename T>
static int transfer_impl(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout,
                         void* ws, int* status, hipStream_t s, void* bulk = nullptr) {
  const size_t HW = size_t(P->d.height) * P->d.width;
  const size_t n = size_t(P->d.n_rays) * P->d.n_wavelengths * 4 * HW;
  if (P->d.diffraction < 0) {
    if (bulk) return fail(DTMMX_E_UNSUPPORTED, "bulk output is not built for diffraction < 0");
    return transfer_full<T>(P, S, (const cpx<T>*)fin, (cpx<T>*)fout, ws, s);
  }
  if (P->d.npass > 1 && P->d.method == DTMMX_METHOD_2X2)
    return transfer_2x2_npass<T>(P, S, (cpx<T>*)fin, (cpx<T>*)fout, ws, status, s, (cpx<T>*)bulk);
  if (P->d.npass > 1) return transfer_npass<T>(P, S, (cpx<T>*)fin, (cpx<T>*)fout, ws, status, s);
  if (P->d.method == DTMMX_METHOD_2X2 && P->d.reflection == 2)
    return transfer_2x2_refl2<T>(P, S, (const cpx<T>*)fin, (cpx<T>*)fout, ws, s, (cpx<T>*)bulk);
  if (P->d.diffraction == 0)
    return transfer_nodiff<T>(P, S, (const cpx<T>*)fin, (cpx<T>*)fout, ws, s, (cpx<T>*)bulk);
  if (P->d.method == DTMMX_METHOD_4X4 && P->d.reflection == 2)
    return transfer_refl2<T>(P, S, (const cpx<T>*)fin, (cpx<T>*)fout, ws, s, (cpx<T>*)bulk);
  if (P->d.npass > 1) return transfer_npass<T>(P, S, (cpx<T>*)fin, (cpx<T>*)fout, ws, status, s);
  HIP_TRY(hipMemcpyAsync(fout, fin, n * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
  return propagate<T>(P, S, (cpx<T>*)fout, ws, status, s, 1, (cpx<T>*)bulk);
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
  Batch b{batch, 1, 0, 1};
  int r = row_pass<T>(&P, (cpx<T>*)data, b, 1, inverse ? OP_INV : OP_FWD, s);
  if (r == DTMMX_OK) r = col_pass<T>(&P, (cpx<T>*)data, b, 1, inverse ? OP_INV : OP_FWD, s);
  hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(P.tw_w);
  (void)hipFree(P.tw_h);
  if (r != DTMMX_OK) return r;
  if (e != hipSuccess) return fail(DTMMX_E_HIP, std::string("fft2: ") + hipGetErrorString(e));
  return DTMMX_OK;
}

// Steps (sub-steps of all layers) of a stack: one tape field each.
static size_t total_steps(const dtmmx_stack* S) {
  size_t n = 0;
  for (int l = 0; l < S->n_layers; ++l) n += S->n_steps ? size_t(S->n_steps[l]) : 1;
  return n;
}

static bool adjoint_supported(const dtmmx_plan* P) {
  return P->d.method == DTMMX_METHOD_4X4 && P->wn <= 1 && P->d.npass == 1 && P->d.diffraction == 1 &&
         P->d.reflection == 1;
}

// Forward 4x4 transfer (diffraction 1, npass 1) that records the row-pass
// input of every step on the tape for the adjoint.
static int ckpt_interval(const dtmmx_plan* P) { return P->d.checkpoint > 1 ? P->d.checkpoint : 1; }

// The propagation from the row-pass output of step l to the row-pass input
// of step l + 1 (one beam per ray, 4x4): the diagonal column pass between
// equal effective media, else the dense boundary operator in k-space.
template <typename T>
static int step_transition(const dtmmx_plan* P, cpx<T>* f, Batch b, const std::vector<Step>& st,
                           int l, double scale, hipStream_t s) {
  if (eq(st[l].e, st[l + 1].e)) {
    ColArgs<T> ca{};
    ca.da = st[l].d;
    ca.db = st[l + 1].d;
    ca.ea = st[l].e;
    ca.eb = st[l + 1].e;
    ca.kca = 1.;
    ca.kcb = 1.;
    ca.scale = scale;
    ca.dense_sign = 1;
    return col_pass<T>(P, f, b, 4, OP_DIAG, s, &ca);
  }
  TRY(col_pass<T>(P, f, b, 4, OP_FWD, s));
  KArgs<T> bd{};
  bd.kop = KOP_BOUND4;
  bd.d0 = st[l].d;
  bd.d1 = st[l + 1].d;
  bd.e0 = st[l].e;
  bd.e1 = st[l + 1].e;
  bd.scale = scale;
  TRY(kspace<T>(P, f, f, b, bd, s));
  return col_pass<T>(P, f, b, 4, OP_INV, s);
}

template <typename T>
static int record_impl(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout,
                       void* tape, int* status, hipStream_t s) {
  const int H = P->d.height, W = P->d.width;
  const size_t HW = size_t(H) * W;
  const int nrays = P->d.n_rays, nlam = P->d.n_wavelengths;
  const size_t fsz = size_t(nrays) * nlam * 4 * HW;
  const double scale = 1.0 / double(HW);
  cpx<T>* out = (cpx<T>*)fout;
  cpx<T>* tp = (cpx<T>*)tape;
  HIP_TRY(hipMemcpyAsync(out, fin, fsz * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
  const std::vector<Step> st = expand_steps(S, 1);
  const int L = int(st.size());
  if (L == 0) return DTMMX_OK;
  Batch b{nrays, nlam, 0, nlam};
  TRY(row_pass<T>(P, out, b, 4, OP_FWD, s));
  TRY(col_pass<T>(P, out, b, 4, OP_FWD, s));
  KArgs<T> k{};
  k.kop = KOP_HEAD4;
  k.d0 = st[0].d;
  k.e0 = st[0].e;
  k.scale = scale;
  TRY(kspace<T>(P, out, out, b, k, s));
  TRY(col_pass<T>(P, out, b, 4, OP_INV, s));
  const int K = ckpt_interval(P);
  for (int l = 0; l < L; ++l) {
    RowArgs<T> la = layer_args<T>(S, HW, st[l].l, st[l].d, 1, status);
    if (l % K == 0) la.save = tp + size_t(l / K) * fsz;  // checkpoint: the row kernel keeps its input
    TRY(row_pass<T>(P, out, b, 4, OP_ROP, s, &la));
    if (l + 1 < L) TRY(step_transition<T>(P, out, b, st, l, scale, s));
  }
  TRY(col_pass<T>(P, out, b, 4, OP_FWD, s));
  KArgs<T> t{};
  t.kop = KOP_TAIL4;
  t.d0 = st[L - 1].d;
  t.e0 = st[L - 1].e;
  t.scale = scale;
  TRY(kspace<T>(P, out, out, b, t, s));
  TRY(col_pass<T>(P, out, b, 4, OP_INV, s));
  TRY(row_pass<T>(P, out, b, 4, OP_INV, s));
  return DTMMX_OK;
}

// Adjoint of record_impl's propagation applied to `seed` = dL/dconj(F_out)
// (real space), accumulating dL/dtheta and dL/dphi of every voxel into
// grad [2][n_layers][H][W] (dL = 2 Re <seed, dF_out>).  Every operator is
// replaced by its conjugate transpose, in reverse order: unnormalised FFTs
// by the opposite transform, diagonal k-space factors by their conjugates,
// dense k-space operators by M^H, the per-pixel operator by R^H.
template <typename T>
static int adjoint_impl(const dtmmx_plan* P, const dtmmx_stack* S, const void* tape,
                        const void* seed, double* grad, void* ws, int* status, hipStream_t s) {
  const int H = P->d.height, W = P->d.width;
  const size_t HW = size_t(H) * W;
  const int nrays = P->d.n_rays, nlam = P->d.n_wavelengths;
  const size_t fsz = size_t(nrays) * nlam * 4 * HW;
  const double scale = 1.0 / double(HW);
  const std::vector<Step> st = expand_steps(S, 1);
  const int L = int(st.size());
  if (L == 0) return DTMMX_OK;
  const int nbeam = nrays / P->d.rays_per_beam;
  cpx<T>* adj = (cpx<T>*)ws;
  double* gpart = (double*)((char*)ws + (fsz * sizeof(cpx<T>) + 255) / 256 * 256);
  cpx<T>* table =
      (cpx<T>*)((char*)gpart + (size_t(nrays) * nlam * 2 * HW * sizeof(double) + 255) / 256 * 256);
  // checkpointed tape: wk holds the recomputed row input of a block's last
  // step, blk those of the steps between the checkpoint and the last
  const int K = ckpt_interval(P);
  cpx<T>* wk = (cpx<T>*)((char*)table +
                         (size_t(GRAD_COMPS) * nbeam * HW * sizeof(cpx<T>) + 255) / 256 * 256);
  cpx<T>* blk = wk + fsz;
  const cpx<T>* tp = (const cpx<T>*)tape;
  HIP_TRY(hipMemcpyAsync(adj, seed, fsz * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
  Batch b{nrays, nlam, 0, nlam};
  TRY(row_pass<T>(P, adj, b, 4, OP_FWD, s));
  TRY(col_pass<T>(P, adj, b, 4, OP_FWD, s));
  KArgs<T> t{};
  t.kop = KOP_TAIL4;
  t.d0 = st[L - 1].d;
  t.e0 = st[L - 1].e;
  t.scale = scale;
  t.adjoint = 1;
  TRY(kspace<T>(P, adj, adj, b, t, s));
  TRY(col_pass<T>(P, adj, b, 4, OP_INV, s));
  const int nl = S->n_layers;
  for (int l = L - 1; l >= 0; --l) {
    if (l + 1 < L) {
      if (eq(st[l].e, st[l + 1].e)) {
        ColArgs<T> ca{};
        ca.da = st[l].d;
        ca.db = st[l + 1].d;
        ca.ea = st[l].e;
        ca.eb = st[l + 1].e;
        ca.kca = 1.;
        ca.kcb = 1.;
        ca.scale = scale;
        ca.dense_sign = 1;
        ca.conj_k = 1;
        TRY(col_pass<T>(P, adj, b, 4, OP_DIAG, s, &ca));
      } else {
        TRY(col_pass<T>(P, adj, b, 4, OP_FWD, s));
        KArgs<T> bd{};
        bd.kop = KOP_BOUND4;
        bd.d0 = st[l].d;
        bd.d1 = st[l + 1].d;
        bd.e0 = st[l].e;
        bd.e1 = st[l + 1].e;
        bd.scale = scale;
        bd.adjoint = 1;
        TRY(kspace<T>(P, adj, adj, b, bd, s));
        TRY(col_pass<T>(P, adj, b, 4, OP_INV, s));
      }
    }
    const cpx<T>* row_in = tp + size_t(l) * fsz;  // row-pass input of step l
    if (K > 1) {
      const int l0 = l / K * K;
      cpx<T>* ck = const_cast<cpx<T>*>(tp) + size_t(l / K) * fsz;
      if (l == L - 1 || (l + 1) % K == 0) {
        // last step of a block: recompute the block's forward from its checkpoint
        for (int m = l0; m < l; ++m) {
          RowArgs<T> lm = layer_args<T>(S, HW, st[m].l, st[m].d, 1, status);
          if (m == l0) {
            lm.acc = wk;  // reads the checkpoint, writes wk
            lm.acc_mode = 1;
            TRY(row_pass<T>(P, ck, b, 4, OP_ROP, s, &lm));
          } else {
            lm.save = blk + size_t(m - l0 - 1) * fsz;
            TRY(row_pass<T>(P, wk, b, 4, OP_ROP, s, &lm));
          }
          TRY(step_transition<T>(P, wk, b, st, m, scale, s));
        }
      }
      row_in = l == l0 ? ck : ((l == L - 1 || (l + 1) % K == 0) ? wk : blk + size_t(l - l0 - 1) * fsz);
    }
    const RowArgs<T> la = layer_args<T>(S, HW, st[l].l, st[l].d, 1, status);
    if (l + 1 == L || st[l + 1].l != st[l].l) {  // sub-steps of a layer share the table
      EigArgs<T> e{};
      e.eps_re = la.eps_re;
      e.eps_im = la.eps_im;
      e.ang = la.ang;
      e.comp_stride = la.comp_stride;
      e.HW = HW;
      e.nbeam = nbeam;
      e.table = table;
      e.status = status;
      e.bm = P->bm;
      hipLaunchKernelGGL((eig_grad_kernel<T>), dim3(unsigned((HW + NT - 1) / NT), nbeam), dim3(NT), 0,
                         s, e);
      HIP_TRY(hipGetLastError());
    }
    AdjArgs<T> a{};
    a.adj = adj;
    a.table = table;
    a.nbeam = nbeam;
    a.tape = row_in;
    a.H = H;
    a.W = W;
    a.b = b;
    a.tw = (const cpx<T>*)P->tw_w;
    a.eps_re = la.eps_re;
    a.eps_im = la.eps_im;
    a.ang = la.ang;
    a.comp_stride = la.comp_stride;
    a.d = st[l].d;
    a.status = status;
    a.gpart = gpart;
    a.bm = P->bm;
    timer_start(0, s);
    TRY(launch_row_adj<T>(W, a, H * nrays * nlam, s));
    timer_stop(s);
    hipLaunchKernelGGL(grad_sum_kernel, dim3(unsigned((HW + NT - 1) / NT)), dim3(NT), 0, s,
                       (const double*)gpart, nrays * nlam, HW, grad + size_t(st[l].l) * HW,
                       grad + (size_t(nl) + st[l].l) * HW);
    HIP_TRY(hipGetLastError());
  }
  return DTMMX_OK;
}

// Precision-specific entry points (defined in the part that builds them).
int transfer_c64(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout, void* ws,
                 int* status, hipStream_t s);
int transfer_c128(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout, void* ws,
                  int* status, hipStream_t s);
int bulk_c64(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout, void* bulk, void* ws,
             int* status, hipStream_t s);
int bulk_c128(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout, void* bulk, void* ws,
              int* status, hipStream_t s);
int fft2_c64(void* data, int batch, int H, int W, int inverse, hipStream_t s);
int fft2_c128(void* data, int batch, int H, int W, int inverse, hipStream_t s);
int layer_c64(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout, cd e_in,
              void* ws, int* status, hipStream_t s);
int layer_c128(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout, cd e_in,
               void* ws, int* status, hipStream_t s);
int record_c64(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout, void* tape,
               int* status, hipStream_t s);
int record_c128(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout, void* tape,
                int* status, hipStream_t s);
int adjoint_c64(const dtmmx_plan* P, const dtmmx_stack* S, const void* tape, const void* seed,
                double* grad, void* ws, int* status, hipStream_t s);
int adjoint_c128(const dtmmx_plan* P, const dtmmx_stack* S, const void* tape, const void* seed,
                 double* grad, void* ws, int* status, hipStream_t s);

// Device tables of one polarizing-microscope specter call, uploaded by the
// host part into the caller's workspace.
struct PomJob {
  const dtmmx_pom_desc* d;
  const void* field;
  void* spec;
  void* wfield;        // combined / copied fields (diffraction only)
  void* tw_w;
  void* tw_h;
  int* amax;           // [nlam][H] betamax mask rows
  const void* pmat;    // device [npm][4][4] complex double (npm > 1), or null
  const double* pm0;   // host [4][4] complex: the shared analyser (npm == 1), or null
};

struct ColorJob {
  const void* spec;
  void* out;
  long long n;
  int nlam;
  const double* cmf;   // device [nlam][3] (nlam > COLOR_MAXL), else null
  const double* cmf_h; // host [nlam][3] (nlam <= COLOR_MAXL: kernel arguments)
  double* part;        // device [nblocks]
  int norm_mode;
  double norm;
  int gamma_mode;
  double gamma;
  int gray;
};

int pom_c64(const PomJob& j, hipStream_t s);
int illum_c64(const dtmmx_plan* P, const dtmmx_illum_desc* d, void* f, char* ws, hipStream_t s);
int illum_c128(const dtmmx_plan* P, const dtmmx_illum_desc* d, void* f, char* ws, hipStream_t s);
int pom_c128(const PomJob& j, hipStream_t s);
int color_c64(const ColorJob& j, hipStream_t s);
int color_c128(const ColorJob& j, hipStream_t s);

// Specter of dtmmx_pom_specter: (Jones combination) -> FFT -> F P F^-1 of
// the medium (with 1/HW) -> IFFT -> analyser and Poynting per pixel.
// The field diffraction F P(d) F^-1 of an isotropic medium eps on the x
// spectrum of 4-component planes (after the forward row pass, before the
// inverse one): one fused column pass for a real medium with eps > betamax^2
// (every passing mode propagates), else forward column pass, the dense
// k-space operator and the inverse column pass.
template <typename T>
static int field_diffraction(const dtmmx_plan* P, cpx<T>* f, Batch b, cd eps, double d,
                             int fmode, double scale, hipStream_t s) {
  if (eps.y == 0. && eps.x > P->d.betamax * P->d.betamax * (1. + 1e-9)) {
    ColArgs<T> ca{};
    ca.da = d;
    ca.ea = eps;
    ca.scale = scale;
    ca.fmode = fmode;
    return col_pass<T>(P, f, b, 4, OP_FIELD, s, &ca);
  }
  TRY(col_pass<T>(P, f, b, 4, OP_FWD, s));
  KArgs<T> k{};
  k.kop = KOP_FIELD4;
  k.d0 = d;
  k.e0 = eps;
  k.scale = scale;
  k.fmode = fmode;
  TRY(kspace<T>(P, f, f, b, k, s));
  return col_pass<T>(P, f, b, 4, OP_INV, s);
}

template <typename T>
static int pom_impl(const PomJob& j, hipStream_t s) {
  const dtmmx_pom_desc* d = j.d;
  const int H = d->height, W = d->width, nlam = d->n_wavelengths, nf = d->n_fields;
  const long long HW = (long long)H * W;
  PomArgs<T> a{};
  a.HW = HW;
  a.nlam = nlam;
  a.nf = nf;
  a.npol = d->n_pol;
  a.jc = mk<double>(d->jones[0], d->jones[1]);
  a.js = mk<double>(d->jones[2], d->jones[3]);
  a.pmat = (const cd*)j.pmat;
  a.npm = d->pmat ? d->n_pmat : 0;
  if (j.pm0)
    for (int i = 0; i < 16; ++i) a.pm0[i] = mk<double>(j.pm0[2 * i], j.pm0[2 * i + 1]);
  a.spec = (T*)j.spec;
  const unsigned gx = unsigned((HW + NT - 1) / NT);
  if (!d->diffraction) {
    a.src = (const cpx<T>*)j.field;
    hipLaunchKernelGGL((pom_specter_kernel<T>), dim3(gx), dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
    return DTMMX_OK;
  }
  cpx<T>* F = (cpx<T>*)j.wfield;
  const size_t nplanes = size_t(nf) * nlam * 4;
  dtmmx_plan P{};
  P.d.height = H;
  P.d.width = W;
  P.d.n_wavelengths = nlam;
  P.d.betamax = d->betamax;
  P.tw_w = j.tw_w;
  P.tw_h = j.tw_h;
  P.amax = j.amax;
  P.bm.nlam = nlam;
  P.bm.rays_per_beam = nf;  // one (unused) beam for all fields
  for (int l = 0; l < nlam; ++l) P.bm.k0[l] = d->k0[l];
  Batch b{nf, nlam, 0, nlam};
  if (d->n_pol == 2) {
    // the Jones combination of the two polarised fields, read on the way
    // into the forward x transforms
    RowArgs<T> rc{};
    rc.csrc = (const cpx<T>*)j.field;
    rc.cjc = a.jc;
    rc.cjs = a.js;
    TRY(row_pass<T>(&P, F, b, 4, OP_COMB, s, &rc));
  } else {
    HIP_TRY(hipMemcpyAsync(F, j.field, nplanes * size_t(HW) * sizeof(cpx<T>),
                           hipMemcpyDeviceToDevice, s));
    TRY(row_pass<T>(&P, F, b, 4, OP_FWD, s));
  }
  TRY(field_diffraction<T>(&P, F, b, mk<double>(d->eps_re, d->eps_im), d->d, d->mode,
                           1.0 / double(HW), s));
  // the rays a row work-group takes (row_pass: two complex64 rays up to 512
  // points when the count is even)
  const int nr_row = (sizeof(T) == 4 && W <= 512 && nf % 2 == 0) ? 2 : 1;
  if (nf == nr_row) {
    // inverse x transforms with the specter of each pixel in the same pass
    RowArgs<T> ra{};
    ra.spec = (T*)j.spec;
    ra.spec_npm = a.npm;
    ra.spec_pmat = a.pmat;
    for (int i = 0; i < 16; ++i) ra.spec_pm0[i] = a.pm0[i];
    return row_pass<T>(&P, F, b, 4, OP_SPEC, s, &ra);
  }
  TRY(row_pass<T>(&P, F, b, 4, OP_INV, s));
  a.src = F;
  a.npol = 1;
  hipLaunchKernelGGL((pom_specter_kernel<T>), dim3(gx), dim3(NT), 0, s, a);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

template <typename T>
static int color_impl(const ColorJob& j, hipStream_t s) {
  ColorArgs<T> a{};
  a.spec = (const T*)j.spec;
  a.out = (T*)j.out;
  a.n = j.n;
  a.nlam = j.nlam;
  a.cmf = j.cmf;
  if (!j.cmf)
    for (int i = 0; i < 3 * j.nlam; ++i) a.cmf_v[i] = j.cmf_h[i];
  a.part = j.part;
  a.nblocks = int((j.n + NT - 1) / NT);
  a.norm = j.norm_mode == 1 ? j.norm : 0.;
  a.norm_auto = j.norm_mode == 2;
  a.gray = j.gray;
  a.gamma_mode = j.gamma_mode;
  a.gamma = j.gamma;
  const dim3 grid(unsigned(a.nblocks));
  if (a.norm_auto) {
    hipLaunchKernelGGL((color_kernel<T, 0>), grid, dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((color_kernel<T, 1>), dim3(1), dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL((color_kernel<T, 2>), grid, dim3(NT), 0, s, a);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

// ---------------------------------------------------------------------------
// Illumination field (dtmm/field.py:586-713)
// ---------------------------------------------------------------------------

struct BeamValues {
  double v[MAXB];
};

// Power normalisation factors of the illumination (field.py:699-711):
// sqrt(intensity of the beam / beams), times sqrt(power before / power
// after) of the diffraction-limited projection.
static __global__ __launch_bounds__(NT) void illum_factor_kernel2(const double* p1, const double* p2,
                                                            BeamValues inten, int nplanes,
                                                            int nlam, int rays_per_beam,
                                                            int nbeams, double* out) {
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q >= nplanes) return;
  const int beam = (q / nlam) / rays_per_beam;
  double f = sqrt(inten.v[beam] / nbeams);
  if (p2) f *= sqrt(p1[q] / p2[q]);
  out[q] = f;
}

template <typename T>
static int illum_impl(const dtmmx_plan* P, const dtmmx_illum_desc* d, cpx<T>* field, char* ws,
                      hipStream_t s) {
  const int H = P->d.height, W = P->d.width;
  const size_t HW = size_t(H) * W;
  const int nlam = P->d.n_wavelengths, nrays = P->d.n_rays;
  const int planes = nrays * nlam;
  const int nblk = int((HW + NT - 1) / NT);
  const int nbeams = nrays / P->d.rays_per_beam;
  double* part1 = (double*)ws;
  double* part2 = part1 + size_t(planes) * nblk;
  double* sum1 = part2 + size_t(planes) * nblk;
  double* sum2 = sum1 + planes;
  double* fac = sum2 + planes;
  IllumArgs<T> a{};
  a.field = field;
  a.H = H;
  a.W = W;
  a.nlam = nlam;
  a.nrays = nrays;
  a.rays_per_beam = P->d.rays_per_beam;
  a.n = d->n;
  a.mode = d->mode;
  a.polarized = d->polarized;
  a.jones[0] = mk<double>(d->jones_re[0], d->jones_im[0]);
  a.jones[1] = mk<double>(d->jones_re[1], d->jones_im[1]);
  a.window = d->window;
  a.part = d->diffraction ? part1 : nullptr;
  a.bm = P->bm;
  const dim3 grid((unsigned)nblk, (unsigned)planes);
  hipLaunchKernelGGL((illum_kernel<T, ILL_BUILD>), grid, dim3(NT), 0, s, a);
  HIP_TRY(hipGetLastError());
  BeamValues iv{};
  for (int b = 0; b < nbeams; ++b) iv.v[b] = d->intensity[b];
  if (d->diffraction) {
    // x transforms; y transforms with the field diffraction operator of the
    // cover medium between them; inverse x transforms
    const Batch b{nrays, nlam, 0, nlam};
    TRY(row_pass<T>(P, field, b, 4, OP_FWD, s));
    TRY(field_diffraction<T>(P, field, b, mk<double>(d->n * d->n), -d->focus,
                             d->mode > 0 ? 1 : -1, 1.0 / double(HW), s));
    TRY(row_pass<T>(P, field, b, 4, OP_INV, s));
    IllumArgs<T> a2 = a;
    a2.part = part2;
    hipLaunchKernelGGL((illum_kernel<T, ILL_POWER>), grid, dim3(NT), 0, s, a2);
    HIP_TRY(hipGetLastError());
    TRY(sum_rows(part1, planes, nblk, 1., sum1, s));
    TRY(sum_rows(part2, planes, nblk, 1., sum2, s));
    hipLaunchKernelGGL(illum_factor_kernel2, dim3(unsigned((planes + NT - 1) / NT)), dim3(NT), 0, s,
                       (const double*)sum1, (const double*)sum2, iv, planes, nlam,
                       P->d.rays_per_beam, nbeams, fac);
  } else {
    hipLaunchKernelGGL(illum_factor_kernel2, dim3(unsigned((planes + NT - 1) / NT)), dim3(NT), 0, s,
                       (const double*)nullptr, (const double*)nullptr, iv, planes, nlam,
                       P->d.rays_per_beam, nbeams, fac);
  }
  HIP_TRY(hipGetLastError());
  IllumArgs<T> a3 = a;
  a3.part = nullptr;
  a3.scale = fac;
  hipLaunchKernelGGL((illum_kernel<T, ILL_SCALE>), grid, dim3(NT), 0, s, a3);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

#if DTMMX_C64_PART
int illum_c64(const dtmmx_plan* P, const dtmmx_illum_desc* d, void* f, char* ws, hipStream_t s) {
  return illum_impl<float>(P, d, (cpx<float>*)f, ws, s);
}
#endif
#if DTMMX_C128_PART
int illum_c128(const dtmmx_plan* P, const dtmmx_illum_desc* d, void* f, char* ws, hipStream_t s) {
  return illum_impl<double>(P, d, (cpx<double>*)f, ws, s);
}
#endif
#if DTMMX_C64_PART
int pom_c64(const PomJob& j, hipStream_t s) { return pom_impl<float>(j, s); }
int color_c64(const ColorJob& j, hipStream_t s) { return color_impl<float>(j, s); }
int transfer_c64(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout, void* ws,
                 int* status, hipStream_t s) {
  return transfer_impl<float>(P, S, fin, fout, ws, status, s);
}
int bulk_c64(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout, void* bulk, void* ws,
             int* status, hipStream_t s) {
  return transfer_impl<float>(P, S, fin, fout, ws, status, s, bulk);
}
int fft2_c64(void* data, int batch, int H, int W, int inverse, hipStream_t s) {
  return fft2_impl<float>(data, batch, H, W, inverse, s);
}
int layer_c64(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout, cd e_in,
              void* ws, int* status, hipStream_t s) {
  return propagate_layer_k<float>(P, S, (const cpx<float>*)fin, (cpx<float>*)fout, e_in, ws,
                                  status, s);
}
int record_c64(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout, void* tape,
               int* status, hipStream_t s) {
  return record_impl<float>(P, S, fin, fout, tape, status, s);
}
int adjoint_c64(const dtmmx_plan* P, const dtmmx_stack* S, const void* tape, const void* seed,
                double* grad, void* ws, int* status, hipStream_t s) {
  return adjoint_impl<float>(P, S, tape, seed, grad, ws, status, s);
}
#endif
#if DTMMX_C128_PART
int pom_c128(const PomJob& j, hipStream_t s) { return pom_impl<double>(j, s); }
int color_c128(const ColorJob& j, hipStream_t s) { return color_impl<double>(j, s); }
int transfer_c128(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout, void* ws,
                  int* status, hipStream_t s) {
  return transfer_impl<double>(P, S, fin, fout, ws, status, s);
}
int bulk_c128(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout, void* bulk, void* ws,
              int* status, hipStream_t s) {
  return transfer_impl<double>(P, S, fin, fout, ws, status, s, bulk);
}
int fft2_c128(void* data, int batch, int H, int W, int inverse, hipStream_t s) {
  return fft2_impl<double>(data, batch, H, W, inverse, s);
}
int layer_c128(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout, cd e_in,
               void* ws, int* status, hipStream_t s) {
  return propagate_layer_k<double>(P, S, (const cpx<double>*)fin, (cpx<double>*)fout, e_in, ws,
                                   status, s);
}
int record_c128(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout, void* tape,
                int* status, hipStream_t s) {
  return record_impl<double>(P, S, fin, fout, tape, status, s);
}
int adjoint_c128(const dtmmx_plan* P, const dtmmx_stack* S, const void* tape, const void* seed,
                 double* grad, void* ws, int* status, hipStream_t s) {
  return adjoint_impl<double>(P, S, tape, seed, grad, ws, status, s);
}
#endif

#if DTMMX_HOST_PART
// numpy.mod of two doubles: the remainder with the sign of b.
static double np_mod(double a, double b) {
  double m = std::fmod(a, b);
  if (m != 0.) {
    if ((b < 0.) != (m < 0.)) m += b;
  } else {
    m = std::copysign(0., b);
  }
  return m;
}

// Tent windows of dtmm/mode.py:10-139 for diffraction = n > 1 on the host,
// in float64 as the reference computes them: window (i, j) is
// cx_i(beta_x) * cy_j(beta_y), zero where |beta| >= betamax, with centres
// numpy.linspace(-betamax, betamax, n) shifted by the beam's
// (beta_x, beta_y) reduced modulo the spacing to the nearest multiple; the
// mean beam of a window is the window^2-weighted mean (beta_x, beta_y) over
// wavelengths and k; all-zero windows (over beams and wavelengths) are
// dropped.  The window factors are separable, so the device keeps cx and cy.
struct WinTables {
  std::vector<int> wi, wj;             // non-zero windows
  std::vector<double> beta, phi;       // [window][beam]
  std::vector<double> cx, cy;          // [beam][lam][n][W], [beam][lam][n][H]
  std::vector<int> mask;               // [lam][H]
};

// betamax mask per (wavelength, ky): the largest passing |kx| index (-1 if
// none) with beta >= betamax cut (dtmm/diffract.py:31, :55, :86) and beta
// computed as in dtmm/wave.py:77-78.
static std::vector<int> betamax_rows(const dtmmx_plan_desc* d) {
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
  return amax;
}

static void build_windows(const dtmmx_plan_desc* d, const std::vector<int>& amax, WinTables& t) {
#pragma clang fp contract(off)
  const int n = d->diffraction, H = d->height, W = d->width, nl = d->n_wavelengths;
  const int nb = d->n_rays / d->rays_per_beam;
  const double bm = d->betamax;
  const double sp = 2 * bm / (n - 1);
  std::vector<double> ax(n);
  {
    const double start = -bm, step = (bm - start) / double(n - 1);
    for (int i = 0; i < n; ++i) ax[i] = double(i) * step + start;
    ax[n - 1] = bm;
  }
  const double st = ax[1] - ax[0];
  auto tent = [&](double f, double c, int typ) {
    const double dx = (f - c) / st;
    const double v = 1. - std::fabs(dx);
    if (v < 0.) return 0.;
    if ((typ < 0 && dx < 0.) || (typ > 0 && dx > 0.)) return 1.;
    return v;
  };
  auto typ = [&](int i) { return i == 0 ? -1 : (i == n - 1 ? 1 : 0); };
  // largest |kx| index with |beta| < betamax on the window's beta grid
  std::vector<int> wmax(size_t(nl) * H);
  std::vector<int>& mask = t.mask;
  mask.assign(size_t(nl) * H, -1);
  for (int l = 0; l < nl; ++l) {
    const double lf = 2 * M_PI / d->k0[l];
    for (int y = 0; y < H; ++y) {
      const double by = lf * fftfreq(y, H);
      int best = -1;
      for (int m = 0; m <= W / 2; ++m) {
        const double bx = lf * fftfreq(m, W);
        if (std::sqrt(bx * bx + by * by) >= bm) break;
        best = m;
      }
      wmax[size_t(l) * H + y] = best;
      mask[size_t(l) * H + y] = std::min(best, amax[size_t(l) * H + y]);
    }
  }
  std::vector<double>& cx = t.cx;
  std::vector<double>& cy = t.cy;
  cx.assign(size_t(nb) * nl * n * W, 0.);
  cy.assign(size_t(nb) * nl * n * H, 0.);
  std::vector<double> SX(size_t(nb) * n * n, 0.), SY(SX.size(), 0.), NN(SX.size(), 0.);
  std::vector<double> sx(n), nx(n);
  for (int q = 0; q < nb; ++q) {
    const double ox = d->beta[q] * std::cos(d->phi[q]), oy = d->beta[q] * std::sin(d->phi[q]);
    double xo = np_mod(ox, sp), xm = np_mod(ox, -sp);
    if (std::fabs(xo) > std::fabs(xm)) xo = xm;
    double yo = np_mod(oy, sp), ym = np_mod(oy, -sp);
    if (std::fabs(yo) > std::fabs(ym)) yo = ym;
    for (int l = 0; l < nl; ++l) {
      const double lf = 2 * M_PI / d->k0[l];
      double* cxl = cx.data() + (size_t(q) * nl + l) * n * W;
      double* cyl = cy.data() + (size_t(q) * nl + l) * n * H;
      for (int i = 0; i < n; ++i) {
        for (int x = 0; x < W; ++x) cxl[i * W + x] = tent(lf * fftfreq(x, W), ax[i] + xo, typ(i));
        for (int y = 0; y < H; ++y) cyl[i * H + y] = tent(lf * fftfreq(y, H), ax[i] + yo, typ(i));
      }
      for (int y = 0; y < H; ++y) {
        const int thr = wmax[size_t(l) * H + y];
        const double by = lf * fftfreq(y, H);
        for (int i = 0; i < n; ++i) {
          double a = 0., c = 0.;
          for (int x = 0; x < W; ++x) {
            if (mode_index(x, W) > thr) continue;
            const double w2 = cxl[i * W + x] * cxl[i * W + x];
            a += w2 * (lf * fftfreq(x, W));
            c += w2;
          }
          sx[i] = a;
          nx[i] = c;
        }
        for (int j = 0; j < n; ++j) {
          const double wy = cyl[j * H + y] * cyl[j * H + y];
          if (wy == 0.) continue;
          for (int i = 0; i < n; ++i) {
            const size_t o = (size_t(q) * n + i) * n + j;
            SX[o] += wy * sx[i];
            SY[o] += wy * nx[i] * by;
            NN[o] += wy * nx[i];
          }
        }
      }
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      bool any = false;
      for (int q = 0; q < nb; ++q) any = any || NN[(size_t(q) * n + i) * n + j] > 0.;
      if (!any) continue;
      t.wi.push_back(i);
      t.wj.push_back(j);
      for (int q = 0; q < nb; ++q) {
        const size_t o = (size_t(q) * n + i) * n + j;
        const double bx = NN[o] > 0. ? SX[o] / NN[o] : 0., by = NN[o] > 0. ? SY[o] / NN[o] : 0.;
        t.beta.push_back(std::sqrt(bx * bx + by * by));
        t.phi.push_back(std::atan2(by, bx));
      }
    }
}

static int upload_windows(dtmmx_plan* P, const std::vector<int>& amax) {
  WinTables t;
  build_windows(&P->d, amax, t);
  P->wn = P->d.diffraction;
  P->win_i = t.wi;
  P->win_j = t.wj;
  P->win_beta = t.beta;
  P->win_phi = t.phi;
  const std::vector<double>& cx = t.cx;
  const std::vector<double>& cy = t.cy;
  const std::vector<int>& mask = t.mask;
  hipError_t e = hipMalloc((void**)&P->wcx, cx.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&P->wcy, cy.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&P->wmask, mask.size() * sizeof(int));
  if (e == hipSuccess)
    e = hipMemcpy(P->wcx, cx.data(), cx.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(P->wcy, cy.data(), cy.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(P->wmask, mask.data(), mask.size() * sizeof(int), hipMemcpyHostToDevice);
  // tent-row order: the n windows of one row read the same window-row lines
  const int nwin = int(t.wi.size());
  const int nb = P->d.n_rays / P->d.rays_per_beam;
  std::vector<int> perm(nwin);
  for (int w = 0; w < nwin; ++w) perm[w] = w;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
    return t.wj[a] != t.wj[b] ? t.wj[a] < t.wj[b] : t.wi[a] < t.wi[b];
  });
  std::vector<int> ord(2 * size_t(nwin));
  std::vector<double> bp(2 * size_t(nwin) * nb);
  for (int k = 0; k < nwin; ++k) {
    const int w = perm[k];
    ord[k] = t.wi[w];
    ord[nwin + k] = t.wj[w];
    for (int q = 0; q < nb; ++q) {
      bp[size_t(k) * nb + q] = t.beta[size_t(w) * nb + q];
      bp[(size_t(nwin) + k) * nb + q] = t.phi[size_t(w) * nb + q];
    }
  }
  if (nwin > 0) {
    if (e == hipSuccess) e = hipMalloc((void**)&P->win_ord, ord.size() * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&P->win_bp, bp.size() * sizeof(double));
    if (e == hipSuccess)
      e = hipMemcpy(P->win_ord, ord.data(), ord.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(P->win_bp, bp.data(), bp.size() * sizeof(double), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? DTMMX_E_NOMEM : DTMMX_E_HIP,
                std::string("window tables: ") + hipGetErrorString(e));
  return DTMMX_OK;
}

#endif  // DTMMX_HOST_PART
}  // namespace dtmmx

#if DTMMX_HOST_PART
// Workspace of dtmmx_transfer_noalloc, the transfer for callers that cannot
// allocate.  Each GPU the library runs on holds its own copy of the array
// for the life of the process; in host code `scratch` only names the
// symbol, and its device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[DTMMX_NOALLOC_SCRATCH_BYTES];

namespace dtmmx {
constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on the device of `stream` (the current device
// for the null stream), looked up once per device with that device current.
static int scratch_address(hipStream_t stream, void** out) {
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  int dev = cur;
  if (stream) HIP_TRY(hipStreamGetDevice(stream, &dev));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
    return fail(DTMMX_E_ARG, "device index " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    if (dev != cur) HIP_TRY(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (dev != cur) HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(DTMMX_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DTMMX_OK;
}
}  // namespace dtmmx

extern "C" {

int dtmmx_abi_version(void) { return DTMMX_ABI_VERSION; }

const char* dtmmx_last_error(void) { return g_err.c_str(); }

// Mode lists of exact per-mode propagation (dtmm/propagate_4x4.py:530-546):
// per wavelength the FFT modes with beta < betamax (strict), beta computed
// as dtmm/wave.py:54-81 does, in ascending beta with ties in row-major
// order (Python's stable sort); and the chunk count that gives a launch
// about 2^18 threads.
static int upload_full_modes(dtmmx_plan* P) {
  const int H = P->d.height, W = P->d.width, nl = P->d.n_wavelengths;
  const size_t HW = size_t(H) * W;
  std::vector<int> modes(size_t(nl) * HW, 0), counts(nl, 0);
  std::vector<std::pair<double, int>> b;
  for (int l = 0; l < nl; ++l) {
    b.clear();
    const double k0 = std::fabs(P->d.k0[l]);
    for (int ky = 0; ky < H; ++ky)
      for (int kx = 0; kx < W; ++kx) {
        const double fx = fftfreq(kx, W), fy = fftfreq(ky, H);
        const double beta = (2 * M_PI / k0) * std::sqrt(fx * fx + fy * fy);
        if (beta < P->d.betamax) b.push_back({beta, ky * W + kx});
      }
    std::stable_sort(b.begin(), b.end(),
                     [](const std::pair<double, int>& u, const std::pair<double, int>& v) {
                       return u.first < v.first;
                     });
    counts[l] = int(b.size());
    for (size_t i = 0; i < b.size(); ++i) modes[size_t(l) * HW + i] = b[i].second;
  }
  const int ngroup = (P->d.n_rays + FULL_RAYS - 1) / FULL_RAYS;
  const long long threads = (long long)HW * nl * ngroup;
  long long ch = (262144 + threads - 1) / threads;
  P->full_chunks = int(ch < 1 ? 1 : (ch > 32 ? 32 : ch));
  hipError_t e = hipMalloc((void**)&P->full_modes, modes.size() * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void**)&P->full_nmodes, counts.size() * sizeof(int));
  if (e == hipSuccess)
    e = hipMemcpy(P->full_modes, modes.data(), modes.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(P->full_nmodes, counts.data(), counts.size() * sizeof(int),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? DTMMX_E_NOMEM : DTMMX_E_HIP,
                std::string("mode lists: ") + hipGetErrorString(e));
  return DTMMX_OK;
}

static int validate_desc(const dtmmx_plan_desc* d) {
  if (d->height <= 0 || d->width <= 0) return fail(DTMMX_E_ARG, "empty grid");
  // without diffraction every pixel is an independent stack and no transform
  // runs (the 4x4 passes' normalisations aside): any grid, down to the single
  // pixel of the reference's 1D examples
  const bool no_fft = d->diffraction == 0 && !(d->method == DTMMX_METHOD_4X4 && d->npass > 1);
  if (!no_fft && (!size_supported(d->height) || !size_supported(d->width)))
    return fail(DTMMX_E_UNSUPPORTED, "grid " + std::to_string(d->height) + "x" +
                                         std::to_string(d->width) + " not built");
  if (size_t(d->height) * size_t(d->width) > (size_t(1) << 28))
    return fail(DTMMX_E_ARG, "grid above 2^28 pixels");
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
  // diffraction 0: the per-pixel modes (4x4 reflection 2, 2x2 reflection 0),
  // one pass
  const bool nodiff = d->diffraction == 0 &&
                      ((d->method == DTMMX_METHOD_4X4 && d->reflection == 2) ||
                       (d->method == DTMMX_METHOD_2X2 &&
                        ((d->reflection == 0 && d->npass == 1) || d->reflection == 2)));
  // 4x4 reflection mode 2 (any npass; diffraction > 1: tent windows)
  const bool refl2 = d->method == DTMMX_METHOD_4X4 && d->reflection == 2;
  // diffraction < 0: exact per-mode propagation (4x4, one pass)
  const bool full = d->diffraction < 0;
  if (full && (d->method != DTMMX_METHOD_4X4 || d->npass != 1))
    return fail(DTMMX_E_UNSUPPORTED,
                "diffraction < 0 (exact per-mode propagation) is built for the 4x4 method "
                "with npass 1");
  if (d->diffraction < 1 && !nodiff && !full)
    return fail(DTMMX_E_UNSUPPORTED,
                "diffraction must be >= 1, or 0 for the 4x4 method in reflection mode 2 and the "
                "2x2 method in reflection mode 0 with npass 1");
  if (d->diffraction > 1 && d->method != DTMMX_METHOD_4X4 &&
      !(d->npass == 1 && (d->reflection == 0 || d->reflection == 1)))
    return fail(DTMMX_E_UNSUPPORTED, "diffraction > 1 is built for the 4x4 method and the 2x2 "
                                     "method in reflection modes 0 and 1 with npass 1");
  if (d->diffraction > 64) return fail(DTMMX_E_ARG, "diffraction above 64");
  if (d->npass < 1) return fail(DTMMX_E_ARG, "npass must be >= 1");
  // 2x2 with npass > 1: reflection mode 1 (diffraction 1) or 2 (diffraction 0
  // or 1); its passes have no normalisation (dtmm/transfer.py:1105-1178)
  if (d->npass > 1 && d->method == DTMMX_METHOD_2X2 &&
      !((d->reflection == 1 && d->diffraction == 1) ||
        (d->reflection == 2 && d->diffraction <= 1)))
    return fail(DTMMX_E_UNSUPPORTED, "npass > 1 with the 2x2 method is built for reflection "
                                     "mode 1 with diffraction 1 and mode 2 with diffraction 0 "
                                     "or 1");
  if (d->npass > 1 && d->method == DTMMX_METHOD_4X4 &&
      (d->norm & ~(DTMMX_NORM_FFT | DTMMX_NORM_REF)) != 0)
    return fail(DTMMX_E_ARG, "norm must combine DTMMX_NORM_FFT and DTMMX_NORM_REF");
  if (d->npass > 1 && d->method == DTMMX_METHOD_4X4 && (d->norm & DTMMX_NORM_REF) &&
      d->diffraction > 1)
    return fail(DTMMX_E_UNSUPPORTED, "DTMMX_NORM_REF is built with diffraction 0 or 1");
  if (d->checkpoint < 0) return fail(DTMMX_E_ARG, "checkpoint interval must be >= 0");
  if (d->method == DTMMX_METHOD_4X4 && d->reflection != 1 && !refl2 && !full)
    return fail(DTMMX_E_UNSUPPORTED,
                "the 4x4 method is built for reflection modes 1 and 2");
  if (d->method == DTMMX_METHOD_2X2 && d->reflection != 0 && d->reflection != 1 &&
      !(d->reflection == 2 && d->diffraction <= 1))
    return fail(DTMMX_E_UNSUPPORTED, "the 2x2 method is built for reflection modes 0 and 1, "
                                     "and 2 with diffraction 0 or 1");
  return DTMMX_OK;
}

int dtmmx_plan_create(dtmmx_plan** out, const dtmmx_plan_desc* d) {
  if (!out || !d) return fail(DTMMX_E_ARG, "null argument");
  *out = nullptr;
  {
    const int r = validate_desc(d);
    if (r != DTMMX_OK) return r;
  }
  const int nbeam = d->n_rays / d->rays_per_beam;
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
  const int H = d->height, W = d->width;
  const std::vector<int> amax = betamax_rows(d);
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
  if (d->diffraction > 1) {
    const int r = upload_windows(P, amax);
    if (r != DTMMX_OK) {
      dtmmx_plan_destroy(P);
      return r;
    }
  }
  if (d->diffraction < 0) {
    const int r = upload_full_modes(P);
    if (r != DTMMX_OK) {
      dtmmx_plan_destroy(P);
      return r;
    }
  }
  *out = P;
  return DTMMX_OK;
}

int dtmmx_plan_destroy(dtmmx_plan* P) {
  if (!P) return DTMMX_OK;
  if (P->tw_w) (void)hipFree(P->tw_w);
  if (P->tw_h) (void)hipFree(P->tw_h);
  if (P->amax) (void)hipFree(P->amax);
  if (P->wcx) (void)hipFree(P->wcx);
  if (P->wcy) (void)hipFree(P->wcy);
  if (P->wmask) (void)hipFree(P->wmask);
  if (P->win_ord) (void)hipFree(P->win_ord);
  if (P->win_bp) (void)hipFree(P->win_bp);
  if (P->full_modes) (void)hipFree(P->full_modes);
  if (P->full_nmodes) (void)hipFree(P->full_nmodes);
  delete P;
  return DTMMX_OK;
}

size_t dtmmx_plan_workspace_size(const dtmmx_plan* P) {
  if (!P) return 0;
  return ws_layout(P).total;
}

size_t dtmmx_transfer_workspace_size(const dtmmx_plan* P, const dtmmx_stack* S) {
  if (!P) return 0;
  return ws_layout(P).total + stack_ws_bytes(P, S);
}

static int check_transfer(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin,
                          const void* fout, const void* ws, size_t ws_bytes, bool need_ws = true) {
  if (!P || !S || !fin || !fout) return fail(DTMMX_E_ARG, "null argument");
  if (S->n_layers < 0) return fail(DTMMX_E_ARG, "negative layer count");
  if (S->n_layers > 0 && (!S->thickness || !S->eps_eff_re || !S->eps_re || !S->angles))
    return fail(DTMMX_E_ARG, "null stack array");
  if (S->n_steps)
    for (int l = 0; l < S->n_layers; ++l)
      if (S->n_steps[l] < 1) return fail(DTMMX_E_ARG, "n_steps must be >= 1");
  const size_t need = dtmmx_transfer_workspace_size(P, S);
  if (need_ws && (ws_bytes < need || (need && !ws)))
    return fail(DTMMX_E_ARG, "workspace too small: " + std::to_string(need) + " bytes needed");
  if (fin == fout) return fail(DTMMX_E_ARG, "field_in and field_out must not alias");
  if (S->eff_eps_re && S->n_layers > 0) {
    if (!S->eff_angles) return fail(DTMMX_E_ARG, "anisotropic effective media need eff_angles");
    const bool k4 = P->d.method == DTMMX_METHOD_4X4 && P->d.reflection == 1 &&
                    P->d.diffraction == 1 && P->wn <= 1 &&
                    !(P->d.npass > 1 && (P->d.norm & DTMMX_NORM_REF));
    if (!(P->d.diffraction == 0 || k4))
      return fail(DTMMX_E_UNSUPPORTED,
                  "anisotropic effective media are built for the 4x4 method in reflection mode 1 "
                  "with diffraction 1 (without DTMMX_NORM_REF), and for diffraction 0");
  }
  return DTMMX_OK;
}

// status bit 2: the output holds NaN or infinite values
static int check_finite(const dtmmx_plan* P, const void* f, size_t nfields, int* status,
                        hipStream_t s) {
  if (!status) return DTMMX_OK;
  const size_t n = nfields * 2 * size_t(P->d.n_rays) * P->d.n_wavelengths * 4 *
                   size_t(P->d.height) * P->d.width;
  const unsigned nb = unsigned(std::min<size_t>((n + NT - 1) / NT, 4096));
  if (P->d.precision == DTMMX_C64)
    hipLaunchKernelGGL((finite_kernel<float>), dim3(nb), dim3(NT), 0, s, (const float*)f, n,
                       status);
  else
    hipLaunchKernelGGL((finite_kernel<double>), dim3(nb), dim3(NT), 0, s, (const double*)f, n,
                       status);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

size_t dtmmx_tape_size(const dtmmx_plan* P, const dtmmx_stack* S) {
  if (!P || !S) return 0;
  const size_t cb = P->d.precision == DTMMX_C64 ? sizeof(cpx<float>) : sizeof(cpx<double>);
  const size_t fsz = size_t(P->d.n_rays) * P->d.n_wavelengths * 4 * P->d.height * P->d.width;
  const size_t K = size_t(ckpt_interval(P));
  return (total_steps(S) + K - 1) / K * fsz * cb;
}

size_t dtmmx_adjoint_workspace_size(const dtmmx_plan* P) {
  if (!P) return 0;
  const size_t cb = P->d.precision == DTMMX_C64 ? sizeof(cpx<float>) : sizeof(cpx<double>);
  const size_t HW = size_t(P->d.height) * P->d.width;
  const size_t f = size_t(P->d.n_rays) * P->d.n_wavelengths * 4 * HW * cb;
  const size_t nbeam = size_t(P->d.n_rays / P->d.rays_per_beam);
  const size_t K = size_t(ckpt_interval(P));
  return (f + 255) / 256 * 256 +
         (size_t(P->d.n_rays) * P->d.n_wavelengths * 2 * HW * sizeof(double) + 255) / 256 * 256 +
         (size_t(GRAD_COMPS) * nbeam * HW * cb + 255) / 256 * 256 + (K > 1 ? (K - 1) * f : 0);
}

int dtmmx_transfer_record(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout,
                          void* tape, size_t tape_bytes, int* status, void* stream) {
  {
    const int r = check_transfer(P, S, fin, fout, nullptr, 0, false);
    if (r != DTMMX_OK) return r;
  }
  if (!adjoint_supported(P))
    return fail(DTMMX_E_UNSUPPORTED, "the adjoint is built for 4x4, diffraction 1, npass 1");
  if (S->eff_eps_re)
    return fail(DTMMX_E_UNSUPPORTED, "the adjoint is built for isotropic effective media");
  if (!tape || tape_bytes < dtmmx_tape_size(P, S)) return fail(DTMMX_E_ARG, "tape too small");
  hipStream_t s = (hipStream_t)stream;
  if (P->d.precision == DTMMX_C64) return record_c64(P, S, fin, fout, tape, status, s);
  return record_c128(P, S, fin, fout, tape, status, s);
}

int dtmmx_transfer_adjoint(const dtmmx_plan* P, const dtmmx_stack* S, const void* tape,
                           const void* seed, double* grad, void* ws, size_t ws_bytes, int* status,
                           void* stream) {
  if (!P || !S || !tape || !seed || !grad) return fail(DTMMX_E_ARG, "null argument");
  if (!adjoint_supported(P))
    return fail(DTMMX_E_UNSUPPORTED, "the adjoint is built for 4x4, diffraction 1, npass 1");
  if (S->eff_eps_re)
    return fail(DTMMX_E_UNSUPPORTED, "the adjoint is built for isotropic effective media");
  if (!ws || ws_bytes < dtmmx_adjoint_workspace_size(P))
    return fail(DTMMX_E_ARG, "workspace too small");
  if (S->n_layers > 0 && (!S->thickness || !S->eps_eff_re || !S->eps_re || !S->angles))
    return fail(DTMMX_E_ARG, "null stack array");
  hipStream_t s = (hipStream_t)stream;
  if (P->d.precision == DTMMX_C64) return adjoint_c64(P, S, tape, seed, grad, ws, status, s);
  return adjoint_c128(P, S, tape, seed, grad, ws, status, s);
}

int dtmmx_plan_set_tiles(dtmmx_plan* P, int col_lines, int row_rays) {
  if (!P) return fail(DTMMX_E_ARG, "null plan");
  if (!(col_lines == 0 || col_lines == 1 || col_lines == 2 || col_lines == 4 || col_lines == 8))
    return fail(DTMMX_E_ARG, "col_lines must be 0, 1, 2, 4 or 8");
  if (row_rays < 0 || row_rays > 2) return fail(DTMMX_E_ARG, "row_rays must be 0, 1 or 2");
  P->col_lines = col_lines;
  P->row_rays = row_rays;
  return DTMMX_OK;
}

int dtmmx_plan_get_tiles(const dtmmx_plan* P, int* col_lines, int* row_rays) {
  if (!P || !col_lines || !row_rays) return fail(DTMMX_E_ARG, "null argument");
  *col_lines = P->col_lines;
  *row_rays = P->row_rays;
  return DTMMX_OK;
}

int dtmmx_plan_set_row_split(dtmmx_plan* P, int row_split) {
  if (!P) return fail(DTMMX_E_ARG, "null plan");
  if (row_split < 0 || row_split > 8) return fail(DTMMX_E_ARG, "row_split must be 0 to 8");
  P->row_split = row_split;
  return DTMMX_OK;
}

int dtmmx_plan_get_row_split(const dtmmx_plan* P, int* row_split) {
  if (!P || !row_split) return fail(DTMMX_E_ARG, "null argument");
  *row_split = P->row_split;
  return DTMMX_OK;
}

int dtmmx_transfer_batch(const dtmmx_job* jobs, int n, int* status, void* stream) {
  if (n < 0 || (n > 0 && !jobs)) return fail(DTMMX_E_ARG, "bad job array");
  for (int i = 0; i < n; ++i) {
    const dtmmx_job& j = jobs[i];
    const int r = check_transfer(j.plan, j.stack, j.field_in, j.field_out, j.workspace,
                                 j.workspace_bytes);
    if (r != DTMMX_OK) return fail(r, "job " + std::to_string(i) + ": " + g_err);
  }
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n; ++i) {
    const dtmmx_job& j = jobs[i];
    const int r = j.plan->d.precision == DTMMX_C64
                      ? transfer_c64(j.plan, j.stack, j.field_in, j.field_out, j.workspace, status, s)
                      : transfer_c128(j.plan, j.stack, j.field_in, j.field_out, j.workspace, status,
                                      s);
    if (r != DTMMX_OK) return fail(r, "job " + std::to_string(i) + ": " + g_err);
    TRY(check_finite(j.plan, j.field_out, 1, status, s));
  }
  return DTMMX_OK;
}

int dtmmx_transfer_bulk(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout,
                        void* bulk, void* ws, size_t ws_bytes, int* status, void* stream) {
  {
    const int r = check_transfer(P, S, fin, fout, ws, ws_bytes);
    if (r != DTMMX_OK) return r;
  }
  if (!bulk) return fail(DTMMX_E_ARG, "null bulk buffer");
  if ((P->d.npass != 1 && P->d.method != DTMMX_METHOD_2X2) || P->d.diffraction > 1)
    return fail(DTMMX_E_UNSUPPORTED, "bulk output is built with diffraction 0 or 1 and npass 1 "
                                     "(4x4) or any npass (2x2)");
  hipStream_t s = (hipStream_t)stream;
  const int r = P->d.precision == DTMMX_C64 ? bulk_c64(P, S, fin, fout, bulk, ws, status, s)
                                           : bulk_c128(P, S, fin, fout, bulk, ws, status, s);
  if (r != DTMMX_OK) return r;
  return check_finite(P, fout, 1, status, s);
}

int dtmmx_transfer(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout,
                   void* ws, size_t ws_bytes, int* status, void* stream) {
  {
    const int r = check_transfer(P, S, fin, fout, ws, ws_bytes);
    if (r != DTMMX_OK) return r;
  }
  hipStream_t s = (hipStream_t)stream;
  const int r = P->d.precision == DTMMX_C64 ? transfer_c64(P, S, fin, fout, ws, status, s)
                                           : transfer_c128(P, S, fin, fout, ws, status, s);
  if (r != DTMMX_OK) return r;
  return check_finite(P, fout, 1, status, s);
}

int dtmmx_propagate_layer(const dtmmx_plan* P, const dtmmx_stack* S, const void* fin, void* fout,
                          double e_in_re, double e_in_im, void* ws, size_t ws_bytes, int* status,
                          void* stream) {
  {
    const int r = check_transfer(P, S, fin, fout, ws, ws_bytes);
    if (r != DTMMX_OK) return r;
  }
  if (S->n_layers < 1) return fail(DTMMX_E_ARG, "the layer step needs a layer");
  if (S->eff_eps_re)
    return fail(DTMMX_E_UNSUPPORTED, "the layer step is built for isotropic effective media");
  if (P->d.diffraction != 1 || P->d.npass != 1 || P->wn > 1)
    return fail(DTMMX_E_UNSUPPORTED, "the layer step is built for diffraction 1 and npass 1");
  if ((P->d.method == DTMMX_METHOD_4X4 && P->d.reflection != 1) ||
      (P->d.method == DTMMX_METHOD_2X2 && P->d.reflection > 1))
    return fail(DTMMX_E_UNSUPPORTED, "the layer step is built for 4x4 reflection 1 and 2x2 "
                                     "reflection 0 or 1");
  hipStream_t s = (hipStream_t)stream;
  const cd e_in = mk<double>(e_in_re, e_in_im);
  if (P->d.precision == DTMMX_C64) return layer_c64(P, S, fin, fout, e_in, ws, status, s);
  return layer_c128(P, S, fin, fout, e_in, ws, status, s);
}

int dtmmx_transfer_noalloc(const dtmmx_plan* P, const dtmmx_stack* S, void* fin, void* fout,
                           int* status, void* stream) {
  const size_t need = dtmmx_transfer_workspace_size(P, S);
  if (P && need > DTMMX_NOALLOC_SCRATCH_BYTES)
    return fail(DTMMX_E_ARG, "the transfer's workspace (" + std::to_string(need) +
                                 " bytes) exceeds the no-allocation scratch (" +
                                 std::to_string(size_t(DTMMX_NOALLOC_SCRATCH_BYTES)) +
                                 " bytes): use dtmmx_transfer with a workspace");
  hipStream_t s = (hipStream_t)stream;
  void* ws = nullptr;
  {
    const int r = check_transfer(P, S, fin, fout, nullptr, 0, false);
    if (r != DTMMX_OK) return r;
  }
  if (need) TRY(scratch_address(s, &ws));
  const int r = P->d.precision == DTMMX_C64 ? transfer_c64(P, S, fin, fout, ws, status, s)
                                           : transfer_c128(P, S, fin, fout, ws, status, s);
  if (r != DTMMX_OK) return r;
  return check_finite(P, fout, 1, status, s);
}

int dtmmx_timing_begin(int capacity, int stride) {
  if (capacity <= 0) return fail(DTMMX_E_ARG, "capacity must be positive");
  if (stride <= 0) return fail(DTMMX_E_ARG, "stride must be positive");
  g_timer.stride = stride;
  for (int& c : g_timer.seen) c = 0;
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
  if (precision == DTMMX_C64) return fft2_c64(data, batch, height, width, inverse, s);
  if (precision == DTMMX_C128) return fft2_c128(data, batch, height, width, inverse, s);
  return fail(DTMMX_E_ARG, "bad precision");
}

static int validate_pom(const dtmmx_pom_desc* d) {
  if (!d) return fail(DTMMX_E_ARG, "null descriptor");
  if (d->height <= 0 || d->width <= 0) return fail(DTMMX_E_ARG, "empty grid");
  if (d->diffraction && (!size_supported(d->height) || !size_supported(d->width)))
    return fail(DTMMX_E_UNSUPPORTED, "grid " + std::to_string(d->height) + "x" +
                                         std::to_string(d->width) + " not built");
  if (d->n_wavelengths <= 0 || d->n_wavelengths > MAXL)
    return fail(DTMMX_E_ARG, "n_wavelengths out of range");
  if (!d->k0) return fail(DTMMX_E_ARG, "null k0");
  for (int l = 0; l < d->n_wavelengths; ++l)
    if (!(d->k0[l] > 0.) || !std::isfinite(d->k0[l])) return fail(DTMMX_E_ARG, "k0 must be > 0");
  if (d->n_fields <= 0) return fail(DTMMX_E_ARG, "n_fields must be >= 1");
  // grid y of the combination and k-space kernels: one block row per plane
  if ((long long)d->n_fields * d->n_wavelengths * 4 > 65535)
    return fail(DTMMX_E_ARG, "too many fields x wavelengths");
  if (d->n_pol != 1 && d->n_pol != 2) return fail(DTMMX_E_ARG, "n_pol must be 1 or 2");
  if (d->mode < -1 || d->mode > 1) return fail(DTMMX_E_ARG, "mode must be -1, 0 or 1");
  if (d->pmat && d->n_pmat != 1 && d->n_pmat != d->n_wavelengths)
    return fail(DTMMX_E_ARG, "n_pmat must be 1 or n_wavelengths");
  if (d->precision != DTMMX_C64 && d->precision != DTMMX_C128)
    return fail(DTMMX_E_ARG, "bad precision");
  if (!std::isfinite(d->d) || !std::isfinite(d->betamax) || !std::isfinite(d->eps_re) ||
      !std::isfinite(d->eps_im))
    return fail(DTMMX_E_ARG, "non-finite parameter");
  return DTMMX_OK;
}

struct PomLayout {
  size_t pmat = 0, field = 0, tw_w = 0, tw_h = 0, amax = 0, total = 0;
};

static PomLayout pom_layout(const dtmmx_pom_desc* d) {
  PomLayout L;
  const size_t cs = d->precision == DTMMX_C64 ? 8 : 16;
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t r = o;
    o += (n + 255) / 256 * 256;
    return r;
  };
  L.pmat = take(d->pmat ? size_t(d->n_pmat) * 16 * 16 : 0);
  if (d->diffraction) {
    const size_t HW = size_t(d->height) * d->width;
    L.field = take(size_t(d->n_fields) * d->n_wavelengths * 4 * HW * cs);
    L.tw_w = take(size_t(d->width) * cs);
    L.tw_h = take(size_t(d->height) * cs);
    L.amax = take(size_t(d->n_wavelengths) * d->height * sizeof(int));
  }
  L.total = o;
  return L;
}

// Device tables of the polarizing-microscope diffraction (twiddles of both
// sides, betamax rows) kept per device, grid, wavelengths and betamax for
// the life of the process, so that a repeated call enqueues its kernels
// without uploading anything.  At most POM_TABLES sets; beyond that a call
// uploads its tables into the workspace and waits for the upload.
struct PomTables {
  int dev, prec, H, W;
  std::vector<double> k0;
  double betamax;
  void* tw_w;
  void* tw_h;
  int* amax;
};
static std::mutex g_pom_mu;
static std::vector<PomTables> g_pom_tables;
constexpr size_t POM_TABLES = 16;

// 0: tables found or made, 1: the cache is full, < 0: error.
static int pom_tables(const dtmmx_pom_desc* d, void** tw_w, void** tw_h, int** amax) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::vector<double> k0(d->k0, d->k0 + d->n_wavelengths);
  std::lock_guard<std::mutex> lock(g_pom_mu);
  for (const PomTables& t : g_pom_tables)
    if (t.dev == dev && t.prec == d->precision && t.H == d->height && t.W == d->width &&
        t.k0 == k0 && t.betamax == d->betamax) {
      *tw_w = t.tw_w;
      *tw_h = t.tw_h;
      *amax = t.amax;
      return 0;
    }
  if (g_pom_tables.size() >= POM_TABLES) return 1;
  PomTables t{dev, d->precision, d->height, d->width, k0, d->betamax, nullptr, nullptr, nullptr};
  const size_t cs = d->precision == DTMMX_C64 ? sizeof(cpx<float>) : sizeof(cpx<double>);
  dtmmx_plan_desc pd{};
  pd.height = d->height;
  pd.width = d->width;
  pd.n_wavelengths = d->n_wavelengths;
  pd.k0 = d->k0;
  pd.betamax = d->betamax;
  const std::vector<int> rows = betamax_rows(&pd);
  auto release = [&]() {
    (void)hipFree(t.tw_w);
    (void)hipFree(t.tw_h);
    (void)hipFree(t.amax);
  };
  if (hipMalloc(&t.tw_w, size_t(d->width) * cs) != hipSuccess ||
      hipMalloc(&t.tw_h, size_t(d->height) * cs) != hipSuccess ||
      hipMalloc((void**)&t.amax, rows.size() * sizeof(int)) != hipSuccess) {
    release();
    return fail(DTMMX_E_NOMEM, "pom tables: device allocation failed");
  }
  hipError_t e;
  if (d->precision == DTMMX_C64) {
    auto a = twiddles<float>(d->width), b = twiddles<float>(d->height);
    e = hipMemcpy(t.tw_w, a.data(), a.size() * cs, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.tw_h, b.data(), b.size() * cs, hipMemcpyHostToDevice);
  } else {
    auto a = twiddles<double>(d->width), b = twiddles<double>(d->height);
    e = hipMemcpy(t.tw_w, a.data(), a.size() * cs, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.tw_h, b.data(), b.size() * cs, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess)
    e = hipMemcpy(t.amax, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    release();
    return fail(DTMMX_E_HIP, std::string("pom tables: ") + hipGetErrorString(e));
  }
  *tw_w = t.tw_w;
  *tw_h = t.tw_h;
  *amax = t.amax;
  g_pom_tables.push_back(std::move(t));
  return 0;
}

size_t dtmmx_pom_workspace_size(const dtmmx_pom_desc* d) {
  if (validate_pom(d) != DTMMX_OK) return 0;
  return pom_layout(d).total;
}

int dtmmx_pom_specter(const dtmmx_pom_desc* d, const void* field, void* specter, void* workspace,
                      size_t workspace_bytes, void* stream) {
  TRY(validate_pom(d));
  if (!field || !specter) return fail(DTMMX_E_ARG, "null field or specter");
  const PomLayout L = pom_layout(d);
  if (L.total && (!workspace || workspace_bytes < L.total))
    return fail(DTMMX_E_ARG, "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  char* w = (char*)workspace;
  PomJob j{};
  j.d = d;
  j.field = field;
  j.spec = specter;
  bool upload = false;
  if (d->pmat && d->n_pmat == 1) {
    j.pm0 = d->pmat;  // passed by value in the kernel arguments
  } else if (d->pmat) {
    HIP_TRY(hipMemcpyAsync(w + L.pmat, d->pmat, size_t(d->n_pmat) * 16 * 16,
                           hipMemcpyHostToDevice, s));
    j.pmat = w + L.pmat;
    upload = true;
  }
  std::vector<cpx<float>> tf_w, tf_h;
  std::vector<cpx<double>> td_w, td_h;
  std::vector<int> amax;
  int cached = 1;
  if (d->diffraction) {
    j.wfield = w + L.field;
    cached = pom_tables(d, &j.tw_w, &j.tw_h, &j.amax);
    if (cached < 0) return cached;
  }
  if (d->diffraction && cached == 1) {
    j.tw_w = w + L.tw_w;
    j.tw_h = w + L.tw_h;
    j.amax = (int*)(w + L.amax);
    if (d->precision == DTMMX_C64) {
      tf_w = twiddles<float>(d->width);
      tf_h = twiddles<float>(d->height);
      HIP_TRY(hipMemcpyAsync(j.tw_w, tf_w.data(), tf_w.size() * sizeof(cpx<float>),
                             hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(j.tw_h, tf_h.data(), tf_h.size() * sizeof(cpx<float>),
                             hipMemcpyHostToDevice, s));
    } else {
      td_w = twiddles<double>(d->width);
      td_h = twiddles<double>(d->height);
      HIP_TRY(hipMemcpyAsync(j.tw_w, td_w.data(), td_w.size() * sizeof(cpx<double>),
                             hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(j.tw_h, td_h.data(), td_h.size() * sizeof(cpx<double>),
                             hipMemcpyHostToDevice, s));
    }
    dtmmx_plan_desc pd{};
    pd.height = d->height;
    pd.width = d->width;
    pd.n_wavelengths = d->n_wavelengths;
    pd.k0 = d->k0;
    pd.betamax = d->betamax;
    amax = betamax_rows(&pd);
    HIP_TRY(hipMemcpyAsync(j.amax, amax.data(), amax.size() * sizeof(int), hipMemcpyHostToDevice,
                           s));
    upload = true;
  }
  // the host tables above must outlive their copies
  if (upload) HIP_TRY(hipStreamSynchronize(s));
  return d->precision == DTMMX_C64 ? pom_c64(j, s) : pom_c128(j, s);
}

size_t dtmmx_specter2color_workspace_size(long long n_pixels, int n_wavelengths) {
  if (n_pixels <= 0 || n_wavelengths <= 0) return 0;
  const size_t nb = size_t((n_pixels + NT - 1) / NT);
  return (size_t(n_wavelengths) * 3 * sizeof(double) + 255) / 256 * 256 + nb * sizeof(double);
}

int dtmmx_specter2color(const void* specter, void* rgb, long long n_pixels, int n_wavelengths,
                        const double* cmf, int norm_mode, double norm, int gamma_mode,
                        double gamma, int gray, int precision, void* workspace,
                        size_t workspace_bytes, void* stream) {
  if (!specter || !rgb || !cmf) return fail(DTMMX_E_ARG, "null argument");
  if (n_pixels <= 0 || n_wavelengths <= 0) return fail(DTMMX_E_ARG, "empty specter");
  if (norm_mode < 0 || norm_mode > 2) return fail(DTMMX_E_ARG, "norm_mode must be 0, 1 or 2");
  if (gamma_mode < 0 || gamma_mode > 2) return fail(DTMMX_E_ARG, "gamma_mode must be 0, 1 or 2");
  if (precision != DTMMX_C64 && precision != DTMMX_C128) return fail(DTMMX_E_ARG, "bad precision");
  const size_t need = dtmmx_specter2color_workspace_size(n_pixels, n_wavelengths);
  if (!workspace || workspace_bytes < need) return fail(DTMMX_E_ARG, "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  char* w = (char*)workspace;
  ColorJob j{};
  if (n_wavelengths > COLOR_MAXL) {
    HIP_TRY(hipMemcpyAsync(w, cmf, size_t(n_wavelengths) * 3 * sizeof(double),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    j.cmf = (const double*)w;
  } else {
    j.cmf_h = cmf;
  }
  j.spec = specter;
  j.out = rgb;
  j.n = n_pixels;
  j.nlam = n_wavelengths;
  j.part = (double*)(w + (size_t(n_wavelengths) * 3 * sizeof(double) + 255) / 256 * 256);
  j.norm_mode = norm_mode;
  j.norm = norm;
  j.gamma_mode = gamma_mode;
  j.gamma = gamma;
  j.gray = gray;
  return precision == DTMMX_C64 ? color_c64(j, s) : color_c128(j, s);
}

long long dtmmx_director2data_blocks(long long layer_size) {
  return layer_size > 0 ? (layer_size + NT - 1) / NT : 0;
}

int dtmmx_director2data(const double* director, const unsigned char* mask, long long layer_size,
                        int n_layers, double no, double ne, double nhost, double scale_factor,
                        double* eps_re, double* angles, double* layer_sums, void* stream) {
  if (!director || !eps_re || !angles || !layer_sums) return fail(DTMMX_E_ARG, "null argument");
  if (layer_size <= 0 || n_layers <= 0) return fail(DTMMX_E_ARG, "empty director field");
  if (n_layers > 65535) return fail(DTMMX_E_ARG, "too many layers");
  const long long nb = dtmmx_director2data_blocks(layer_size);
  if (nb > 0x7fffffffLL) return fail(DTMMX_E_ARG, "layer too large");
  if (!std::isfinite(no) || !std::isfinite(ne) || !std::isfinite(scale_factor) ||
      scale_factor == 0.)
    return fail(DTMMX_E_ARG, "bad refractive indices or scale factor");
  if (mask && !std::isfinite(nhost)) return fail(DTMMX_E_ARG, "nhost is required with a mask");
  DirArgs a{};
  a.director = director;
  a.mask = mask;
  a.layer_size = layer_size;
  a.n = layer_size * n_layers;
  a.eo = no * no;
  a.ee = ne * ne;
  a.eh = nhost * nhost;
  a.scale = scale_factor;
  a.eps_re = eps_re;
  a.angles = angles;
  a.sums = layer_sums;
  hipLaunchKernelGGL(director2data_kernel, dim3(unsigned(nb), unsigned(n_layers)), dim3(NT), 0,
                     (hipStream_t)stream, a);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

int dtmmx_q2data(const double* qtensor, const unsigned char* mask, long long layer_size,
                 int n_layers, double no, double ne, double nhost, double scale_factor,
                 int biaxial, double* eps_re, double* angles, double* layer_sums, void* stream) {
  if (!qtensor || !eps_re || !angles || !layer_sums) return fail(DTMMX_E_ARG, "null argument");
  if (layer_size <= 0 || n_layers <= 0) return fail(DTMMX_E_ARG, "empty Q tensor field");
  if (n_layers > 65535) return fail(DTMMX_E_ARG, "too many layers");
  const long long nb = dtmmx_director2data_blocks(layer_size);
  if (nb > 0x7fffffffLL) return fail(DTMMX_E_ARG, "layer too large");
  if (!std::isfinite(no) || !std::isfinite(ne) || !std::isfinite(scale_factor) ||
      scale_factor == 0.)
    return fail(DTMMX_E_ARG, "bad refractive indices or scale factor");
  if (mask && !std::isfinite(nhost)) return fail(DTMMX_E_ARG, "nhost is required with a mask");
  QArgs a{};
  a.q = qtensor;
  a.mask = mask;
  a.layer_size = layer_size;
  a.n = layer_size * n_layers;
  a.epsa = (ne * ne - no * no) / scale_factor;
  a.epsm = (no * no * 2. + ne * ne) / 3.;
  a.eh = nhost * nhost;
  a.biaxial = biaxial ? 1 : 0;
  a.eps_re = eps_re;
  a.angles = angles;
  a.sums = layer_sums;
  hipLaunchKernelGGL(q2data_kernel, dim3(unsigned(nb), unsigned(n_layers)), dim3(NT), 0,
                     (hipStream_t)stream, a);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

static int tmm_run(const dtmmx_tmm_desc* d, int method, double sbeta, double sphi,
                   void* fvec_in, void* fvec_out, int has_out, void* cmat, void* stream) {
  if (!d) return fail(DTMMX_E_ARG, "null descriptor");
  if (d->n_elements <= 0 || d->n_layers <= 0) return fail(DTMMX_E_ARG, "empty stack");
  if (!d->kd || !d->eps_re || !d->angles) return fail(DTMMX_E_ARG, "null layer data");
  if (!fvec_in && !cmat) return fail(DTMMX_E_ARG, "nothing to compute");
  if (fvec_in && !fvec_out) return fail(DTMMX_E_ARG, "fvec_out is required with fvec_in");
  if (method < DTMMX_TMM_4X4 || method > DTMMX_TMM_2X2_1) return fail(DTMMX_E_ARG, "unknown method");
  if (!std::isfinite(d->beta) || !std::isfinite(d->phi) || !std::isfinite(d->n_in) ||
      !std::isfinite(d->n_out) || !std::isfinite(sbeta) || !std::isfinite(sphi))
    return fail(DTMMX_E_ARG, "non-finite beam or media");
  if (method == DTMMX_TMM_4X4_2 || method == DTMMX_TMM_4X4_R) {
    if (!std::isfinite(d->cfact)) return fail(DTMMX_E_ARG, "non-finite cfact");
  }
  if ((d->n_elements + NT - 1) / NT > 0x7fffffffLL) return fail(DTMMX_E_ARG, "too many stacks");
  TmmArgs a{};
  a.n = d->n_elements;
  a.L = d->n_layers;
  a.kd = d->kd;
  a.eps_re = d->eps_re;
  a.eps_im = d->eps_im;
  a.ang = d->angles;
  a.beta = sbeta;
  a.phi = sphi;
  a.cfact = d->cfact;
  a.mbeta = d->beta;
  a.mphi = d->phi;
  a.reflect_in = d->reflect_in ? 1 : 0;
  a.reflect_out = d->reflect_out ? 1 : 0;
  // f_iso of the media at the beam (beta, phi) (tmm.py:1074-1091) and their
  // inverses, on the host
  auto media = [&](double n, cd* f, cd* fi) {
    Modes m;
    iso_modes(d->beta, mk<double>(n * n), m);
    rotate_rows(m, std::cos(d->phi), std::sin(d->phi));
    cd inv[4][4];
    inv4(m.F, inv);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        f[i * 4 + j] = m.F[i][j];
        fi[i * 4 + j] = inv[i][j];
      }
  };
  media(d->n_in, a.fin, a.fini);
  media(d->n_out, a.fout, a.fouti);
  a.fvec_in = (cd*)fvec_in;
  a.fvec_out = (cd*)fvec_out;
  a.has_out = has_out ? 1 : 0;
  a.cmat = (cd*)cmat;
  const dim3 grid(unsigned((d->n_elements + NT - 1) / NT));
  hipStream_t s = (hipStream_t)stream;
  switch (method) {
    case DTMMX_TMM_4X4: hipLaunchKernelGGL((tmm_kernel<0>), grid, dim3(NT), 0, s, a); break;
    case DTMMX_TMM_4X4_1: hipLaunchKernelGGL((tmm_kernel<1>), grid, dim3(NT), 0, s, a); break;
    case DTMMX_TMM_4X4_2: hipLaunchKernelGGL((tmm_kernel<2>), grid, dim3(NT), 0, s, a); break;
    case DTMMX_TMM_4X4_R: hipLaunchKernelGGL((tmm_kernel<3>), grid, dim3(NT), 0, s, a); break;
    case DTMMX_TMM_2X2: hipLaunchKernelGGL((tmm2_kernel<false>), grid, dim3(NT), 0, s, a); break;
    default: hipLaunchKernelGGL((tmm2_kernel<true>), grid, dim3(NT), 0, s, a); break;
  }
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

int dtmmx_tmm_transfer4x4(const dtmmx_tmm_desc* d, void* fvec_in, void* fvec_out, int has_out,
                          void* cmat, void* stream) {
  if (!d) return fail(DTMMX_E_ARG, "null descriptor");
  return tmm_run(d, DTMMX_TMM_4X4, d->beta, d->phi, fvec_in, fvec_out, has_out, cmat, stream);
}

int dtmmx_tmm_transfer(const dtmmx_tmm_desc* d, void* fvec_in, void* fvec_out, int has_out,
                       void* cmat, void* stream) {
  if (!d) return fail(DTMMX_E_ARG, "null descriptor");
  return tmm_run(d, d->method, d->stack_beta, d->stack_phi, fvec_in, fvec_out, has_out, cmat,
                 stream);
}

size_t dtmmx_illumination_workspace_size(const dtmmx_plan* P) {
  if (!P) return 0;
  const size_t HW = size_t(P->d.height) * P->d.width;
  const size_t nblk = (HW + NT - 1) / NT;
  const size_t planes = size_t(P->d.n_rays) * P->d.n_wavelengths;
  return (2 * planes * nblk + 3 * planes) * sizeof(double) + 256;
}

int dtmmx_illumination(const dtmmx_plan* P, const dtmmx_illum_desc* d, void* field,
                       void* workspace, size_t workspace_bytes, void* stream) {
  if (!P || !d || !field || !workspace || !d->intensity) return fail(DTMMX_E_ARG, "null argument");
  if (workspace_bytes < dtmmx_illumination_workspace_size(P))
    return fail(DTMMX_E_ARG, "workspace too small");
  if (!(d->n > 0.) || !std::isfinite(d->n) || !std::isfinite(d->focus))
    return fail(DTMMX_E_ARG, "bad medium index or focus");
  if (d->mode != 1 && d->mode != -1) return fail(DTMMX_E_ARG, "mode must be +1 or -1");
  if (!d->polarized && P->d.rays_per_beam != 2)
    return fail(DTMMX_E_ARG, "unpolarised light needs two rays (x, y) per beam");
  if (d->polarized && P->d.rays_per_beam != 1)
    return fail(DTMMX_E_ARG, "a Jones vector needs one ray per beam");
  if (d->polarized) {
    const double nj = d->jones_re[0] * d->jones_re[0] + d->jones_im[0] * d->jones_im[0] +
                      d->jones_re[1] * d->jones_re[1] + d->jones_im[1] * d->jones_im[1];
    if (!(nj > 0.) || !std::isfinite(nj)) return fail(DTMMX_E_ARG, "bad Jones vector");
  }
  if (d->diffraction && P->d.diffraction != 1)
    return fail(DTMMX_E_ARG, "diffraction needs a plan with diffraction 1");
  if (d->diffraction && !size_supported(P->d.height))
    return fail(DTMMX_E_UNSUPPORTED, "grid size has no transform");
  const int nbeams = P->d.n_rays / P->d.rays_per_beam;
  for (int b = 0; b < nbeams; ++b)
    if (!(d->intensity[b] >= 0.) || !std::isfinite(d->intensity[b]))
      return fail(DTMMX_E_ARG, "bad intensity");
  // 256-byte aligned start inside the workspace
  char* ws = (char*)(((uintptr_t)workspace + 255) & ~uintptr_t(255));
  hipStream_t s = (hipStream_t)stream;
  if (P->d.precision == DTMMX_C64) return illum_c64(P, d, field, ws, s);
  return illum_c128(P, d, field, ws, s);
}

// ---------------------------------------------------------------------------
// 2D non-iterative method (dtmm/tmm2d.py), one wavelength per call
// ---------------------------------------------------------------------------

static size_t al256(size_t b) { return (b + 255) / 256 * 256; }
static char* align_ws(void* ws) { return (char*)(((uintptr_t)ws + 255) & ~uintptr_t(255)); }

// Layers per chunk: all layers of a chunk are built in one launch; their
// per-pixel blocks take at most T2D_CHUNK_BYTES.
static constexpr size_t T2D_CHUNK_BYTES = size_t(256) << 20;

static int t2d_chunk(const dtmmx_tmm2d_desc* d) {
  const size_t per = size_t(d->n_modes) * d->width * 16 * sizeof(cd);
  size_t c = std::max<size_t>(1, T2D_CHUNK_BYTES / std::max<size_t>(per, 1));
  c = std::min<size_t>(c, size_t(T2D_MAXL));
  return int(std::min<size_t>(c, size_t(d->n_layers)));
}

size_t dtmmx_tmm2d_workspace_size(const dtmmx_tmm2d_desc* d) {
  if (!d || d->n_modes <= 0 || d->width <= 0 || d->n_layers <= 0) return 0;
  const size_t N = 4 * size_t(d->n_modes);
  const size_t cb = sizeof(cd);
  const size_t lc = size_t(t2d_chunk(d));
  return 256 + al256(lc * d->n_modes * d->width * 16 * cb) + al256(size_t(d->width) * cb) +
         al256(lc * N * N * cb) + al256((lc + 1) / 2 * N * N * cb) + 2 * al256(N * N * cb);
}

int dtmmx_tmm2d_stack(const dtmmx_tmm2d_desc* d, void* cmat, void* ws, size_t ws_bytes,
                      void* stream) {
  if (!d || !cmat || !ws) return fail(DTMMX_E_ARG, "null argument");
  if (d->width <= 0 || d->n_layers <= 0 || d->n_modes <= 0 || !d->mode_index || !d->thickness ||
      !d->eps_re || !d->angles)
    return fail(DTMMX_E_ARG, "empty or incomplete 2D stack");
  if (d->method != DTMMX_TMM_4X4 && d->method != DTMMX_TMM_4X4_1)
    return fail(DTMMX_E_ARG, "Unsupported method");
  if (!(d->k0 > 0.) || !std::isfinite(d->betay)) return fail(DTMMX_E_ARG, "bad wavenumber or betay");
  if (ws_bytes < dtmmx_tmm2d_workspace_size(d)) return fail(DTMMX_E_ARG, "workspace too small");
  const int W = d->width, n = d->n_modes, N = 4 * n, L = d->n_layers;
  const int lc = t2d_chunk(d);
  const size_t cb = sizeof(cd), NN = size_t(N) * N;
  char* w = align_ws(ws);
  cd* M = (cd*)w;
  w += al256(size_t(lc) * n * W * 16 * cb);
  cd* tw = (cd*)w;
  w += al256(size_t(W) * cb);
  cd* A = (cd*)w;  // the chunk's layer matrices, then tree levels
  w += al256(size_t(lc) * NN * cb);
  cd* B = (cd*)w;
  w += al256(size_t(lc + 1) / 2 * NN * cb);
  cd* P = (cd*)w;  // product of the chunks so far
  w += al256(NN * cb);
  cd* Q = (cd*)w;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(t2d_tw_kernel, dim3(unsigned((W + NT - 1) / NT)), dim3(NT), 0, s, tw, W);
  HIP_TRY(hipGetLastError());
  const size_t stride = size_t(L) * W;
  const dim3 gm(unsigned((N + 15) / 16), unsigned((N + 15) / 16));
  bool have = false;
  for (int l0 = 0; l0 < L; l0 += lc) {
    const int nl = std::min(lc, L - l0);
    const bool last = l0 + nl == L;
    T2dArgs a{};
    a.W = W;
    a.n = n;
    a.idx = d->mode_index;
    a.k0 = d->k0;
    a.betay = d->betay;
    for (int l = 0; l < nl; ++l) a.kd[l] = d->k0 * d->thickness[l0 + l];
    a.nl = nl;
    a.eps_re = d->eps_re + size_t(l0) * W;
    a.eps_im = d->eps_im ? d->eps_im + size_t(l0) * W : nullptr;
    a.ang = d->angles + size_t(l0) * W;
    a.stride = stride;
    a.method = d->method == DTMMX_TMM_4X4_1 ? 1 : 0;
    a.M = M;
    a.tw = tw;
    a.C = A;
    hipLaunchKernelGGL(t2d_mode_kernel, dim3(unsigned((W + NT - 1) / NT), unsigned(n), unsigned(nl)),
                       dim3(NT), 0, s, a);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(t2d_dft_kernel, dim3(unsigned(size_t(nl) * n * n)), dim3(64), 0, s, a);
    HIP_TRY(hipGetLastError());
    // ordered product M_l0 . M_l0+1 ... of the chunk as a tree of pairwise
    // products of neighbours (matrix products are associative)
    cd* src = A;
    cd* dst = B;
    int cnt = nl;
    while (cnt > 1) {
      const int half = cnt / 2;
      hipLaunchKernelGGL(zgemm_kernel, dim3(gm.x, gm.y, unsigned(half)), dim3(256), 0, s,
                         (const cd*)src, 2 * NN, (const cd*)(src + NN), 2 * NN, dst, NN, N);
      HIP_TRY(hipGetLastError());
      if (cnt & 1)
        HIP_TRY(hipMemcpyAsync(dst + size_t(half) * NN, src + size_t(cnt - 1) * NN, NN * cb,
                               hipMemcpyDeviceToDevice, s));
      cnt = half + (cnt & 1);
      std::swap(src, dst);
    }
    // src holds the chunk's product
    if (!have) {
      HIP_TRY(hipMemcpyAsync(last ? (cd*)cmat : P, src, NN * cb, hipMemcpyDeviceToDevice, s));
      have = true;
    } else {
      cd* out = last ? (cd*)cmat : Q;
      hipLaunchKernelGGL(zgemm_kernel, gm, dim3(256), 0, s, (const cd*)P, size_t(0),
                         (const cd*)src, size_t(0), out, size_t(0), N);
      HIP_TRY(hipGetLastError());
      std::swap(P, Q);
    }
  }
  return DTMMX_OK;
}

int dtmmx_tmm2d_system(int n, const void* fmatin, const void* cmat, const void* fmatout,
                       void* smat, void* ws, size_t ws_bytes, void* stream) {
  if (n <= 0 || !fmatin || !cmat || !fmatout || !smat || !ws) return fail(DTMMX_E_ARG, "null argument");
  if (ws_bytes < 256 + al256(size_t(n) * 16 * sizeof(cd))) return fail(DTMMX_E_ARG, "workspace too small");
  cd* fini = (cd*)align_ws(ws);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(t2d_inv_kernel, dim3(unsigned((n + NT - 1) / NT)), dim3(NT), 0, s,
                     (const cd*)fmatin, fini, n);
  HIP_TRY(hipGetLastError());
  const size_t NN = size_t(4 * n) * (4 * n);
  hipLaunchKernelGGL(t2d_system_kernel, dim3(unsigned((NN + NT - 1) / NT)), dim3(NT), 0, s,
                     (const cd*)fini, (const cd*)cmat, (const cd*)fmatout, (cd*)smat, n);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

size_t dtmmx_tmm2d_reflection_workspace_size(int n) {
  if (n <= 0) return 0;
  const size_t N = 4 * size_t(n);
  return 256 + al256(N * 2 * N * sizeof(cd)) + al256(N * sizeof(cd));
}

int dtmmx_tmm2d_reflection(int n, const void* smat, void* rmat, void* ws, size_t ws_bytes,
                           void* stream) {
  if (n <= 0 || !smat || !rmat || !ws) return fail(DTMMX_E_ARG, "null argument");
  if (ws_bytes < dtmmx_tmm2d_reflection_workspace_size(n)) return fail(DTMMX_E_ARG, "workspace too small");
  const int N = 4 * n;
  char* w = align_ws(ws);
  cd* G = (cd*)w;
  w += al256(size_t(N) * 2 * N * sizeof(cd));
  cd* fac = (cd*)w;
  hipStream_t s = (hipStream_t)stream;
  if (N <= GJ_SMALL) {
    hipLaunchKernelGGL(gj_small_kernel, dim3(1), dim3(256), 0, s, (const cd*)smat, (cd*)rmat, N);
    HIP_TRY(hipGetLastError());
    return DTMMX_OK;
  }
  const size_t NG = size_t(N) * 2 * N;
  hipLaunchKernelGGL(gj_init_kernel, dim3(unsigned((NG + NT - 1) / NT)), dim3(NT), 0, s,
                     (const cd*)smat, G, N);
  HIP_TRY(hipGetLastError());
  for (int k = 0; k < N; ++k) {
    hipLaunchKernelGGL(gj_pivot_kernel, dim3(1), dim3(NT), 0, s, G, N, k, fac);
    hipLaunchKernelGGL(gj_elim_kernel, dim3(unsigned((NG + NT - 1) / NT)), dim3(NT), 0, s, G, N, k,
                       (const cd*)fac);
  }
  HIP_TRY(hipGetLastError());
  const size_t NN = size_t(N) * N;
  hipLaunchKernelGGL(gj_extract_kernel, dim3(unsigned((NN + NT - 1) / NT)), dim3(NT), 0, s,
                     (const cd*)G, (cd*)rmat, N);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

int dtmmx_tmm2d_reflect(int n, long long nvec, void* fvec_in, const void* fmatin,
                        const void* rmat, const void* fmatout, void* fvec_out, int has_out,
                        void* ws, size_t ws_bytes, void* stream) {
  if (n <= 0 || nvec <= 0 || !fvec_in || !fmatin || !rmat || !fmatout || !fvec_out || !ws)
    return fail(DTMMX_E_ARG, "null argument");
  const size_t N = 4 * size_t(n);
  const size_t need = 256 + 2 * al256(size_t(n) * 16 * sizeof(cd)) + al256(2 * nvec * N * sizeof(cd));
  if (ws_bytes < need) return fail(DTMMX_E_ARG, "workspace too small");
  char* w = align_ws(ws);
  cd* fini = (cd*)w;
  w += al256(size_t(n) * 16 * sizeof(cd));
  cd* fouti = (cd*)w;
  w += al256(size_t(n) * 16 * sizeof(cd));
  cd* tmp = (cd*)w;
  hipStream_t s = (hipStream_t)stream;
  const dim3 gi(unsigned((n + NT - 1) / NT));
  hipLaunchKernelGGL(t2d_inv_kernel, gi, dim3(NT), 0, s, (const cd*)fmatin, fini, n);
  hipLaunchKernelGGL(t2d_inv_kernel, gi, dim3(NT), 0, s, (const cd*)fmatout, fouti, n);
  HIP_TRY(hipGetLastError());
  const dim3 gv(unsigned((nvec * N + NT - 1) / NT));
  hipLaunchKernelGGL(t2d_reflect_kernel, gv, dim3(NT), 0, s, (cd*)fvec_in, (const cd*)fmatin,
                     (const cd*)fini, (const cd*)rmat, (const cd*)fmatout, (const cd*)fouti,
                     (cd*)fvec_out, has_out ? 1 : 0, n, nvec, tmp);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(t2d_reflect_out_kernel, gv, dim3(NT), 0, s, (cd*)fvec_in, (const cd*)fmatin,
                     (const cd*)fmatout, (cd*)fvec_out, n, nvec, (const cd*)tmp);
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

int dtmmx_tmm2d_modes(int W, int n, const int* idx, long long nvec, const void* field,
                      void* modes, int inverse, void* ws, size_t ws_bytes, void* stream) {
  if (W <= 0 || n <= 0 || nvec <= 0 || !idx || !field || !modes || !ws)
    return fail(DTMMX_E_ARG, "null argument");
  if (ws_bytes < 256 + al256(size_t(W) * sizeof(cd))) return fail(DTMMX_E_ARG, "workspace too small");
  cd* tw = (cd*)align_ws(ws);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(t2d_tw_kernel, dim3(unsigned((W + NT - 1) / NT)), dim3(NT), 0, s, tw, W);
  HIP_TRY(hipGetLastError());
  if (!inverse) {
    const size_t ne = size_t(nvec) * n * 4;
    hipLaunchKernelGGL(t2d_field2modes_kernel, dim3(unsigned((ne + NT - 1) / NT)), dim3(NT), 0, s,
                       (const cd*)field, (cd*)modes, idx, (const cd*)tw, W, n, nvec);
  } else {
    const size_t ne = size_t(nvec) * 4 * W;
    hipLaunchKernelGGL(t2d_modes2field_kernel, dim3(unsigned((ne + NT - 1) / NT)), dim3(NT), 0, s,
                       (const cd*)modes, (cd*)field, idx, (const cd*)tw, W, n, nvec);
  }
  HIP_TRY(hipGetLastError());
  return DTMMX_OK;
}

int dtmmx_host_windows(const dtmmx_plan_desc* d, int max_windows, int* n_windows, int* wi,
                       int* wj, double* beta, double* phi) {
  if (!d || !n_windows) return fail(DTMMX_E_ARG, "null argument");
  {
    const int r = validate_desc(d);
    if (r != DTMMX_OK) return r;
  }
  if (d->diffraction < 2) return fail(DTMMX_E_ARG, "diffraction must be > 1");
  WinTables t;
  build_windows(d, betamax_rows(d), t);
  const int n = int(t.wi.size());
  *n_windows = n;
  if (n > max_windows) return fail(DTMMX_E_ARG, "max_windows too small");
  const int nb = d->n_rays / d->rays_per_beam;
  for (int w = 0; w < n; ++w) {
    if (wi) wi[w] = t.wi[w];
    if (wj) wj[w] = t.wj[w];
    for (int q = 0; q < nb; ++q) {
      if (beta) beta[size_t(w) * nb + q] = t.beta[size_t(w) * nb + q];
      if (phi) phi[size_t(w) * nb + q] = t.phi[size_t(w) * nb + q];
    }
  }
  return DTMMX_OK;
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
                angles[3 * i], angles[3 * i + 1], angles[3 * i + 2], m, kind);
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

int dtmmx_host_rop_grad(int n, const double* beta, const double* phi, const double* eps_re,
                        const double* eps_im, const double* angles, double* alpha, double* proj,
                        double* dalpha, double* dproj) {
  if (n < 0 || !beta || !phi || !eps_re || !angles || !alpha || !proj || !dalpha || !dproj)
    return fail(DTMMX_E_ARG, "null argument");
  int rc = DTMMX_OK;
  auto put = [](double* o, cd z) {
    o[0] = z.x;
    o[1] = z.y;
  };
  for (int i = 0; i < n; ++i) {
    cd e[3];
    for (int c = 0; c < 3; ++c) e[c] = mk<double>(eps_re[3 * i + c], eps_im ? eps_im[3 * i + c] : 0.);
    RopGrad g;
    rop_grad(beta[i], std::cos(phi[i]), std::sin(phi[i]), phi[i], e[0], e[1], e[2],
             angles[3 * i + 1], angles[3 * i + 2], g);
    if (g.kind == KIND_BIAXIAL) rc = fail(DTMMX_E_UNSUPPORTED, "biaxial voxel");
    for (int j = 0; j < 4; ++j) {
      put(alpha + 8 * i + 2 * j, g.a[j]);
      for (int k = 0; k < 2; ++k) put(dalpha + 16 * i + 8 * k + 2 * j, g.da[k][j]);
    }
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        put(proj + 16 * i + 4 * a + 2 * b, g.P0[a][b]);
        put(proj + 16 * i + 8 + 4 * a + 2 * b, g.P1[a][b]);
        for (int k = 0; k < 2; ++k) {
          put(dproj + 32 * i + 16 * k + 4 * a + 2 * b, g.dP0[k][a][b]);
          put(dproj + 32 * i + 16 * k + 8 + 4 * a + 2 * b, g.dP1[k][a][b]);
        }
      }
  }
  return rc;
}

int dtmmx_host_sincos(int n, const float* x, float* s, float* c) {
  if (n < 0 || (n > 0 && (!x || !s || !c))) return fail(DTMMX_E_ARG, "bad sincos arguments");
  for (int i = 0; i < n; ++i) sincos_cw(x[i], s + i, c + i);
  return DTMMX_OK;
}

int dtmmx_host_rop(int n, const double* beta, const double* phi, const double* eps_re,
                   const double* eps_im, const double* angles, double* alpha, double* proj,
                   int* ok) {
  if (n < 0 || !beta || !phi || !eps_re || !angles || !alpha || !proj || !ok)
    return fail(DTMMX_E_ARG, "null argument");
  int rc = DTMMX_OK;
  for (int i = 0; i < n; ++i) {
    cd e[3];
    for (int c = 0; c < 3; ++c) e[c] = mk<double>(eps_re[3 * i + c], eps_im ? eps_im[3 * i + c] : 0.);
    RopData r;
    rop_full(beta[i], std::cos(phi[i]), std::sin(phi[i]), phi[i], e[0], e[1], e[2], angles[3 * i],
             angles[3 * i + 1], angles[3 * i + 2], r);
    for (int j = 0; j < 4; ++j) {
      alpha[8 * i + 2 * j] = r.a[j].x;
      alpha[8 * i + 2 * j + 1] = r.a[j].y;
    }
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        proj[16 * i + 4 * a + 2 * b] = r.P0[a][b].x;
        proj[16 * i + 4 * a + 2 * b + 1] = r.P0[a][b].y;
        proj[16 * i + 8 + 4 * a + 2 * b] = r.P1[a][b].x;
        proj[16 * i + 8 + 4 * a + 2 * b + 1] = r.P1[a][b].y;
      }
    ok[2 * i] = r.ok4 ? 1 : 0;
    ok[2 * i + 1] = r.ok2 ? 1 : 0;
  }
  return rc;
}

}  // extern "C"
#endif  // DTMMX_HOST_PART
