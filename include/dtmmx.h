/*
 * dtmmx -- C ABI of the MI355X diffractive transfer-matrix engine.
 *
 * The reference (IJSComplexMatter/dtmm) exposes this path only as Python:
 * dtmm/transfer.py:322 transfer_field -> :618 transfer_4x4 / :1009
 * transfer_2x2, whose layer loops (transfer.py:752-804, :1117-1174) call
 * dtmm/propagate_4x4.py:376 propagate_4x4_effective_1 and
 * dtmm/propagate_2x2.py:193 propagate_2x2_effective_1.  This header is the
 * boundary those Python functions bind through ctypes.  Plain C types only:
 * device buffers are raw pointers (e.g. torch.Tensor.data_ptr()), streams
 * are hipStream_t passed as void*.
 *
 * Conventions
 *   - Field buffers are complex (interleaved re, im), complex64 when
 *     precision == DTMMX_C64 and complex128 when DTMMX_C128, laid out as
 *     [n_rays][n_wavelengths][4][height][width] (the reference's
 *     (..., wavelengths, 4, H, W) field_data layout, components
 *     Ex, Hy, Ey, Hx; tmm.py:13).
 *   - Material buffers are float64 structure-of-arrays
 *     [3][n_layers][height][width]: eps_re / eps_im are the eigenvalues
 *     of the dielectric tensor (eps_im may be NULL), angles are
 *     (psi, theta, phi) in radians (data.py:534-618 block layout,
 *     transposed component-first).
 *   - Every call returns 0 on success or a negative DTMMX_E* code; the
 *     text of the last error of the calling thread is available from
 *     dtmmx_last_error().  Calls are asynchronous on the given stream.
 *   - A plan is immutable once its launch shapes are set
 *     (dtmmx_plan_set_tiles) and may be used from several threads on
 *     different streams, provided each call gets its own workspace.
 */
#ifndef DTMMX_H
#define DTMMX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DTMMX_ABI_VERSION 5

enum {
  DTMMX_OK = 0,
  DTMMX_E_ARG = -1,         /* invalid argument                          */
  DTMMX_E_UNSUPPORTED = -2, /* valid in the reference, not built here    */
  DTMMX_E_HIP = -3,         /* HIP runtime error                         */
  DTMMX_E_NOMEM = -4        /* device allocation failed                  */
};

enum { DTMMX_C64 = 0, DTMMX_C128 = 1 };
enum { DTMMX_METHOD_2X2 = 2, DTMMX_METHOD_4X4 = 4 };

#define DTMMX_NORM_FFT 1
#define DTMMX_NORM_REF 2
#define DTMMX_MAX_WAVELENGTHS 64
#define DTMMX_MAX_BEAMS 64

/* Static description of a transfer problem: grid, wavelengths and beams.
 * Rays: the field has n_rays outer planes (polarisations x beams);
 * consecutive groups of rays_per_beam rays share one beam (beta, phi). */
typedef struct {
  int height, width;
  int n_wavelengths;
  int n_rays;
  int rays_per_beam;
  int method;     /* DTMMX_METHOD_2X2 or DTMMX_METHOD_4X4              */
  int precision;  /* DTMMX_C64 or DTMMX_C128                            */
  double betamax; /* transfer.py:440, dtmm.ini:6                        */
  const double* k0;   /* [n_wavelengths] host: 2 pi pixelsize / lambda  */
  const double* beta; /* [n_rays / rays_per_beam] host                  */
  const double* phi;  /* [n_rays / rays_per_beam] host                  */
  int diffraction;    /* < 0 (4x4, npass 1): exact per-mode propagation
                         (propagate_4x4.py:515-564; the reference's -1 or
                         inf), every FFT mode with beta < betamax
                         propagated at each pixel with its own beam;
                         0 (npass 1; 4x4 reflection 2, 2x2 reflection 0):
                         no diffraction, per-pixel stacks in real space;
                         1: one mean beam per ray (propagate_4x4.py:382);
                         n > 1 (4x4 only): the k-space field is split by
                         the non-zero n x n tent windows of dtmm/mode.py,
                         each propagated with its window's mean beam
                         (propagate_4x4.py:390-419)                      */
  int npass;          /* >= 1; > 1: alternating forward and backward
                         passes (reflections); 4x4 with the
                         normalisations of transfer.py:815-872, 2x2
                         (reflection 1 with diffraction 1, reflection 2
                         with diffraction 0 or 1) with the reflected wave
                         of every interface kept between passes
                         (transfer.py:1076-1178)                          */
  int norm;           /* 4x4, npass > 1: DTMMX_NORM_FFT (1, per-mode rescaling,
                         transfer.py:78-91), 0 (total intensity,
                         transfer.py:206-221), DTMMX_NORM_REF (2, per-pixel
                         "local" rescaling, transfer.py:103-138) or
                         DTMMX_NORM_FFT | DTMMX_NORM_REF; with
                         DTMMX_NORM_REF the intensities come from a
                         reference field propagated with the 2x2 method
                         in the forward passes (transfer.py:769-775,
                         :805-835; diffraction 0 or 1)                  */
  double smooth;      /* npass > 1: notch width of dtmm/denoise.py (dtmm.ini:8,
                         0.1); negative values count as 0                 */
  int reflection;     /* 4x4: 1, or 2 (real space, transfer.py:719-725; any
                         npass; diffraction > 1: the tent windows' parts,
                         propagate_4x4.py:176-204).
                         2x2: 0; 2 (diffraction 0 or 1: real space with
                         the per-pixel interface transmission into every
                         layer, propagate_2x2.py:91-190); or 1 =
                         Fresnel transmission at
                         interfaces between different effective media
                         (propagate_2x2.py:28-30, diffract.py:215-225)    */
  int checkpoint;     /* adjoint tape (no counterpart in the reference):
                         the forward keeps the row-pass input of every
                         checkpoint-th propagation step and the adjoint
                         recomputes the steps between; 0 or 1: every step */
} dtmmx_plan_desc;

/* A layer stack.  Layer l is propagated as n_steps[l] sub-steps of
 * thickness[l] / n_steps[l] (transfer.py:933-934: (n, (t/n, ...))). */
typedef struct {
  int n_layers;
  const double* thickness;   /* [n_layers] host, pixel units            */
  const double* eps_eff_re;  /* [n_layers] host: isotropic effective eps */
  const double* eps_eff_im;  /* [n_layers] host or NULL                  */
  const double* eps_re;      /* device [3][n_layers][H][W]               */
  const double* eps_im;      /* device [3][n_layers][H][W] or NULL       */
  const double* angles;      /* device [3][n_layers][H][W]               */
  double n_in, n_out;        /* input/output media (transfer.py:936-939) */
  const int* n_steps;        /* [n_layers] host, >= 1; NULL: all 1       */
  const int* biaxial;        /* [n_layers] host: nonzero where the layer
                                has voxels with eps0 != eps1 (numeric
                                eigensolve, dtmm/tmm.py:395-403); NULL:
                                none.  A biaxial voxel in an unflagged
                                layer sets status bit 0 and yields NaN. */
  /* Anisotropic effective media (eff_data 1, 2, per-layer symmetries or a
   * custom effective block; dtmm/transfer.py:926-976, dtmm/data.py:1364-1419).
   * NULL eff_eps_re: the isotropic eps_eff above.  Otherwise layer l's
   * effective medium has eigenvalues eff_eps_re/im[l][0..2] (eff_eps_im may
   * be NULL) and Euler angles eff_angles[l][0..2] (psi, theta, phi), and its
   * diffraction step the thickness eff_thickness[l] (NULL: thickness[l]), all
   * host arrays.  Built for the 4x4 method in reflection mode 1 with
   * diffraction 1 (any npass without DTMMX_NORM_REF, bulk output) and for
   * diffraction 0, where no effective medium enters; other plans fail with
   * DTMMX_E_UNSUPPORTED, as do the adjoint and dtmmx_propagate_layer. */
  const double* eff_eps_re;
  const double* eff_eps_im;
  const double* eff_angles;
  const double* eff_thickness;
} dtmmx_stack;

typedef struct dtmmx_plan dtmmx_plan;

int dtmmx_abi_version(void);
const char* dtmmx_last_error(void);

/* Plan: validates the description, builds twiddle and betamax-mask
 * tables on the current device.  Replaces the cached k-space matrix
 * builders of dtmm/matrix.py:141-174 and dtmm/diffract.py:47-93. */
int dtmmx_plan_create(dtmmx_plan** plan, const dtmmx_plan_desc* desc);
int dtmmx_plan_destroy(dtmmx_plan* plan);

/* Device scratch bytes one transfer call on this plan needs: the npass,
 * window and 2x2 fields where the plan uses them, and the per-pixel
 * operator table of one layer (13 complex values per pixel and beam) that
 * layers flagged biaxial are solved into. */
size_t dtmmx_plan_workspace_size(const dtmmx_plan* plan);

/* Device scratch bytes of a transfer of this stack: the plan's part plus
 * what depends on the stack, the reflected 2-component wave of each of the
 * n_layers + 1 interfaces that the 2x2 method with npass > 1 keeps between
 * passes (the refl array of transfer.py:1081-1083).  Equal to
 * dtmmx_plan_workspace_size for every other plan.  dtmmx_transfer,
 * dtmmx_transfer_bulk and dtmmx_transfer_batch check the workspace against
 * it. */
size_t dtmmx_transfer_workspace_size(const dtmmx_plan* plan, const dtmmx_stack* stack);

/* Launch shapes of the plan's per-layer kernels: col_lines = field lines
 * per column work-group (1, 2, 4, or 8 when two rays share a beam; lines
 * that do not divide the work fall back to the automatic choice),
 * row_rays = rays per row work-group (1, or 2 sharing the per-pixel solve);
 * 0 = chosen from the grid.  Results do not depend on them.  The reference
 * has no counterpart (its numba kernels have no launch shape); they are the
 * knobs of dtmm_amd.tune, which times candidates on the caller's workload. */
int dtmmx_plan_set_tiles(dtmmx_plan* plan, int col_lines, int row_rays);
int dtmmx_plan_get_tiles(const dtmmx_plan* plan, int* col_lines, int* row_rays);

/* Row work-groups per (row, ray group) of the plan's row kernels, each
 * taking a contiguous chunk of the wavelengths (1 to 8; 0 = chosen from the
 * grid and the device's compute units, so that small grids and the
 * two-component 2x2 lines still fill every SIMD).  Results do not depend on
 * it.  No counterpart in the reference; a knob of dtmm_amd.tune. */
int dtmmx_plan_set_row_split(dtmmx_plan* plan, int row_split);
int dtmmx_plan_get_row_split(const dtmmx_plan* plan, int* row_split);

/* Full transfer through the stack.
 *   4x4 (reflection mode 1, diffraction >= 1, npass >= 1): replaces
 *       transfer.py:694-875 with propagate_4x4.py:376-444 and :80-101 per
 *       layer.
 *   4x4 (reflection mode 2, diffraction 1, npass 1): real-space sub-steps
 *       e P/2 F^-1, k-space E diffraction (corrected_Epn_diffraction_matrix,
 *       matrix.py:38-43, :62-68), F P/2 e^-1, replacing
 *       propagate_4x4.py:156-175 with :18-51.
 *   4x4 (reflection mode 2, diffraction 0, npass 1) and 2x2 (reflection
 *       mode 0, diffraction 0, npass 1): per-pixel stacks in real space,
 *       replacing propagate_4x4.py:156-175 with :18-51 (dmatpn None) and
 *       propagate_2x2.py:286-305 with :91-190 (dmat None), the 2x2 input
 *       projection transfer.py:94-101 and output E2H_mat (tmm.py:1051-1072).
 *   2x2 (reflection mode 0 or 1, diffraction 1, npass 1): replaces
 *       transfer.py:1071-1176 with propagate_2x2.py:193-211 and :19-87 per
 *       interface.
 *   2x2 (npass > 1; reflection mode 1 with diffraction 1, mode 2 with
 *       diffraction 0 or 1): the pass loop of transfer.py:1105-1178 with
 *       propagate_2x2.py:19-87 (mode 1, k space) or :91-190 (mode 2, real
 *       space) and the reflected wave of every interface kept between
 *       passes; backward passes propagate the backward modes of each voxel.
 *       Reflection mode 2 keeps field - t.field at each interface, minus
 *       the reflected wave, as the reference does, so its reflected field
 *       has the opposite sign of mode 1's.
 * field_in and field_out are device buffers of n_rays*n_wavelengths*4*H*W
 * complex values; they may not alias.  With npass == 1 field_in is not
 * modified; with npass > 1 it receives the field at the input side
 * (input plus reflection), as the reference writes it into the caller's
 * array (transfer.py:843, :862-866).
 * status (device int[4], may be NULL) receives flags (set, never cleared):
 * status[0] bit 0 = a biaxial voxel was met in a layer not flagged biaxial
 * (its output is NaN); bit 2 = the output holds NaN or infinite values
 * (checked after the transfer, one read of field_out; the reference has no
 * such check). */
int dtmmx_transfer(const dtmmx_plan* plan, const dtmmx_stack* stack, void* field_in,
                   void* field_out, void* workspace, size_t workspace_bytes, int* status,
                   void* stream);

/* One layer step on a k-space field, the layer-level functions of the
 * reference: dtmm/propagate_4x4.py:376 propagate_4x4_effective_1 (4x4,
 * reflection 1; field [n_rays][n_wavelengths][4][H][W]) and
 * dtmm/propagate_2x2.py:193 propagate_2x2_effective_1 (2x2, mode +1, no
 * reflection storage; the E field [n_rays][n_wavelengths][2][H][W]), with
 * diffraction 1 and npass 1.  The stack holds the layer (its sub-steps are
 * n_steps, its isotropic effective medium eps_eff); field_in and field_out
 * are the numpy fft2 of the real-space field (same convention in and out)
 * and may not alias.  2x2 reflection 1: the first sub-step starts with the
 * Fresnel transmission of the tangential E field from the isotropic medium
 * e_in (the effective medium of the layer before, propagate_2x2.py:28-30).
 * workspace: dtmmx_plan_workspace_size bytes. */
int dtmmx_propagate_layer(const dtmmx_plan* plan, const dtmmx_stack* stack, const void* field_in,
                          void* field_out, double e_in_re, double e_in_im, void* workspace,
                          size_t workspace_bytes, int* status, void* stream);

/* dtmmx_transfer for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It has no counterpart in the reference.  The same kernels run with their
 * workspace in a static device array of DTMMX_NOALLOC_SCRATCH_BYTES that the
 * library holds on every GPU it runs on until the process ends; the call
 * allocates nothing, and the caller passes only the plan, the stack, the
 * fields and the stream.  It fails with DTMMX_E_ARG, before any work, when
 * dtmmx_transfer_workspace_size(plan, stack) exceeds the array.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (concurrent
 * streams or threads) race and must be serialised by the caller.  No kernel
 * relies on the array's contents at the start of a call. */
#define DTMMX_NOALLOC_SCRATCH_BYTES (128u << 20)
int dtmmx_transfer_noalloc(const dtmmx_plan* plan, const dtmmx_stack* stack, void* field_in,
                           void* field_out, int* status, void* stream);

/* dtmmx_transfer that also writes the field after every stack layer:
 * bulk is device memory [n_layers][n_rays][n_wavelengths][4][H][W] (field
 * layout, real space, components Ex, Hy, Ey, Hx).  Replaces the ret_bulk
 * output of transfer_4x4 (transfer.py:670-690, :777-778, :876-877), whose
 * bulk_out[1 .. n_layers] these slots are; bulk_out[0] and bulk_out[-1] are
 * field_in and field_out.  The 2x2 method's slots (transfer_2x2,
 * transfer.py:1045-1050, propagate_2x2.py:80-85) hold E after each layer and
 * the H of each voxel's forward modes; with npass > 1 each slot sums that
 * over the forward passes and, over the backward passes, the field entering
 * the layer from the next one with the H of its backward modes
 * (propagate_2x2.py:54-57, :136-142).  Built with diffraction 0 or 1 and
 * npass 1 (4x4: reflection 1 or 2; 2x2: reflection 0, 1 or 2) or any npass
 * (2x2 as dtmmx_transfer). */
int dtmmx_transfer_bulk(const dtmmx_plan* plan, const dtmmx_stack* stack, void* field_in,
                        void* field_out, void* bulk, void* workspace, size_t workspace_bytes,
                        int* status, void* stream);

/* One independent transfer of a batch: its own plan (grid, wavelengths,
 * beams, mode), stack, fields and workspace. */
typedef struct {
  const dtmmx_plan* plan;
  const dtmmx_stack* stack;
  void* field_in;
  void* field_out;
  void* workspace;
  size_t workspace_bytes;
} dtmmx_job;

/* Batched execute for samples of uneven size and depth: runs the n jobs in
 * order on `stream`, as n dtmmx_transfer calls would (the reference has no
 * batch entry point; its users loop transfer_field over samples,
 * transfer.py:322).  Every job is validated before any work is queued;
 * status (device int[4], may be NULL) collects the flags of all jobs. */
int dtmmx_transfer_batch(const dtmmx_job* jobs, int n, int* status, void* stream);

/* Adjoint (reverse-mode) transfer for gradients with respect to the
 * director angles.  The reference has no gradients; this is the adjoint of
 * the 4x4 transfer above (diffraction 1, npass 1), for a real loss
 * L(field_out) of the output field.
 *
 * dtmmx_transfer_record: the forward transfer, keeping the row-pass input
 *   of every desc.checkpoint-th propagation step on `tape` (dtmmx_tape_size
 *   bytes: ceil(steps / checkpoint) fields); the adjoint recomputes the
 *   steps between two checkpoints from the first of them.
 * dtmmx_transfer_adjoint: from seed = dL/dconj(field_out) (device, field
 *   layout; dL = 2 Re <seed, d field_out>) and the tape, adds dL/dtheta and
 *   dL/dphi of every voxel to grad (device float64 [2][n_layers][H][W]:
 *   theta = angles[1], phi = angles[2] of the stack), summed over rays,
 *   wavelengths and sub-steps.  workspace: dtmmx_adjoint_workspace_size. */
size_t dtmmx_tape_size(const dtmmx_plan* plan, const dtmmx_stack* stack);
size_t dtmmx_adjoint_workspace_size(const dtmmx_plan* plan);
int dtmmx_transfer_record(const dtmmx_plan* plan, const dtmmx_stack* stack, const void* field_in,
                          void* field_out, void* tape, size_t tape_bytes, int* status,
                          void* stream);
int dtmmx_transfer_adjoint(const dtmmx_plan* plan, const dtmmx_stack* stack, const void* tape,
                           const void* seed, double* grad, void* workspace,
                           size_t workspace_bytes, int* status, void* stream);

/* In-place unnormalised 2D FFT of `batch` planes of H x W complex values
 * (forward exp(-i..) when inverse == 0), the transform of dtmm/fft.py:277
 * fft2 / :306 ifft2 without the 1/(HW) factor. */
int dtmmx_fft2(void* data, int batch, int height, int width, int inverse, int precision,
               void* stream);

/* Polarizing-microscope specter of transferred fields, the step after the
 * transfer: dtmm/field_viewer.py:70-143 calculate_pom_field (input_fft =
 * output_fft = False, window None) summed over fields as in
 * FieldViewer._calculate_specter (:1262-1289), with the specter of
 * dtmm/field.py:800-845 field2specter / field2spectersum.
 * Per field: the two polarised fields are combined with the Jones weights
 * (n_pol 2, field_viewer.py:122-127); the combination is diffracted by
 * field_diffraction_matrix (dtmm/diffract.py:107-119: F P(k0 d) F^-1 of the
 * isotropic medium eps, zero where beta >= betamax; mode 0 both directions,
 * +1 forward, -1 backward modes only) when diffraction != 0; the analyser
 * matrix pmat (per wavelength or shared) is applied per pixel
 * (field_viewer.py:140-141); the Poynting z component of the result is
 * summed over the n_fields fields into specter[H][W][n_wavelengths] (real:
 * float for DTMMX_C64, double for DTMMX_C128; overwritten). */
typedef struct {
  int height, width;
  int n_wavelengths;  /* <= DTMMX_MAX_WAVELENGTHS                          */
  const double* k0;   /* [n_wavelengths] host: 2 pi pixelsize / lambda     */
  int n_fields;       /* fields summed into the specter                    */
  int n_pol;          /* 1: field [n_fields][n_wavelengths][4][H][W];
                         2: field [n_fields][2][n_wavelengths][4][H][W]     */
  double jones[4];    /* n_pol 2: re, im of the x weight, re, im of the y  */
  int diffraction;    /* 1: apply the diffraction matrix; 0: none          */
  int mode;           /* 0, +1 or -1 (see above)                           */
  double d;           /* focus: propagation distance, pixel units          */
  double eps_re, eps_im; /* isotropic medium (n^2)                         */
  double betamax;
  const double* pmat; /* host [n_pmat][4][4] complex (re, im), or NULL     */
  int n_pmat;         /* 1 (shared) or n_wavelengths                       */
  int precision;      /* DTMMX_C64 or DTMMX_C128                            */
} dtmmx_pom_desc;

size_t dtmmx_pom_workspace_size(const dtmmx_pom_desc* desc);
/* Enqueues its kernels on `stream` and returns.  The first call for a
 * device, grid, set of k0, betamax and precision allocates and uploads its
 * FFT twiddles and betamax rows and keeps them for the life of the process
 * (at most 16 such sets; further sets are uploaded into the workspace and
 * waited for).  A shared analyser travels in the kernel arguments; one per
 * wavelength (n_pmat > 1) is uploaded into the workspace and waited for. */
int dtmmx_pom_specter(const dtmmx_pom_desc* desc, const void* field, void* specter,
                      void* workspace, size_t workspace_bytes, void* stream);

/* Specter to colour (dtmm/color.py:133-200 specter2color): XYZ = specter .
 * cmf (cmf: host [n_wavelengths][3]); norm_mode 0: no normalisation, 1:
 * XYZ / norm, 2: XYZ / max of all XYZ values (norm = True); gray: Y in all
 * three channels, else the sRGB matrix of D65 white; gamma_mode 0: none, 1:
 * sRGB curve (gamma = True), 2: v^(1/gamma) clipped to [0, 1].
 * specter [n_pixels][n_wavelengths] -> rgb [n_pixels][3], real of the
 * given precision (float for DTMMX_C64, double for DTMMX_C128).  The cmf
 * travels in the kernel arguments for up to 128 wavelengths (no upload);
 * a longer one is uploaded into the workspace and waited for. */
size_t dtmmx_specter2color_workspace_size(long long n_pixels, int n_wavelengths);
int dtmmx_specter2color(const void* specter, void* rgb, long long n_pixels, int n_wavelengths,
                        const double* cmf, int norm_mode, double norm, int gamma_mode,
                        double gamma, int gray, int precision, void* workspace,
                        size_t workspace_bytes, void* stream);

/* Optical data of a director field on the device: dtmm/data.py:277-320
 * director2data (order sqrt(|n|) / scale_factor, uniaxial_order of
 * (no, no, ne), host index where mask is 0) and director2angles (psi 0,
 * theta = atan2(sqrt(x^2 + y^2), z), phi = atan2(y, x)).  director: device
 * [n_layers * layer_size][3] double (the reference's (..., 3) layout); mask:
 * device [n_layers * layer_size] bytes or NULL; eps_re and angles: device
 * [3][n_layers * layer_size] (the dtmmx_stack layout).  layer_sums: device
 * [n_layers][dtmmx_director2data_blocks(layer_size)] partial sums of the
 * isotropic part (e0 + e1 + e2) / 3, whose mean is the effective medium of
 * dtmm/transfer.py:34-41.  Enqueued on `stream`. */
long long dtmmx_director2data_blocks(long long layer_size);
int dtmmx_director2data(const double* director, const unsigned char* mask, long long layer_size,
                        int n_layers, double no, double ne, double nhost, double scale_factor,
                        double* eps_re, double* angles, double* layer_sums, void* stream);

/* Optical data of a Q-tensor field on the device: dtmm/data.py:322-373
 * Q2data with Q2eps (data.py:432-487), eps2epsva (data.py:489-513) and the
 * eigenvalue order of tensor_eig (dtmm/linalg.py:172-222, the most distinct
 * eigenvalue at index 2); biaxial = 0 applies uniaxial_order(1, ...).
 * qtensor: device [n_layers * layer_size][6] double (Q00, Q11, Q22, Q01,
 * Q02, Q12, the reference's (..., 6) layout); the other arguments as in
 * dtmmx_director2data (layer_sums has dtmmx_director2data_blocks(layer_size)
 * entries per layer).  The angles describe the same eigenframe as the
 * reference's up to the sign of each eigenvector.  Enqueued on `stream`. */
int dtmmx_q2data(const double* qtensor, const unsigned char* mask, long long layer_size,
                 int n_layers, double no, double ne, double nhost, double scale_factor,
                 int biaxial, double* eps_re, double* angles, double* layer_sums, void* stream);

/* Non-iterative 1D transfer matrix method, method "4x4" of
 * dtmm/tmm.py:1565-1673 transfer4x4 for n_elements independent stacks of
 * n_layers homogeneous layers each (every stack has its own kd, eps and
 * angles per layer), one beam (beta, phi) between isotropic input and output
 * media n_in, n_out: the characteristic matrix C = M_1 ... M_n with
 * M = F P(-kd) F^-1 (layer_mat, stack_mat, :1112-1296) and the boundary
 * solve of transmit4x4 (:1491-1563, system_mat :1298, reflection_mat
 * :1338).  fvec_in [n][4] complex128 (device): the incident field, replaced
 * by incident + reflected; fvec_out [n][4] complex128 (device): the
 * transmitted field, whose backward waves are read first as incoming from
 * the output side when has_out != 0; cmat [n][4][4] complex128 (device) or
 * NULL receives C; fvec_in may be NULL when only cmat is wanted.  fp64
 * throughout.  Enqueued on `stream`. */
typedef struct {
  long long n_elements;
  int n_layers;
  const double* kd;      /* device [n_layers][n_elements]: k0 * thickness   */
  const double* eps_re;  /* device [3][n_layers][n_elements]               */
  const double* eps_im;  /* device [3][n_layers][n_elements] or NULL       */
  const double* angles;  /* device [3][n_layers][n_elements]: psi, theta, phi */
  double beta, phi;
  double n_in, n_out;
  /* read by dtmmx_tmm_transfer only */
  int method;                 /* DTMMX_TMM_4X4 ... DTMMX_TMM_2X2_1 */
  double cfact;               /* 4x4_2, 4x4_r: backward-mode damping (stack_mat's 0.01) */
  double stack_beta, stack_phi;  /* beam of the layers (the reference's transfer
                                    builds them at its default 0, 0) */
  int reflect_in, reflect_out;   /* 2x2 methods: Fresnel at the first / last interface */
} dtmmx_tmm_desc;

int dtmmx_tmm_transfer4x4(const dtmmx_tmm_desc* desc, void* fvec_in, void* fvec_out, int has_out,
                          void* cmat, void* stream);

/* The non-iterative 1D methods of dtmm/tmm.py (replaces layer_mat
 * :1112-1213, stack_mat :1215-1296, transfer4x4 :1565-1673, transfer2x2
 * :1675-1695 and transfer :1697-1748 for every element of the inputs'
 * broadcast shape).  desc->method selects:
 *   DTMMX_TMM_4X4, _4X4_1 (backward phases zeroed), _4X4_2 (backward modes
 *   damped by cfact), _4X4_R (8x8 characteristic matrix): transmit4x4 as
 *   dtmmx_tmm_transfer4x4, with the layers at (stack_beta, stack_phi) and the
 *   isotropic media at (beta, phi); cmat [n][4][4] ([n][8][8] for _4X4_R);
 *   _2X2, _2X2_1 (layers entered through Eti_mat): cmat [n][2][2] is
 *   M_n ... M_1; with fvec_in, fvec_out [n][4] receives E2fvec(C fvec2E(fvec_in))
 *   in the output medium, with the Fresnel t_mat of the first / last
 *   interface when reflect_in / reflect_out (those layers' modes at the
 *   media beam); fvec_in is read only and has_out is ignored.
 * fp64 throughout, enqueued on `stream`. */
#define DTMMX_TMM_4X4 0
#define DTMMX_TMM_4X4_1 1
#define DTMMX_TMM_4X4_2 2
#define DTMMX_TMM_4X4_R 3
#define DTMMX_TMM_2X2 4
#define DTMMX_TMM_2X2_1 5
int dtmmx_tmm_transfer(const dtmmx_tmm_desc* desc, void* fvec_in, void* fvec_out, int has_out,
                       void* cmat, void* stream);

/* Illumination field on the device (replaces dtmm/field.py:586-713
 * illumination_data, with illumination_waves :459-475, waves2field2
 * :566-582, planewave wave.py:376-409 and wave2eigenwave wave.py:432-476).
 * For every ray (beam x polarisation) and wavelength of `plan`: the beam's
 * plane wave, normalised to 1 at pixel (0, 0), times `window`, or, when
 * window is NULL, snapped to the strongest FFT eigen-wave of the grid (unit
 * amplitude); carried by the normalised forward (mode +1) or backward (-1)
 * field vector of the isotropic medium of index n with the ray's Jones
 * vector (polarized: `jones` for every ray; otherwise x for the first and y
 * for the second ray of each beam, rays_per_beam 2).  With diffraction the
 * field is then multiplied per FFT mode by F P(-focus) F^-1 of the medium on
 * the modes of its direction (zero for beta >= the plan's betamax) and
 * rescaled to the power it had before.  Every plane is finally scaled by
 * sqrt(intensity[beam] / n_beams).  The plan's method and reflection are not
 * used; diffraction needs a plan with diffraction 1 (its transforms).
 * field [n_rays][n_wavelengths][4][H][W] in the plan's precision (device),
 * window [H][W] float64 (device), intensity on the host.  Enqueued on
 * `stream`; the workspace is dtmmx_illumination_workspace_size bytes. */
typedef struct {
  double n;
  double focus;
  int mode;
  int diffraction;
  int polarized;
  double jones_re[2], jones_im[2];
  const double* intensity;  /* host [n_beams] */
  const double* window;     /* device [H][W] or NULL */
} dtmmx_illum_desc;

size_t dtmmx_illumination_workspace_size(const dtmmx_plan* plan);
int dtmmx_illumination(const dtmmx_plan* plan, const dtmmx_illum_desc* desc, void* field,
                       void* workspace, size_t workspace_bytes, void* stream);

/* The non-iterative 4x4 method for 2D data (replaces dtmm/tmm2d.py:20-263)
 * for one wavelength, in the space of the n masked 1D FFT modes of the x
 * axis (mode_index: their FFT bins, ascending, on the device).  Block
 * matrices are flat 4n x 4n complex128 row-major: element (4 i + a, 4 j + b)
 * is component (a, b) of block (i, j) of the reference's (n, n, 4, 4).
 * fp64 throughout; every call is enqueued on `stream`.
 *
 * dtmmx_tmm2d_stack: stack_mat2d (:94-119) with layer_mat2d (:20-92): per
 * layer, block column j is the masked FFT of F P(-k0 d) F^-1 (at mode j's
 * beam: beta_x of its bin, betay) times mode j's eigen-wave, and the stack is
 * M_1 . M_2 ... M_n; method DTMMX_TMM_4X4 or DTMMX_TMM_4X4_1.
 * dtmmx_tmm2d_system: system_mat2d (:144-159), fmatin/fmatout [n][4][4].
 * dtmmx_tmm2d_reflection: reflection_mat2d (:161-182) of a system matrix.
 * dtmmx_tmm2d_reflect: reflect2d (:184-232) for nvec sets of mode vectors
 * [nvec][n][4]; fvec_in receives incident + reflected, fvec_out the
 * transmitted field (its backward waves are read first when has_out).
 * dtmmx_tmm2d_modes: field2modes1 (dtmm/field.py:216-250) of a field
 * [nvec][4][W] into modes [nvec][n][4], or with inverse modes2field1
 * (:307-340). */
typedef struct {
  int width;
  int n_layers;
  int n_modes;
  const int* mode_index;     /* device [n_modes] */
  double k0;
  double betay;
  const double* thickness;   /* host [n_layers] */
  const double* eps_re;      /* device [3][n_layers][width] */
  const double* eps_im;      /* device [3][n_layers][width] or NULL */
  const double* angles;      /* device [3][n_layers][width] */
  int method;
} dtmmx_tmm2d_desc;

size_t dtmmx_tmm2d_workspace_size(const dtmmx_tmm2d_desc* desc);
int dtmmx_tmm2d_stack(const dtmmx_tmm2d_desc* desc, void* cmat, void* workspace,
                      size_t workspace_bytes, void* stream);
int dtmmx_tmm2d_system(int n_modes, const void* fmatin, const void* cmat, const void* fmatout,
                       void* smat, void* workspace, size_t workspace_bytes, void* stream);
size_t dtmmx_tmm2d_reflection_workspace_size(int n_modes);
int dtmmx_tmm2d_reflection(int n_modes, const void* smat, void* rmat, void* workspace,
                           size_t workspace_bytes, void* stream);
int dtmmx_tmm2d_reflect(int n_modes, long long nvec, void* fvec_in, const void* fmatin,
                        const void* rmat, const void* fmatout, void* fvec_out, int has_out,
                        void* workspace, size_t workspace_bytes, void* stream);
int dtmmx_tmm2d_modes(int width, int n_modes, const int* mode_index, long long nvec,
                      const void* field, void* modes, int inverse, void* workspace,
                      size_t workspace_bytes, void* stream);

/* Optional kernel timing (no counterpart in the reference, which has no
 * profiling hooks): between begin and end every stride-th launch of each
 * per-layer kernel class is bracketed by HIP events on its stream (a stride
 * above 1 keeps the event overhead out of the timed work).  end() waits for
 * the last event and returns, per kernel class, the summed milliseconds and
 * the count of the timed launches
 * (class 0: row layer kernel including its per-pixel fp64 solve, class 1:
 * column k-space kernel, class 3: dense k-space head/tail kernel; class 2
 * is unused).  Not thread-safe; for benchmarking. */
int dtmmx_timing_begin(int capacity, int stride);
int dtmmx_timing_end(double* ms, int* count, int nclass);

/* Host evaluation of the k-space windows a plan with diffraction > 1 uses
 * (dtmm/mode.py:119-139 fft_mask with the beam offsets of
 * propagate_4x4.py:401-402), for checks without a GPU: the non-zero
 * windows in the reference's order, as tent indices (wi, wj) of the
 * diffraction x diffraction grid, and their mean beams beta, phi
 * [window][beam].  *n_windows receives the count; fails with DTMMX_E_ARG if
 * it exceeds max_windows.  Output arrays may be NULL. */
int dtmmx_host_windows(const dtmmx_plan_desc* desc, int max_windows, int* n_windows, int* wi,
                       int* wj, double* beta, double* phi);

/* Host evaluation of the per-pixel eigen-solution used on the device
 * (dtmm/tmm.py:435 alphaf), for checks without a GPU.  Inputs are n
 * voxels (angles = psi, theta, phi); outputs alpha[n][4] and fmat[n][4][4]
 * as interleaved complex128.  Biaxial voxels: numeric eigenvectors of unit
 * 2-norm with the largest component real (numpy.linalg.eig's scaling);
 * within a direction the two modes follow the eigenvalue order of the
 * solver, which the layer operator does not depend on. */
int dtmmx_host_alphaf(int n, const double* beta, const double* phi, const double* eps_re,
                      const double* eps_im, const double* angles, double* alpha, double* fmat);

/* Host evaluation of the per-pixel data of the real-space layer operator
 * E diag(exp(i kd alpha)) E^-1 (dtmm/propagate_4x4.py:88-97,
 * dtmm/propagate_2x2.py:33-39) written as p2 I + (p0 - p2) P0 on
 * (E+x, E+y) and p3 I + (p1 - p3) P1 on (E-x, E-y): alpha[n][4] and
 * proj[n][2][2][2] (P0, P1) as interleaved complex128, ok[n][2] = E
 * invertible for the 4x4 / 2x2 method.  The same routines run in the row
 * kernel (isotropic, uniaxial) and in the biaxial layer-table kernel; this
 * entry point exists for checks without a GPU. */
int dtmmx_host_rop(int n, const double* beta, const double* phi, const double* eps_re,
                   const double* eps_im, const double* angles, double* alpha, double* proj,
                   int* ok);

/* Host evaluation of the float sine and cosine the complex64 layer kernels
 * use for every phase factor exp(i kd alpha) (dtmm/tmm.py:582-585 phase_mat
 * computes them in the reference's precision): s[i], c[i] of x[i] for n
 * values, for checks without a GPU. */
int dtmmx_host_sincos(int n, const float* x, float* s, float* c);

/* dtmmx_host_rop with the theta and phi derivatives (no counterpart in the
 * reference): dalpha[n][2][4] and dproj[n][2][2][2][2] ([direction][P0|P1])
 * as interleaved complex128; the routine the adjoint kernels run.
 * Isotropic and uniaxial voxels only (DTMMX_E_UNSUPPORTED for biaxial). */
int dtmmx_host_rop_grad(int n, const double* beta, const double* phi, const double* eps_re,
                        const double* eps_im, const double* angles, double* alpha, double* proj,
                        double* dalpha, double* dproj);

#ifdef __cplusplus
}
#endif

#endif /* DTMMX_H */
