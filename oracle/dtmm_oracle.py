"""Complex128 NumPy restatement of the reference diffractive transfer-matrix path.

TEST INFRASTRUCTURE ONLY -- the checker for ``dtmm_amd``; never shipped or
called by the product.  All citations are to the reference package
(IJSComplexMatter/dtmm v0.7.0.dev), paths relative to its repository root.

Scope (what the GPU engine implements and this module checks):

* wave-vector grids ``betaphi``/``betaxy``/``k0``          (dtmm/wave.py)
* per-pixel Berreman eigen-solution ``alphaf``             (dtmm/tmm.py, rotation.py)
* isotropic-medium k-space diffraction/correction matrices (dtmm/diffract.py, matrix.py)
* ``transfer_4x4`` with reflection=1, diffraction >= 1, npass >= 1 (dtmm/transfer.py,
  dtmm/propagate_4x4.py); diffraction > 1 uses the tent windows of dtmm/mode.py;
  npass > 1 the normalisations of transfer.py and dtmm/denoise.py
* ``transfer_4x4`` with reflection=2, any finite diffraction and npass:
  real-space sub-steps of dtmm/propagate_4x4.py _transfer_ray_4x4_2 (per-pixel
  f.P.f^-1 without diffraction; e.P/2.f^-1, k-space E diffraction,
  f.P/2.e^-1 with it; the tent windows' real-space parts for diffraction > 1)
  and the pass normalisations in real space
* ``transfer_2x2`` with npass=1: reflection 0 or 1 with diffraction 1,
  reflection 0 with diffraction 0, and reflection 2 (real space, the
  interface transmission Eti_mat into every layer) with diffraction 0 or 1
  (dtmm/transfer.py, dtmm/propagate_2x2.py); reflection 1 applies the Fresnel
  interface transmission of dtmm/tmm.py tr_mat; ret_bulk for both methods
* the illumination rays ``illumination_rays`` / ``aperture2rays`` /
  ``aperture``                                             (dtmm/field.py, window.py)
* the polarizing-microscope specter after the transfer: ``calculate_pom_field``,
  ``field2specter``, ``pom_specter`` (dtmm/field_viewer.py, field.py), the Jones
  4x4 matrices (dtmm/jones.py, jones4.py) and ``specter2color`` (dtmm/color.py)
* the non-iterative 1D transfer matrix method (method "4x4"): ``layer_mat``,
  ``stack_mat``, ``system_mat``, ``reflection_mat``, ``transmit4x4``,
  ``transfer4x4``                                          (dtmm/tmm.py)
* ``illumination_data`` and ``field2intensity``            (dtmm/field.py),
  with the eigen-wave snapping of ``window=None``          (dtmm/wave.py)

Pinning of each function is done in ``tests/test_oracle_*.py`` (golden
vectors of dtmm/test/test_wave.py, closed forms, invariants, brute force).
The biaxial eigen branch (``_alphaf_biaxial``) is pinned only by the Berreman
eigen-equation and the Poynting-sort invariants: the reference's own tests
hold no vector for it.
"""
import numpy as np

__all__ = [
    "BETAMAX", "N_COVER", "k0", "betaphi", "betaxy", "alphaf", "alphaffi",
    "inv", "dotmm", "dotmf", "dotmdm", "dotmdmf", "dotmv", "phase_mat", "E_mat", "E2H_mat",
    "normalize_f", "field2intensity", "fft2", "ifft2", "refind2eps",
    "uniaxial_order", "validate_optical_data", "create_layers", "effective_block",
    "eig_symmetry", "eps2epsva",
    "director2Q", "Q2eps", "tensor_eig", "rotation_angles", "Q2data", "epsva2eps",
    "projection_matrix", "field_diffraction_matrix", "diffract",
    "transmitted_field", "first_corrected_Epn_diffraction_matrix",
    "second_corrected_Epn_diffraction_matrix", "first_E_diffraction_matrix",
    "second_E_diffraction_matrix", "propagate_4x4_effective_1",
    "propagate_2x2_effective_1", "transfer_4x4", "transfer_2x2",
    "transfer_field", "illumination_data", "blackman", "planewave",
    "jonesvec", "eigenmask", "eigenindices", "eigenwave", "wave2eigenwave", "avec", "fvec", "auxiliary_matrix", "poynting", "set_fft_workers",
    "fft_window", "fft_windows", "fft_betaxy_mean", "fft_mask",
    "DTMM_NORM_FFT", "DTMM_NORM_REF", "normalize_f", "fvec2avec", "f_iso",
    "polarizer", "as4x4", "ray_jonesmat4x4", "jones2H", "project_normalized_local", "S_mat", "Etri_mat", "tr_mat", "E_tr_matrix", "exp_notch_filter", "denoise_fftfield", "denoise_field",
    "normalize_field", "project_normalized_fft", "project_normalized_total", "total_intensity",
    "polarizer", "as4x4", "jonesmat4x4", "ray_jonesmat4x4", "field2specter",
    "calculate_pom_field", "pom_specter", "XYZ2RGBD65", "apply_gamma", "apply_srgb_gamma",
    "specter2color", "f_iso", "layer_mat", "stack_mat", "system_mat", "reflection_mat",
    "transmit4x4", "transfer4x4", "transfer2x2", "transfer1d", "projection_mat", "t_mat",
    "fvec2E", "E2fvec", "TMM_METHODS", "betax1", "eigenmask1", "eigenbetax1", "eigenindices1",
    "eigenwave1", "field2modes1", "modes2field1", "bdotmm", "bdotmd", "bdotdm", "layer_mat2d",
    "stack_mat2d", "f_iso2d", "system_mat2d", "reflection_mat2d", "reflect2d", "transfer2d",
    "Eti_mat", "corrected_E_diffraction_matrix",
    "tukey", "aperture", "illumination_aperture", "aperture2rays", "illumination_rays",
    "cauchy2eps", "sellmeier2eps", "is_optical_data_dispersive", "mean_betaphi",
    "propagate_4x4_full",
    "field2betaphi",
]

# dtmm/dtmm.ini:6 (betamax = 0.8) and dtmm/conf.py:148, :503-510 (nin/nout
# default to n_cover = 1.5, dtmm.ini "n_cover = 1.5").
BETAMAX = 0.8
N_COVER = 1.5
SMOOTH = 0.1

CDTYPE = np.complex128
FDTYPE = np.float64


# ---------------------------------------------------------------------------
# wave.py
# ---------------------------------------------------------------------------

def k0(wavelength, pixelsize=1.):
    """Vacuum wave number in pixel units (dtmm/wave.py:280-296)."""
    return np.asarray(2 * np.pi / np.asarray(wavelength) * pixelsize, dtype=FDTYPE)


def betaphi(shape, k0):
    """beta, phi of the FFT eigen-waves (dtmm/wave.py:54-81).

    beta = (2 pi / |k0|) * sqrt(fx^2 + fy^2) with fftfreq grids, phi = atan2(fy, fx).
    """
    k0 = np.abs(np.asarray(k0, FDTYPE)[..., None, None])
    ay, ax = [np.asarray(np.fft.fftfreq(x), dtype=FDTYPE) for x in shape[-2:]]
    xx, yy = np.meshgrid(ax, ay, copy=False, indexing="xy")
    beta = (2 * np.pi / k0) * np.sqrt(xx ** 2 + yy ** 2)
    phi = np.arctan2(yy, xx)
    return beta, phi


def betaxy(shape, k0):
    """betax, betay of the FFT eigen-waves (dtmm/wave.py:84-110)."""
    k0 = np.asarray(k0, dtype=FDTYPE)[..., None, None]
    ay, ax = [np.asarray(np.fft.fftfreq(x), dtype=FDTYPE) for x in shape[-2:]]
    xx, yy = np.meshgrid(ax, ay, copy=False, indexing="xy")
    l = 2 * np.pi / k0
    return l * xx, l * yy


def planewave(shape, k0, beta, phi):
    """Plane wave normalised to 1 at pixel (0,0) (dtmm/wave.py:376-409)."""
    k0 = np.asarray(k0)[..., None, None]
    beta = np.asarray(beta)[..., None, None]
    phi = np.asarray(phi)[..., None, None]
    ay, ax = [np.arange(-l // 2 + 1., l // 2 + 1.) for l in shape[-2:]]
    xx, yy = np.meshgrid(ax, ay, indexing="xy", copy=False)
    kx = np.asarray(k0 * beta * np.cos(phi), dtype=FDTYPE)
    ky = np.asarray(k0 * beta * np.sin(phi), dtype=FDTYPE)
    out = np.exp(1j * (kx * xx + ky * yy))
    return out / out[..., 0, 0][..., None, None]


def eigenmask(shape, k0, betamax=BETAMAX):
    """FFT modes with beta <= betamax (dtmm/wave.py:256-277)."""
    b, _ = betaphi(shape, k0)
    return b <= betamax


def eigenindices(shape, k0, betamax=BETAMAX):
    """(row, column) indices of the modes of eigenmask, one array per wave
    number (dtmm/wave.py:201-225)."""
    mask = eigenmask(shape, k0, betamax)
    if mask.ndim == 2:
        return np.argwhere(mask)
    return tuple(np.argwhere(m) for m in mask)


def eigenwave(shape, i, j, amplitude=None):
    """Plane wave of FFT mode (i, j); by default of unit amplitude in real
    space (dtmm/wave.py:140-171)."""
    f = np.zeros(shape, CDTYPE)
    if amplitude is None:
        amplitude = np.multiply.reduce(shape[-2:])
    f[..., i, j] = amplitude
    return np.fft.ifft2(f)


def wave2eigenwave(wave):
    """Each 2D wave replaced by the eigen-wave of its strongest FFT mode,
    with the same total power and a real positive coefficient
    (dtmm/wave.py:432-476)."""
    wave = np.asarray(wave)
    shape = wave.shape
    flat = wave.reshape((-1,) + shape[-2:])
    out = np.empty(flat.shape, CDTYPE)
    for k, a in enumerate(np.abs(np.fft.fft2(flat)) ** 2):
        ii, jj = np.unravel_index(np.argmax(a), a.shape)
        out[k] = eigenwave(a.shape, ii, jj, amplitude=a.sum() ** 0.5)
    return out.reshape(shape)


# ---------------------------------------------------------------------------
# window.py
# ---------------------------------------------------------------------------

def blackman(shape):
    """2D radial Blackman window (dtmm/window.py:17-54)."""
    ny, nx = [l / 2 for l in shape]
    ay, ax = [np.arange(-l / 2. + .5, l / 2. + .5) for l in shape]
    xx, yy = np.meshgrid(ax, ay, indexing="xy")
    r = ((xx / nx) ** 2 + (yy / ny) ** 2) ** 0.5
    out = 0.42 + 0.5 * np.cos(np.pi * r) + 0.08 * np.cos(2 * np.pi * r)
    out[r >= 1.] = 0.
    return np.asarray(out, FDTYPE)


def mean_betaphi(field, k0):
    """Power-weighted mean beam of a field (..., 4, H, W): the spectrum of
    the Blackman-windowed field weights the (beta_x, beta_y) of every FFT
    mode, summed over the components; beta = |mean|, phi = its angle
    (dtmm/field.py:849-872, _fft_betaphi).  k0 broadcasts against the
    leading axes (one per wavelength)."""
    field = np.asarray(field)
    f = fft2(field * blackman(field.shape[-2:]))
    bx, by = betaxy(field.shape[-2:], k0)
    s = (f * np.conj(f)).real
    sx = (s * bx[..., None, :, :]).sum(axis=(-3, -2, -1))
    sy = (s * by[..., None, :, :]).sum(axis=(-3, -2, -1))
    return np.sqrt(sx ** 2 + sy ** 2) / s.sum(axis=(-3, -2, -1)), np.arctan2(sy, sx)


def field2betaphi(field_in, ks, multiray=False):
    """The beam transfer_field uses when beta and phi are not given
    (dtmm/field.py:492-500, transfer.py:662-664): the mean of mean_betaphi
    over every wavelength and plane, or per ray (first axis) with multiray
    and more than one leading axis."""
    beta, phi = mean_betaphi(field_in, ks)
    field_in = np.asarray(field_in)
    if field_in.ndim > 4 and multiray:
        axes = tuple(range(1, field_in.ndim - 3))
        return beta.mean(axis=axes), phi.mean(axis=axes)
    return beta.mean(), phi.mean()


# ---------------------------------------------------------------------------
# mode.py: tent windows in (beta_x, beta_y) for diffraction > 1
# ---------------------------------------------------------------------------

def fft_window(fbetax, fbetay, betax, betay, step, xtyp, ytyp, betamax):
    """One tent window on the FFT grid (dtmm/mode.py:10-50).

    The tent is 1 - |d| with d = (beta_x - centre) / step, clipped at 0; the
    first (xtyp < 0) and last (xtyp > 0) tents are flat (1) on their outer
    side within one step; zero where beta >= betamax.  `betax`, `betay` (the
    tent centres) broadcast over leading axes like the reference's gufunc
    loop dimensions."""
    bx = np.asarray(betax, FDTYPE)[..., None, None]
    by = np.asarray(betay, FDTYPE)[..., None, None]

    def tent(f, c, typ):
        dd = (f - c) / step
        cc = 1. - np.abs(dd)
        if typ < 0:
            cc = np.where(dd < 0, 1., cc)
        elif typ > 0:
            cc = np.where(dd > 0, 1., cc)
        return np.where(1. - np.abs(dd) < 0., 0., cc)

    cx = tent(fbetax, bx, xtyp)
    cy = tent(fbetay, by, ytyp)
    beta = (fbetax ** 2 + fbetay ** 2) ** 0.5
    coeff = np.where(beta >= betamax, 0., cx * cy)
    return np.where(coeff > 0., coeff, 0.)


def fft_windows(betax, betay, n, betax_off=0., betay_off=0., betamax=BETAMAX):
    """n*n tent windows, index i*n + j for centre (ax[i], ay[j]) shifted by the
    beam offset reduced to the nearest multiple of the spacing
    (dtmm/mode.py:52-102)."""
    d = 2 * betamax / (n - 1)

    def reduce(off):
        o = np.mod(off, d)
        om = np.mod(off, -d)
        return np.where(np.abs(o) > np.abs(om), om, o)

    xoffset = reduce(np.asarray(betax_off, FDTYPE))
    yoffset = reduce(np.asarray(betay_off, FDTYPE))
    ax = np.linspace(-betamax, betamax, n)
    ay = np.linspace(-betamax, betamax, n)
    step = ax[1] - ax[0]
    out = None
    for i, bx in enumerate(ax):
        xtyp = -1 if i == 0 else (1 if i == n - 1 else 0)
        for j, by in enumerate(ay):
            ytyp = -1 if j == 0 else (1 if j == n - 1 else 0)
            w = fft_window(betax, betay, bx + xoffset, by + yoffset, step, xtyp, ytyp, betamax)
            if out is None:
                out = np.empty((n * n,) + w.shape, w.dtype)
            out[i * n + j] = w
    return out


def fft_betaxy_mean(betax, betay, windows):
    """Intensity (window squared) weighted mean beta_x, beta_y of each window,
    summed over every axis of the beta grid, wavelengths included
    (dtmm/mode.py:108-117); 0 for an all-zero window."""
    axis = tuple(range(-betax.ndim, 0))
    w2 = windows ** 2
    bx = (betax * w2).sum(axis=axis)
    by = (betay * w2).sum(axis=axis)
    norm = w2.sum(axis=axis)
    m = norm > 0
    return np.divide(bx, norm, where=m, out=bx), np.divide(by, norm, where=m, out=by)


def fft_mask(shape, ks, n, betax_off=0., betay_off=0., betamax=BETAMAX):
    """Non-zero windows and their mean (beta, phi) (dtmm/mode.py:119-139)."""
    betax, betay = betaxy(shape[-2:], np.asarray(ks, FDTYPE)[..., None])
    windows = fft_windows(betax, betay, n, betax_off, betay_off, betamax)
    bx, by = fft_betaxy_mean(betax, betay, windows)
    beta = (bx ** 2 + by ** 2) ** 0.5
    phi = np.arctan2(by, bx)
    nonzero = np.asarray([not np.all(w == 0.) for w in windows])
    return windows[nonzero], (beta[nonzero], phi[nonzero])


# ---------------------------------------------------------------------------
# linalg.py (as mathematics, not loops)
# ---------------------------------------------------------------------------

def inv(mat):
    """Batched 4x4/2x2 inverse that returns zeros where det == 0
    (dtmm/linalg.py:428-457 _inv2x2, :460-518 _inv4x4, :521-550 inv)."""
    mat = np.asarray(mat, CDTYPE)
    out = np.zeros_like(mat)
    flat = mat.reshape((-1,) + mat.shape[-2:])
    o = out.reshape(flat.shape)
    det = np.linalg.det(flat)
    ok = det != 0
    if np.any(ok):
        o[ok] = np.linalg.inv(flat[ok])
    return out


def dotmm(a, b):
    """Matrix-matrix product (dtmm/linalg.py:1046)."""
    return np.matmul(a, b)


def dotmv(a, v):
    """Matrix-vector product over the last axis (dtmm/linalg.py dotmv)."""
    return np.matmul(a, np.asarray(v)[..., None])[..., 0]


def dotmdm(a, d, b):
    """a . diag(d) . b (dtmm/linalg.py:1233)."""
    return np.matmul(a * np.asarray(d)[..., None, :], b)


def dotmd(a, d):
    """a . diag(d) (dtmm/linalg.py:1207)."""
    return a * np.asarray(d)[..., None, :]


def _f2v(f):
    return np.moveaxis(f, -3, -1)[..., None]


def _v2f(v):
    return np.moveaxis(v[..., 0], -1, -3)


def dotmf(a, f):
    """Per-pixel matrix times field (..., C, H, W) (dtmm/linalg.py:833-862, :979-991)."""
    return _v2f(np.matmul(a, _f2v(f)))


def dotmdmf(a, d, b, f):
    """a . diag(d) . b times field (dtmm/linalg.py:905-942, :1002-1018)."""
    v = np.matmul(b, _f2v(f))
    v = v * np.asarray(d)[..., None]
    return _v2f(np.matmul(a, v))


_FFT_WORKERS = None


def set_fft_workers(n):
    """Use scipy.fft with `n` threads for the 2D FFTs (the reference's default
    backend, dtmm/fft.py:178-200 with workers = core count, dtmm/conf.py:661);
    None selects numpy.fft on one thread."""
    global _FFT_WORKERS
    _FFT_WORKERS = n


def fft2(f):
    """2D FFT over the last two axes (dtmm/fft.py:277)."""
    if _FFT_WORKERS:
        import scipy.fft
        return scipy.fft.fft2(f, workers=_FFT_WORKERS)
    return np.fft.fft2(f)


def ifft2(f):
    """2D inverse FFT over the last two axes (dtmm/fft.py:306)."""
    if _FFT_WORKERS:
        import scipy.fft
        return scipy.fft.ifft2(f, workers=_FFT_WORKERS)
    return np.fft.ifft2(f)


# ---------------------------------------------------------------------------
# rotation.py
# ---------------------------------------------------------------------------

def _rotation_matrix_uniaxial(theta, phi):
    """R = R_phi . R_theta (dtmm/rotation.py:176-193)."""
    ct, st, cf, sf = np.cos(theta), np.sin(theta), np.cos(phi), np.sin(phi)
    z = np.zeros_like(ct)
    return np.stack([np.stack([ct * cf, -sf, cf * st], -1),
                     np.stack([ct * sf, cf, st * sf], -1),
                     np.stack([-st, z, ct], -1)], -2)


def _rotation_matrix(psi, theta, phi):
    """R = R_phi . R_theta . R_psi (dtmm/rotation.py:125-150)."""
    cp, sp = np.cos(psi), np.sin(psi)
    ct, st = np.cos(theta), np.sin(theta)
    cf, sf = np.cos(phi), np.sin(phi)
    return np.stack([
        np.stack([ct * cf * cp - sf * sp, -ct * cf * sp - sf * cp, cf * st], -1),
        np.stack([ct * sf * cp + cf * sp, cf * cp - ct * sf * sp, st * sf], -1),
        np.stack([-cp * st, st * sp, ct], -1)], -2)


def _rotate_diagonal_tensor(R, d):
    """(eps_xx, eps_yy, eps_zz, eps_xy, eps_xz, eps_yz) of R.diag(d).R^T
    (dtmm/rotation.py:285-292)."""
    out = [sum(d[..., k] * R[..., i, k] * R[..., i, k] for k in range(3)) for i in range(3)]
    out.append(sum(d[..., k] * R[..., 0, k] * R[..., 1, k] for k in range(3)))
    out.append(sum(d[..., k] * R[..., 0, k] * R[..., 2, k] for k in range(3)))
    out.append(sum(d[..., k] * R[..., 1, k] * R[..., 2, k] for k in range(3)))
    return np.stack(out, -1)


def _dotr2m(c, s, F):
    """Rotate the (Ex, Hy, Ey, Hx) rows of F by the beam azimuth
    (dtmm/linalg.py:674-713)."""
    c = c[..., None]
    s = s[..., None]
    r0 = F[..., 0, :] * c - F[..., 2, :] * s
    r1 = F[..., 1, :] * c + F[..., 3, :] * s
    r2 = F[..., 0, :] * s + F[..., 2, :] * c
    r3 = -F[..., 1, :] * s + F[..., 3, :] * c
    return np.stack([r0, r1, r2, r3], -2)


# ---------------------------------------------------------------------------
# tmm.py -- Berreman eigen-solution
# ---------------------------------------------------------------------------

def auxiliary_matrix(beta, eps):
    """Berreman Delta matrix for the (Ex, Hy, Ey, Hx) vector
    (dtmm/tmm.py:99-121).  eps = (xx, yy, zz, xy, xz, yz)."""
    beta = np.asarray(beta, FDTYPE)
    e2m = 1. / eps[..., 2]
    e4 = eps[..., 4] * e2m
    e5 = eps[..., 5] * e2m
    z = np.zeros(np.broadcast_shapes(beta.shape, e2m.shape), CDTYPE)
    L00 = -beta * e4 + z
    L01 = 1. - beta * beta * e2m + z
    L02 = -beta * e5 + z
    L10 = eps[..., 0] - eps[..., 4] * e4 + z
    L12 = eps[..., 3] - eps[..., 5] * e4 + z
    L32 = beta * beta + eps[..., 5] * e5 - eps[..., 1] + z
    return np.stack([np.stack([L00, L01, L02, z], -1),
                     np.stack([L10, L00, L12, z], -1),
                     np.stack([z, z, z, z - 1.], -1),
                     np.stack([-L12, -L02, L32, z], -1)], -2)


def poynting(fvec):
    """z Poynting component Re(Ex Hy*) - Re(Ey Hx*) (dtmm/tmm.py:292-297)."""
    f = np.asarray(fvec)
    return (f[..., 0] * np.conj(f[..., 1])).real - (f[..., 2] * np.conj(f[..., 3])).real


def _alphaf_iso(beta, eps):
    """Isotropic modes, F not normalised (dtmm/tmm.py:123-154)."""
    n = beta.shape[0]
    alpha = np.zeros((n, 4), CDTYPE)
    F = np.zeros((n, 4, 4), CDTYPE)
    aout = np.sqrt(eps[:, 0] - beta ** 2)
    ok = aout != 0.
    a = aout[ok]
    gp = eps[ok, 0] / a
    gs = -a
    alpha[ok] = np.stack([a, -a, a, -a], -1)
    F[ok, 0, 0] = 0.5
    F[ok, 0, 1] = 0.5
    F[ok, 1, 0] = 0.5 * gp
    F[ok, 1, 1] = -0.5 * gp
    F[ok, 2, 2] = 0.5
    F[ok, 2, 3] = 0.5
    F[ok, 3, 2] = 0.5 * gs
    F[ok, 3, 3] = -0.5 * gs
    return alpha, F


def _alphaf_uniaxial(beta, eps, theta, phi):
    """Uniaxial modes with unit-norm columns (dtmm/tmm.py:156-288), with the
    rotation of dtmm/rotation.py:390-395 (phi already relative to the beam)."""
    n = beta.shape[0]
    R = _rotation_matrix_uniaxial(theta, phi)
    ct = R[:, 2, 2]
    st = -R[:, 2, 0]
    sf = -R[:, 0, 1]
    cf = R[:, 1, 1]
    eps11 = eps[:, 0]
    delta = eps[:, 2] - eps11
    alpha = np.zeros((n, 4), CDTYPE)
    F = np.zeros((n, 4, 4), CDTYPE)
    z = beta == 0.
    nz = ~z
    # beta == 0 branch (tmm.py:172-179, :203-233)
    if np.any(z):
        e11 = eps11[z]
        ev02 = e11
        evs = np.sqrt(ev02)
        u = e11 + delta[z] * ct[z] ** 2
        w = e11 * (ev02 + delta[z])
        sq = np.sqrt(u * w) / u
        evpp, evpm = sq, -sq
        alpha[z] = np.stack([evpp, evpm, evs, -evs], -1)
        s, c = sf[z], cf[z]
        F[z, :, 2] = np.stack([-evs * s, -e11 * s, evs * c, -e11 * c], -1)
        F[z, :, 3] = np.stack([-evs * s, e11 * s, evs * c, e11 * c], -1)
        F[z, :, 0] = np.stack([c + 0j, evpp * c, s + 0j, -evpp * s], -1)
        F[z, :, 1] = np.stack([c + 0j, evpm * c, s + 0j, -evpm * s], -1)
    # general branch (tmm.py:181-194, :235-276)
    if np.any(nz):
        b = beta[nz]
        e11 = eps11[nz]
        d = delta[nz]
        ct_, st_, cf_ = ct[nz], st[nz], cf[nz]
        ev02 = e11 - b * b
        evs = np.sqrt(ev02)
        u = e11 + d * ct_ ** 2
        gama = b * cf_
        v = gama * d * 2 * st_ * ct_
        w = gama * gama * (d * st_ ** 2) - e11 * (ev02 + d)
        sq = np.sqrt(v * v - 4 * u * w) / 2 / u
        v = v / 2 / u
        evpp = -v + sq
        evpm = -v - sq
        alpha[nz] = np.stack([evpp, evpm, evs, -evs], -1)
        sfst = R[nz, 1, 2]
        cfst = R[nz, 0, 2]
        ctbeta = ct_ * b
        ctbetaeps11 = ctbeta / e11
        e11sfst = e11 * sfst
        evssfst = evs * sfst
        evscfst = evs * cfst
        evsctbeta = evs * ctbeta
        ev02cfst = ev02 * cfst
        ev02cfsteps11 = ev02cfst / e11
        F[nz, :, 2] = np.stack([evssfst, e11sfst, ctbeta - evscfst, ev02cfst - evsctbeta], -1)
        F[nz, :, 3] = np.stack([-evssfst, e11sfst, evscfst + ctbeta, ev02cfst + evsctbeta], -1)
        F[nz, :, 0] = np.stack([evpp * ctbetaeps11 - ev02cfsteps11, ctbeta - evpp * cfst,
                                -sfst + 0j, evpp * sfst], -1)
        F[nz, :, 1] = np.stack([-evpm * ctbetaeps11 + ev02cfsteps11, evpm * cfst - ctbeta,
                                sfst + 0j, -evpm * sfst], -1)
    norm = np.sqrt((F.real ** 2 + F.imag ** 2).sum(-2))
    F = F / norm[:, None, :]
    return alpha, F


def _alphaf_biaxial(beta, eps, psi, theta, phi):
    """General (biaxial) modes by numeric eigen-decomposition of the Berreman
    matrix, sorted by Poynting sign (+ modes to columns 0, 2)
    (dtmm/tmm.py:395-403, :299-369)."""
    n = beta.shape[0]
    R = _rotation_matrix(psi, theta, phi)
    e6 = _rotate_diagonal_tensor(R, eps)
    L = auxiliary_matrix(beta, e6)
    a0, F0 = np.linalg.eig(L)
    alpha = np.zeros((n, 4), CDTYPE)
    F = np.zeros((n, 4, 4), CDTYPE)
    p = poynting(np.swapaxes(F0, -1, -2))  # (n, 4): one value per column
    for m in range(n):
        i, j, ok = 0, 1, True
        for k in range(4):
            if p[m, k] >= 0.:
                if i >= 4:
                    ok = False
                if ok:
                    alpha[m, i] = a0[m, k]
                    F[m, :, i] = F0[m, :, k]
                i += 2
            else:
                if j >= 4:
                    ok = False
                if ok:
                    alpha[m, j] = a0[m, k]
                    F[m, :, j] = F0[m, :, k]
                j += 2
        if not ok:
            alpha[m] = np.nan
            F[m] = 0.
    return alpha, F


def alphaf(beta=0., phi=0., epsv=(1., 1., 1.), epsa=(0., 0., 0.)):
    """Eigenvalues alpha (..., 4) and field matrix F (..., 4, 4)
    (dtmm/tmm.py:372-403 _alphaf_vec, :435-464 alphaf).

    Branch selection by exact equality: isotropic if eps0 == eps1 == eps2,
    uniaxial if eps0 == eps1, biaxial otherwise (tmm.py:384-389).
    """
    beta = np.asarray(0. if beta is None else beta, FDTYPE)
    phi = np.asarray(0. if phi is None else phi, FDTYPE)
    epsv = np.asarray(epsv, CDTYPE)
    epsa = np.asarray(epsa, FDTYPE)
    shape = np.broadcast_shapes(beta.shape, phi.shape, epsv.shape[:-1], epsa.shape[:-1])
    b = np.broadcast_to(beta, shape).ravel()
    p = np.broadcast_to(phi, shape).ravel()
    ev = np.broadcast_to(epsv, shape + (3,)).reshape(-1, 3)
    ea = np.broadcast_to(epsa, shape + (3,)).reshape(-1, 3)
    n = b.shape[0]
    alpha = np.zeros((n, 4), CDTYPE)
    F = np.zeros((n, 4, 4), CDTYPE)
    iso = (ev[:, 0] == ev[:, 1]) & (ev[:, 1] == ev[:, 2])
    uni = (~iso) & (ev[:, 0] == ev[:, 1])
    bia = ~(iso | uni)
    if np.any(iso):
        alpha[iso], F[iso] = _alphaf_iso(b[iso], ev[iso])
    if np.any(uni):
        alpha[uni], F[uni] = _alphaf_uniaxial(b[uni], ev[uni], ea[uni, 1], ea[uni, 2] - p[uni])
    if np.any(bia):
        alpha[bia], F[bia] = _alphaf_biaxial(b[bia], ev[bia], ea[bia, 0], ea[bia, 1],
                                             ea[bia, 2] - p[bia])
    F = _dotr2m(np.cos(p), np.sin(p), F)
    return alpha.reshape(shape + (4,)), F.reshape(shape + (4, 4))


def alphaffi(beta=0., phi=0., epsv=(1., 1., 1.), epsa=(0., 0., 0.)):
    """alphaf plus inverse of F (dtmm/tmm.py:466-505)."""
    a, f = alphaf(beta, phi, epsv, epsa)
    return a, f, inv(f)


def phase_mat(alpha, kd, mode=None):
    """exp(i kd alpha); mode +1/-1 keeps the even/odd modes
    (dtmm/tmm.py:582-646)."""
    alpha = np.asarray(alpha)
    kd = np.asarray(kd, FDTYPE)[..., None]
    if mode == +1:
        alpha = alpha[..., ::2]
    elif mode == -1:
        alpha = alpha[..., 1::2]
    return np.exp(1j * kd * alpha)


def E_mat(fmat, mode=None):
    """E-field part of F: +1 -> F[::2, ::2], -1 -> F[::2, 1::2], None -> the
    interleaved 4x4 block matrix (dtmm/tmm.py:974-997)."""
    if mode == +1:
        return fmat[..., ::2, ::2].copy()
    if mode == -1:
        return fmat[..., ::2, 1::2].copy()
    out = np.zeros_like(fmat)
    out[..., ::2, ::2] = fmat[..., ::2, ::2]
    out[..., 1::2, 1::2] = fmat[..., ::2, 1::2]
    return out


def E2H_mat(fmat, mode=+1):
    """H = B.A^-1 E for the forward (+1) or backward (-1) modes
    (dtmm/tmm.py:1051-1072)."""
    if mode == +1:
        A, B = fmat[..., ::2, ::2], fmat[..., 1::2, ::2]
    else:
        A, B = fmat[..., ::2, 1::2], fmat[..., 1::2, 1::2]
    return np.matmul(B, inv(A))


def normalize_f(fmat):
    """Scale columns to unit |Poynting| (dtmm/tmm.py:699-724)."""
    fmat = np.asarray(fmat, CDTYPE)
    p = poynting(np.swapaxes(fmat, -1, -2))
    n = np.abs(p) ** 0.5
    with np.errstate(divide="ignore"):
        s = np.where(n == 0., 0., 1. / np.where(n == 0., 1., n))
    return fmat * s[..., None, :]


def jonesvec(pol, phi=0.):
    """Normalised Jones vector seen in a frame rotated by phi
    (dtmm/jones.py:55-98)."""
    pol = np.asarray(pol, CDTYPE)
    phi = np.asarray(phi)
    norm = (pol[..., 0] * pol[..., 0].conj() + pol[..., 1] * pol[..., 1].conj()) ** 0.5
    pol = pol / norm[..., None]
    c, s = np.cos(-phi), np.sin(-phi)
    r = np.empty(c.shape + (2, 2), FDTYPE)
    r[..., 0, 0] = c
    r[..., 1, 1] = c
    r[..., 0, 1] = -s
    r[..., 1, 0] = s
    return dotmv(r, pol)


def avec(jones=(1, 0), amplitude=1., mode=+1):
    """Mode amplitude vector (dtmm/tmm.py:1750-1806)."""
    jones = np.asarray(jones)
    amplitude = np.asarray(amplitude)
    c, s = jones[..., 0], jones[..., 1]
    shape = np.broadcast(c, amplitude).shape + (4,)
    out = np.zeros(shape, CDTYPE)
    if mode == +1:
        out[..., 0] = c
        out[..., 2] = s
    else:
        out[..., 1] = c
        out[..., 3] = s
    return out * amplitude[..., None]


def fvec(fmat, jones=(1, 0), amplitude=1., mode=+1):
    """Field vector F_norm . avec (dtmm/tmm.py:1808-1843, :1877-1904)."""
    a = avec(jones, amplitude, mode)
    return dotmv(normalize_f(fmat), a)


# ---------------------------------------------------------------------------
# data.py
# ---------------------------------------------------------------------------

def refind2eps(refind):
    """eps = n^2 (dtmm/data.py:1059-1062)."""
    return np.asarray(refind) ** 2


def uniaxial_order(order, eps):
    """Scale a diagonal tensor to the given uniaxial order
    (dtmm/data.py:1215-1233)."""
    eps = np.asarray(eps)
    order = np.asarray(order, FDTYPE)
    m = (eps[..., 0] + eps[..., 1] + eps[..., 2]) / 3.
    delta = eps[..., 2] - (eps[..., 0] + eps[..., 1]) / 2.
    o = np.broadcast_to(order, m.shape)
    e1 = np.where(o == 0., m, m - 1. / 3. * o * delta)
    e3 = np.where(o == 0., m, m + 2. / 3. * o * delta)
    e1 = np.where(o >= 0, e1, eps[..., 0])
    e2 = np.where(o >= 0, e1, eps[..., 1])
    e3 = np.where(o >= 0, e3, eps[..., 2])
    return np.stack([e1, e2, e3], -1)


def director2Q(director, order=1.):
    """Q tensor (..., 6) of a director field (dtmm/data.py:376-407); the
    order S is the square root of the director length (director2order,
    data.py:963-973)."""
    n = np.asarray(director, FDTYPE)
    S = np.sqrt(np.sqrt(np.sum(n * n, axis=-1)))
    d = S / 3 * order
    return np.stack([order * n[..., 0] ** 2 - d, order * n[..., 1] ** 2 - d,
                     order * n[..., 2] ** 2 - d, order * n[..., 0] * n[..., 1],
                     order * n[..., 0] * n[..., 2], order * n[..., 1] * n[..., 2]], -1)


def Q2eps(q, no=1.5, ne=1.6, scale_factor=1.):
    """eps (..., 6) of a Q tensor (..., 6) (dtmm/data.py:432-487)."""
    q = np.asarray(q, FDTYPE)
    epsa = (ne ** 2 - no ** 2) / scale_factor
    epsm = (no ** 2 * 2 + ne ** 2) / 3.
    out = np.empty_like(q)
    out[..., 0:3] = epsa * q[..., 0:3] + epsm
    out[..., 3:] = epsa * q[..., 3:]
    return out


def tensor_eig(t6):
    """Eigenvalues and eigenvectors (columns) of a symmetric tensor (..., 6),
    ordered as dtmm/linalg.py:172-222 orders them: ascending with a
    right-handed frame, then the most distinct eigenvalue moved to index 2 by
    a cyclic permutation.  A tensor whose off-diagonal elements, scaled by
    its largest element, have a squared norm of at most 1e-12 is taken as
    diagonal: its eigenvalues keep the input order with the identity frame
    before the cyclic permutation (linalg.py:251-318).  numpy.linalg.eigh
    replaces the reference's closed form (linalg.py:225-315); eigenvectors
    agree up to sign."""
    t6 = np.asarray(t6, FDTYPE)
    m = np.empty(t6.shape[:-1] + (3, 3))
    for k, (i, j) in enumerate(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
        m[..., i, j] = m[..., j, i] = t6[..., k]
    w, v = np.linalg.eigh(m)
    v[..., :, 2] = np.cross(v[..., :, 0], v[..., :, 1])
    scale = np.abs(t6).max(axis=-1)
    sm = np.divide(1., scale, out=np.zeros_like(scale), where=scale > 0)
    diagonal = ~(((t6[..., 3:] * sm[..., None]) ** 2).sum(-1) > 1e-12)
    w = np.where(diagonal[..., None], t6[..., :3] * sm[..., None] * scale[..., None], w)
    v = np.where(diagonal[..., None, None], np.eye(3), v)
    m2 = np.abs(w[..., 1] - w[..., 0])
    m1 = np.abs(w[..., 2] - w[..., 0])
    m0 = np.abs(w[..., 2] - w[..., 1])
    move = (m2 > m1) | (m2 > m0)
    perm = np.where((move & (m1 > m0))[..., None], [1, 2, 0],
                    np.where(move[..., None], [2, 0, 1], [0, 1, 2]))
    w = np.take_along_axis(w, perm, -1)
    v = np.take_along_axis(v, perm[..., None, :], -1)
    return w, v


def rotation_angles(R):
    """Euler angles (psi, theta, phi) of R = R_phi R_theta R_psi
    (dtmm/rotation.py:154-171)."""
    R = np.asarray(R, FDTYPE)
    theta = np.arccos(np.clip(R[..., 2, 2], -1, 1))
    flat = ((R[..., 1, 2] == 0.) & (R[..., 0, 2] == 0.)) | (theta == 0.)
    phi = np.where(flat, np.arccos(np.clip(R[..., 1, 1], -1, 1)),
                   np.arctan2(R[..., 1, 2], R[..., 0, 2]))
    psi = np.where(flat, 0., np.arctan2(R[..., 2, 1], -R[..., 2, 0]))
    return np.stack([psi, theta, phi], -1)


def Q2data(q, mask=None, no=1.5, ne=1.6, nhost=None, scale_factor=1., biaxial=False):
    """(thickness, epsv, epsa) of a Q-tensor field (L, H, W, 6)
    (dtmm/data.py:322-373 with eps2epsva, data.py:489-513)."""
    q = np.asarray(q, FDTYPE)
    w, v = tensor_eig(Q2eps(q, no, ne, scale_factor))
    ang = rotation_angles(v)
    if not biaxial:
        w = uniaxial_order(1., w)
    if mask is not None:
        w[np.logical_not(mask), :] = nhost ** 2
    return np.ones(q.shape[0]), w, ang


def epsva2eps(epsv, epsa):
    """(..., 6) tensor R diag(epsv) R^T of eigenvalues and Euler angles
    (dtmm/data.py:515-534)."""
    epsa = np.asarray(epsa, FDTYPE)
    R = _rotation_matrix(epsa[..., 0], epsa[..., 1], epsa[..., 2])
    return _rotate_diagonal_tensor(R, np.asarray(epsv))


def cauchy2eps(coefficients, wavelength):
    """eps = (c0 + c1 / w^2 + c2 / w^4 + ...)^2 with w in microns, per row of
    coefficients (..., n) (dtmm/data.py:1110-1129, EpsilonCauchy.__call__
    :1183-1193 divides the nanometre wavelength by 1000)."""
    c = np.asarray(coefficients, FDTYPE)
    w = wavelength / 1000.
    n = c[..., 0].copy()
    for i in range(1, c.shape[-1]):
        n += c[..., i] / w ** (2 * i)
    return n ** 2


def sellmeier2eps(coefficients, wavelength):
    """eps = c0 + sum_i w^2 c_(2i+1) / (w^2 - c_(2i+2)) with w in microns
    (dtmm/data.py:1072-1093, EpsilonSellmeier.__call__ :1195-1204)."""
    c = np.asarray(coefficients, FDTYPE)
    w = wavelength / 1000.
    eps = c[..., 0].copy()
    for i in range((c.shape[-1] - 1) // 2):
        eps += c[..., 1 + 2 * i] * w ** 2 / (w ** 2 - c[..., 2 + 2 * i])
    return eps


def is_optical_data_dispersive(data):
    """Whether any block's eps is a callable of the wavelength
    (dtmm/data.py:1131-1142)."""
    blocks = data if isinstance(data, list) else [data]
    return any(callable(b[1]) for b in blocks)


def validate_optical_block(block, wavelength=None):
    """Cast and add singleton axes so that a block broadcasts to
    (L, H, W, 3) (dtmm/data.py:534-618).  A callable eps is evaluated at
    `wavelength` (nm), as validate_optical_data(..., wavelength=w) does for
    the per-wavelength transfers of split_wavelengths."""
    d, material, angles = block
    if callable(material):
        if wavelength is None:
            raise ValueError("Epsilon is a callable. You must provide wavelength!")
        material = material(wavelength)
    d = np.asarray(d, FDTYPE)
    material = np.asarray(material)
    material = np.asarray(material, CDTYPE if np.iscomplexobj(material) else FDTYPE)
    angles = np.asarray(angles, FDTYPE)
    if d.ndim == 0:
        d, material, angles = d[None], material[None], angles[None]
    if material.ndim == 2:
        material = material[:, None, None, :]
    elif material.ndim == 3:
        material = material[:, None, :, :]
    return d, material, angles


def validate_optical_data(data, wavelength=None):
    """A list of validated blocks (dtmm/data.py:620-643)."""
    if isinstance(data, list):
        return [validate_optical_block(b, wavelength) for b in data]
    return [validate_optical_block(data, wavelength)]


def _isotropic_effective_data(block):
    """Per-layer isotropic effective medium: the volume mean of the
    isotropic part of eps (dtmm/transfer.py:34-41)."""
    d, material, angles = block
    n = len(d)
    epseff = uniaxial_order(0., material).mean(axis=(0, 1, 2))
    epseff = np.broadcast_to(epseff, (n, 3)).copy()
    aeff = np.zeros((n, 3), FDTYPE)
    return d, epseff, aeff


_SYMMETRY = {"isotropic": 0, "uniaxial": 1, "biaxial": 2}


def _symmetry_int(arg):
    """(dtmm/data.py:1343-1352)."""
    if isinstance(arg, str):
        try:
            return _SYMMETRY[arg]
        except KeyError:
            raise ValueError("Invalid symmetry argument.")
    if arg in (0, 1, 2, -1):
        return arg
    raise ValueError("Invalid symmetry argument.")


def _parse_symmetry(arg):
    """A symmetry or a sequence of per-layer symmetries (dtmm/data.py:1354-1362)."""
    if isinstance(arg, str):
        return _symmetry_int(arg)
    try:
        return tuple(_parse_symmetry(a) for a in arg)
    except TypeError:
        return _symmetry_int(arg)


def eig_symmetry(order, eig):
    """Isotropic (0), uniaxial (1) or unchanged (2, and -1) eigenvalues
    (dtmm/data.py:1276-1306): orders above 1 become -1 for uniaxial_order."""
    order = np.array(order, int)
    order[order > 1] = -1
    return uniaxial_order(order, eig)


def eps2epsva(eps):
    """Eigenvalues and Euler angles of tensors (..., 6) (dtmm/data.py:489-513)
    for real tensors; the reference's complex branch warns that its
    eigenvectors are then unreliable, and the oracle does not cover it."""
    eps = np.asarray(eps)
    if np.iscomplexobj(eps):
        if np.any(eps.imag != 0):
            raise NotImplementedError("oracle covers real effective tensors")
        eps = eps.real
    w, v = tensor_eig(eps)
    return w, rotation_angles(v)


def effective_block(block, symmetry=0):
    """Effective layers of a block (dtmm/data.py:1364-1419): symmetry 0 is
    the mean of the isotropic parts over the block; otherwise the
    laboratory-frame tensors are averaged (over the whole block, or over each
    layer for a sequence of per-layer symmetries), diagonalised, and made
    uniaxial (1) or kept biaxial (2)."""
    d, epsv, epsa = block
    axis = list(range(epsv.ndim - 1))
    if isinstance(symmetry, (str, int)) and symmetry in ("isotropic", 0):
        ev = eig_symmetry(0, epsv).mean(tuple(axis))
        ea = np.zeros_like(ev, FDTYPE)
    else:
        order = np.asarray(_parse_symmetry(symmetry), int)
        if order.ndim != 0:
            if len(order) != len(d):
                raise ValueError("Shape of the symmetry argument is incompatible with block size.")
            axis.pop(0)
        shape = (len(d),) + tuple(np.broadcast_shapes(epsv.shape[1:], epsa.shape[1:]))
        eps = epsva2eps(np.broadcast_to(epsv, shape), np.broadcast_to(epsa, shape))
        eps = eps.mean(tuple(axis))
        ev, ea = eps2epsva(eps)
        ev = eig_symmetry(order, ev)
    if ev.ndim == 1:
        ev = np.asarray((ev,) * len(d)).copy()
        ea = np.asarray((ea,) * len(d)).copy()
    return d, ev, ea


def _is_block(data):
    """A custom effective block (d, epsv, epsa), as opposed to a symmetry
    identifier or a sequence of per-layer symmetries."""
    if not isinstance(data, (tuple, list)) or len(data) != 3:
        return False
    return np.ndim(data[1]) >= 1 and np.ndim(data[2]) >= 1


def _effective_layers(block, eff_data):
    """(d, epsv, epsa) of a block's effective layers, each epsv/epsa (L, 3)
    (dtmm/transfer.py:941-948): a custom block as given, else
    effective_block, the default (None) the isotropic mean."""
    d = block[0]
    if eff_data is None:
        return _isotropic_effective_data(block)
    if _is_block(eff_data):
        de, ee, ae = eff_data
        de = np.atleast_1d(np.asarray(de, FDTYPE))
        ee = np.asarray(ee)
        ee = np.asarray(ee, CDTYPE if np.iscomplexobj(ee) else FDTYPE)
        ee = np.broadcast_to(ee.reshape(-1, 3) if ee.ndim > 1 else ee, (len(de), 3)).copy()
        ae = np.broadcast_to(np.asarray(ae, FDTYPE).reshape(-1, 3) if np.ndim(ae) > 1
                             else np.asarray(ae, FDTYPE), (len(de), 3)).copy()
        if len(de) != len(d):
            raise ValueError("the effective block needs one layer per layer of the block")
        return de, ee, ae
    de, ee, ae = effective_block(block, eff_data)
    return de, np.asarray(ee).reshape(len(de), 3), np.asarray(ae, FDTYPE).reshape(len(de), 3)


def create_layers(optical_data, nin, nout, nstep, eff_data=None):
    """Layer and effective-layer lists with the input/output media as
    zero-thickness layers (dtmm/transfer.py:926-1005): eff_data None (the
    isotropic mean), a symmetry 0/1/2 ("isotropic", "uniaxial", "biaxial"),
    a sequence of per-layer symmetries or a custom effective block, one for
    all blocks or a list with one entry per block."""
    if is_optical_data_dispersive(optical_data):
        raise ValueError("You are using dispersive data, so you must use "
                         "`split_wavelengths=True`")
    blocks = validate_optical_data(optical_data)
    nb = len(blocks)
    if isinstance(optical_data, list) and not (eff_data is None or isinstance(eff_data, (str, int))):
        if len(eff_data) != nb:
            raise ValueError("Number of elements in eff_data must match length of optical_data.")
        effs = list(eff_data)
    else:
        effs = [eff_data] * nb
    layers, eff_layers = [], []
    for ib, (d, epsv, epsa) in enumerate(blocks):
        substeps = np.broadcast_to(np.asarray(nstep), (len(d),))
        ev_b = np.broadcast_to(epsv, (len(d),) + epsv.shape[1:])
        ea_b = np.broadcast_to(epsa, (len(d),) + epsa.shape[1:])
        ls = [(int(n), (t / n, ev, ea)) for n, t, ev, ea in zip(substeps, d, ev_b, ea_b)]
        de, ee, ae = _effective_layers((d, epsv, epsa), effs[ib])
        es = [(int(n), (t / n, ev, ea)) for n, t, ev, ea in zip(substeps, de, ee, ae)]
        if ib == 0:
            ls.insert(0, (1, (0., np.broadcast_to(refind2eps([nin] * 3), ev_b[0].shape),
                              np.zeros(3, FDTYPE))))
            es.insert(0, (1, (0., refind2eps([nin] * 3), np.zeros(3, FDTYPE))))
        if ib == nb - 1:
            ls.append((1, (0., np.broadcast_to(refind2eps([nout] * 3), ev_b[0].shape),
                           np.zeros(3, FDTYPE))))
            es.append((1, (0., refind2eps([nout] * 3), np.zeros(3, FDTYPE))))
        layers += ls
        eff_layers += es
    return layers, eff_layers


# ---------------------------------------------------------------------------
# diffract.py / matrix.py (isotropic or general effective media)
# ---------------------------------------------------------------------------

def _mask(beta, betamax):
    return beta >= betamax


def diffraction_alphaffi(shape, ks, epsv, epsa, betamax):
    """Eigen data on the FFT grid, zeroed where beta >= betamax
    (dtmm/diffract.py:47-76)."""
    ks = np.abs(np.asarray(ks))
    beta, phi = betaphi(shape, ks)
    alpha, f = alphaf(beta, phi, epsv, epsa)
    fi = inv(f)
    m = _mask(beta, betamax)
    fi[m] = 0.
    f[m] = 0.
    alpha[m] = 0.
    return alpha, f, fi


def diffraction_alphaf(shape, ks, epsv, epsa, betamax):
    """(dtmm/diffract.py:22-42)."""
    ks = np.abs(np.asarray(ks))
    beta, phi = betaphi(shape, ks)
    alpha, f = alphaf(beta, phi, epsv, epsa)
    m = _mask(beta, betamax)
    f[m] = 0.
    alpha[m] = 0.
    return alpha, f


def _phase_matrix(alpha, kd, mode=None):
    """(dtmm/diffract.py:96-105)."""
    kd = np.asarray(kd, FDTYPE)
    out = phase_mat(alpha, kd[..., None, None])
    if mode in ("t", +1):
        out[..., 1::2] = 0.
    elif mode in ("r", -1):
        out[..., ::2] = 0.
    return out


def field_diffraction_matrix(shape, ks, d=1., epsv=(1, 1, 1), epsa=(0, 0, 0.), mode="b",
                             betamax=BETAMAX):
    """F . P(kd) . F^-1 on the FFT grid (dtmm/diffract.py:107-119)."""
    ks = np.asarray(ks, FDTYPE)
    alpha, f, fi = diffraction_alphaffi(shape, ks, epsv, epsa, betamax)
    pmat = _phase_matrix(alpha, ks * d, mode=mode)
    return dotmdm(f, pmat, fi)


def projection_matrix(shape, ks, epsv=(1, 1, 1), epsa=(0, 0, 0.), mode=+1, betamax=BETAMAX):
    """F . diag(1 on +/- modes) . F^-1 (dtmm/diffract.py:242-251)."""
    ks = np.asarray(ks, FDTYPE)
    alpha, f, fi = diffraction_alphaffi(shape, ks, epsv, epsa, betamax)
    pmat = _phase_matrix(alpha, np.zeros_like(ks), mode=mode)
    return dotmdm(f, pmat, fi)


def diffract(field, dmat):
    """FFT -> per-k matrix -> IFFT (dtmm/diffract.py:254-299)."""
    return ifft2(dotmf(dmat, fft2(field)))


def transmitted_field(field, ks, n=1., betamax=BETAMAX):
    """Forward-propagating part of a field in a medium of index n
    (dtmm/transfer.py:238-242, :270-293)."""
    pmat = projection_matrix(field.shape[-2:], ks, epsv=refind2eps([n] * 3),
                             epsa=(0., 0., 0.), mode=+1, betamax=betamax)
    return diffract(field, pmat)


def _epn_correction_matrix(beta, phi, ks, d, epsv, epsa):
    """E . P(-kd) . E^-1 for the mean beam (dtmm/matrix.py:132-139)."""
    alpha, f = alphaf(beta, phi, epsv, epsa)
    kd = -np.asarray(ks) * d
    pmat = phase_mat(alpha, kd[..., None, None])
    e = E_mat(f)
    return dotmdm(e, pmat, inv(e))


def _first_epn_diffraction_matrix(shape, ks, d, epsv, epsa, betamax):
    """E . P(kd) . F^-1 on the FFT grid (dtmm/matrix.py:109-118)."""
    ks = np.asarray(ks, FDTYPE)
    alpha, f, fi = diffraction_alphaffi(shape, ks, epsv, epsa, betamax)
    kd = ks * d
    pmat = phase_mat(alpha, kd[..., None, None])
    return dotmdm(E_mat(f), pmat, fi)


def _second_epn_diffraction_matrix(shape, ks, d, epsv, epsa, betamax):
    """F . P(kd) . E^-1 on the FFT grid (dtmm/matrix.py:120-130)."""
    ks = np.asarray(ks, FDTYPE)
    alpha, f = diffraction_alphaf(shape, ks, epsv, epsa, betamax)
    kd = ks * d
    ei = inv(E_mat(f))
    pmat = phase_mat(alpha, kd[..., None, None])
    return dotmdm(f, pmat, ei)


def first_corrected_Epn_diffraction_matrix(shape, ks, beta, phi, d, epsv, epsa, betamax=BETAMAX):
    """C . D1 (dtmm/matrix.py:141-157, window=None)."""
    dmat = _first_epn_diffraction_matrix(shape, ks, d, epsv, epsa, betamax)
    cmat = _epn_correction_matrix(beta, phi, ks, d, epsv, epsa)
    return dotmm(cmat, dmat)


def second_corrected_Epn_diffraction_matrix(shape, ks, beta, phi, d, epsv, epsa, betamax=BETAMAX):
    """D2 . C (dtmm/matrix.py:159-174, window=None)."""
    dmat = _second_epn_diffraction_matrix(shape, ks, d, epsv, epsa, betamax)
    cmat = _epn_correction_matrix(beta, phi, ks, d, epsv, epsa)
    return dotmm(dmat, cmat)


def _alphaEEi(beta, phi, epsv, epsa, mode):
    """(dtmm/tmm.py:509-579)."""
    alpha, f = alphaf(beta, phi, epsv, epsa)
    e = E_mat(f, mode=mode)
    alpha = alpha[..., ::2] if mode == +1 else alpha[..., 1::2]
    return alpha, e, inv(e)


def _E_diffraction_matrix(shape, ks, d, epsv, epsa, mode, betamax):
    """J . P(kd) . J^-1 on the FFT grid (dtmm/diffract.py:78-93, :172-180)."""
    ks = np.asarray(ks, FDTYPE)
    kabs = np.abs(ks)
    beta, phi = betaphi(shape, kabs)
    alpha, j, ji = _alphaEEi(beta, phi, epsv, epsa, mode)
    m = _mask(beta, betamax)
    ji[m] = 0.
    j[m] = 0.
    alpha[m] = 0.
    pmat = _phase_matrix(alpha, ks * d)
    return dotmdm(j, pmat, ji)


def _E_correction_matrix(beta, phi, ks, d, epsv, epsa, mode):
    """(dtmm/matrix.py:31-36)."""
    alpha, j, ji = _alphaEEi(beta, phi, epsv, epsa, mode)
    kd = -np.asarray(ks) * d
    pmat = _phase_matrix(alpha, kd)
    return dotmdm(j, pmat, ji)


def first_E_diffraction_matrix(shape, ks, beta, phi, d, epsv, epsa, mode=+1, betamax=BETAMAX):
    """C . D (dtmm/matrix.py:53-58)."""
    dmat = _E_diffraction_matrix(shape, ks, d, epsv, epsa, mode, betamax)
    cmat = _E_correction_matrix(beta, phi, ks, d, epsv, epsa, mode)
    return dotmm(cmat, dmat)


def corrected_E_diffraction_matrix(shape, ks, beta, phi, d, epsv, epsa, mode=+1,
                                   betamax=BETAMAX):
    """C(d/2) . D(d) . C(d/2) (dtmm/matrix.py:38-43)."""
    dmat = _E_diffraction_matrix(shape, ks, d, epsv, epsa, mode, betamax)
    cmat = _E_correction_matrix(beta, phi, ks, d / 2., epsv, epsa, mode)
    return dotmm(cmat, dotmm(dmat, cmat))


def second_E_diffraction_matrix(shape, ks, beta, phi, d, epsv, epsa, mode=+1, betamax=BETAMAX):
    """D . C (dtmm/matrix.py:46-51)."""
    dmat = _E_diffraction_matrix(shape, ks, d, epsv, epsa, mode, betamax)
    cmat = _E_correction_matrix(beta, phi, ks, d, epsv, epsa, mode)
    return dotmm(dmat, cmat)


# ---------------------------------------------------------------------------
# propagate_4x4.py / propagate_2x2.py
# ---------------------------------------------------------------------------

class _Cache:
    """Memo for the per-layer k-space matrices; the reference caches the same
    pure functions (dtmm/conf.py:224-330), so this changes no number."""

    def __init__(self):
        self.d = {}

    def get(self, key, fn):
        if key not in self.d:
            if len(self.d) > 8:
                self.d.clear()
            self.d[key] = fn()
        return self.d[key]


def _key(*args):
    out = []
    for a in args:
        a = np.asarray(a)
        out.append((a.shape, a.dtype.str, a.tobytes()))
    return tuple(out)


def _transfer_ray_4x4_1(field, ks, layer, dmat1, dmat2, beta, phi, nsteps):
    """D1 -> IFFT -> e.P.e^-1 -> FFT -> D2 per substep
    (dtmm/propagate_4x4.py:80-101)."""
    d, epsv, epsa = layer
    kd = ks * d
    alpha, f = alphaf(beta, phi, epsv, epsa)
    p = phase_mat(alpha, kd[..., None, None])
    e = E_mat(f)
    ei = inv(e)
    for _ in range(nsteps):
        field = dotmf(dmat1, field)
        field = ifft2(field)
        field = dotmdmf(e, p, ei, field)
        field = fft2(field)
        field = dotmf(dmat2, field)
    return field


def corrected_Epn_diffraction_matrix(shape, ks, beta, phi, d, epsv, epsa, betamax=BETAMAX):
    """C(d/2) . D(d) . C(d/2) of the forward (+1) and backward (-1) E modes
    (dtmm/matrix.py:38-43, :62-68)."""
    out = []
    for mode in (+1, -1):
        dmat = _E_diffraction_matrix(shape, ks, d, epsv, epsa, mode, betamax)
        cmat = _E_correction_matrix(beta, phi, ks, d / 2., epsv, epsa, mode)
        out.append(dotmm(cmat, dotmm(dmat, cmat)))
    return out


def _transfer_ray_4x4_2(field, ks, layer, beta, phi, nsteps, dmatpn=None):
    """Reflection mode 2 in real space (dtmm/propagate_4x4.py:18-51).
    Without diffraction (dmatpn None) every sub-step applies
    f diag(exp(i k0 d alpha)) f^-1 to each pixel.  With diffraction it
    applies e diag(exp(i k0 d alpha / 2)) f^-1 (to the E parts of the
    forward and backward modes), diffracts the forward and backward parts
    with dmatpn, and returns with f diag(exp(i k0 d alpha / 2)) e^-1."""
    d, epsv, epsa = layer
    alpha, f, fi = alphaffi(beta, phi, epsv, epsa)
    if dmatpn is None:
        p = phase_mat(alpha, (ks * d)[..., None, None])
        for _ in range(nsteps):
            field = dotmdmf(f, p, fi, field)
        return field
    dmatp, dmatn = dmatpn
    p = phase_mat(alpha, (ks * d / 2)[..., None, None])
    e = E_mat(f)
    ei = inv(e)
    for _ in range(nsteps):
        field = dotmdmf(e, p, fi, field)
        field[..., 0::2, :, :] = diffract(field[..., 0::2, :, :], dmatp)
        field[..., 1::2, :, :] = diffract(field[..., 1::2, :, :], dmatn)
        field = dotmdmf(f, p, ei, field)
    return field


def propagate_4x4_effective_2(field, ks, layer, effective_layer, beta=0, phi=0, nsteps=1,
                              diffraction=1, betamax=BETAMAX, cache=None):
    """Layer step of the 4x4 method, reflection mode 2, diffraction 0 or 1
    (dtmm/propagate_4x4.py:156-175)."""
    if diffraction == 0:
        return _transfer_ray_4x4_2(field, ks, layer, beta, phi, nsteps)
    d_eff, epsv_eff, epsa_eff = effective_layer
    shape = field.shape[-2:]
    if diffraction > 1:
        # the real-space field split by the k-space tent windows, each part
        # through all sub-steps with its window's mean beam, summed
        # (propagate_4x4.py:176-204)
        if diffraction == np.inf:
            raise NotImplementedError("oracle covers finite diffraction")
        beta = np.asarray(beta, FDTYPE)
        phi = np.asarray(phi, FDTYPE)
        if beta.ndim > 0:
            bshape = beta.shape
            beta = beta[..., 0]
            phi = phi[..., 0]
        else:
            bshape = ()
        windows, (betas, phis) = fft_mask(field.shape, ks, int(diffraction),
                                          beta * np.cos(phi), beta * np.sin(phi), betamax)
        nw = len(windows)
        betas = betas.reshape((nw,) + bshape)
        phis = phis.reshape((nw,) + bshape)
        dmatps, dmatns = corrected_Epn_diffraction_matrix(shape, ks, betas, phis, d_eff,
                                                          epsv_eff, epsa_eff, betamax)
        fk = fft2(field)
        fout = np.zeros_like(field)
        for window, b, p, dmp, dmn in zip(windows, betas, phis, dmatps, dmatns):
            fout += _transfer_ray_4x4_2(ifft2(fk * window), ks, layer, b, p, nsteps, (dmp, dmn))
        return fout
    if diffraction != 1:
        raise NotImplementedError("oracle covers reflection mode 2 with diffraction >= 0")
    cache = cache or _Cache()
    key = _key(shape, ks, beta, phi, d_eff, epsv_eff, epsa_eff, betamax)
    dmatpn = cache.get(("pn",) + key, lambda: corrected_Epn_diffraction_matrix(
        shape, ks, beta, phi, d_eff, epsv_eff, epsa_eff, betamax))
    return _transfer_ray_4x4_2(field, ks, layer, beta, phi, nsteps, dmatpn)


def propagate_4x4_effective_1(field, ks, layer, effective_layer, beta=0, phi=0, nsteps=1,
                              diffraction=1, betamax=BETAMAX, cache=None):
    """Layer step of the 4x4 method, reflection mode 1
    (dtmm/propagate_4x4.py:376-444).

    diffraction == 1: one mean beam.  diffraction = n > 1: the k-space field
    is split by the non-zero n*n tent windows (fft_mask), each part is
    propagated through all `nsteps` sub-steps with its window's mean beam, and
    the parts are summed (propagate_4x4.py:390-419)."""
    d_eff, epsv_eff, epsa_eff = effective_layer
    shape = field.shape[-2:]
    cache = cache or _Cache()
    if diffraction == 1:
        key = _key(shape, ks, beta, phi, d_eff, epsv_eff, epsa_eff, betamax)
        dmat1 = cache.get(("d1",) + key, lambda: first_corrected_Epn_diffraction_matrix(
            shape, ks, beta, phi, d_eff / 2, epsv_eff, epsa_eff, betamax))
        dmat2 = cache.get(("d2",) + key, lambda: second_corrected_Epn_diffraction_matrix(
            shape, ks, beta, phi, d_eff / 2, epsv_eff, epsa_eff, betamax))
        return _transfer_ray_4x4_1(field, ks, layer, dmat1, dmat2, beta, phi, nsteps)
    if not diffraction > 1:
        raise NotImplementedError("oracle covers diffraction >= 1 for the 4x4 method")
    beta = np.asarray(beta, FDTYPE)
    phi = np.asarray(phi, FDTYPE)
    if beta.ndim > 0:
        bshape = beta.shape
        beta = beta[..., 0]
        phi = phi[..., 0]
    else:
        bshape = ()
    windows, (betas, phis) = fft_mask(field.shape, ks, int(diffraction),
                                      beta * np.cos(phi), beta * np.sin(phi), betamax)
    nw = len(windows)
    betas = betas.reshape((nw,) + bshape)
    phis = phis.reshape((nw,) + bshape)
    # the window matrices are pure functions of the effective layer and the
    # window beams; the reference caches them (dtmm/conf.py:224-330)
    key = _key(shape, ks, betas, phis, d_eff, epsv_eff, epsa_eff, betamax)
    dmats1 = cache.get(("w1",) + key, lambda: first_corrected_Epn_diffraction_matrix(
        shape, ks, betas, phis, d_eff / 2, epsv_eff, epsa_eff, betamax))
    dmats2 = cache.get(("w2",) + key, lambda: second_corrected_Epn_diffraction_matrix(
        shape, ks, betas, phis, d_eff / 2, epsv_eff, epsa_eff, betamax))
    fout = np.zeros_like(field)
    for window, b, p, dm1, dm2 in zip(windows, betas, phis, dmats1, dmats2):
        fout += _transfer_ray_4x4_1(field * window, ks, layer, dm1, dm2, b, p, nsteps)
    return fout


def propagate_4x4_full(field, ks, layer, nsteps=1, betamax=BETAMAX):
    """Exact per-mode propagation through one layer (dtmm/propagate_4x4.py:
    515-564): for every step and wavelength the real-space field (..., 4, H,
    W) is split into its FFT modes with beta < betamax (strict, :537), each
    mode's plane wave is multiplied at every pixel by F P(kd) F^-1 of that
    pixel's material at the mode's own (beta, phi), and the results are
    summed in ascending order of beta (stable, :546).  kd = k0 d / nsteps
    with d the thickness in `layer` (:522); transfer_4x4 passes the
    sub-step thickness t / nsteps there, so each of the nsteps steps
    propagates t / nsteps^2 (restated as the reference computes it)."""
    d, epsv, epsa = layer
    ks = np.asarray(ks, FDTYPE)
    shape = field.shape[-2:]
    kd = ks * d / nsteps
    ii, jj = np.meshgrid(range(shape[0]), range(shape[1]), indexing="ij")
    out = np.array(field, CDTYPE)
    for _ in range(nsteps):
        src = out
        out = np.empty_like(src)
        for i in range(len(ks)):
            ff = np.fft.fft2(src[..., i, :, :, :])
            acc = np.zeros_like(src[..., i, :, :, :])
            b, p = betaphi(shape, ks[i])
            mask = b < betamax
            amp = ff[..., mask]
            betas, phis, iind, jind = b[mask], p[mask], ii[mask], jj[mask]
            for j in sorted(range(len(betas)), key=lambda q: betas[q]):
                alpha, f, fi = alphaffi(betas[j], phis[j], epsv, epsa)
                pm = np.exp(1j * kd[i] * alpha)
                w = eigenwave(amp.shape[:-1] + shape, iind[j], jind[j], amplitude=amp[..., j])
                acc += dotmdmf(f, pm, fi, w)
            out[..., i, :, :, :] = acc
    return out


def S_mat(fmatin, fmatout, fmatini=None, mode=+1):
    """Blocks of F_in^-1 F_out coupling forward output modes to the input
    modes of the same / opposite direction (dtmm/tmm.py:850-878)."""
    if fmatini is None:
        fmatini = inv(fmatin)
    S = dotmm(fmatini, fmatout)
    if mode == +1:
        return S[..., ::2, ::2], S[..., 1::2, 0::2]
    return S[..., 1::2, 1::2], S[..., 0::2, 1::2]


def Eti_mat(fmatin, fmatout, fmatini=None, mode=+1):
    """E transmission into the modes of fmatout, no reflections
    (dtmm/tmm.py:999-1022)."""
    A = E_mat(fmatin, mode=mode)
    St, _ = S_mat(fmatin, fmatout, fmatini, mode)
    return dotmm(inv(St), inv(A))


def _transfer_ray_2x2_2(field, ks, in_layer, out_layer, dmat, beta, phi, nsteps, mode=+1,
                        bulk=None, reflection=2, refl=None):
    """2x2 layer step in real space (dtmm/propagate_2x2.py:91-190).  With
    `reflection` the first half step carries the interface from the in_layer
    voxel modes: without `refl` the transmission Eti (:172-177); with `refl`
    (npass > 1, :120-160) the tangential E transmission t of tr_mat, the
    stored wave refl <- field - t.field (that is, minus the reflected
    tangential E, because t = 1 + r) and field <- t.field + (the old refl),
    the backward passes (mode -1) adding that field and its H to `bulk`
    before the layer.  Then the E diffraction dmat and a plain half step
    e P e^-1; forward passes add E and H to `bulk` after the layer.  Returns
    (field, refl)."""
    d, epsv, epsa = out_layer
    alpha, fmat = alphaf(beta, phi, epsv, epsa)
    e = E_mat(fmat, mode=mode)
    p = phase_mat(alpha, (np.asarray(ks) * d / 2)[..., None, None], mode=mode)
    ei0 = inv(e)
    for j in range(nsteps):
        if j == 0 and reflection != 0:
            _, epsv_in, epsa_in = in_layer
            _, fmat_in = alphaf(beta, phi, epsv_in, epsa_in)
            if refl is not None:
                tmat, _ = tr_mat(fmat_in, fmat, mode=mode)
                trans = refl.copy()
                tf = dotmf(tmat, field)
                refl[...] = field - tf
                field = tf + trans
                if mode == -1 and bulk is not None:
                    bulk[..., ::2, :, :] += field
                    bulk[..., 1::2, :, :] += dotmf(E2H_mat(fmat, mode=mode), field)
                field = dotmf(ei0, field)
                field = dotmf(dotmd(e, p), field) if d != 0. else dotmf(e, field)
            else:
                field = dotmdmf(e, p, Eti_mat(fmat_in, fmat, mode=mode), field)
        elif d != 0.:
            field = dotmdmf(e, p, ei0, field)
        if dmat is not None:
            field = diffract(field, dmat)
        if d != 0.:
            field = dotmdmf(e, p, ei0, field)
    if mode == +1 and bulk is not None:
        bulk[..., 1::2, :, :] += dotmf(E2H_mat(fmat, mode=mode), field)
        bulk[..., ::2, :, :] += field
    return field, refl


def Etri_mat(fmatin, fmatout, fmatini=None, mode=+1):
    """(dtmm/tmm.py:1025-1049)."""
    A = E_mat(fmatin, mode=mode)
    Ai = inv(A)
    St, Sr = S_mat(fmatin, fmatout, fmatini, mode)
    ei = dotmm(inv(St), Ai)
    return ei, dotmm(Sr, ei)


def tr_mat(fmatin, fmatout, fmatini=None, mode=+1):
    """2x2 transmission and reflection matrices of the tangential E field at
    an interface (dtmm/tmm.py:925-950)."""
    er = E_mat(fmatin, mode=-mode)
    et = E_mat(fmatout, mode=mode)
    eti, eri = Etri_mat(fmatin, fmatout, fmatini, mode)
    return dotmm(et, eti), dotmm(er, eri)


def E_tr_matrix(shape, ks, epsv_in=(1., 1., 1.), epsa_in=(0., 0., 0.), epsv_out=(1., 1., 1.),
                epsa_out=(0., 0., 0.), mode=+1, betamax=BETAMAX):
    """Interface transmission/reflection on the FFT grid
    (dtmm/diffract.py:215-225)."""
    _, fin, fini = diffraction_alphaffi(shape, ks, epsv_in, epsa_in, betamax)
    _, fout, _ = diffraction_alphaffi(shape, ks, epsv_out, epsa_out, betamax)
    return tr_mat(fin, fout, fini, mode)


def _transfer_ray_2x2_1(fft_field, ks, layer, dmat1, dmat2, beta, phi, nsteps, mode, bulk,
                        tmat=None, rmat=None, refl=None):
    """2x2 layer step on the k-space E field (dtmm/propagate_2x2.py:19-87):
    reflection mode 1 passes the Fresnel transmission `tmat` of the
    effective-medium interface before the first sub-step.  With `refl`
    (npass > 1, :45-58) the reflected wave rmat.field replaces refl in place
    and the old refl is added to the transmitted field; backward passes
    (mode -1) add that field (real space) and its H to `bulk` before the
    layer, forward passes after it."""
    d, epsv, epsa = layer
    alpha, fmat = alphaf(beta, phi, epsv, epsa)
    e = E_mat(fmat, mode=mode)
    ei = inv(e)
    kd = ks * d
    p = phase_mat(alpha, kd[..., None, None], mode=mode)
    for j in range(nsteps):
        if j == 0 and tmat is not None:
            if refl is not None:
                trans = refl.copy()
                refl[...] = dotmf(rmat, fft_field)
                fft_field = dotmf(tmat, fft_field) + trans
                if mode == -1 and bulk is not None:
                    field = ifft2(fft_field)
                    bulk[..., ::2, :, :] += field
                    bulk[..., 1::2, :, :] += dotmf(E2H_mat(fmat, mode=mode), field)
            else:
                fft_field = dotmf(tmat, fft_field)
        if dmat1 is not None:
            fft_field = dotmf(dmat1, fft_field)
        field = ifft2(fft_field)
        field = dotmdmf(e, p, ei, field)
        fft_field = fft2(field)
        if dmat2 is not None:
            fft_field = dotmf(dmat2, fft_field)
    if mode == +1 and bulk is not None:
        field = ifft2(fft_field)
        e2h = E2H_mat(fmat, mode=mode)
        bulk[..., 1::2, :, :] += dotmf(e2h, field)
        bulk[..., ::2, :, :] += field
    return fft_field


def propagate_2x2_effective_1(field, ks, layer_out, effective_layer_out, beta=0, phi=0,
                              nsteps=1, diffraction=1, betamax=BETAMAX, mode=+1, bulk=None,
                              cache=None, reflection=0, effective_layer_in=None, refl=None):
    """Layer step of the 2x2 method, reflection mode 0 or 1
    (dtmm/propagate_2x2.py:193-211; diffraction in {0, 1}).  `refl`
    (npass > 1) is updated in place."""
    d_out, epsv_out, epsa_out = effective_layer_out
    shape = field.shape[-2:]
    tmat = rmat = None
    if reflection:
        _, epsv_in, epsa_in = effective_layer_in
        tmat, rmat = E_tr_matrix(shape, ks, epsv_in, epsa_in, epsv_out, epsa_out, mode, betamax)
    if diffraction == 1:
        cache = cache or _Cache()
        key = _key(shape, ks, beta, phi, d_out, epsv_out, epsa_out, betamax, mode)
        dmat1 = cache.get(("e1",) + key, lambda: first_E_diffraction_matrix(
            shape, ks, beta, phi, d_out / 2, epsv_out, epsa_out, mode, betamax))
        dmat2 = cache.get(("e2",) + key, lambda: second_E_diffraction_matrix(
            shape, ks, beta, phi, d_out / 2, epsv_out, epsa_out, mode, betamax))
    elif diffraction == 0:
        if refl is not None:
            # the reference applies dmat1 unconditionally after storing the
            # reflection (propagate_2x2.py:58), which fails without diffraction
            raise NotImplementedError("reflection mode 1 with npass > 1 needs diffraction 1")
        dmat1 = dmat2 = None
    elif diffraction > 1 and diffraction != np.inf:
        return _propagate_2x2_windows(field, ks, layer_out, effective_layer_out, beta, phi,
                                      nsteps, int(diffraction), betamax, mode, bulk, cache,
                                      tmat, rmat, refl)
    else:
        raise NotImplementedError("oracle covers finite diffraction >= 0 for the 2x2 method")
    return _transfer_ray_2x2_1(field, ks, layer_out, dmat1, dmat2, beta, phi, nsteps, mode, bulk,
                               tmat, rmat, refl)


def _propagate_2x2_windows(field, ks, layer_out, effective_layer_out, beta, phi, nsteps, n,
                           betamax, mode, bulk, cache, tmat, rmat, refl):
    """diffraction = n > 1 (dtmm/propagate_2x2.py:212-257): the k-space E
    field (and the stored reflection) is split by the n*n tent windows of
    fft_mask, each part goes through the layer step with its window's mean
    beam and that beam's E diffraction matrices, and the parts (and their
    reflections) are summed.  `bulk` receives every part's contribution."""
    d_out, epsv_out, epsa_out = effective_layer_out
    shape = field.shape[-2:]
    beta = np.asarray(beta, FDTYPE)
    phi = np.asarray(phi, FDTYPE)
    if beta.ndim > 0:
        bshape = beta.shape
        beta = beta[..., 0]
        phi = phi[..., 0]
    else:
        bshape = ()
    windows, (betas, phis) = fft_mask(field.shape, ks, n, beta * np.cos(phi),
                                      beta * np.sin(phi), betamax)
    nw = len(windows)
    betas = betas.reshape((nw,) + bshape)
    phis = phis.reshape((nw,) + bshape)
    key = _key(shape, ks, betas, phis, d_out, epsv_out, epsa_out, betamax, mode)
    dmats1 = cache.get(("ew1",) + key, lambda: first_E_diffraction_matrix(
        shape, ks, betas, phis, d_out / 2, epsv_out, epsa_out, mode, betamax))
    dmats2 = cache.get(("ew2",) + key, lambda: second_E_diffraction_matrix(
        shape, ks, betas, phis, d_out / 2, epsv_out, epsa_out, mode, betamax))
    fout = np.zeros_like(field)
    refl_sum = None if refl is None else np.zeros_like(refl)
    for window, b, p, dm1, dm2 in zip(windows, betas, phis, dmats1, dmats2):
        rpart = None if refl is None else refl * window
        fout += _transfer_ray_2x2_1(field * window, ks, layer_out, dm1, dm2, b, p, nsteps, mode,
                                    bulk, tmat, rmat, rpart)
        if refl is not None and tmat is not None:
            refl_sum += rpart
    if refl is not None:
        refl[...] = refl_sum
    return fout


# ---------------------------------------------------------------------------
# transfer.py
# ---------------------------------------------------------------------------

def _validate_betaphi(beta, phi, extendeddim=0):
    """(dtmm/transfer.py:43-65)."""
    if beta is None or phi is None:
        raise ValueError("Both beta and phi must be defined!")
    beta = np.asarray(beta, FDTYPE)
    phi = np.asarray(phi, FDTYPE)
    if beta.ndim != phi.ndim:
        raise ValueError("Beta nad phi should have same dimensions!")
    if beta.ndim == 1:
        if len(beta) != len(phi):
            raise ValueError("Beta nad phi should have same length!")
        for _ in range(extendeddim):
            beta = beta[..., None]
            phi = phi[..., None]
    elif beta.ndim != 0:
        raise ValueError("Only length 1 or scalar values are supported for beta and phi")
    else:
        for _ in range(extendeddim + 1):
            beta = beta[..., None]
            phi = phi[..., None]
    return beta, phi


# norm flags (dtmm/transfer.py:30-31)
DTMM_NORM_FFT = 1 << 0
DTMM_NORM_REF = 1 << 1


def exp_notch_filter(x, x0, sigma):
    """(dtmm/denoise.py:21-22)."""
    return np.asarray((1 - np.exp(-np.abs(x - x0).clip(0, x0) / sigma)) /
                      (1 - np.exp(-x0 / sigma)), FDTYPE)


def denoise_fftfield(ffield, ks, beta, smooth=1.):
    """Attenuates the k-modes around |k| = beta (dtmm/denoise.py:30-39)."""
    b, _ = betaphi(ffield.shape[-2:], ks)
    f = exp_notch_filter(b, beta, smooth)
    return ffield * f[..., None, :, :]


def denoise_field(field, ks, beta, smooth=1.):
    """(dtmm/denoise.py:24-28)."""
    return ifft2(denoise_fftfield(fft2(field), ks, beta, smooth))


def normalize_field(field, intensity_in, intensity_out):
    """Per-pixel amplitude factor |clip(I_in / I_out, -1, 1)|, 0 where
    I_out == 0 (dtmm/transfer.py:67-75)."""
    i_out = np.array(intensity_out, copy=True)
    i_in = np.array(intensity_in, copy=True)
    m = i_out == 0.
    i_out[m] = 1.
    i_in[m] = 0.
    fact = (i_in / i_out)[..., None, :, :]
    return field * np.abs(np.clip(fact, -1, 1))


def project_normalized_fft(field, dmat, ref=None):
    """Projection in k-space, each mode rescaled to the intensity it had
    before, or that the reference field has (dtmm/transfer.py:78-91,
    window=None)."""
    f1 = fft2(field)
    i1 = field2intensity(f1 if ref is None else fft2(ref))
    f2 = dotmf(dmat, f1)
    i2 = field2intensity(f2)
    return ifft2(normalize_field(f2, i1, i2))


def total_intensity(field):
    """Intensity summed over pixels (dtmm/transfer.py:200-204)."""
    i = field2intensity(field)
    return i.sum(axis=(-2, -1))


def normalize_field_total(field, i1, i2):
    """(dtmm/transfer.py:189-198)."""
    i2 = np.array(i2, copy=True)
    i1 = np.array(i1, copy=True)
    m = i2 == 0.
    i2[m] = 1.
    i1[m] = 0.
    fact = np.abs(np.clip(i1 / i2, -1, 1))
    return field * fact[..., None, None, None]


def project_normalized_total(field, dmat):
    """(dtmm/transfer.py:206-221, ref=None, window=None)."""
    i1 = total_intensity(field)
    out = ifft2(dotmf(dmat, fft2(field)))
    i2 = total_intensity(out)
    return normalize_field_total(out, i1, i2)


def normalize_f(fmat):
    """Columns of a field matrix scaled to unit |Poynting z|; zero columns
    stay zero (dtmm/tmm.py:701-725)."""
    fmat = np.asarray(fmat, CDTYPE)
    p = (fmat[..., 0, :] * np.conj(fmat[..., 1, :])).real - \
        (fmat[..., 2, :] * np.conj(fmat[..., 3, :])).real
    n = np.abs(p) ** 0.5
    nz = n != 0.
    scale = np.zeros_like(n)
    scale[nz] = 1. / n[nz]
    return fmat * scale[..., None, :]


def fvec2avec(fvec, fmat):
    """Mode amplitudes of field vectors in the normalised modes of fmat
    (dtmm/tmm.py:1851-1873)."""
    return dotmv(inv(normalize_f(fmat)), fvec)


def f_iso(beta=0., phi=0., n=1.):
    """Field matrix of an isotropic medium (dtmm/tmm.py:1074-1091)."""
    return alphaf(beta, phi, refind2eps([n] * 3), np.zeros(3))[1]


def polarizer(jones):
    """Jones polarizer j j^H (dtmm/jones.py:236-279)."""
    j = np.asarray(jones)
    return j[..., :, None] * np.conj(j[..., None, :])


def as4x4(jmat):
    """Jones matrix on the forward modes, its conjugate on the backward ones
    (dtmm/jones.py:371-398)."""
    jmat = np.asarray(jmat)
    out = np.zeros(jmat.shape[:-2] + (4, 4), jmat.dtype)
    out[..., 0::2, 0::2] = jmat
    out[..., 1::2, 1::2] = np.conj(jmat)
    return out


def ray_jonesmat4x4(jmat, epsv=(1., 1., 1.)):
    """4x4 field-basis matrix of a Jones matrix at normal incidence
    (dtmm/jones4.py:205-246 with beta = phi = 0, so the frame rotation is the
    identity; jonesmat4x4 :75-103)."""
    f = normalize_f(alphaf(0., 0., epsv, np.zeros(3))[1])
    return dotmm(f, dotmm(as4x4(jmat), inv(f)))


def jones2H(jones, ks, n=1., betamax=BETAMAX, mode=+1):
    """(Ex, Ey) -> (Hy, Hx) of the modes of an isotropic medium on the FFT
    grid, zero above betamax (dtmm/transfer.py:255-268)."""
    alpha, f, fi = diffraction_alphaffi(jones.shape[-2:], ks, refind2eps([n] * 3),
                                        np.zeros(3), betamax)
    return diffract(jones, E2H_mat(f, mode=mode))


def project_normalized_local(field, dmat, nout, ref=None):
    """Forward projection in the output medium, each pixel rescaled to the
    intensity the reference field has in that pixel's polarisation, plus the
    reference's part of the other polarisation (dtmm/transfer.py:103-138,
    window None).  The Jones vector of each pixel comes from the normal-
    incidence forward amplitudes of the projected field."""
    f = ifft2(dotmf(dmat, fft2(field)))
    a = fvec2avec(np.moveaxis(f, -3, -1), f_iso(n=nout))
    with np.errstate(invalid="ignore", divide="ignore"):
        jmat = polarizer(jonesvec(a[..., 0::2]))  # a zero vector gives NaN as there
    pmat1 = ray_jonesmat4x4(jmat, epsv=refind2eps([nout] * 3))
    pmat2 = -pmat1
    for i in range(4):
        pmat2[..., i, i] += 1.
    src = field if ref is None else ref
    intensity1 = field2intensity(dotmf(pmat1, src))
    ref = dotmf(pmat2, src)
    intensity2 = field2intensity(f)
    return normalize_field(f, intensity1, intensity2) + ref


def transfer_4x4(field_data, optical_data, beta=0., phi=0., nin=1., nout=1., npass=1,
                 nstep=1, diffraction=1, reflection=1, betamax=BETAMAX, norm=DTMM_NORM_FFT,
                 smooth=SMOOTH, layer_callback=None, ret_bulk=False, eff_data=None):
    """4x4 transfer, reflection mode 1 (dtmm/transfer.py:618-884).

    The input is not projected onto forward modes and the input/output media
    are not propagated (transfer.py:694, :700).  npass > 1 alternates
    forward and backward passes (reversed layers, negative thickness,
    transfer.py:744-804) with the normalisations of transfer.py:815-872
    (norm DTMM_NORM_FFT, 0 "total", DTMM_NORM_REF "local" or FFT|REF, the
    last two with a reference field propagated by the 2x2 method);
    like the reference, the input array of field_data (when complex128)
    then receives the field at the input side, reflections included.
    ret_bulk (npass 1) returns the fields of every layer index instead: the
    input, the real-space field after each layer and the output
    (transfer.py:670-690, :777-778, :876-877).
    """
    if ret_bulk and npass != 1:
        raise NotImplementedError("oracle covers ret_bulk with npass=1")
    if diffraction < 0 or diffraction == np.inf:
        # exact per-mode propagation of every layer in real space
        # (dtmm/transfer.py:719-725 work_in_fft False, :782-803), npass 1
        if npass != 1 or ret_bulk:
            raise NotImplementedError("oracle covers full diffraction with npass 1")
        field_in, wavelengths, pixelsize = field_data
        field = np.array(field_in, CDTYPE)
        layers, _ = create_layers(optical_data, nin, nout, nstep, eff_data)
        ks = k0(wavelengths, pixelsize)
        for j in range(1, len(layers) - 1):
            nsub, layer = layers[j]
            field = propagate_4x4_full(field, ks, layer, nsub, betamax)
        return field, wavelengths, pixelsize
    if reflection == 2 and diffraction >= 0 and npass == 1:
        # dtmm/transfer.py:694-700, :719-725 (work_in_fft False), :782-795
        # (propagate_4x4_effective_2), :874: the input is neither projected
        # nor transformed, input and output media are skipped
        field_in, wavelengths, pixelsize = field_data
        field = np.array(field_in, CDTYPE)
        layers, eff_layers = create_layers(optical_data, nin, nout, nstep, eff_data)
        ks = k0(wavelengths, pixelsize)
        beta, phi = _validate_betaphi(beta, phi, extendeddim=field.ndim - 2)
        cache = _Cache()
        bulk = np.zeros((len(layers),) + field.shape, CDTYPE) if ret_bulk else None
        for j in range(1, len(layers) - 1):
            nsub, layer = layers[j]
            _, eff = eff_layers[j]
            field = propagate_4x4_effective_2(field, ks, layer, eff, beta, phi, nsub, diffraction,
                                              betamax, cache)
            if bulk is not None:
                bulk[j] = field
            if layer_callback is not None:
                layer_callback(j)
        if bulk is not None:
            bulk[0] = field_in
            bulk[-1] = field
            return bulk, wavelengths, pixelsize
        return field, wavelengths, pixelsize
    if reflection not in (1, 2):
        raise NotImplementedError("oracle covers reflection modes 1 and 2 for the 4x4 method")
    # reflection mode 2 works in real space (dtmm/transfer.py:719-725)
    work_in_fft = reflection != 2
    # DTMM_NORM_REF (transfer.py:631-639): a reference field, propagated with
    # the 2x2 method (reflection 0) in the forward passes, sets the
    # intensities the normalisations restore; without DTMM_NORM_FFT the
    # normalisation is the per-pixel "local" one.  (The block of
    # transfer.py:705-708 compares the renamed norm string with 2 and never
    # runs.)
    calc_reference = bool(norm & DTMM_NORM_REF)
    if calc_reference and diffraction > 1:
        raise NotImplementedError("oracle covers DTMM_NORM_REF with diffraction 0 or 1")
    norm = "fft" if norm & DTMM_NORM_FFT else ("local" if calc_reference else "total")
    if smooth < 0:
        smooth = 0.
    field_in, wavelengths, pixelsize = field_data
    field_in = np.asarray(field_in, CDTYPE)
    layers, eff_layers = create_layers(optical_data, nin, nout, nstep, eff_data)
    ks = k0(wavelengths, pixelsize)
    n = len(layers)
    beta, phi = _validate_betaphi(beta, phi, extendeddim=field_in.ndim - 2)
    cache = _Cache()
    field_out = np.zeros_like(field_in)
    indices = list(range(1, n - 1))
    field0 = transmitted_field(field_in, ks, n=nin, betamax=betamax)
    i0 = total_intensity(field0)
    ref = field_in.copy() if calc_reference else None
    field = fft2(field_in.copy()) if work_in_fft else field_in.copy()
    if npass > 1:
        pin_mat = projection_matrix(field.shape[-2:], ks, epsv=refind2eps([nin] * 3), mode=+1,
                                    betamax=betamax)
        pout_mat = projection_matrix(field.shape[-2:], ks, epsv=refind2eps([nout] * 3), mode=+1,
                                     betamax=betamax)
    bulk = np.zeros((n,) + field_in.shape, CDTYPE) if ret_bulk else None
    for i in range(npass):
        direction = (-1) ** i
        for j in indices:
            nsub, (thickness, ev, ea) = layers[j]
            _, (d, e, a) = eff_layers[j]
            if calc_reference and i % 2 == 0:
                # the reference's E field through this layer
                # (transfer.py:769-775, propagate_2x2.py:284-305)
                dmat = None
                if diffraction:
                    key = _key(field.shape[-2:], ks, beta, phi, d * direction, e, a, betamax,
                               direction)
                    dmat = cache.get(("e2r",) + key, lambda: corrected_E_diffraction_matrix(
                        field.shape[-2:], ks, beta, phi, d * direction, e, a, direction,
                        betamax))
                ref[..., ::2, :, :], _ = _transfer_ray_2x2_2(
                    ref[..., ::2, :, :], ks, layers[j - 1][1], (thickness * direction, ev, ea),
                    dmat, beta, phi, nsub, direction, None, 0, None)
            if reflection == 2:
                field = propagate_4x4_effective_2(field, ks, (thickness * direction, ev, ea),
                                                  (d * direction, e, a), beta, phi, nsub,
                                                  diffraction, betamax, cache)
            else:
                field = propagate_4x4_effective_1(field, ks, (thickness * direction, ev, ea),
                                                  (d * direction, e, a), beta, phi, nsub,
                                                  diffraction, betamax, cache)
            if bulk is not None:
                bulk[j] = ifft2(field) if work_in_fft else field
            if layer_callback is not None:
                layer_callback(j)
        if calc_reference and i % 2 == 0:
            # H of the forward modes of the output medium (transfer.py:805-806)
            ref[..., 1::2, :, :] = jones2H(ref[..., ::2, :, :], ks, n=nout, betamax=betamax)
        indices.reverse()
        if work_in_fft:
            field = ifft2(field)
        if npass == 1:
            field_out[...] = field
            if bulk is not None:
                bulk[0] = field_in
                bulk[-1] = field_out
                return bulk, wavelengths, pixelsize
            break
        sigma = smooth * (npass - i) / npass
        if i % 2 == 0:
            if i == 0:
                field = field - field_out
            if i != npass - 1:
                if calc_reference:
                    ref = ref * (nin / nout) ** 0.5
                if norm == "fft":
                    field = project_normalized_fft(field, pout_mat, ref=ref)
                elif norm == "local":
                    field = project_normalized_local(field, pout_mat, nout, ref=ref)
                else:
                    field = project_normalized_total(field, pout_mat)
                field = denoise_field(field, ks, nin, sigma)
            field_out += field
            field = field_out.copy()
        else:
            field_in[...] = field
            if i != npass - 1:
                ffield = dotmf(pin_mat, fft2(field))
                field = ifft2(denoise_fftfield(ffield, ks, nin, sigma))
                fact = (i0 / total_intensity(field))[..., None, None, None]
                field = field * fact
                field_out *= fact
                field_in *= fact
                field = field0 - field
                field_in += field
                if calc_reference:
                    ref = field.copy()
        if work_in_fft:
            field = fft2(field)
    return field_out, wavelengths, pixelsize


def transfer_2x2(field_data, optical_data, beta=0., phi=0., nin=1., nout=1., npass=1, nstep=1,
                 diffraction=1, reflection=0, betamax=BETAMAX, layer_callback=None,
                 ret_bulk=False, eff_data=None):
    """2x2 (E-field) transfer (dtmm/transfer.py:1009-1183): reflection mode 0
    (diffraction 0 or 1), 1 (diffraction 1; diffraction 0 with npass 1) and
    2 (diffraction 0 or 1), any npass with reflection 1 or 2.

    npass > 1 (transfer.py:1076-1083, :1105-1178): the caller's input array
    is overwritten with its forward part and collects the backward field of
    every backward pass; the reflected waves of every interface are kept
    between passes; passes alternate direction, the odd ones run the layers
    in reverse with negative thickness and mode -1.  ret_bulk returns
    (n_layers + 2, ...): the input side, the field after every layer (E and
    the H of each voxel's modes, summed over passes) and the output
    (transfer.py:1045-1050, :1130-1140)."""
    if reflection not in (0, 1, 2):
        raise ValueError("Invalid reflection. The 2x2 method supports reflection mode 0,1 or 2.")
    if npass > 1 and reflection == 0:
        raise ValueError("Reflection mode `0` not compatible with npass > 1")
    user_in, wavelengths, pixelsize = field_data
    field_in = np.asarray(user_in, CDTYPE)
    layers, eff_layers = create_layers(optical_data, nin, nout, nstep, eff_data)
    ks = k0(wavelengths, pixelsize)
    n = len(layers) - 1
    beta, phi = _validate_betaphi(beta, phi, extendeddim=field_in.ndim - 2)
    if ret_bulk:
        bulk_out = np.zeros((n + 1,) + field_in.shape, CDTYPE)
        bulk_out[0] = field_in
        field_in = bulk_out[0]
        field_out = bulk_out[-1]
    else:
        bulk_out = None
        field_out = np.zeros_like(field_in)
    if diffraction > 0:
        field0 = transmitted_field(field_in, ks, n=nin, betamax=betamax)
    else:
        field0 = _transmitted_field_direct(field_in, beta, phi, n=nin)
    field = field0[..., ::2, :, :].copy()
    if npass > 1:
        field_in[...] = field0
        refl = np.zeros((n + 1,) + field.shape, CDTYPE)
    else:
        refl = [None] * n
    if reflection == 2 and diffraction > 1:
        raise NotImplementedError("oracle covers 2x2 reflection 2 with diffraction 0 or 1")
    if not (diffraction >= 0 and diffraction != np.inf):
        raise NotImplementedError("oracle covers finite diffraction >= 0 for the 2x2 method")
    # k-space for reflection 0/1 with diffraction, and reflection 1 always
    # (transfer.py:1085-1093); real space otherwise
    work_in_fft = reflection == 1 or (reflection == 0 and diffraction > 0)
    if work_in_fft:
        field = fft2(field)
    shape = field.shape[-2:]
    cache = _Cache()
    indices = list(range(n))
    for i in range(npass):
        if i > 0:
            field[...] = 0.
        direction = (-1) ** i
        for j in indices:
            jin = j + (1 - direction) // 2
            jout = j + (1 + direction) // 2
            _, input_layer = layers[jin]
            nsub, (t_out, ev_out, ea_out) = layers[jout]
            output_layer = (t_out * direction, ev_out, ea_out)
            _, (d_in, e_in, a_in) = eff_layers[jin]
            _, (d_out, e_out, a_out) = eff_layers[jout]
            eff_in = (d_in * direction, e_in, a_in)
            eff_out = (d_out * direction, e_out, a_out)
            if jout == 0:
                bulk = field_in
            elif jout == n:
                bulk = field_out
            else:
                bulk = None if bulk_out is None else bulk_out[jout]
            if work_in_fft:
                field = propagate_2x2_effective_1(field, ks, output_layer, eff_out, beta, phi,
                                                  nsub, diffraction, betamax, direction, bulk,
                                                  cache, reflection, eff_in, refl[j])
            else:
                dmat = None
                if diffraction:
                    key = _key(shape, ks, beta, phi, eff_out[0], e_out, a_out, betamax, direction)
                    dmat = cache.get(("e2",) + key, lambda: corrected_E_diffraction_matrix(
                        shape, ks, beta, phi, eff_out[0], e_out, a_out, direction, betamax))
                field, _ = _transfer_ray_2x2_2(field, ks, input_layer, output_layer, dmat, beta,
                                               phi, nsub, direction, bulk, reflection, refl[j])
            if layer_callback is not None:
                layer_callback(j)
        indices.reverse()
    if npass > 1 and not ret_bulk and field_in is not user_in:
        # the reference writes into the caller's array (transfer.py:1080)
        try:
            user_in[...] = field_in
        except (TypeError, ValueError):
            pass
    return (bulk_out if ret_bulk else field_out), wavelengths, pixelsize


def _transmitted_field_direct(field, beta, phi, n=1.):
    """(dtmm/transfer.py:94-101)."""
    ev = refind2eps([n] * 3)
    alpha, fmat = alphaf(beta, phi, ev, np.zeros_like(ev))
    diag = np.zeros(fmat.shape[:-1], CDTYPE)
    diag[..., 0::2] = 1.
    pmat = dotmdm(fmat, diag, inv(fmat))
    return _v2f(np.matmul(pmat, _f2v(field)))


def transfer_field(field_data, optical_data, beta=None, phi=None, nin=None, nout=None,
                   npass=None, nstep=1, diffraction=None, reflection=None, method=None,
                   betamax=None, norm=DTMM_NORM_FFT, smooth=SMOOTH, layer_callback=None,
                   ret_bulk=False, split_wavelengths=False, multiray=False, eff_data=None):
    """Defaults and dispatch of dtmm.transfer_field (dtmm/transfer.py:322-615)
    for the covered modes.  The configured default method is the quoted string
    '"2x2"' (dtmm/dtmm.ini:23 read by configparser), which makes the automatic
    reflection mode 1 (transfer.py:459).

    Per-wavelength transfers (transfer.py:486-528): with split_wavelengths
    (and no ret_bulk) the (..., wavelengths, 4, H, W) input is cut into one
    field per wavelength and the result put back together; a tuple of
    per-wavelength fields (each (..., 4, H, W)) returns a tuple of
    (field, wavelength, pixelsize) results.  Each wavelength is an
    independent transfer with the optical data evaluated there (a callable
    eps is called with the wavelength), so dispersive data is allowed only
    on this route (transfer.py:535-536)."""
    field_in, wavelengths, pixelsize = field_data
    kw = dict(beta=beta, phi=phi, nin=nin, nout=nout, npass=npass, nstep=nstep,
              diffraction=diffraction, reflection=reflection, method=method, betamax=betamax,
              norm=norm, smooth=smooth, layer_callback=layer_callback, ret_bulk=ret_bulk,
              multiray=multiray, eff_data=eff_data)
    if split_wavelengths and not isinstance(field_in, tuple) and not ret_bulk:
        field_in = np.asarray(field_in)
        parts = tuple(field_in[..., i, :, :, :] for i in range(len(wavelengths)))
        res = transfer_field((parts, wavelengths, pixelsize), optical_data, **kw)
        return np.stack([r[0] for r in res], -4), wavelengths, pixelsize
    if isinstance(field_in, tuple):
        out = []
        for f, w in zip(field_in, wavelengths):
            # one wavelength: a wavelength axis of length 1 computes what the
            # reference's scalar wave number does
            f1 = np.expand_dims(f, -4)
            opt = validate_optical_data(optical_data, wavelength=w)
            r = _transfer_field(f1, np.asarray([w], FDTYPE), pixelsize, opt, **kw)[0]
            out.append((np.squeeze(r, -4), w, pixelsize))
        return tuple(out)
    return _transfer_field(field_in, wavelengths, pixelsize, optical_data, **kw)


def _transfer_field(field_in, wavelengths, pixelsize, optical_data, beta=None, phi=None,
                    nin=None, nout=None, npass=None, nstep=1, diffraction=None, reflection=None,
                    method=None, betamax=None, norm=DTMM_NORM_FFT, smooth=SMOOTH,
                    layer_callback=None, ret_bulk=False, multiray=False, eff_data=None):
    """The single transfer of transfer_field (dtmm/transfer.py:554-615).
    Without beta and phi the beam is inferred from the input field
    (transfer.py:662-664, :1040-1042)."""
    field_data = (field_in, wavelengths, pixelsize)
    if beta is None and phi is None:
        beta, phi = field2betaphi(field_in, k0(wavelengths, pixelsize), multiray)
    nin = N_COVER if nin is None else nin
    nout = N_COVER if nout is None else nout
    method = '"2x2"' if method is None else method
    npass = 1 if npass is None else npass
    diffraction = 1 if diffraction is None else diffraction
    betamax = BETAMAX if betamax is None else betamax
    if reflection is None:
        reflection = 0 if method == "2x2" else 1
        if method == "4x4" and diffraction == 0:
            reflection = 2
        if npass > 1:
            reflection = 1
            if diffraction > 1 and method == "2x2":
                reflection = 2
    if method == "4x4":
        return transfer_4x4(field_data, optical_data, beta, phi, nin, nout, npass, nstep,
                            diffraction, reflection, betamax, norm, smooth, layer_callback,
                            ret_bulk, eff_data)
    return transfer_2x2(field_data, optical_data, beta, phi, nin, nout, npass, nstep,
                        diffraction, reflection, betamax, layer_callback, ret_bulk, eff_data)


# ---------------------------------------------------------------------------
# field.py
# ---------------------------------------------------------------------------

def field2intensity(field):
    """Poynting z per pixel of a (..., 4, H, W) field (dtmm/field.py:771-799)."""
    f = np.asarray(field)
    return (f[..., 0, :, :] * np.conj(f[..., 1, :, :])).real - \
        (f[..., 2, :, :] * np.conj(f[..., 3, :, :])).real


def _diffracted_field(field, ks, d=0., n=1., mode="t", betamax=BETAMAX):
    """(dtmm/diffract.py:301-304)."""
    pmat = field_diffraction_matrix(field.shape[-2:], ks, d=d, epsv=refind2eps([n] * 3),
                                    epsa=(0., 0., 0.), mode=mode, betamax=betamax)
    return diffract(field, pmat)


def illumination_data(shape, wavelengths, pixelsize=1., beta=0., phi=0., intensity=1.,
                      n=None, focus=0., window=None, jones=None, diffraction=True,
                      betamax=BETAMAX, backdir=False):
    """Forward (backward with backdir) propagating input field data
    (dtmm/field.py:586-713); without a window the plane waves are snapped to
    FFT eigen-waves (field.py:686-687); with diffraction the field is
    diffracted by -focus in the cover medium (:703-711)."""
    n = N_COVER if n is None else n
    wavelengths = np.asarray(wavelengths)
    wavenumbers = 2 * np.pi / wavelengths * pixelsize
    if jones is None:
        intensity = intensity / 2.
    mode = -1 if backdir else +1
    _beta = np.asarray(beta, FDTYPE)
    _phi = np.asarray(phi, FDTYPE)
    _intensity = np.asarray(intensity, FDTYPE)
    nrays = len(_beta) if _beta.ndim > 0 else 1
    b = _beta[:, None] if _beta.ndim == 1 else _beta
    p = _phi[:, None] if _phi.ndim == 1 else _phi
    bb = b if b.ndim >= 1 else b[..., None]
    pp = p if p.ndim >= 1 else p[..., None]
    waves = planewave(shape, wavenumbers, bb, pp)
    if window is None:
        waves = wave2eigenwave(waves)
    else:
        waves = waves * window
    b = _beta[:, None, None, None] if _beta.ndim == 1 else _beta[..., None, None]
    p = _phi[:, None, None, None] if _phi.ndim == 1 else _phi[..., None, None]
    inten = _intensity[..., None, None, None, None]
    alpha, fmat = alphaf(b, p, refind2eps([n] * 3), np.zeros(3))
    if jones is None:
        f1 = np.moveaxis(fvec(fmat, jonesvec((1, 0), p), waves, mode), -1, -3)
        f2 = np.moveaxis(fvec(fmat, jonesvec((0, 1), p), waves, mode), -1, -3)
        shp = list(f1.shape)
        shp.insert(-4, 2)
        field = np.empty(shp, CDTYPE)
        field[..., 0, :, :, :, :] = f1
        field[..., 1, :, :, :, :] = f2
    else:
        field = np.moveaxis(fvec(fmat, jonesvec(jones, p), waves, mode), -1, -3).copy()
    i1 = field2intensity(field)
    norm = np.ones_like(i1)
    norm[:, ...] = (inten / nrays) ** 0.5
    if diffraction:
        field = _diffracted_field(field, wavenumbers, d=-focus, n=n, mode=mode, betamax=betamax)
        i2 = field2intensity(field)
        ratio = (i1.sum((-2, -1)) / i2.sum((-2, -1))) ** 0.5
        norm = norm * ratio[..., None, None]
    field = norm[..., None, :, :] * field
    return field, wavelengths, pixelsize


# ---------------------------------------------------------------------------
# jones.py, jones4.py, field_viewer.py, color.py (the step after the transfer)
# ---------------------------------------------------------------------------

def polarizer(jones):
    """j j^H (dtmm/jones.py:236-278)."""
    j = np.asarray(jones, CDTYPE)
    return j[..., :, None] * np.conj(j[..., None, :])


def as4x4(jmat):
    """J on the forward, conj(J) on the backward modes (dtmm/jones.py:371-397)."""
    jmat = np.asarray(jmat, CDTYPE)
    out = np.zeros(jmat.shape[:-2] + (4, 4), CDTYPE)
    out[..., 0::2, 0::2] = jmat
    out[..., 1::2, 1::2] = np.conj(jmat)
    return out


def jonesmat4x4(jmat, fmat):
    """F as4x4(J) F^-1 with F normalised (dtmm/jones4.py:75-101)."""
    f = normalize_f(fmat)
    return dotmm(f, dotmm(as4x4(jmat), inv(f)))


def ray_jonesmat4x4(jmat, beta=0., phi=0., epsv=(1., 1., 1.), epsa=(0., 0., 0.)):
    """(dtmm/jones4.py:205-246)."""
    _, f = alphaf(beta, phi, epsv, epsa)
    c, s = np.cos(phi), np.sin(phi)
    r = np.asarray([[c, -s], [s, c]], CDTYPE)
    return jonesmat4x4(np.swapaxes(r, -1, -2) @ np.asarray(jmat, CDTYPE) @ r, f)


def field2specter(field):
    """Poynting z per wavelength with the wavelength axis last:
    (..., nlam, 4, H, W) -> (..., H, W, nlam) (dtmm/field.py:800-841)."""
    return np.moveaxis(field2intensity(field), -3, -1)


def calculate_pom_field(field, jvec=None, pmat=None, dmat=None):
    """Jones combination of the x/y fields, diffraction by dmat, analyser
    pmat in real space (dtmm/field_viewer.py:70-143, input_fft = output_fft =
    False, window None)."""
    field = np.asarray(field, CDTYPE)
    if jvec is not None:
        if field.shape[-5] != 2:
            raise ValueError("Invalid field shape.")
        c, s = jvec
        field = field[..., 0, :, :, :, :] * c + field[..., 1, :, :, :, :] * s
    if dmat is not None:
        field = diffract(field, dmat)
    if pmat is not None:
        field = dotmf(pmat, field)
    return field


def pom_specter(field, ks, jvec=None, pmat=None, focus=0., n=1., mode=None,
                betamax=BETAMAX, diffraction=True):
    """FieldViewer.calculate_specter (dtmm/field_viewer.py:1240-1330): the
    diffraction matrix of the immersion medium (:1017-1029), then per leading
    field calculate_pom_field and field2specter, summed."""
    field = np.asarray(field, CDTYPE)
    ks = np.asarray(ks, FDTYPE)
    dmat = None
    if diffraction:
        dmat = field_diffraction_matrix(field.shape[-2:], ks, d=focus,
                                        epsv=refind2eps([n] * 3), epsa=(0., 0., 0.),
                                        mode=mode, betamax=betamax)
    if pmat is not None:
        pmat = np.asarray(pmat, CDTYPE)
        if pmat.ndim == 3:  # one per wavelength, broadcast over pixels (:969-975)
            pmat = pmat[:, None, None, :, :]
    nd = 5 if jvec is None else 6
    tmp = field.reshape((-1,) + field.shape[-(nd - 1):])
    spec = 0.
    for data in tmp:
        spec = spec + field2specter(calculate_pom_field(data, jvec, pmat, dmat))
    return spec


# sRGB of D65 white (dtmm/color.py:39-50): the IEC 61966-2-1 matrix and curve.
XYZ2RGBD65 = np.array([[3.2404542, -1.5371385, -0.4985314],
                       [-0.9692660, 1.8760108, 0.0415560],
                       [0.0556434, -0.2040259, 1.0572252]])


def apply_gamma(value, gamma):
    """(dtmm/color.py:52-69)."""
    v = np.asarray(value, FDTYPE)
    return np.where(v > 1., 1., np.where(v < 0, 0., np.abs(v) ** (1. / gamma)))


def apply_srgb_gamma(value):
    """(dtmm/color.py:71-85)."""
    v = np.asarray(value, FDTYPE)
    lo = np.where(v < 0., 0., 12.92 * v)
    hi = np.where(v > 1., 1., (1 + 0.055) * np.abs(v) ** (1 / 2.4) - 0.055)
    return np.where(v < 0.0031308, lo, hi)


def specter2color(spec, cmf, norm=False, gamma=True, gray=False):
    """(dtmm/color.py:111-200)."""
    cmf = np.asarray(cmf, FDTYPE)
    if cmf.shape[-1] != 3:
        raise ValueError("Grayscale cmf! Cannot convert to color.")
    out = np.asarray(spec, FDTYPE) @ cmf
    if norm is True:
        out = out / out.max()
    elif norm != 0:
        out = out / norm
    if gray:
        out = np.repeat(out[..., 1:2], 3, axis=-1)
    else:
        out = out @ XYZ2RGBD65.T
    if gamma is True:
        out = apply_srgb_gamma(out)
    elif gamma is not False:
        out = apply_gamma(out, gamma)
    return out


# ---------------------------------------------------------------------------
# tmm.py: the non-iterative 1D transfer matrix method (method "4x4")
# ---------------------------------------------------------------------------

def f_iso(beta=0., phi=0., n=1.):
    """Field matrix of an isotropic medium, not normalised (dtmm/tmm.py:1074-1091)."""
    return alphaf(beta, phi, refind2eps([n] * 3), np.zeros(3))[1]


TMM_METHODS = ("4x4", "4x4_1", "4x4_2", "4x4_r", "2x2", "2x2_1")


def layer_mat(kd, epsv, epsa, beta=0., phi=0., cfact=0.1, method="4x4", fmatin=None,
              retfmat=False):
    """Characteristic matrix of one layer (dtmm/tmm.py:1112-1213).

    "4x4": F P(-kd) F^-1 with F normalised; "4x4_1": the backward phases
    zeroed; "4x4_2": the backward modes damped, alpha += 2i cfact Re(alpha);
    "4x4_r": the 8x8 matrix f1 pmat f1^-1 with f1 = diag(F, F) and pmat
    holding the forward-undamped, backward-damped phases (rows 0-3) and the
    plain phases copied to rows 4-7; "2x2": E P(kd) E^-1 of the forward
    modes (F not normalised); "2x2_1": the same with E^-1 replaced by the
    transmission from the previous layer's modes `fmatin` (Eti_mat)."""
    if method in ("2x2", "2x2_1"):
        alpha, fmat = alphaf(beta, phi, epsv, epsa)
        f = E_mat(fmat, mode=+1)
        if fmatin is not None and method == "2x2_1":
            fi = Eti_mat(fmatin, fmat, mode=+1)
        else:
            fi = inv(f)
        pmat = phase_mat(alpha[..., ::2], kd)
    elif method in ("4x4", "4x4_1", "4x4_r", "4x4_2"):
        alpha, fmat = alphaf(beta, phi, epsv, epsa)
        fmat = normalize_f(fmat)
        fi = inv(fmat)
        f = fmat
        kd = np.asarray(kd, FDTYPE)
        if method == "4x4_r":
            alpha1 = alpha.copy()
            alpha1[..., 1::2] = alpha1[..., 1::2] + alpha1[..., 1::2].real * 2j * cfact
            alpha = alpha.copy()
            alpha[..., ::2] = alpha[..., ::2] - alpha[..., ::2].real * 2j * cfact
            pmat1 = phase_mat(alpha1, -kd)
            pmat2 = phase_mat(alpha, -kd)
            pmat = np.zeros(pmat1.shape[:-1] + (8, 8), CDTYPE)
            for i in range(4):
                pmat[..., i, i] = pmat1[..., i]
            pmat[..., 4, 0] = pmat1[..., 0]
            pmat[..., 6, 2] = pmat1[..., 2]
            pmat[..., 5, 1] = pmat2[..., 1]
            pmat[..., 7, 3] = pmat2[..., 3]
            f1 = np.zeros_like(pmat)
            f1[..., 0:4, 0:4] = f
            f1[..., 4:8, 4:8] = f
            out = dotmm(f1, dotmm(pmat, inv(f1)))
            return (fmat, out) if retfmat else out
        if method == "4x4_2":
            alpha = alpha.copy()
            alpha[..., 1::2] = alpha[..., 1::2] + alpha[..., 1::2].real * 2j * cfact
        pmat = phase_mat(alpha, -kd)
        if method == "4x4_1":
            pmat[..., 1::2] = 0.
    else:
        raise ValueError("Unknown method!")
    out = dotmdm(f, pmat, fi)
    return (fmat, out) if retfmat else out


def stack_mat(kd, epsv, epsa, beta=0., phi=0., cfact=0.01, method="4x4"):
    """M_1 . M_2 ... M_n for the 4x4 methods, M_n ... M_2 . M_1 for the 2x2
    ones; "2x2_1" enters each layer from the previous one's modes
    (dtmm/tmm.py:1215-1296)."""
    out = None
    fmat = None
    for i in range(len(kd)):
        if method == "2x2_1":
            fmat, m = layer_mat(kd[i], epsv[i], epsa[i], beta, phi, cfact, method, fmatin=fmat,
                                retfmat=True)
        else:
            m = layer_mat(kd[i], epsv[i], epsa[i], beta, phi, cfact, method)
        if out is None:
            out = m.copy()
        elif method.startswith("2x2"):
            out = dotmm(m, out)
        else:
            out = dotmm(out, m)
    return out


def system_mat(cmat, fmatin=None, fmatout=None):
    """Fin^-1 . C . Fout (dtmm/tmm.py:1298-1336).  An 8x8 matrix ("4x4_r")
    has its two diagonal blocks transformed and is reduced to 4x4 by the
    reference's selection matrices: rows 0, 5, 2, 7 of its first four
    columns."""
    if fmatin is None:
        fmatin = f_iso()
    if fmatout is None:
        fmatout = fmatin
    fmatini = inv(fmatin)
    if cmat.shape[-1] == 8:
        c = np.array(cmat, CDTYPE)
        c[..., 0:4, 0:4] = dotmm(dotmm(fmatini, c[..., 0:4, 0:4]), fmatout)
        c[..., 4:8, 4:8] = dotmm(dotmm(fmatini, c[..., 4:8, 4:8]), fmatout)
        m1 = np.zeros((8, 4))
        m1[np.arange(4), np.arange(4)] = 1.
        m0 = np.zeros((4, 8))
        m0[0, 0] = m0[1, 5] = m0[2, 2] = m0[3, 7] = 1.
        return np.matmul(m0, np.matmul(c, m1))
    return dotmm(dotmm(fmatini, cmat), fmatout)


def reflection_mat(smat):
    """(dtmm/tmm.py:1338-1357)."""
    smat = np.asarray(smat, CDTYPE)
    m1 = np.zeros_like(smat)
    m2 = np.zeros_like(smat)
    for i in range(smat.shape[-1] // 2):
        m1[..., 2 * i + 1, 2 * i + 1] = 1.
        m2[..., 2 * i, 2 * i] = -1.
    m1[..., :, 0::2] = -smat[..., :, 0::2]
    m2[..., :, 1::2] = smat[..., :, 1::2]
    return dotmm(inv(m1), m2)


def transmit4x4(fvec_in, cmat, fmatin=None, fmatout=None, fvec_out=None):
    """Output field of the stack; returns (fvec_out, fvec_in with the
    reflected waves added) (dtmm/tmm.py:1491-1563)."""
    fvec_in = np.asarray(fvec_in, CDTYPE)
    if fmatin is None:
        fmatin = f_iso()
    if fmatout is None:
        fmatout = fmatin
    fmatini, fmatouti = inv(fmatin), inv(fmatout)
    shape = np.broadcast(fvec_in[..., None], cmat[..., 0:4, 0:4], fmatin, fmatout).shape[:-1]
    smat = system_mat(cmat, fmatin, fmatout)
    avec = dotmv(fmatini, fvec_in)
    a = np.zeros(shape, CDTYPE)
    a[..., 0::2] = avec[..., 0::2]
    avec = a.copy()
    if fvec_out is not None:
        bvec = dotmv(fmatouti, fvec_out)
        a[..., 1::2] = bvec[..., 1::2]
    else:
        bvec = np.zeros_like(avec)
    out = dotmv(reflection_mat(smat), a)
    avec[..., 1::2] = out[..., 1::2]
    bvec[..., ::2] = out[..., ::2]
    return dotmv(fmatout, bvec), dotmv(fmatin, avec)


def transfer4x4(fvec_in, kd, epsv, epsa, beta=0., phi=0., nin=1., nout=1., fvec_out=None,
                method="4x4"):
    """(dtmm/tmm.py:1565-1673): returns (fvec_out, fvec_in with
    reflections).  As in the reference, the stack matrix is built by
    stack_mat(kd, epsv, epsa, method=method) with its default beam (beta 0,
    phi 0, :1655) and cfact 0.01, while the media use beta and phi."""
    if method not in ("4x4", "4x4_1", "4x4_r", "4x4_2"):
        raise ValueError("Unknown method '{}'!".format(method))
    fmatin = f_iso(beta, phi, nin)
    fmatout = f_iso(beta, phi, nout)
    cmat = stack_mat(kd, epsv, epsa, method=method)
    return transmit4x4(fvec_in, cmat, fmatin, fmatout, fvec_out)


def projection_mat(fmat, mode=+1):
    """F diag(1 on the modes of `mode`) F^-1 (dtmm/tmm.py:742-768)."""
    diag = np.zeros(fmat.shape[:-1], CDTYPE)
    if mode == +1:
        diag[..., 0::2] = 1.
    else:
        diag[..., 1::2] = 1.
    return dotmdm(fmat, diag, inv(fmat))


def t_mat(fmatin, fmatout, mode=+1):
    """E transmission from the modes of fmatin into those of fmatout
    (dtmm/tmm.py:952-972)."""
    return dotmm(E_mat(fmatout, mode=mode), Eti_mat(fmatin, fmatout, mode=mode))


def fvec2E(fvec, fmat=None, mode=+1):
    """(dtmm/tmm.py:1359-1388)."""
    if fmat is None:
        fmat = f_iso()
    return dotmv(projection_mat(fmat, mode), np.asarray(fvec, CDTYPE))[..., ::2]


def E2fvec(evec, fmat=None, mode=+1):
    """(dtmm/tmm.py:1390-1412)."""
    evec = np.asarray(evec, CDTYPE)
    if fmat is None:
        fmat = f_iso()
    out = np.empty(evec.shape[:-1] + (4,), CDTYPE)
    out[..., ::2] = evec
    out[..., 1::2] = dotmv(E2H_mat(fmat, mode), evec)
    return out


def transfer2x2(evec, kd, epsv, epsa, beta=0., phi=0., nin=None, nout=None, method="2x2"):
    """(dtmm/tmm.py:1675-1695): optional Fresnel transmission from the
    isotropic input medium into the first layer and from the last layer into
    the output medium (beta, phi); the stack matrix with its default beam."""
    if method not in ("2x2", "2x2_1"):
        raise ValueError("Unknown method!")
    evec = np.asarray(evec, CDTYPE)
    if nin is not None:
        fin = f_iso(beta, phi, nin)
        _, fout = alphaf(beta, phi, epsv[0], epsa[0])
        evec = dotmv(t_mat(fin, fout), evec)
    evec = dotmv(stack_mat(kd, epsv, epsa, method=method), evec)
    if nout is not None:
        fout = f_iso(beta, phi, nout)
        _, fin = alphaf(beta, phi, epsv[-1], epsa[-1])
        evec = dotmv(t_mat(fin, fout), evec)
    return evec


def transfer1d(fvec_in, kd, epsv, epsa, beta=0., phi=0., nin=1., nout=1., method="2x2",
               reflect_in=False, reflect_out=False, fvec_out=None):
    """``transfer`` (dtmm/tmm.py:1697-1748).  The 2x2 methods return only the
    transmitted field (fvec_in untouched); the 4x4 methods return (fvec_out,
    fvec_in with reflections) as transfer4x4 does."""
    if method.startswith("2x2"):
        fin = f_iso(beta, phi, nin)
        fout = f_iso(beta, phi, nout)
        evec = fvec2E(fvec_in, fin, +1)
        evec = transfer2x2(evec, kd, epsv, epsa, beta, phi, nin if reflect_in else None,
                           nout if reflect_out else None, method)
        return E2fvec(evec, fout, +1)
    return transfer4x4(fvec_in, kd, epsv, epsa, beta, phi, nin, nout, fvec_out, method)


# ---------------------------------------------------------------------------
# tmm2d.py: the non-iterative 4x4 method for 2D data (x, z), in the space of
# the 1D FFT modes of the x axis
# ---------------------------------------------------------------------------

def betax1(n, k0):
    """(dtmm/wave.py:507-527)."""
    k0 = np.abs(np.asarray(k0, FDTYPE)[..., None])
    return (2 * np.pi / k0) * np.asarray(np.fft.fftfreq(n), FDTYPE)


def eigenmask1(n, k0, betamax=BETAMAX):
    """Modes with |beta_x| <= betamax (dtmm/wave.py:530-552)."""
    return np.abs(betax1(n, k0)) <= betamax


def _per_k(mask, k0, fn):
    k0 = np.asarray(k0)
    if k0.ndim == 0:
        return fn(mask, k0)
    return tuple(fn(mask[i], k0[i]) for i in range(len(k0)))


def eigenbetax1(n, k0, betamax=BETAMAX):
    """(dtmm/wave.py:609-632)."""
    b = betax1(n, k0)
    m = np.abs(b) <= betamax
    if m.ndim == 2:
        return tuple(b[i][m[i]] for i in range(m.shape[0]))
    return b[m]


def eigenindices1(n, k0, betamax=BETAMAX):
    """(dtmm/wave.py:647-671)."""
    ii = np.arange(n)
    m = eigenmask1(n, k0, betamax)
    if m.ndim == 2:
        return tuple(ii[m[i]] for i in range(m.shape[0]))
    return ii[m]


def eigenwave1(n, i, amplitude=None):
    """ifft of a unit spike (amplitude n by default) at mode i
    (dtmm/wave.py:482-504)."""
    f = np.zeros((n,), CDTYPE)
    f[..., i] = n if amplitude is None else amplitude
    return np.fft.ifft(f)


def field2modes1(field, k0, betamax=BETAMAX):
    """(mask, modes (..., n, 4)) of a (..., [lambda], 4, W) field
    (dtmm/field.py:216-250); one modes array per wavelength for several."""
    if isinstance(field, tuple):
        out = tuple(field2modes1(field[i], k0[i], betamax) for i in range(len(field)))
        return tuple(o[0] for o in out), tuple(o[1] for o in out)
    f = np.fft.fft(np.asarray(field, CDTYPE))
    k0 = np.asarray(k0)
    mask = eigenmask1(f.shape[-1], k0, betamax)
    if k0.ndim == 0:
        return mask, np.moveaxis(f[..., mask], -2, -1)
    return mask, tuple(np.moveaxis(f[..., i, :, :][..., mask[i]], -2, -1)
                       for i in range(len(k0)))


def modes2field1(mask, modes):
    """Inverse of field2modes1 (dtmm/field.py:307-340)."""
    if isinstance(mask, tuple):
        return tuple(modes2field1(mask[i], modes[i]) for i in range(len(mask)))
    if mask.ndim == 1:
        out = np.zeros(modes.shape[:-2] + (4,) + mask.shape[-1:], CDTYPE)
        out[..., mask] = np.moveaxis(modes, -1, -2)
        return np.fft.ifft(out)
    out = np.zeros(modes[0].shape[:-2] + (len(mask), 4) + mask.shape[-1:], CDTYPE)
    for i, (mode, m) in enumerate(zip(modes, mask)):
        o = out[..., i, :, :]
        o[..., m] = np.moveaxis(mode, -1, -2)
    return np.fft.ifft(out)


def bdotmm(m1, m2):
    """Product of two n x n block matrices of 4x4 blocks (dtmm/linalg.py:1101-1116)."""
    return np.einsum("ikab,kjbc->ijac", m1, m2)


def bdotmd(m1, m2):
    """Full times block diagonal (n, 4, 4) (dtmm/linalg.py:1134-1145)."""
    return np.einsum("ijab,jbc->ijac", m1, m2)


def bdotdm(m1, m2):
    """Block diagonal (n, 4, 4) times full (dtmm/linalg.py:1156-1168)."""
    return np.einsum("iab,ijbc->ijac", m1, m2)


def _layer_mat2d(k0, d, epsv, epsa, mask, betaxs, betay, indices, method):
    """Column j: the FFT (restricted to the mask) of F P(-k0 d) F^-1 at mode
    j's beam times mode j's eigen-wave (dtmm/tmm2d.py:61-92)."""
    n = len(betaxs)
    kd = k0 * d
    W = epsv.shape[-2]
    out = np.empty((n, n, 4, 4), CDTYPE)
    beta = np.sqrt(betaxs ** 2 + betay ** 2)
    phi = np.arctan2(betay, betaxs)
    for j, (b, p) in enumerate(zip(beta, phi)):
        alpha, f, fi = alphaffi(b, p, epsv, epsa)
        pmat = phase_mat(alpha, -kd)
        if method == "4x4_1":
            pmat[..., 1::2] = 0.
        elif method != "4x4":
            raise ValueError("Unsupported method!")
        wave = eigenwave1(W, indices[j], amplitude=1.)
        m = dotmdm(f, pmat, fi)
        mf = np.fft.fft(m * wave[..., None, None], axis=-3)
        out[:, j, :, :] = mf[mask, ...]
    return out


def layer_mat2d(k0, d, epsv, epsa, betay=0., method="4x4", mask=None):
    """(dtmm/tmm2d.py:20-59): one (n, n, 4, 4) block matrix per wavelength."""
    if method not in ("4x4", "4x4_1"):
        raise ValueError("Unsupported method: '{}'".format(method))
    k0 = np.asarray(k0)
    W = epsv.shape[-2]
    if mask is None:
        mask = eigenmask1(W, k0)
    betas = _per_k(mask, k0, lambda m, k: betax1(W, k)[m])
    indices = _per_k(mask, k0, lambda m, k: np.arange(W)[m])
    if k0.ndim == 0:
        return _layer_mat2d(k0, d, epsv, epsa, mask, betas, betay, indices, method)
    return tuple(_layer_mat2d(k0[i], d, epsv, epsa, mask[i], betas[i], betay, indices[i],
                              method) for i in range(len(k0)))


def stack_mat2d(k, d, epsv, epsa, betay=0., method="4x4", mask=None):
    """M_1 . M_2 ... M_n of the block matrices (dtmm/tmm2d.py:94-119)."""
    out = None
    for i in range(len(d)):
        mat = layer_mat2d(k, d[i], epsv[i], epsa[i], betay=betay, mask=mask, method=method)
        if out is None:
            out = tuple(m.copy() for m in mat) if isinstance(mat, tuple) else mat.copy()
        elif isinstance(mat, tuple):
            out = tuple(bdotmm(o, m) for o, m in zip(out, mat))
        else:
            out = bdotmm(out, mat)
    return out


def f_iso2d(shape, k0, n=1., betay=0., betamax=BETAMAX):
    """Field matrices of the isotropic medium for every mode (dtmm/tmm2d.py:121-134)."""
    k0 = np.asarray(k0)
    betax = eigenbetax1(shape, k0, betamax)

    def one(bx):
        return f_iso(np.sqrt(bx ** 2 + betay ** 2), np.arctan2(betay, bx), n)
    if k0.ndim == 0:
        return one(betax)
    return tuple(one(b) for b in betax)


def system_mat2d(fmatin, cmat, fmatout):
    """Fin^-1 . C . Fout block-wise (dtmm/tmm2d.py:144-159)."""
    if isinstance(fmatin, tuple):
        return tuple(system_mat2d(fi, c, fo) for fi, c, fo in zip(fmatin, cmat, fmatout))
    return bdotmd(bdotdm(inv(fmatin), cmat), fmatout)


def reflection_mat2d(smat):
    """(dtmm/tmm2d.py:161-182): the 4n x 4n reflection matrix."""
    if isinstance(smat, tuple):
        return tuple(reflection_mat2d(s) for s in smat)
    n = smat.shape[-4]
    sm = np.moveaxis(smat, -2, -3).reshape(smat.shape[:-4] + (4 * n, 4 * n))
    return reflection_mat(sm)


def reflect2d(fvecin, fmatin, rmat, fmatout, fvecout=None):
    """Transmitted mode vectors; fvecin receives incident + reflected
    (dtmm/tmm2d.py:184-232).  Returns (fvec_out, fvec_in)."""
    if isinstance(fvecin, tuple):
        res = [reflect2d(fvecin[i], fmatin[i], rmat[i], fmatout[i],
                         None if fvecout is None else fvecout[i]) for i in range(len(fvecin))]
        return tuple(r[0] for r in res), tuple(r[1] for r in res)
    avec = dotmv(inv(fmatin), fvecin)
    a = np.zeros(avec.shape, CDTYPE)
    a[..., 0::2] = avec[..., 0::2]
    if fvecout is not None:
        bvec = dotmv(inv(fmatout), fvecout)
        a[..., 1::2] = bvec[..., 1::2]
    else:
        bvec = np.zeros_like(avec)
    av = a.reshape(a.shape[:-2] + (a.shape[-2] * a.shape[-1],))
    out = dotmv(rmat, av).reshape(a.shape)
    avec[..., 1::2] = out[..., 1::2]
    bvec[..., ::2] = out[..., ::2]
    return dotmv(fmatout, bvec), dotmv(fmatin, avec)


def transfer2d(field_data_in, optical_data, betay=0., nin=1., nout=1., method="4x4",
               betamax=BETAMAX, field_out=None):
    """(dtmm/tmm2d.py:234-263): returns (field_out, wavelengths, pixelsize)
    and writes incident + reflected into the input field array."""
    f, w, p = field_data_in
    shape = f.shape[-1]
    d, epsv, epsa = optical_data
    ks = k0(w, p)
    fmode_out = None
    if field_out is not None:
        _, fmode_out = field2modes1(field_out, ks, betamax=betamax)
    mask, fmode_in = field2modes1(f, ks, betamax=betamax)
    fmatin = f_iso2d(shape, ks, nin, betay, betamax)
    fmatout = f_iso2d(shape, ks, nout, betay, betamax)
    cmat = stack_mat2d(ks, d, epsv, epsa, betay=betay, mask=mask, method=method)
    smat = system_mat2d(fmatin, cmat, fmatout)
    rmat = reflection_mat2d(smat)
    fmode_out, fmode_in = reflect2d(fmode_in, fmatin, rmat, fmatout, fmode_out)
    out = modes2field1(mask, fmode_out)
    f[...] = modes2field1(mask, fmode_in)
    return out, w, p


# ---------------------------------------------------------------------------
# window.py aperture and field.py illumination rays
# ---------------------------------------------------------------------------

def _radius(shape, scale=1.):
    """Radius of the half-pixel-centred grid, 1 at the edge (dtmm/window.py:17-23)."""
    ny, nx = [l / 2 for l in shape]
    ay, ax = [np.arange(-l / 2. + .5, l / 2. + .5) for l in shape]
    xx, yy = np.meshgrid(ax, ay, indexing="xy")
    return ((xx / (nx * scale)) ** 2 + (yy / (ny * scale)) ** 2) ** 0.5


def tukey(r, alpha=0.1, rmax=1.):
    """(dtmm/window.py:143-156)."""
    r = np.asarray(r, FDTYPE)
    out = np.ones(r.shape, FDTYPE)
    alpha = alpha * rmax
    if alpha > 0.:
        m = r > rmax - alpha
        out[m] = (0.5 * (1 - np.cos(np.pi * (r - rmax) / alpha)))[m]
    out[r >= rmax] = 0.
    return out


def aperture(shape, diameter=1., smooth=0.05):
    """(Annular) aperture window (dtmm/window.py:107-141)."""
    try:
        d1, d2 = diameter
    except TypeError:
        d1, d2 = 0, diameter
    a = tukey(_radius(shape, d2), smooth)
    if d1 == 0:
        return a
    return a * (1 - tukey(_radius(shape, d1), smooth))


def illumination_aperture(diameter=5., smooth=0.1):
    """(dtmm/field.py:356-369)."""
    try:
        inner, outer = diameter
        inner, outer = int(round(inner)), int(round(outer))
        if inner >= outer:
            raise ValueError("Invalid dimater, rounded inner diameter should be lower that "
                             "outer diameter.")
    except TypeError:
        inner, outer = 0, int(round(diameter))
    return aperture((outer, outer), (inner / outer, 1.), min(max(smooth, 0), 1))


def aperture2rays(diaphragm, betastep=0.1, norm=True):
    """beta, phi, intensity of the lit pixels of a diaphragm image
    (dtmm/field.py:342-354)."""
    shape = diaphragm.shape
    ay, ax = [np.arange(-l / 2. + .5, l / 2. + .5) for l in shape]
    xx, yy = np.meshgrid(ax, ay, indexing="xy")
    phi = np.arctan2(yy, xx)
    beta = np.sqrt(xx ** 2 + yy ** 2) * betastep
    mask = diaphragm > 0.
    intensity = diaphragm[mask]
    if norm:
        intensity = intensity / intensity.sum() * len(intensity)
    return (np.asarray(beta[mask], FDTYPE), np.asarray(phi[mask], FDTYPE),
            np.asarray(intensity, FDTYPE))


def illumination_rays(NA, diameter=5., smooth=0.1):
    """(dtmm/field.py:371-404)."""
    try:
        low, high = diameter
    except TypeError:
        low, high = 0, diameter
    a = illumination_aperture((low, high), smooth)
    return aperture2rays(a, betastep=2. * NA / (high - 1), norm=True)
