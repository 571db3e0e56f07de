"""Field-data model: illumination construction and Poynting intensity.

Host-side NumPy input builders with the reference's names and argument
meanings (dtmm/field.py, dtmm/wave.py, dtmm/window.py).
"""
import os

import numpy as np
import scipy.fft

from .data import refind2eps

FDTYPE = np.float64
CDTYPE = np.complex128
BETAMAX = 0.8
N_COVER = 1.5


def k0(wavelength, pixelsize=1.):
    """Vacuum wave number in pixel units (dtmm/wave.py:280-296)."""
    return np.asarray(2 * np.pi / np.asarray(wavelength) * pixelsize, dtype=FDTYPE)


def betaphi(shape, k0):
    """beta, phi of the FFT eigen-waves (dtmm/wave.py:54-81)."""
    k0 = np.abs(np.asarray(k0, FDTYPE)[..., None, None])
    ay, ax = [np.asarray(np.fft.fftfreq(x), dtype=FDTYPE) for x in shape[-2:]]
    xx, yy = np.meshgrid(ax, ay, copy=False, indexing="xy")
    beta = (2 * np.pi / k0) * np.sqrt(xx ** 2 + yy ** 2)
    phi = np.arctan2(yy, xx)
    return beta, phi


def blackman(shape):
    """2D radial Blackman window (dtmm/window.py:33-54)."""
    ny, nx = [l / 2 for l in shape]
    ay, ax = [np.arange(-l / 2. + .5, l / 2. + .5) for l in shape]
    xx, yy = np.meshgrid(ax, ay, indexing="xy")
    r = ((xx / nx) ** 2 + (yy / ny) ** 2) ** 0.5
    out = 0.42 + 0.5 * np.cos(np.pi * r) + 0.08 * np.cos(2 * np.pi * r)
    out[r >= 1.] = 0.
    return np.asarray(out, FDTYPE)


def planewave(shape, k0, beta, phi):
    """Plane wave normalised to 1 at pixel (0, 0) (dtmm/wave.py:376-409)."""
    k0 = np.asarray(k0)[..., None, None]
    beta = np.asarray(beta)[..., None, None]
    phi = np.asarray(phi)[..., None, None]
    ay, ax = [np.arange(-l // 2 + 1., l // 2 + 1.) for l in shape[-2:]]
    xx, yy = np.meshgrid(ax, ay, indexing="xy", copy=False)
    kx = np.asarray(k0 * beta * np.cos(phi), dtype=FDTYPE)
    ky = np.asarray(k0 * beta * np.sin(phi), dtype=FDTYPE)
    out = np.exp(1j * (kx * xx + ky * yy))
    return out / out[..., 0, 0][..., None, None]


def wave2eigenwave(wave):
    """Snap each 2D wave to its strongest FFT eigen-wave with the same power
    (dtmm/wave.py:432-476, eigenwave :140)."""
    wave = np.asarray(wave)
    shape = wave.shape
    flat = wave.reshape((-1,) + shape[-2:])
    out = np.empty(flat.shape, CDTYPE)
    s = np.abs(np.fft.fft2(flat)) ** 2
    for i, a in enumerate(s):
        avg = a.sum()
        ii, jj = np.unravel_index(np.argmax(a), a.shape)
        f = np.zeros(a.shape, CDTYPE)
        f[ii, jj] = avg ** 0.5
        out[i] = np.fft.ifft2(f)
    return out.reshape(shape)


def _iso_fmat(beta, phi, n):
    """Field matrix of an isotropic medium, rotated by phi
    (dtmm/tmm.py:123-154 and dtmm/linalg.py:674-713)."""
    beta = np.asarray(beta, FDTYPE)
    phi = np.asarray(phi, FDTYPE)
    eps = complex(refind2eps(n))
    a = np.sqrt(eps - beta ** 2 + 0j)
    F = np.zeros(np.broadcast_shapes(beta.shape, phi.shape) + (4, 4), CDTYPE)
    with np.errstate(divide="ignore", invalid="ignore"):
        gp = eps / a
        F[..., 1, 0] = 0.5 * gp
        F[..., 1, 1] = -0.5 * gp
    F[..., 0, 0] = 0.5
    F[..., 0, 1] = 0.5
    F[..., 2, 2] = 0.5
    F[..., 2, 3] = 0.5
    F[..., 3, 2] = -0.5 * a
    F[..., 3, 3] = 0.5 * a
    c, s = np.cos(phi)[..., None], np.sin(phi)[..., None]
    r0 = F[..., 0, :] * c - F[..., 2, :] * s
    r1 = F[..., 1, :] * c + F[..., 3, :] * s
    r2 = F[..., 0, :] * s + F[..., 2, :] * c
    r3 = -F[..., 1, :] * s + F[..., 3, :] * c
    return np.stack([r0, r1, r2, r3], -2)


def _normalize_f(fmat):
    """Unit |Poynting| columns (dtmm/tmm.py:699-724)."""
    p = (fmat[..., 0, :] * np.conj(fmat[..., 1, :])).real - \
        (fmat[..., 2, :] * np.conj(fmat[..., 3, :])).real
    n = np.abs(p) ** 0.5
    s = np.where(n == 0., 0., 1. / np.where(n == 0., 1., n))
    return fmat * s[..., None, :]


def _jonesvec(pol, phi):
    """Normalised Jones vector in the frame rotated by phi (dtmm/jones.py:55-98)."""
    pol = np.asarray(pol, CDTYPE)
    norm = (pol[..., 0] * pol[..., 0].conj() + pol[..., 1] * pol[..., 1].conj()) ** 0.5
    pol = pol / norm[..., None]
    c, s = np.cos(-phi), np.sin(-phi)
    return np.stack([c * pol[..., 0] - s * pol[..., 1], s * pol[..., 0] + c * pol[..., 1]], -1)


def _fvec(fmat, jones, amplitude, mode=+1):
    """Field vectors of the given polarisation and amplitude
    (dtmm/tmm.py:1750-1843, :1877-1904)."""
    c, s = jones[..., 0], jones[..., 1]
    shape = np.broadcast(c, amplitude).shape + (4,)
    a = np.zeros(shape, CDTYPE)
    i0, i1 = (0, 2) if mode == +1 else (1, 3)
    a[..., i0] = c
    a[..., i1] = s
    a = a * amplitude[..., None]
    return np.matmul(_normalize_f(fmat), a[..., None])[..., 0]


def _projection_fft(field, ks, n, mode, betamax, d=0.):
    """Apply F diag(exp(i k0 d alpha) on the modes of direction `mode`, 0 on
    the others) F^-1 per FFT mode in a medium of index n, zero for
    beta >= betamax: the field diffraction matrix of dtmm/diffract.py:107-119
    (d = 0: the projection onto one direction), applied as in :254-299.
    alpha = +sqrt(n^2 - beta^2) for the forward modes (columns 0, 2) and its
    negative for the backward ones (columns 1, 3)."""
    shape = field.shape[-2:]
    beta, phi = betaphi(shape, ks)
    F = _iso_fmat(beta, phi, n)
    mask = beta >= betamax
    ok = ~mask
    Fi = np.zeros_like(F)
    Fi[ok] = np.linalg.inv(F[ok])
    F[mask] = 0.
    a = np.sqrt(n ** 2 - beta ** 2 + 0j)
    kd = np.asarray(ks, FDTYPE)[:, None, None] * d
    ph = np.zeros(beta.shape + (4,), np.complex128)
    if mode == +1:
        ph[..., 0] = ph[..., 2] = np.exp(1j * kd * a)
    else:
        ph[..., 1] = ph[..., 3] = np.exp(-1j * kd * a)
    P = np.matmul(F * ph[..., None, :], Fi)
    workers = len(os.sched_getaffinity(0))
    ff = scipy.fft.fft2(field, workers=workers)
    v = np.moveaxis(ff, -3, -1)[..., None]
    v = np.matmul(P, v)[..., 0]
    return scipy.fft.ifft2(np.moveaxis(v, -1, -3), workers=workers)


def field2intensity(field):
    """Poynting z per pixel of a (..., 4, H, W) field (dtmm/field.py:771-799).
    Accepts numpy arrays or torch tensors."""
    if hasattr(field, "is_cuda"):
        return (field[..., 0, :, :] * field[..., 1, :, :].conj()).real - \
            (field[..., 2, :, :] * field[..., 3, :, :].conj()).real
    f = np.asarray(field)
    return (f[..., 0, :, :] * np.conj(f[..., 1, :, :])).real - \
        (f[..., 2, :, :] * np.conj(f[..., 3, :, :])).real


def illumination_data(shape, wavelengths, pixelsize=1., beta=0., phi=0., intensity=1.,
                      n=None, focus=0., window=None, backdir=False, jones=None,
                      diffraction=True, eigenmodes=None, betamax=BETAMAX, device=None,
                      precision="c128"):
    """Forward (or backward) propagating input field data
    (dtmm/field.py:586-713).  Returns (field, wavelengths, pixelsize) with
    field of shape (..., [2 pol], n_wavelengths, 4, H, W).  The plane waves
    are snapped to FFT eigen-waves exactly when ``window`` is None, as in the
    reference (field.py:686-687); ``eigenmodes`` is accepted for signature
    parity and, as there, does not change the field.

    ``device="cuda"`` builds the field on the GPU (``dtmmx_illumination``) and
    returns it as a device tensor of ``precision`` ("c64" or "c128"), which
    ``transfer_field`` takes without a host copy; ``window`` may then be a
    device tensor.  Beams are scalars or 1D sequences there."""
    if device is not None:
        return _illumination_device(shape, wavelengths, pixelsize, beta, phi, intensity, n,
                                    focus, window, backdir, jones, diffraction, betamax, device,
                                    precision)
    n = N_COVER if n is None else n
    wavelengths = np.asarray(wavelengths)
    wavenumbers = 2 * np.pi / wavelengths * pixelsize
    if wavenumbers.ndim != 1:
        raise ValueError("Wavelengths should be 1D array")
    if jones is None:
        intensity = intensity / 2.
    mode = -1 if backdir else +1
    _beta = np.asarray(beta, FDTYPE)
    _phi = np.asarray(phi, FDTYPE)
    _intensity = np.asarray(intensity, FDTYPE)
    nrays = len(_beta) if _beta.ndim > 0 else 1
    b = _beta[:, None] if _beta.ndim == 1 else _beta
    p = _phi[:, None] if _phi.ndim == 1 else _phi
    b = b if b.ndim >= 1 else b[..., None]
    p = p if p.ndim >= 1 else p[..., None]
    waves = planewave(shape, wavenumbers, b, p)
    if window is None:
        waves = wave2eigenwave(waves)
    else:
        waves = waves * window
    b = _beta[:, None, None, None] if _beta.ndim == 1 else _beta[..., None, None]
    p = _phi[:, None, None, None] if _phi.ndim == 1 else _phi[..., None, None]
    inten = _intensity[..., None, None, None, None]
    fmat = _iso_fmat(b, p, n)
    if jones is None:
        f1 = np.moveaxis(_fvec(fmat, _jonesvec((1, 0), p), waves, mode), -1, -3)
        f2 = np.moveaxis(_fvec(fmat, _jonesvec((0, 1), p), waves, mode), -1, -3)
        shp = list(f1.shape)
        shp.insert(-4, 2)
        field = np.empty(shp, CDTYPE)
        field[..., 0, :, :, :, :] = f1
        field[..., 1, :, :, :, :] = f2
    else:
        field = np.moveaxis(_fvec(fmat, _jonesvec(jones, p), waves, mode), -1, -3).copy()
    i1 = field2intensity(field)
    norm = np.ones_like(i1)
    norm[:, ...] = (inten / nrays) ** 0.5
    if diffraction:
        # the beams focus `focus` pixels into the sample: diffracted back by
        # -focus in the cover medium (dtmm/field.py:703-711)
        field = _projection_fft(field, wavenumbers, n, mode, betamax, d=-focus)
        i2 = field2intensity(field)
        ratio = (i1.sum((-2, -1)) / i2.sum((-2, -1))) ** 0.5
        norm = norm * ratio[..., None, None]
    field = np.ascontiguousarray(norm[..., None, :, :] * field)
    return field, wavelengths, pixelsize


def _illumination_device(shape, wavelengths, pixelsize, beta, phi, intensity, n, focus, window,
                         backdir, jones, diffraction, betamax, device, precision):
    """The GPU build of illumination_data (dtmmx_illumination)."""
    import ctypes
    import torch
    from . import _lib
    from .transfer import Plan
    if not torch.cuda.is_available():
        raise RuntimeError("dtmm_amd needs a ROCm GPU (no CPU fallback)")
    if str(device) not in ("cuda", "cuda:0") and not str(device).startswith("cuda"):
        raise ValueError("device must be a CUDA (HIP) device")
    n = N_COVER if n is None else n
    wavelengths = np.asarray(wavelengths)
    ks = k0(wavelengths, pixelsize)
    if ks.ndim != 1:
        raise ValueError("Wavelengths should be 1D array")
    _beta = np.asarray(beta, FDTYPE)
    _phi = np.asarray(phi, FDTYPE)
    if _beta.ndim > 1 or _phi.ndim > 1:
        raise NotImplementedError("the device build takes scalar or 1D beams")
    b, p = np.broadcast_arrays(np.atleast_1d(_beta), np.atleast_1d(_phi))
    nbeams = len(b)
    if np.ndim(intensity) != 0:
        # the reference broadcasts intensity[..., None, None, None, None]
        # into the (beams, wavelengths, H, W) norm (field.py:696, :703),
        # which only a scalar survives
        raise ValueError("intensity must be a scalar")
    inten = np.full((nbeams,), float(intensity))
    if jones is None:
        inten = inten / 2.
    rpb = 1 if jones is not None else 2
    H, W = shape
    # the transforms of the diffraction-limited projection come from a plan
    # with diffraction 1 (the 4x4 method's reflection mode 2 without it)
    plan = Plan(shape, ks, b, p, nbeams * rpb, rpb, "4x4", precision, betamax,
                diffraction=1 if diffraction else 0, reflection=1 if diffraction else 2)
    field = torch.empty((nbeams * rpb, len(ks), 4, H, W), dtype=plan.dtype, device="cuda")
    win = None
    if window is not None:
        win = torch.as_tensor(window, dtype=torch.float64, device="cuda").contiguous()
        if tuple(win.shape) != (H, W):
            raise ValueError("window must have the shape of the field")
    lib = _lib.load()
    ws = torch.empty(int(lib.dtmmx_illumination_workspace_size(plan._h)), dtype=torch.uint8,
                     device="cuda")
    d = _lib.IllumDesc()
    d.n, d.focus = float(n), float(focus)
    d.mode = -1 if backdir else 1
    d.diffraction = 1 if diffraction else 0
    d.polarized = 1 if jones is not None else 0
    if jones is not None:
        j = np.asarray(jones, np.complex128)
        d.jones_re[0], d.jones_re[1] = j.real
        d.jones_im[0], d.jones_im[1] = j.imag
    d.intensity = inten.ctypes.data
    d.window = win.data_ptr() if win is not None else None
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.dtmmx_illumination(plan._h, ctypes.byref(d), ctypes.c_void_p(field.data_ptr()),
                                      ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                      ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    shp = ((nbeams,) if _beta.ndim == 1 or _phi.ndim == 1 else ()) + \
        ((2,) if jones is None else ()) + (len(ks), 4, H, W)
    return field.reshape(shp), wavelengths, pixelsize


def _radius(shape, scale=1.):
    """Radius of the half-pixel-centred grid, 1 at the edge (dtmm/window.py:17-23)."""
    ny, nx = [l / 2 for l in shape]
    ay, ax = [np.arange(-l / 2. + .5, l / 2. + .5) for l in shape]
    xx, yy = np.meshgrid(ax, ay, indexing="xy")
    return np.sqrt((xx / (nx * scale)) ** 2 + (yy / (ny * scale)) ** 2)


def _tukey(r, alpha, rmax=1.):
    """Flat top with a raised-cosine edge of relative width alpha
    (dtmm/window.py:143-156)."""
    out = np.ones(np.shape(r), FDTYPE)
    w = alpha * rmax
    if w > 0.:
        edge = r > rmax - w
        out[edge] = 0.5 * (1 - np.cos(np.pi * (r[edge] - rmax) / w))
    out[r >= rmax] = 0.
    return out


def aperture(shape, diameter=1., smooth=0.05):
    """(Annular) aperture window (dtmm/window.py:107-141)."""
    inner, outer = diameter if np.ndim(diameter) else (0, diameter)
    a = _tukey(_radius(shape, outer), smooth)
    if inner == 0:
        return a
    return a * (1 - _tukey(_radius(shape, inner), smooth))


def illumination_aperture(diameter=5., smooth=0.1):
    """Diaphragm image of an (annular) field aperture (dtmm/field.py:356-369)."""
    if np.ndim(diameter):
        inner, outer = (int(round(x)) for x in diameter)
        if inner >= outer:
            raise ValueError("Invalid dimater, rounded inner diameter should be lower that "
                             "outer diameter.")
    else:
        inner, outer = 0, int(round(diameter))
    return aperture((outer, outer), (inner / outer, 1.), min(max(smooth, 0), 1))


def aperture2rays(diaphragm, betastep=0.1, norm=True):
    """beta, phi and intensity of the lit pixels of a diaphragm image
    (dtmm/field.py:342-354)."""
    diaphragm = np.asarray(diaphragm)
    ay, ax = [np.arange(-l / 2. + .5, l / 2. + .5) for l in diaphragm.shape]
    xx, yy = np.meshgrid(ax, ay, indexing="xy")
    lit = diaphragm > 0.
    intensity = diaphragm[lit]
    if norm:
        intensity = intensity * (len(intensity) / intensity.sum())
    return (np.asarray(np.hypot(xx, yy)[lit] * betastep, FDTYPE),
            np.asarray(np.arctan2(yy, xx)[lit], FDTYPE), np.asarray(intensity, FDTYPE))


def illumination_rays(NA, diameter=5., smooth=0.1):
    """beta, phi, intensity of rays filling a numerical aperture NA
    (dtmm/field.py:371-404); pass them to illumination_data."""
    low, high = diameter if np.ndim(diameter) else (0, diameter)
    return aperture2rays(illumination_aperture((low, high), smooth), betastep=2. * NA / (high - 1))
