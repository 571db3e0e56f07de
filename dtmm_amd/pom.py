"""Polarizing-microscope step after the transfer: specter and colour.

``calculate_pom_specter`` is what ``FieldViewer.calculate_specter`` computes
(dtmm/field_viewer.py:1262-1330): each output field is (optionally) combined
from its x and y polarised parts with the polarizer's Jones vector,
refocused by the diffraction matrix of the immersion medium
(dtmm/diffract.py:107-119), passed through the analyser's 4x4 matrix and
reduced to the Poynting z component per wavelength (dtmm/field.py:800-845),
summed over the fields.  ``specter2color`` is dtmm/color.py:133-200.  Both run
in HIP kernels through ``dtmmx_pom_specter`` / ``dtmmx_specter2color``; the
2x2 and 4x4 Jones helpers below are small host-side matrices, as in the
reference (dtmm/jones.py, dtmm/jones4.py).

Inputs may be numpy arrays (results come back as numpy arrays) or device
tensors (results stay on the device).
"""
import ctypes

import numpy as np

from . import _lib
from .adjoint import _iso_fmat_normal
from .field import BETAMAX, CDTYPE

__all__ = ["polarizer", "as4x4", "rotated_matrix", "jonesmat4x4", "ray_jonesmat4x4",
           "calculate_pom_specter", "field2specter", "field2spectersum", "specter2color"]


def polarizer(jones):
    """2x2 polarizer j j^H of a normalised Jones vector (dtmm/jones.py:236-278)."""
    j = np.asarray(jones)
    c, s = j[..., 0], j[..., 1]
    out = np.empty(j.shape + (2,), CDTYPE)
    out[..., 0, 0] = c * np.conj(c)
    out[..., 0, 1] = c * np.conj(s)
    out[..., 1, 0] = s * np.conj(c)
    out[..., 1, 1] = s * np.conj(s)
    return out


def as4x4(jmat):
    """Jones matrix acting on (+x, -x, +y, -y) mode amplitudes: J on the
    forward, conj(J) on the backward modes (dtmm/jones.py:371-397)."""
    jmat = np.asarray(jmat)
    out = np.zeros(jmat.shape[:-2] + (4, 4), jmat.dtype)
    out[..., 0::2, 0::2] = jmat
    out[..., 1::2, 1::2] = np.conj(jmat)
    return out


def rotated_matrix(jmat, phi):
    """R^T J R with R the 2D rotation by phi (dtmm/jones.py:346-368,
    dtmm/rotation.py:64-76)."""
    c, s = np.cos(phi), np.sin(phi)
    r = np.asarray([[c, -s], [s, c]], CDTYPE)
    return np.swapaxes(r, -1, -2) @ np.asarray(jmat) @ r


def jonesmat4x4(jmat, n=1.):
    """F as4x4(J) F^-1 with F the normalised field matrix of an isotropic
    medium of index n at normal incidence (dtmm/jones4.py:75-101)."""
    f = _iso_fmat_normal(n)
    return f @ as4x4(np.asarray(jmat, CDTYPE)) @ np.linalg.inv(f)


def ray_jonesmat4x4(jmat, n=1.):
    """Laboratory-frame 4x4 matrix of a Jones matrix for a normal-incidence
    ray in an isotropic medium of index n (dtmm/jones4.py:205-246 with
    beta = phi = 0, the viewer's analyser and retarder)."""
    return jonesmat4x4(jmat, n)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("dtmm_amd needs a ROCm GPU (no CPU fallback)")
    return torch


def _to_device(a, torch, cdtype=None):
    if isinstance(a, torch.Tensor):
        if not a.is_cuda:
            raise ValueError("tensor inputs must be device tensors")
        t = a if cdtype is None else a.to(cdtype)
        return t.contiguous(), True
    a = np.asarray(a)
    t = torch.from_numpy(np.ascontiguousarray(a))
    if cdtype is not None:
        t = t.to(cdtype)
    return t.to("cuda"), False


def _precision(t, torch, precision):
    if precision is None:
        precision = "c64" if t.dtype == torch.complex64 else "c128"
    if precision not in ("c64", "c128"):
        raise ValueError("precision must be 'c64' or 'c128'")
    return precision


_MODES = {None: 0, "b": 0, 0: 0, "t": 1, +1: 1, "r": -1, -1: -1}


def calculate_pom_specter(field, wavenumbers, jvec=None, pmat=None, focus=0., n=1.,
                          mode=None, betamax=BETAMAX, diffraction=True, precision=None):
    """Specter (H, W, n_wavelengths) of the polarizing microscope image of
    output fields, summed over the fields.

    field: (..., n_wavelengths, 4, H, W), or with ``jvec`` (the polarizer's
    Jones vector, (c, s)) (..., 2, n_wavelengths, 4, H, W) holding the x and
    y polarised fields (dtmm/field_viewer.py:122-127).  ``pmat``: analyser
    4x4 matrix, shared or one per wavelength ((4, 4) or (n_wavelengths, 4,
    4)), applied in real space after the diffraction (field_viewer.py:140).
    ``focus`` (pixels), ``n`` (index of the immersion medium), ``mode`` and
    ``betamax`` set the diffraction matrix of dtmm/diffract.py:107-119
    (field_viewer.py:1017-1029); ``diffraction=False`` skips it.
    """
    torch = _torch()
    t, on_dev = _to_device(field, torch)
    if not t.is_complex():
        t = t.to(torch.complex128)
    precision = _precision(t, torch, precision)
    t = t.to(torch.complex64 if precision == "c64" else torch.complex128).contiguous()
    ks = np.abs(np.atleast_1d(np.asarray(wavenumbers, np.float64)))
    if t.dim() < 4 or t.shape[-3] != 4:
        raise ValueError("field must have shape (..., n_wavelengths, 4, H, W)")
    nlam, H, W = t.shape[-4], t.shape[-2], t.shape[-1]
    if len(ks) != nlam:
        raise ValueError("wavenumbers do not match the field's wavelengths")
    npol = 1
    lead = tuple(t.shape[:-4])
    if jvec is not None:
        if t.dim() < 5 or t.shape[-5] != 2:
            raise ValueError("Invalid field shape.")
        npol = 2
        lead = tuple(t.shape[:-5])
    nf = int(np.prod(lead)) if lead else 1
    if mode not in _MODES:
        raise ValueError("mode must be None, 'b', 't', 'r', +1 or -1")
    d = _lib.PomDesc()
    d.height, d.width, d.n_wavelengths = H, W, nlam
    d.k0 = _lib.dptr(ks)
    d.n_fields, d.n_pol = nf, npol
    if jvec is not None:
        c, s = np.asarray(jvec, np.complex128)
        d.jones[:] = (c.real, c.imag, s.real, s.imag)
    d.diffraction = 1 if diffraction else 0
    d.mode = _MODES[mode]
    d.d = float(focus)
    eps = complex(n) ** 2
    d.eps_re, d.eps_im = eps.real, eps.imag
    d.betamax = float(betamax)
    pm = None
    if pmat is not None:
        pm = np.asarray(pmat, np.complex128)
        if pm.shape == (4, 4):
            pm = pm[None]
        if pm.ndim != 3 or pm.shape[1:] != (4, 4) or pm.shape[0] not in (1, nlam):
            raise ValueError("pmat must be (4, 4) or (n_wavelengths, 4, 4)")
        pm = np.ascontiguousarray(pm).view(np.float64)
        d.pmat = _lib.dptr(pm)
        d.n_pmat = pm.shape[0]
    d.precision = _lib.DTMMX_C64 if precision == "c64" else _lib.DTMMX_C128
    lib = _lib.load()
    nws = lib.dtmmx_pom_workspace_size(ctypes.byref(d))
    ws = torch.empty(max(int(nws), 1), dtype=torch.uint8, device="cuda")
    rdt = torch.float32 if precision == "c64" else torch.float64
    spec = torch.empty((H, W, nlam), dtype=rdt, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.dtmmx_pom_specter(ctypes.byref(d), ctypes.c_void_p(t.data_ptr()),
                                     ctypes.c_void_p(spec.data_ptr()),
                                     ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(nws),
                                     ctypes.c_void_p(stream)))
    if on_dev:
        return spec
    torch.cuda.synchronize()
    return spec.cpu().numpy()


def field2specter(field):
    """Poynting z per wavelength, the wavelength axis moved last:
    (..., n_wavelengths, 4, H, W) -> (..., H, W, n_wavelengths)
    (dtmm/field.py:800-845)."""
    torch = _torch()
    t, on_dev = _to_device(field, torch)
    lead = tuple(t.shape[:-4])
    nlam, H, W = t.shape[-4], t.shape[-2], t.shape[-1]
    flat = t.reshape((-1,) + tuple(t.shape[-4:]))
    ks = np.ones(nlam)
    outs = [calculate_pom_specter(flat[i], ks, diffraction=False) for i in range(flat.shape[0])]
    out = torch.stack(outs).reshape(lead + (H, W, nlam))
    if on_dev:
        return out
    torch.cuda.synchronize()
    return out.cpu().numpy()


def field2spectersum(field):
    """field2specter summed over the first axis (dtmm/field.py:843-845)."""
    return calculate_pom_specter(field, np.ones(tuple(field.shape)[-4]), diffraction=False)


def specter2color(spec, cmf, norm=False, gamma=True, gray=False):
    """Specter (..., n_wavelengths) -> RGB (..., 3) (dtmm/color.py:133-200).

    cmf: (n_wavelengths, 3) colour matching function.  norm: False (none),
    True (divide by the largest XYZ value) or a number to divide by.  gamma:
    True (sRGB curve), False (none) or a number (v^(1/gamma) clipped to
    [0, 1]).  gray: Y in all three channels.
    """
    torch = _torch()
    t, on_dev = _to_device(spec, torch)
    if t.is_complex():
        raise ValueError("specter must be real")
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    t = t.contiguous()
    cmf = np.ascontiguousarray(np.asarray(cmf, np.float64))
    nlam = t.shape[-1]
    if cmf.shape[-1] != 3:
        raise ValueError("Grayscale cmf! Cannot convert to color.")
    if cmf.shape != (nlam, 3):
        raise ValueError("cmf must have shape (n_wavelengths, 3)")
    if norm is True:
        nm, nv = 2, 0.
    elif norm is False or norm is None or norm == 0:
        nm, nv = 0, 0.
    else:
        nm, nv = 1, float(norm)
    if gamma is True:
        gm, gv = 1, 0.
    elif gamma is False:
        gm, gv = 0, 0.
    else:
        gm, gv = 2, float(gamma)
    npix = int(np.prod(t.shape[:-1])) if t.dim() > 1 else 1
    out = torch.empty(tuple(t.shape[:-1]) + (3,), dtype=t.dtype, device="cuda")
    lib = _lib.load()
    nws = lib.dtmmx_specter2color_workspace_size(npix, nlam)
    ws = torch.empty(max(int(nws), 1), dtype=torch.uint8, device="cuda")
    prec = _lib.DTMMX_C64 if t.dtype == torch.float32 else _lib.DTMMX_C128
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.dtmmx_specter2color(ctypes.c_void_p(t.data_ptr()),
                                       ctypes.c_void_p(out.data_ptr()), npix, nlam,
                                       _lib.dptr(cmf), nm, nv, gm, gv, int(bool(gray)), prec,
                                       ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(nws),
                                       ctypes.c_void_p(stream)))
    if on_dev:
        return out
    torch.cuda.synchronize()
    return out.cpu().numpy()
