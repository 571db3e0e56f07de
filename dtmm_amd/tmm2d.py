"""The non-iterative 4x4 transfer matrix method for 2D data (x, z)
(dtmm/tmm2d.py:20-263) on the GPU.

The field lives in the space of the 1D FFT modes of the x axis that pass
betamax; a layer couples them through the FFT of its per-pixel
characteristic matrices.  ``dtmmx_tmm2d_stack`` builds the block matrices
and their product, ``dtmmx_tmm2d_system`` / ``_reflection`` / ``_reflect``
solve the boundary problem, ``dtmmx_tmm2d_modes`` converts fields to modes
and back; fp64 throughout.  Signatures, argument meanings and the tuple
results of multi-wavelength calls follow the reference.  Block matrices are
returned in the reference's (n, n, 4, 4) layout, the reflection matrix as
its (4n, 4n) array.
"""
import ctypes

import numpy as np

from . import _lib
from .field import _iso_fmat

__all__ = ["layer_mat2d", "stack_mat2d", "f_iso2d", "fi_iso2d", "system_mat2d",
           "reflection_mat2d", "reflect2d", "transfer2d", "field2modes1", "modes2field1",
           "eigenmask1", "eigenbetax1", "eigenindices1"]

BETAMAX = 0.8


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("dtmm_amd needs a ROCm GPU (no CPU fallback)")
    return torch


def _betax1(n, k0):
    k0 = np.abs(np.asarray(k0, np.float64)[..., None])
    return (2 * np.pi / k0) * np.fft.fftfreq(n)


def eigenmask1(n, k0, betamax=BETAMAX):
    """Modes with |beta_x| <= betamax (dtmm/wave.py:530-552)."""
    return np.abs(_betax1(n, k0)) <= betamax


def eigenbetax1(n, k0, betamax=BETAMAX):
    """beta_x of the valid modes (dtmm/wave.py:609-632)."""
    b = _betax1(n, k0)
    m = np.abs(b) <= betamax
    if m.ndim == 2:
        return tuple(b[i][m[i]] for i in range(m.shape[0]))
    return b[m]


def eigenindices1(n, k0, betamax=BETAMAX):
    """FFT bins of the valid modes (dtmm/wave.py:647-671)."""
    ii = np.arange(n)
    m = eigenmask1(n, k0, betamax)
    if m.ndim == 2:
        return tuple(ii[m[i]] for i in range(m.shape[0]))
    return ii[m]


def _dev(a, dtype=None):
    torch = _torch()
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
    return torch.from_numpy(a).to("cuda")


def _ws(nbytes):
    torch = _torch()
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _flat_to_blocks(t, n):
    """(4n, 4n) device matrix -> (n, n, 4, 4) NumPy array."""
    return t.reshape(n, 4, n, 4).permute(0, 2, 1, 3).contiguous().cpu().numpy()


def _blocks_to_flat(b):
    b = np.asarray(b, np.complex128)
    n = b.shape[-4]
    return np.ascontiguousarray(np.moveaxis(b, -2, -3).reshape(4 * n, 4 * n))


class _Layers:
    """Material of the stack on the device: [3][L][W] eps (re, im), angles."""

    def __init__(self, d, epsv, epsa):
        self.d = np.atleast_1d(np.asarray(d, np.float64))
        epsv = np.asarray(epsv)
        epsa = np.asarray(epsa, np.float64)
        if epsv.ndim == 2:
            epsv, epsa = epsv[None], epsa[None]
        L, W = epsv.shape[0], epsv.shape[-2]
        if len(self.d) != L:
            raise ValueError("one thickness per layer")
        self.L, self.W = L, W
        er = np.ascontiguousarray(np.moveaxis(np.broadcast_to(epsv.real, (L, W, 3)), -1, 0))
        ei = np.moveaxis(np.broadcast_to(epsv.imag, (L, W, 3)), -1, 0) \
            if np.iscomplexobj(epsv) and np.any(epsv.imag) else None
        an = np.ascontiguousarray(np.moveaxis(np.broadcast_to(epsa, (L, W, 3)), -1, 0))
        self.er, self.an = _dev(er, np.float64), _dev(an, np.float64)
        self.ei = _dev(ei, np.float64) if ei is not None else None


def _stack_one(layers, k0, idx, betay, method):
    """The (4n, 4n) device stack matrix of one wavelength."""
    torch = _torch()
    lib = _lib.load()
    n = len(idx)
    idx_d = _dev(np.asarray(idx, np.int32))
    d = _lib.Tmm2dDesc()
    d.width, d.n_layers, d.n_modes = layers.W, layers.L, n
    d.mode_index = idx_d.data_ptr()
    d.k0, d.betay = float(k0), float(betay)
    th = np.ascontiguousarray(layers.d, np.float64)
    d.thickness = th.ctypes.data
    d.eps_re, d.angles = layers.er.data_ptr(), layers.an.data_ptr()
    d.eps_im = layers.ei.data_ptr() if layers.ei is not None else None
    d.method = _lib.TMM_METHODS[method]
    cm = torch.empty((4 * n, 4 * n), dtype=torch.complex128, device="cuda")
    ws = _ws(lib.dtmmx_tmm2d_workspace_size(ctypes.byref(d)))
    _lib.check(lib.dtmmx_tmm2d_stack(ctypes.byref(d), ctypes.c_void_p(cm.data_ptr()),
                                     ctypes.c_void_p(ws.data_ptr()), ws.numel(), _stream()))
    return cm


def _check_method(method):
    if method not in ("4x4", "4x4_1"):
        raise ValueError("Unsupported method: '{}'".format(method))


def _masks(W, k0, mask):
    k0 = np.asarray(k0, np.float64)
    if mask is None:
        mask = eigenmask1(W, k0)
    mask = np.asarray(mask, bool)
    if k0.ndim == 0:
        return [np.arange(W)[mask]], [float(k0)], False
    return [np.arange(W)[mask[i]] for i in range(len(k0))], [float(k) for k in k0], True


def stack_mat2d(k, d, epsv, epsa, betay=0., method="4x4", mask=None):
    """M_1 . M_2 ... M_n of the layers' mode-coupling block matrices
    (dtmm/tmm2d.py:94-119), (n, n, 4, 4) per wavelength (a tuple for several)."""
    _check_method(method)
    layers = _Layers(d, epsv, epsa)
    idxs, ks, multi = _masks(layers.W, k, mask)
    out = tuple(_flat_to_blocks(_stack_one(layers, kk, ii, betay, method), len(ii))
                for kk, ii in zip(ks, idxs))
    return out if multi else out[0]


def layer_mat2d(k0, d, epsv, epsa, betay=0., method="4x4", mask=None):
    """Block matrix of one layer (dtmm/tmm2d.py:20-92)."""
    return stack_mat2d(k0, np.atleast_1d(d), np.asarray(epsv)[None], np.asarray(epsa)[None],
                       betay, method, mask)


def f_iso2d(shape, k0, n=1., betay=0., betamax=BETAMAX):
    """Field matrices of the isotropic medium for every mode
    (dtmm/tmm2d.py:121-134)."""
    k0 = np.asarray(k0)
    betax = eigenbetax1(shape, k0, betamax)

    def one(bx):
        return _iso_fmat(np.sqrt(bx ** 2 + betay ** 2), np.arctan2(betay, bx), n)
    if k0.ndim == 0:
        return one(betax)
    return tuple(one(b) for b in betax)


def fi_iso2d(shape, k0, n=1., betay=0., betamax=BETAMAX):
    """Inverses of f_iso2d (dtmm/tmm2d.py:136-142)."""
    f = f_iso2d(shape, k0, n, betay, betamax)
    if isinstance(f, tuple):
        return tuple(np.linalg.inv(x) for x in f)
    return np.linalg.inv(f)


def _system_dev(fin, cm, fout, n):
    torch = _torch()
    lib = _lib.load()
    sm = torch.empty((4 * n, 4 * n), dtype=torch.complex128, device="cuda")
    ws = _ws(256 + 16 * 16 * n + 256)
    _lib.check(lib.dtmmx_tmm2d_system(n, ctypes.c_void_p(fin.data_ptr()),
                                      ctypes.c_void_p(cm.data_ptr()),
                                      ctypes.c_void_p(fout.data_ptr()),
                                      ctypes.c_void_p(sm.data_ptr()),
                                      ctypes.c_void_p(ws.data_ptr()), ws.numel(), _stream()))
    return sm


def system_mat2d(fmatin, cmat, fmatout):
    """Fin^-1 . C . Fout block-wise (dtmm/tmm2d.py:144-159)."""
    if isinstance(fmatin, tuple):
        return tuple(system_mat2d(fi, c, fo) for fi, c, fo in zip(fmatin, cmat, fmatout))
    n = np.shape(fmatin)[-3]
    sm = _system_dev(_dev(fmatin, np.complex128), _dev(_blocks_to_flat(cmat)),
                     _dev(fmatout, np.complex128), n)
    _torch().cuda.synchronize()
    return _flat_to_blocks(sm, n)


def _reflection_dev(sm, n):
    torch = _torch()
    lib = _lib.load()
    rm = torch.empty((4 * n, 4 * n), dtype=torch.complex128, device="cuda")
    ws = _ws(lib.dtmmx_tmm2d_reflection_workspace_size(n))
    _lib.check(lib.dtmmx_tmm2d_reflection(n, ctypes.c_void_p(sm.data_ptr()),
                                          ctypes.c_void_p(rm.data_ptr()),
                                          ctypes.c_void_p(ws.data_ptr()), ws.numel(), _stream()))
    return rm


def reflection_mat2d(smat):
    """The 4n x 4n reflection matrix of a system block matrix
    (dtmm/tmm2d.py:161-182)."""
    if isinstance(smat, tuple):
        return tuple(reflection_mat2d(s) for s in smat)
    n = np.shape(smat)[-4]
    rm = _reflection_dev(_dev(_blocks_to_flat(smat)), n)
    _torch().cuda.synchronize()
    return rm.cpu().numpy()


def _reflect_dev(fvi, fin, rm, fout, fvo, has_out, n, nvec):
    lib = _lib.load()
    ws = _ws(256 + 2 * (16 * 16 * n + 256) + 2 * nvec * 4 * n * 16 + 256)
    _lib.check(lib.dtmmx_tmm2d_reflect(n, nvec, ctypes.c_void_p(fvi.data_ptr()),
                                       ctypes.c_void_p(fin.data_ptr()),
                                       ctypes.c_void_p(rm.data_ptr()),
                                       ctypes.c_void_p(fout.data_ptr()),
                                       ctypes.c_void_p(fvo.data_ptr()), 1 if has_out else 0,
                                       ctypes.c_void_p(ws.data_ptr()), ws.numel(), _stream()))


def reflect2d(fvecin, fmatin, rmat, fmatout, fvecout=None):
    """Transmitted mode vectors (..., n, 4); fvecin receives incident +
    reflected (dtmm/tmm2d.py:184-232)."""
    torch = _torch()
    if isinstance(fvecin, tuple):
        return tuple(reflect2d(fvecin[i], fmatin[i], rmat[i], fmatout[i],
                               None if fvecout is None else fvecout[i])
                     for i in range(len(fvecin)))
    fv = np.asarray(fvecin)
    n = fv.shape[-2]
    nvec = int(np.prod(fv.shape[:-2])) if fv.ndim > 2 else 1
    fvi = _dev(fv.reshape(nvec, n, 4), np.complex128)
    fvo = _dev(np.asarray(fvecout).reshape(nvec, n, 4), np.complex128) if fvecout is not None \
        else torch.zeros((nvec, n, 4), dtype=torch.complex128, device="cuda")
    _reflect_dev(fvi, _dev(fmatin, np.complex128), _dev(rmat, np.complex128),
                 _dev(fmatout, np.complex128), fvo, fvecout is not None, n, nvec)
    torch.cuda.synchronize()
    res_in = fvi.cpu().numpy().reshape(fv.shape)
    if isinstance(fvecin, np.ndarray) and fvecin.flags.writeable:
        fvecin[...] = res_in
    out = fvo.cpu().numpy().reshape(fv.shape)
    if fvecout is not None and isinstance(fvecout, np.ndarray) and fvecout.flags.writeable:
        fvecout[...] = out
        return fvecout
    return out


def _modes_dev(field_d, idx_d, W, n, nvec, inverse, modes_d=None):
    torch = _torch()
    lib = _lib.load()
    if not inverse:
        modes_d = torch.empty((nvec, n, 4), dtype=torch.complex128, device="cuda")
        out = modes_d
    else:
        field_d = torch.empty((nvec, 4, W), dtype=torch.complex128, device="cuda")
        out = field_d
    ws = _ws(256 + W * 16 + 256)
    _lib.check(lib.dtmmx_tmm2d_modes(W, n, ctypes.c_void_p(idx_d.data_ptr()), nvec,
                                     ctypes.c_void_p(field_d.data_ptr()),
                                     ctypes.c_void_p(modes_d.data_ptr()), 1 if inverse else 0,
                                     ctypes.c_void_p(ws.data_ptr()), ws.numel(), _stream()))
    return out


def field2modes1(field, k0, betamax=BETAMAX):
    """(mask, modes (..., n, 4)) of a (..., [lambda], 4, W) field
    (dtmm/field.py:216-250)."""
    if isinstance(field, tuple):
        out = tuple(field2modes1(field[i], k0[i], betamax) for i in range(len(field)))
        return tuple(o[0] for o in out), tuple(o[1] for o in out)
    torch = _torch()
    f = np.asarray(field, np.complex128)
    W = f.shape[-1]
    k0 = np.asarray(k0)
    mask = eigenmask1(W, k0, betamax)
    if k0.ndim == 0:
        idx = np.arange(W)[mask]
        lead = f.shape[:-2]
        nvec = int(np.prod(lead)) if lead else 1
        m = _modes_dev(_dev(f.reshape(nvec, 4, W)), _dev(idx.astype(np.int32)), W, len(idx),
                       nvec, False)
        torch.cuda.synchronize()
        return mask, m.cpu().numpy().reshape(lead + (len(idx), 4))
    out = []
    for i in range(len(k0)):
        fi = np.ascontiguousarray(f[..., i, :, :])
        out.append(field2modes1(fi, k0[i], betamax)[1])
    return mask, tuple(out)


def modes2field1(mask, modes):
    """Inverse of field2modes1 (dtmm/field.py:307-340)."""
    torch = _torch()
    if isinstance(mask, tuple):
        return tuple(modes2field1(mask[i], modes[i]) for i in range(len(mask)))
    mask = np.asarray(mask, bool)
    if mask.ndim == 1:
        W = mask.shape[-1]
        idx = np.arange(W)[mask]
        m = np.asarray(modes, np.complex128)
        lead = m.shape[:-2]
        nvec = int(np.prod(lead)) if lead else 1
        f = _modes_dev(None, _dev(idx.astype(np.int32)), W, len(idx), nvec, True,
                       _dev(m.reshape(nvec, len(idx), 4)))
        torch.cuda.synchronize()
        return f.cpu().numpy().reshape(lead + (4, W))
    parts = [modes2field1(mask[i], modes[i]) for i in range(len(mask))]
    return np.stack(parts, -3)


def transfer2d(field_data_in, optical_data, betay=0., nin=1., nout=1., method="4x4",
               betamax=BETAMAX, field_out=None):
    """Transmitted field data of a 2D stack (dtmm/tmm2d.py:234-263); the
    input field array receives incident + reflected."""
    _check_method(method)
    torch = _torch()
    f, w, p = field_data_in
    W = np.shape(f)[-1]
    d, epsv, epsa = optical_data
    ks = np.asarray(2 * np.pi / np.asarray(w) * p, np.float64)
    layers = _Layers(d, epsv, epsa)
    multi = ks.ndim > 0
    kl = list(ks) if multi else [float(ks)]
    fa = _dev(np.asarray(f, np.complex128))
    fo_all = _dev(np.asarray(field_out, np.complex128)) if field_out is not None else None
    outs, backs = [], []
    for i, kk in enumerate(kl):
        fi = fa[..., i, :, :] if multi else fa
        lead = tuple(fi.shape[:-2])
        nvec = int(np.prod(lead)) if lead else 1
        idx = eigenindices1(W, kk, betamax)
        n = len(idx)
        idx_d = _dev(idx.astype(np.int32))
        mi = _modes_dev(fi.reshape(nvec, 4, W).contiguous(), idx_d, W, n, nvec, False)
        if fo_all is not None:
            fo = fo_all[..., i, :, :] if multi else fo_all
            mo = _modes_dev(fo.reshape(nvec, 4, W).contiguous(), idx_d, W, n, nvec, False)
        else:
            mo = torch.zeros((nvec, n, 4), dtype=torch.complex128, device="cuda")
        fin = _dev(f_iso2d(W, kk, nin, betay, betamax), np.complex128)
        fout = _dev(f_iso2d(W, kk, nout, betay, betamax), np.complex128)
        cm = _stack_one(layers, kk, idx, betay, method)
        rm = _reflection_dev(_system_dev(fin, cm, fout, n), n)
        _reflect_dev(mi, fin, rm, fout, mo, fo_all is not None, n, nvec)
        outs.append(_modes_dev(None, idx_d, W, n, nvec, True, mo).reshape(lead + (4, W)))
        backs.append(_modes_dev(None, idx_d, W, n, nvec, True, mi).reshape(lead + (4, W)))
    out = (torch.stack(outs, -3) if multi else outs[0]).cpu().numpy()
    back = (torch.stack(backs, -3) if multi else backs[0]).cpu().numpy()
    if isinstance(f, np.ndarray) and f.flags.writeable:
        f[...] = back
    if field_out is not None and isinstance(field_out, np.ndarray):
        field_out[...] = out
        out = field_out
    return out, w, p
