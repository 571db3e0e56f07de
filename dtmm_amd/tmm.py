"""Non-iterative 1D transfer matrix method (dtmm/tmm.py:1112-1746, method
"4x4") on the GPU.

Every stack (every element of the broadcast shape of the inputs) is an
independent chain of homogeneous layers: one fp64 GPU thread builds its
characteristic matrix C = M_1 ... M_n, M = F P(-kd) F^-1, and solves the
boundary problem between the isotropic input and output media
(``dtmmx_tmm_transfer4x4``).  The arguments broadcast as in the reference:
``kd[i]``, ``epsv[i]``, ``epsa[i]`` and ``fvec_in`` share one element shape.
``beta`` and ``phi`` are scalars here (one beam per call).
"""
import ctypes

import numpy as np

from . import _lib

__all__ = ["stack_mat", "transfer4x4", "transfer1d"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("dtmm_amd needs a ROCm GPU (no CPU fallback)")
    return torch


def _layers(kd, epsv, epsa, extra_shapes=()):
    kd = [np.asarray(k, np.float64) for k in kd]
    L = len(kd)
    if L == 0:
        raise ValueError("kd must hold at least one layer")
    epsv = [np.asarray(epsv[i]) for i in range(L)]
    epsa = [np.asarray(epsa[i], np.float64) for i in range(L)]
    if any(e.shape[-1:] != (3,) for e in epsv + epsa):
        raise ValueError("epsv and epsa must end in an axis of length 3")
    shape = np.broadcast_shapes(*[k.shape for k in kd], *[e.shape[:-1] for e in epsv],
                                *[a.shape[:-1] for a in epsa], *extra_shapes)
    n = int(np.prod(shape)) if shape else 1
    kda = np.empty((L, n))
    er = np.empty((3, L, n))
    ei = np.zeros((3, L, n))
    an = np.empty((3, L, n))
    for i in range(L):
        kda[i] = np.broadcast_to(kd[i], shape).ravel()
        e = np.broadcast_to(epsv[i], shape + (3,)).reshape(n, 3)
        a = np.broadcast_to(epsa[i], shape + (3,)).reshape(n, 3)
        er[:, i] = e.real.T
        if np.iscomplexobj(e):
            ei[:, i] = e.imag.T
        an[:, i] = a.T
    return shape, n, L, kda, er, (ei if np.any(ei) else None), an


def _scalar(x, name):
    a = np.asarray(x, np.float64)
    if a.ndim != 0:
        raise NotImplementedError("%s must be a scalar (one beam per call)" % name)
    return float(a)


def _run(kd, epsv, epsa, beta, phi, nin, nout, fvec_in=None, fvec_out=None, want_cmat=False):
    torch = _torch()
    extra = [np.shape(fvec_in)[:-1]] if fvec_in is not None else []
    shape, n, L, kda, er, ei, an = _layers(kd, epsv, epsa, extra)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")  # noqa: E731
    kd_d, er_d, an_d = dev(kda), dev(er), dev(an)
    ei_d = dev(ei) if ei is not None else None
    d = _lib.TmmDesc()
    d.n_elements, d.n_layers = n, L
    d.kd, d.eps_re, d.angles = kd_d.data_ptr(), er_d.data_ptr(), an_d.data_ptr()
    d.eps_im = ei_d.data_ptr() if ei_d is not None else None
    d.beta, d.phi = _scalar(beta, "beta"), _scalar(phi, "phi")
    d.n_in, d.n_out = float(nin), float(nout)
    fi = fo = cm = None
    if fvec_in is not None:
        fi = dev(np.broadcast_to(np.asarray(fvec_in, np.complex128), shape + (4,)).reshape(n, 4))
        if fvec_out is not None:
            fo = dev(np.broadcast_to(np.asarray(fvec_out, np.complex128),
                                     shape + (4,)).reshape(n, 4))
        else:
            fo = torch.zeros((n, 4), dtype=torch.complex128, device="cuda")
    if want_cmat:
        cm = torch.empty((n, 4, 4), dtype=torch.complex128, device="cuda")
    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    _lib.check(lib.dtmmx_tmm_transfer4x4(ctypes.byref(d), ptr(fi), ptr(fo),
                                         1 if fvec_out is not None else 0, ptr(cm),
                                         ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    res = {}
    if fi is not None:
        res["fvec_in"] = fi.cpu().numpy().reshape(shape + (4,))
        res["fvec_out"] = fo.cpu().numpy().reshape(shape + (4,))
    if cm is not None:
        res["cmat"] = cm.cpu().numpy().reshape(shape + (4, 4))
    return res


def _check_method(method):
    if method not in ("4x4",):
        if method in ("4x4_1", "4x4_2", "4x4_r", "2x2", "2x2_1"):
            raise NotImplementedError("method %r is not built on the GPU (only '4x4')" % method)
        raise ValueError("Unknown method '{}'!".format(method))


def stack_mat(kd, epsv, epsa, beta=0., phi=0., cfact=0.01, method="4x4", out=None):
    """Characteristic matrix M_1 . M_2 ... M_n of a stack
    (dtmm/tmm.py:1215-1296, method "4x4")."""
    _check_method(method)
    c = _run(kd, epsv, epsa, beta, phi, 1., 1., want_cmat=True)["cmat"]
    if out is not None:
        out[...] = c
        return out
    return c


def transfer4x4(fvec_in, kd, epsv, epsa, beta=0., phi=0., nin=1., nout=1., method="4x4",
                reflect_in=False, reflect_out=False, fvec_out=None):
    """Transmitted field vector of a layered medium between isotropic media
    (dtmm/tmm.py:1565-1673).  As in the reference, a writeable ``fvec_in``
    receives the incident plus reflected field, ``fvec_out`` (if given) is
    the incoming field from the output side and receives the result;
    ``reflect_in`` / ``reflect_out`` are accepted and, as there, unused by
    the 4x4 methods."""
    _check_method(method)
    fvec_in_arr = np.asarray(fvec_in)
    res = _run(kd, epsv, epsa, beta, phi, nin, nout, fvec_in=fvec_in_arr, fvec_out=fvec_out)
    if isinstance(fvec_in, np.ndarray) and fvec_in.flags.writeable and \
            fvec_in.shape == res["fvec_in"].shape:
        fvec_in[...] = res["fvec_in"]
    if fvec_out is not None and isinstance(fvec_out, np.ndarray) and \
            fvec_out.shape == res["fvec_out"].shape:
        fvec_out[...] = res["fvec_out"]
        return fvec_out
    return res["fvec_out"]


def transfer1d(fvec_in, kd, epsv, epsa, beta=0., phi=0., nin=1., nout=1., method="4x4",
               reflect_in=False, reflect_out=False, fvec_out=None):
    """dtmm/tmm.py:1697-1748 ``transfer`` for the "4x4" method."""
    return transfer4x4(fvec_in, kd, epsv, epsa, beta, phi, nin, nout, method, reflect_in,
                       reflect_out, fvec_out)
