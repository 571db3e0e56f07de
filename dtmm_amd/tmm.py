"""Non-iterative 1D transfer matrix method (dtmm/tmm.py:1112-1748) on the GPU.

Every stack (every element of the broadcast shape of the inputs) is an
independent chain of homogeneous layers: one fp64 GPU thread builds its
characteristic matrix and solves the boundary problem between the isotropic
input and output media (``dtmmx_tmm_transfer``).  The methods are the
reference's: "4x4" (C = M_1 ... M_n, M = F P(-kd) F^-1), "4x4_1" (backward
phases zeroed), "4x4_2" (backward modes damped by cfact), "4x4_r" (8x8
characteristic matrix), "2x2" (C = M_n ... M_1 of the forward E modes) and
"2x2_1" (layers entered through the interface transmission).  The arguments
broadcast as in the reference: ``kd[i]``, ``epsv[i]``, ``epsa[i]`` and
``fvec_in`` share one element shape.  ``beta`` and ``phi`` are scalars here
(one beam per call).  As in the reference, ``transfer4x4``, ``transfer2x2``
and ``transfer1d`` build the stack matrix with stack_mat's default beam
(beta 0, phi 0, dtmm/tmm.py:1655, :1687) and use beta, phi for the media.
"""
import ctypes

import numpy as np

from . import _lib

__all__ = ["layer_mat", "stack_mat", "transfer4x4", "transfer2x2", "transfer1d", "transfer"]


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


def _run(kd, epsv, epsa, beta, phi, nin, nout, fvec_in=None, fvec_out=None, want_cmat=False,
         method="4x4", cfact=0.01, stack_beam=None, reflect_in=False, reflect_out=False):
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
    sb, sp = (d.beta, d.phi) if stack_beam is None else stack_beam
    d.stack_beta, d.stack_phi = _scalar(sb, "beta"), _scalar(sp, "phi")
    d.n_in, d.n_out = float(nin), float(nout)
    d.method = _lib.TMM_METHODS[method]
    d.cfact = float(cfact)
    d.reflect_in, d.reflect_out = int(bool(reflect_in)), int(bool(reflect_out))
    two = method.startswith("2x2")
    fi = fo = cm = None
    if fvec_in is not None:
        fi = dev(np.broadcast_to(np.asarray(fvec_in, np.complex128), shape + (4,)).reshape(n, 4))
        if fvec_out is not None and not two:
            fo = dev(np.broadcast_to(np.asarray(fvec_out, np.complex128),
                                     shape + (4,)).reshape(n, 4))
        else:
            fo = torch.zeros((n, 4), dtype=torch.complex128, device="cuda")
    m = 2 if two else (8 if method == "4x4_r" else 4)
    if want_cmat:
        cm = torch.empty((n, m, m), dtype=torch.complex128, device="cuda")
    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    _lib.check(lib.dtmmx_tmm_transfer(ctypes.byref(d), ptr(fi), ptr(fo),
                                      1 if (fvec_out is not None and not two) else 0, ptr(cm),
                                      ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    res = {}
    if fi is not None:
        res["fvec_in"] = fi.cpu().numpy().reshape(shape + (4,))
        res["fvec_out"] = fo.cpu().numpy().reshape(shape + (4,))
    if cm is not None:
        res["cmat"] = cm.cpu().numpy().reshape(shape + (m, m))
    return res


_METHODS = ("4x4", "4x4_1", "4x4_2", "4x4_r", "2x2", "2x2_1")


def _check_method(method, allowed=_METHODS, message="Unknown method!"):
    if method not in allowed:
        raise ValueError(message)


def layer_mat(kd, epsv, epsa, beta=0., phi=0., cfact=0.1, method="4x4", fmatin=None,
              retfmat=False, out=None):
    """Characteristic matrix of one layer (dtmm/tmm.py:1112-1213).  The
    "2x2_1" entry from another medium (fmatin) and the returned field matrix
    (retfmat) are not built: they raise NotImplementedError."""
    _check_method(method)
    if fmatin is not None or retfmat:
        raise NotImplementedError("layer_mat with fmatin or retfmat is not built on the GPU")
    kd = np.asarray(kd, np.float64)
    c = _run([kd], [np.asarray(epsv)], [np.asarray(epsa)], beta, phi, 1., 1., want_cmat=True,
             method=method, cfact=cfact)["cmat"]
    if out is not None:
        out[...] = c
        return out
    return c


def stack_mat(kd, epsv, epsa, beta=0., phi=0., cfact=0.01, method="4x4", out=None):
    """Characteristic matrix of a stack (dtmm/tmm.py:1215-1296): M_1 . M_2
    ... M_n for the 4x4 methods (8x8 for "4x4_r"), M_n ... M_1 for the 2x2
    ones."""
    _check_method(method)
    c = _run(kd, epsv, epsa, beta, phi, 1., 1., want_cmat=True, method=method,
             cfact=cfact)["cmat"]
    if out is not None:
        out[...] = c
        return out
    return c


def transfer4x4(fvec_in, kd, epsv, epsa, beta=0., phi=0., nin=1., nout=1., method="4x4",
                reflect_in=False, reflect_out=False, fvec_out=None):
    """Transmitted field vector of a layered medium between isotropic media
    (dtmm/tmm.py:1565-1673), methods "4x4", "4x4_1", "4x4_2" and "4x4_r".
    As in the reference, a writeable ``fvec_in`` receives the incident plus
    reflected field, ``fvec_out`` (if given) is the incoming field from the
    output side and receives the result; ``reflect_in`` / ``reflect_out``
    are accepted and, as there, unused."""
    _check_method(method, ("4x4", "4x4_1", "4x4_r", "4x4_2"),
                  "Unknown method '{}'!".format(method))
    fvec_in_arr = np.asarray(fvec_in)
    res = _run(kd, epsv, epsa, beta, phi, nin, nout, fvec_in=fvec_in_arr, fvec_out=fvec_out,
               method=method, stack_beam=(0., 0.))
    if isinstance(fvec_in, np.ndarray) and fvec_in.flags.writeable and \
            fvec_in.shape == res["fvec_in"].shape:
        fvec_in[...] = res["fvec_in"]
    if fvec_out is not None and isinstance(fvec_out, np.ndarray) and \
            fvec_out.shape == res["fvec_out"].shape:
        fvec_out[...] = res["fvec_out"]
        return fvec_out
    return res["fvec_out"]


def transfer(fvec_in, kd, epsv, epsa, beta=0., phi=0., nin=1., nout=1., method="2x2",
             reflect_in=False, reflect_out=False, fvec_out=None):
    """dtmm/tmm.py:1697-1748 ``transfer`` (``transfer1d``): the 2x2 methods
    take the forward E part of ``fvec_in`` in the input medium (fvec2E),
    pass it through the stack (with the Fresnel transmission into the first
    and out of the last layer when ``reflect_in`` / ``reflect_out``) and
    return the output medium's field vector (E2fvec, written to
    ``fvec_out`` when given); the 4x4 methods are transfer4x4."""
    if method.startswith("2x2"):
        _check_method(method, ("2x2", "2x2_1"))
        res = _run(kd, epsv, epsa, beta, phi, nin, nout, fvec_in=np.asarray(fvec_in),
                   method=method, stack_beam=(0., 0.), reflect_in=reflect_in,
                   reflect_out=reflect_out)
        out = res["fvec_out"]
        if fvec_out is not None:
            fvec_out[...] = out
            return fvec_out
        return out
    return transfer4x4(fvec_in, kd, epsv, epsa, beta, phi, nin, nout, method, reflect_in,
                       reflect_out, fvec_out)


transfer1d = transfer


def transfer2x2(evec, kd, epsv, epsa, beta=0., phi=0., nin=None, nout=None, method="2x2",
                out=None):
    """dtmm/tmm.py:1675-1695: the E vector through the stack, with the
    Fresnel transmission from an isotropic input medium ``nin`` into the
    first layer and from the last layer into ``nout`` when they are given."""
    _check_method(method, ("2x2", "2x2_1"))
    evec = np.asarray(evec, np.complex128)
    # an E vector is the forward part of the field vector (E, H) of the
    # input medium; E2fvec of the result in the output medium carries it back
    n_in = 1. if nin is None else nin
    n_out = 1. if nout is None else nout
    fv = np.zeros(evec.shape[:-1] + (4,), np.complex128)
    fv[..., ::2] = evec
    fv[..., 1::2] = _e2h(evec, beta, phi, n_in)
    res = _run(kd, epsv, epsa, beta, phi, n_in, n_out, fvec_in=fv, method=method,
               stack_beam=(0., 0.), reflect_in=nin is not None, reflect_out=nout is not None)
    e = res["fvec_out"][..., ::2]
    if out is not None:
        out[...] = e
        return out
    return e


def _e2h(evec, beta, phi, n):
    """H of the forward plane waves of an isotropic medium for the E vector
    (the field vector whose forward projection is exactly evec)."""
    beta = _scalar(beta, "beta")
    phi = _scalar(phi, "phi")
    c, s = np.cos(phi), np.sin(phi)
    a = np.sqrt(n * n - beta * beta + 0j)
    ex, ey = evec[..., 0], evec[..., 1]
    gp, gs = ex * c + ey * s, ey * c - ex * s  # p and s parts in the frame of the beam
    hp, hs = (n * n / a) * gp, -a * gs
    return np.stack([hp * c + hs * s, hs * c - hp * s], -1)
