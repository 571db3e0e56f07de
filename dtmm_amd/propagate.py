"""Layer-level steps of the transfer on k-space fields: the reference's
dtmm/propagate_4x4.py:376 ``propagate_4x4_effective_1`` and
dtmm/propagate_2x2.py:193 ``propagate_2x2_effective_1``, with the same
arguments and return values, run on the GPU by ``dtmmx_propagate_layer``.

Fields are k-space fields (the numpy ``fft2`` of the real-space field):
(..., n_wavelengths, 4, H, W) for the 4x4 method and the E field
(..., n_wavelengths, 2, H, W) for the 2x2 method.  NumPy arrays come back as
NumPy arrays, device tensors stay on the device.  Built for diffraction 1
and an isotropic effective layer of the layer's own thickness (what
transfer_4x4 and transfer_2x2 pass, dtmm/transfer.py:926-976); the 2x2 step
runs mode +1 without reflection storage.  Other arguments raise
``NotImplementedError``.
"""
import ctypes

import numpy as np

from . import _lib
from .transfer import BETAMAX, Plan, Stack, _require_gpu


def _isotropic_eps(effective_layer, what):
    d, epsv, epsa = effective_layer
    e = np.asarray(epsv).reshape(-1)[:3]
    if e.size != 3 or not (e[0] == e[1] == e[2]):
        raise NotImplementedError("%s: only isotropic effective layers are supported" % what)
    return float(np.asarray(d)), complex(e[0])


def _beams(beta, phi, n_rays):
    """Per-beam (beta, phi) arrays and rays per beam from scalars or per-ray
    arrays (extra trailing singleton axes, as dtmm/transfer.py:43-65 adds,
    are dropped)."""
    b = np.asarray(beta, np.float64).reshape(-1) if np.ndim(beta) else np.asarray([beta], float)
    p = np.asarray(phi, np.float64).reshape(-1) if np.ndim(phi) else np.asarray([phi], float)
    if b.shape != p.shape or n_rays % len(b):
        raise ValueError("beta and phi must be scalars or one value per ray")
    return b, p, n_rays // len(b)


def _layer_step(method, field, wavenumbers, layer, d_eff, e_eff, beta, phi, nsteps, diffraction,
                betamax, out, reflection, e_in):
    torch = _require_gpu()
    if diffraction is True:
        diffraction = 1
    if diffraction != 1:
        raise NotImplementedError("the layer step is built for diffraction 1")
    nc = 4 if method == "4x4" else 2
    is_tensor = hasattr(field, "is_cuda")
    ks = np.asarray(wavenumbers, np.float64).reshape(-1)
    shape = tuple(field.shape[-2:])
    if field.shape[-3] != nc or field.shape[-4] != len(ks):
        raise ValueError("field must have shape (..., %d wavelengths, %d, H, W)" % (len(ks), nc))
    lead = tuple(field.shape[:-4])
    n_rays = int(np.prod(lead)) if lead else 1
    d, epsv, epsa = layer
    d = float(np.asarray(d))
    if d_eff != d:
        raise NotImplementedError("the effective layer must have the layer's thickness")
    b, p, rays_per_beam = _beams(beta, phi, n_rays)
    if is_tensor:
        precision = "c64" if field.dtype == torch.complex64 else "c128"
    else:
        precision = "c64" if np.asarray(field).dtype == np.complex64 else "c128"
    plan = Plan(shape, ks, b, p, n_rays, rays_per_beam, method, precision, betamax,
                diffraction=1, npass=1, reflection=reflection)
    epsv = np.broadcast_to(np.asarray(epsv), shape + (3,))
    epsa = np.broadcast_to(np.asarray(epsa, np.float64), shape + (3,))
    stack = Stack([(np.asarray([d * nsteps]), epsv[None], epsa[None])], shape, nstep=int(nsteps))
    stack.eps_eff_re = np.asarray([e_eff.real])
    stack.eps_eff_im = np.asarray([e_eff.imag])
    if is_tensor:
        fin = field.reshape((n_rays, len(ks), nc) + shape).to(device="cuda", dtype=plan.dtype)
        fin = fin.contiguous()
    else:
        fin = torch.from_numpy(np.ascontiguousarray(
            np.asarray(field).reshape((n_rays, len(ks), nc) + shape))).to("cuda", plan.dtype)
    fout = torch.empty_like(fin)
    ws = torch.empty(max(plan.workspace_bytes, 1), dtype=torch.uint8, device="cuda")
    status = torch.zeros(4, dtype=torch.int32, device="cuda")
    s = stack.c_struct()
    lib = _lib.load()
    _lib.check(lib.dtmmx_propagate_layer(
        plan._h, ctypes.byref(s), fin.data_ptr(), fout.data_ptr(), float(e_in.real),
        float(e_in.imag), ws.data_ptr(), plan.workspace_bytes, status.data_ptr(),
        torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    if int(status[0].item()) & 1:
        raise NotImplementedError("a biaxial voxel (eps0 != eps1) was met in a layer not flagged "
                                  "biaxial")
    res = fout.reshape(tuple(field.shape))
    if not is_tensor:
        res = res.cpu().numpy()
    if out is not None:
        out[...] = res
        return out
    return res


def propagate_4x4_effective_1(field, wavenumbers, layer, effective_layer, beta=0, phi=0,
                              nsteps=1, diffraction=True, betamax=BETAMAX, out=None,
                              _reuse=False):
    """One layer of the 4x4 method, reflection mode 1, on a k-space field
    (dtmm/propagate_4x4.py:376-444 with :80-101): for each of `nsteps`
    sub-steps D1, the inverse transform, E P E^-1 per pixel, the forward
    transform and D2."""
    d_eff, e_eff = _isotropic_eps(effective_layer, "propagate_4x4_effective_1")
    return _layer_step("4x4", field, wavenumbers, layer, d_eff, e_eff, beta, phi, nsteps,
                       diffraction, betamax, out, 1, e_eff)


def propagate_2x2_effective_1(field, wavenumbers, layer_in, layer_out, effective_layer_in,
                              effective_layer_out, beta=0, phi=0, nsteps=1, diffraction=True,
                              reflection=True, betamax=BETAMAX, mode=+1, tmpdata=None,
                              split_diffraction=False, refl=None, bulk=None, out=None):
    """One layer of the 2x2 method on a k-space E field
    (dtmm/propagate_2x2.py:193-211 with :19-76): with `reflection` the first
    sub-step starts with the Fresnel transmission from the effective medium
    of `effective_layer_in` into that of `effective_layer_out`.  Returns
    ``(field, refl)`` like the reference; `refl` and `bulk` storage and
    mode -1 (the backward passes of npass > 1) are not built."""
    if mode != +1 or refl is not None or bulk is not None:
        raise NotImplementedError("propagate_2x2_effective_1 is built for mode +1 without "
                                  "refl or bulk storage")
    d_eff, e_eff = _isotropic_eps(effective_layer_out, "propagate_2x2_effective_1")
    e_in = e_eff
    if reflection:
        _, e_in = _isotropic_eps(effective_layer_in, "propagate_2x2_effective_1")
    f = _layer_step("2x2", field, wavenumbers, layer_out, d_eff, e_eff, beta, phi, nsteps,
                    diffraction, betamax, out, 1 if reflection else 0, e_in)
    return f, refl
