"""Drop-in ``transfer_field`` for the diffractive transfer-matrix path.

Mirrors dtmm/transfer.py:322 ``transfer_field`` (same arguments, defaults,
return value and errors) for the modes built on the GPU:

* method "4x4" (dtmm/transfer.py:618-884): reflection mode 1 with
  diffraction >= 1 and any npass (dtmm/propagate_4x4.py:376-444, :80-101);
  reflection mode 2 with diffraction 0, 1 or the tent windows of
  diffraction > 1 and any npass (:156-204); npass > 1 with every norm
  (DTMM_NORM_FFT, 0, and DTMM_NORM_REF alone or with DTMM_NORM_FFT at
  diffraction 0 or 1); diffraction < 0 or inf, the exact per-mode
  propagation (npass 1, :515-564);
* method "2x2" (and the configured default method '"2x2"',
  dtmm/transfer.py:1009-1183): reflection mode 0 (diffraction 0 or 1) and
  1 (diffraction 1) with npass 1 (dtmm/propagate_2x2.py:193-211, :19-87);
  reflection modes 0 and 1 with the tent windows of diffraction > 1 and
  npass 1 (:212-257);
  reflection mode 2 with diffraction 0 or 1 (:91-190); npass > 1 with
  reflection mode 1 (diffraction 1) or 2 (diffraction 0 or 1);
* ret_bulk with diffraction 0 or 1 (4x4: npass 1), per-wavelength
  transfers (split_wavelengths, tuples of fields, dispersive data).

Everything else the reference accepts raises ``NotImplementedError``.  The
propagation runs in ``libdtmmx.so`` (HIP, gfx950); torch provides device
memory and the stream.  There is no CPU fallback.
"""
import copy
import ctypes
import warnings

import numpy as np

from . import _lib
from . import tune as _tune
from .data import (effective_from_tensors, is_optical_data_dispersive, parse_symmetry,
                   uniaxial_order, validate_optical_data)
from .field import blackman, k0 as _k0

DTMM_NORM_FFT = 1 << 0
DTMM_NORM_REF = 1 << 1
SMOOTH = 0.1
BETAMAX = 0.8
N_COVER = 1.5
# dtmm/dtmm.ini:23 holds `method = "2x2"`; configparser returns the quotes,
# so the configured default is this string (it selects the 2x2 method with
# reflection mode 1, dtmm/transfer.py:459).
_DEFAULT_METHOD = '"2x2"'


def _torch():
    import torch
    return torch


def _require_gpu():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("dtmm_amd.transfer_field needs a ROCm GPU (no CPU fallback)")
    _lib.load()
    return torch


def _validate_betaphi(beta, phi):
    """(dtmm/transfer.py:43-65) -- returns 1-D arrays of beams."""
    if beta is None or phi is None:
        raise ValueError("Both beta and phi must be defined!")
    beta = np.asarray(beta, np.float64)
    phi = np.asarray(phi, np.float64)
    if beta.ndim != phi.ndim:
        raise ValueError("Beta nad phi should have same dimensions!")
    if beta.ndim == 1:
        if len(beta) != len(phi):
            raise ValueError("Beta nad phi should have same length!")
        return beta, phi, True
    if beta.ndim != 0:
        raise ValueError("Only length 1 or scalar values are supported for beta and phi")
    return beta[None], phi[None], False


def mean_betaphi(field, k0):
    """Intensity-weighted mean beta, phi of a field (dtmm/field.py:849-872)."""
    field = np.asarray(field)
    shape = field.shape[-2:]
    f = np.fft.fft2(field * blackman(shape))
    k0 = np.asarray(k0, np.float64)[..., None, None]
    ay, ax = [np.fft.fftfreq(x) for x in shape]
    xx, yy = np.meshgrid(ax, ay, indexing="xy")
    bx = (2 * np.pi / k0) * xx
    by = (2 * np.pi / k0) * yy
    s = (np.abs(f) ** 2)
    sx = (s * bx[..., None, :, :]).sum((-3, -2, -1))
    sy = (s * by[..., None, :, :]).sum((-3, -2, -1))
    ss = s.sum((-3, -2, -1))
    return (sx ** 2 + sy ** 2) ** 0.5 / ss, np.arctan2(sy, sx)


def field2betaphi(field_in, ks, multiray=False):
    """Mean beam of the input field (dtmm/field.py:492-500)."""
    beta, phi = mean_betaphi(field_in, ks)
    if field_in.ndim > 4 and multiray:
        axes = tuple(range(1, field_in.ndim - 3))
        return beta.mean(axis=axes), phi.mean(axis=axes)
    return beta.mean(), phi.mean()


class Stack:
    """Layer stack resident on the device: material as float64
    structure-of-arrays [3][L][H][W] plus per-layer thickness, sub-step
    count and isotropic effective eps.  Built from reference optical data
    (dtmm/transfer.py:926-1005, eff_data=0); layer l is propagated in
    nstep sub-steps of thickness d / nstep (transfer.py:933-934).
    Anisotropic effective media (other eff_data) are attached per transfer
    by ``with_effective``."""

    # anisotropic effective media: [L][3] eigenvalues (re, im), Euler angles,
    # thickness of the diffraction step; None: the isotropic eps_eff
    eff_eps_re = None
    eff_eps_im = None
    eff_angles = None
    eff_thickness = None
    # (first layer, layer count) of each optical block
    blocks = None

    def __init__(self, optical_data, shape, nstep=1, nin=N_COVER, nout=N_COVER,
                 split_layers=False, device="cuda"):
        torch = _torch()
        blocks = validate_optical_data(optical_data, shape=shape)
        H, W = shape
        self.shape = (H, W)
        thick, steps, eff, eps_l, biax = [], [], [], [], []
        self.blocks = []
        for d, material, angles in blocks:
            n = len(d)
            if split_layers:
                self.blocks += [(len(thick) + i, 1) for i in range(n)]
            else:
                self.blocks.append((len(thick), n))
            substeps = np.broadcast_to(np.asarray(nstep), (n,))
            mat = np.broadcast_to(material, (n, H, W, 3))
            ang = np.broadcast_to(angles, (n, H, W, 3))
            if split_layers:
                effs = [uniaxial_order(0., mat[i:i + 1]).mean(axis=(0, 1, 2))[0] for i in range(n)]
            else:
                e = uniaxial_order(0., material).mean(axis=(0, 1, 2))
                effs = [e[0]] * n
            for i in range(n):
                ns = int(substeps[i])
                if ns < 1:
                    raise ValueError("nstep must be >= 1")
                thick.append(float(d[i]))
                steps.append(ns)
                eff.append(complex(effs[i]))
                eps_l.append((len(eps_l), mat[i], ang[i]))
                # numeric eigensolve where eps0 != eps1 (dtmm/tmm.py:384-403)
                biax.append(int(np.any(mat[i][..., 0] != mat[i][..., 1])))
        self.n_layers = len(thick)
        self.thickness = np.asarray(thick, np.float64)
        self.n_steps = np.asarray(steps, np.int32)
        self.biaxial = np.asarray(biax, np.int32)
        self.eps_eff_re = np.asarray([e.real for e in eff], np.float64)
        self.eps_eff_im = np.asarray([e.imag for e in eff], np.float64)
        self.nin = float(nin)
        self.nout = float(nout)
        L = self.n_layers
        is_complex = any(np.iscomplexobj(b[1]) for b in blocks)
        eps_re = np.empty((3, L, H, W), np.float64)
        eps_im = np.empty((3, L, H, W), np.float64) if is_complex else None
        angs = np.empty((3, L, H, W), np.float64)
        for l, (_, m, a) in enumerate(eps_l):
            mm = np.moveaxis(m, -1, 0)
            eps_re[:, l] = mm.real
            if eps_im is not None:
                eps_im[:, l] = mm.imag
            angs[:, l] = np.moveaxis(a, -1, 0)
        self.eps_re = torch.from_numpy(eps_re).to(device)
        self.eps_im = torch.from_numpy(eps_im).to(device) if eps_im is not None else None
        self.angles = torch.from_numpy(angs).to(device)
        self.has_imag_eff = bool(np.any(self.eps_eff_im != 0))

    @classmethod
    def from_director(cls, director, mask=None, no=1.5, ne=1.6, nhost=None, scale_factor=1.,
                      thickness=None, nstep=1, nin=N_COVER, nout=N_COVER, split_layers=False):
        """The stack of ``director2data(director, mask, no, ne, nhost,
        scale_factor, thickness)`` (dtmm/data.py:277-320) built on the GPU
        (``dtmmx_director2data``): the director (L, H, W, 3) -- a device
        tensor stays on the device -- becomes the device material arrays and
        the block's effective medium (the mean isotropic eps, per layer with
        ``split_layers``) without a host copy of the material.  Pass the
        result as ``optical_data`` to ``transfer_field``."""
        lib = _lib.load()

        def launch(src, mk, H, W, L, eps_re, angles, sums, stream):
            return lib.dtmmx_director2data(
                src, mk, H * W, L, float(no), float(ne),
                float(nhost) if nhost is not None else float("nan"), float(scale_factor),
                eps_re, angles, sums, stream)
        return cls._from_device_field(director, 3, "director", mask, nhost, launch, False,
                                      thickness, nstep, nin, nout, split_layers)

    @classmethod
    def from_Q(cls, qtensor, mask=None, no=1.5, ne=1.6, nhost=None, scale_factor=1.,
               biaxial=False, thickness=None, nstep=1, nin=N_COVER, nout=N_COVER,
               split_layers=False):
        """The stack of ``Q2data(qtensor, mask, no, ne, nhost, scale_factor,
        biaxial, thickness)`` (dtmm/data.py:322-373) built on the GPU
        (``dtmmx_q2data``): the Q tensor (L, H, W, 6) with elements Q00, Q11,
        Q22, Q01, Q02, Q12, or the full (L, H, W, 3, 3) matrix as Q2eps also
        takes it (data.py:463-472) -- a device tensor stays on the device --
        becomes the device material arrays (eigenvalues of eps, the most
        distinct last) and the Euler angles of the eigenframe, plus the
        effective medium as in ``from_director``."""
        lib = _lib.load()
        torch = _torch()
        if tuple(qtensor.shape[-2:]) == (3, 3) and len(qtensor.shape) == 5:
            rows, cols = [0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2]
            if isinstance(qtensor, torch.Tensor):
                qtensor = qtensor[..., rows, cols]
            else:
                qtensor = np.asarray(qtensor)[..., rows, cols]

        def launch(src, mk, H, W, L, eps_re, angles, sums, stream):
            return lib.dtmmx_q2data(
                src, mk, H * W, L, float(no), float(ne),
                float(nhost) if nhost is not None else float("nan"), float(scale_factor),
                1 if biaxial else 0, eps_re, angles, sums, stream)
        return cls._from_device_field(qtensor, 6, "qtensor", mask, nhost, launch, bool(biaxial),
                                      thickness, nstep, nin, nout, split_layers)

    @classmethod
    def _from_device_field(cls, field, ncomp, name, mask, nhost, launch, biaxial, thickness,
                           nstep, nin, nout, split_layers):
        torch = _torch()
        if isinstance(field, torch.Tensor):
            if not field.is_cuda:
                raise ValueError("tensor inputs must be device tensors")
            dr = field.to(torch.float64).contiguous()
        else:
            dr = torch.from_numpy(np.ascontiguousarray(np.asarray(field, np.float64)))
            dr = dr.to("cuda")
        if dr.dim() != 4 or dr.shape[-1] != ncomp:
            raise ValueError("%s must have shape (layers, height, width, %d)" % (name, ncomp))
        L, H, W = (int(x) for x in dr.shape[:3])
        mk = None
        if mask is not None:
            if isinstance(mask, torch.Tensor):
                mk = mask.to(device="cuda", dtype=torch.uint8).contiguous()
            else:
                mk = torch.from_numpy(np.ascontiguousarray(np.asarray(mask, np.uint8))).to("cuda")
            if tuple(mk.shape) != (L, H, W):
                raise ValueError("mask must have shape (layers, height, width)")
            if nhost is None:
                raise ValueError("nhost is required with a mask")
        lib = _lib.load()
        nb = int(lib.dtmmx_director2data_blocks(H * W))
        eps_re = torch.empty((3, L, H, W), dtype=torch.float64, device="cuda")
        angles = torch.empty((3, L, H, W), dtype=torch.float64, device="cuda")
        sums = torch.empty((L, nb), dtype=torch.float64, device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(launch(
            ctypes.c_void_p(dr.data_ptr()), ctypes.c_void_p(mk.data_ptr()) if mk is not None
            else None, H, W, L, ctypes.c_void_p(eps_re.data_ptr()),
            ctypes.c_void_p(angles.data_ptr()), ctypes.c_void_p(sums.data_ptr()),
            ctypes.c_void_p(stream)))
        layer_sum = sums.sum(dim=1).cpu().numpy()
        if split_layers:
            eff = layer_sum / (H * W)
        else:
            eff = np.full(L, layer_sum.sum() / (L * H * W))
        self = cls.__new__(cls)
        self.shape = (H, W)
        self.n_layers = L
        self.thickness = (np.ones(L) if thickness is None
                          else np.asarray(thickness, np.float64).reshape(L))
        steps = np.broadcast_to(np.asarray(nstep), (L,))
        if np.any(steps < 1):
            raise ValueError("nstep must be >= 1")
        self.n_steps = np.asarray(steps, np.int32)
        # director data and uniaxial Q data have eps0 == eps1 in every voxel
        self.biaxial = np.full(L, 1 if biaxial else 0, np.int32)
        self.eps_eff_re = np.asarray(eff, np.float64)
        self.eps_eff_im = np.zeros(L, np.float64)
        self.nin = float(nin)
        self.nout = float(nout)
        self.eps_re = eps_re
        self.eps_im = None
        self.angles = angles
        self.has_imag_eff = False
        self.blocks = [(i, 1) for i in range(L)] if split_layers else [(0, L)]
        return self

    def layer_tensors(self):
        """Mean laboratory-frame eps tensor (xx, yy, zz, xy, xz, yz) of every
        layer, (L, 6) complex, from the device material: R diag(eps) R^T per
        voxel with R = Rz(phi) Ry(theta) Rz(psi) (dtmm/data.py:515-532),
        averaged over the layer."""
        torch = _torch()
        out = np.empty((self.n_layers, 6), np.complex128)
        for l in range(self.n_layers):
            a = self.angles[:, l]
            cps, sps = torch.cos(a[0]), torch.sin(a[0])
            cth, sth = torch.cos(a[1]), torch.sin(a[1])
            cph, sph = torch.cos(a[2]), torch.sin(a[2])
            r = ((cph * cth * cps - sph * sps, -cph * cth * sps - sph * cps, cph * sth),
                 (sph * cth * cps + cph * sps, cph * cps - sph * cth * sps, sph * sth),
                 (-sth * cps, sth * sps, cth))
            e = [self.eps_re[c, l].to(torch.complex128) for c in range(3)]
            if self.eps_im is not None:
                e = [e[c] + 1j * self.eps_im[c, l] for c in range(3)]
            for k, (i, j) in enumerate(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
                t = r[i][0] * r[j][0] * e[0] + r[i][1] * r[j][1] * e[1] + r[i][2] * r[j][2] * e[2]
                out[l, k] = complex(t.mean().item())
        return out

    def with_effective(self, eff_data, per_block=False):
        """This stack with the effective media of ``eff_data``
        (dtmm/transfer.py:926-1005): None or a symmetry 0 keeps the isotropic
        block means; a symmetry 1 or 2 ("uniaxial", "biaxial") or a sequence
        of per-layer symmetries averages the laboratory-frame tensors
        (dtmm/data.py:1364-1419); a custom effective block (d_eff, epsv,
        epsa) is taken as given.  ``per_block``: eff_data holds one entry per
        optical block.  Effective media that come out isotropic with the
        layer thickness stay on the isotropic path."""
        blocks = self.blocks or [(0, self.n_layers)]
        if per_block:
            if len(eff_data) != len(blocks):
                raise ValueError("Number of elements in eff_data must match length of "
                                 "optical_data.")
            effs = list(eff_data)
        else:
            effs = [eff_data] * len(blocks)

        def default(e):
            return e is None or (isinstance(e, (int, str)) and parse_symmetry(e) == 0)
        if all(default(e) for e in effs):
            return self
        L = self.n_layers
        ev = np.empty((L, 3), np.complex128)
        ea = np.zeros((L, 3))
        de = np.array(self.thickness, np.float64)
        tensors = None
        for (start, n), e in zip(blocks, effs):
            sl = slice(start, start + n)
            if default(e):
                ev[sl] = (self.eps_eff_re[sl] + 1j * self.eps_eff_im[sl])[:, None]
            elif _is_effective_block(e):
                d_e, v_e, a_e = e
                d_e = np.atleast_1d(np.asarray(d_e, np.float64))
                if len(d_e) != n:
                    raise ValueError("the effective block needs one layer per layer of its block")
                ev[sl] = np.broadcast_to(np.asarray(v_e).reshape(-1, 3), (n, 3))
                ea[sl] = np.broadcast_to(np.asarray(a_e, np.float64).reshape(-1, 3), (n, 3))
                de[sl] = d_e
            else:
                if tensors is None:
                    tensors = self.layer_tensors()
                w, a = effective_from_tensors(tensors[sl], e)
                ev[sl] = w
                ea[sl] = a
        out = copy.copy(self)
        iso = (ev[:, 0] == ev[:, 1]) & (ev[:, 1] == ev[:, 2])
        if iso.all() and np.array_equal(de, np.asarray(self.thickness, np.float64)):
            out.eps_eff_re = np.ascontiguousarray(ev[:, 0].real)
            out.eps_eff_im = np.ascontiguousarray(ev[:, 0].imag)
            out.has_imag_eff = bool(np.any(out.eps_eff_im != 0))
            return out
        out.eff_eps_re = np.ascontiguousarray(ev.real)
        out.eff_eps_im = np.ascontiguousarray(ev.imag) if np.any(ev.imag != 0) else None
        out.eff_angles = np.ascontiguousarray(ea)
        out.eff_thickness = np.ascontiguousarray(de)
        return out

    @property
    def anisotropic_effective(self):
        return self.eff_eps_re is not None

    def c_struct(self):
        s = _lib.Stack()
        s.n_layers = self.n_layers
        s.thickness = _lib.dptr(self.thickness)
        s.eps_eff_re = _lib.dptr(self.eps_eff_re)
        s.eps_eff_im = _lib.dptr(self.eps_eff_im)
        s.eps_re = self.eps_re.data_ptr()
        s.eps_im = self.eps_im.data_ptr() if self.eps_im is not None else None
        s.angles = self.angles.data_ptr()
        s.n_in = self.nin
        s.n_out = self.nout
        s.n_steps = _lib.iptr(self.n_steps)
        s.biaxial = _lib.iptr(self.biaxial) if self.biaxial.any() else None
        if self.eff_eps_re is not None:
            s.eff_eps_re = _lib.dptr(self.eff_eps_re)
            s.eff_eps_im = _lib.dptr(self.eff_eps_im) if self.eff_eps_im is not None else None
            s.eff_angles = _lib.dptr(self.eff_angles)
            s.eff_thickness = _lib.dptr(self.eff_thickness)
        return s


# size of the library's per-GPU scratch array used by Plan.execute_noalloc
# (DTMMX_NOALLOC_SCRATCH_BYTES of include/dtmmx.h)
NOALLOC_SCRATCH_BYTES = 128 << 20


class Plan:
    """Owner of a ``dtmmx_plan``: grid, wavelengths, beams, method, precision."""

    def __init__(self, shape, ks, beta, phi, n_rays, rays_per_beam, method, precision="c64",
                 betamax=BETAMAX, diffraction=1, npass=1, norm=DTMM_NORM_FFT, smooth=SMOOTH,
                 reflection=None, checkpoint=1):
        lib = _lib.load()
        self.shape = tuple(shape)
        self.ks = np.ascontiguousarray(ks, np.float64)
        self.beta = np.ascontiguousarray(beta, np.float64)
        self.phi = np.ascontiguousarray(phi, np.float64)
        self.n_rays = int(n_rays)
        self.rays_per_beam = int(rays_per_beam)
        self.method = method
        self.precision = precision
        self.diffraction = int(diffraction)
        self.npass = int(npass)
        self.norm = int(norm)
        self.betamax = float(betamax)
        self.reflection = int(reflection if reflection is not None else
                              (1 if method == "4x4" else 0))
        d = _lib.PlanDesc()
        d.height, d.width = self.shape
        d.n_wavelengths = len(self.ks)
        d.n_rays = self.n_rays
        d.rays_per_beam = self.rays_per_beam
        d.method = _lib.DTMMX_METHOD_4X4 if method == "4x4" else _lib.DTMMX_METHOD_2X2
        d.precision = _lib.DTMMX_C64 if precision == "c64" else _lib.DTMMX_C128
        d.betamax = float(betamax)
        d.k0 = _lib.dptr(self.ks)
        d.beta = _lib.dptr(self.beta)
        d.phi = _lib.dptr(self.phi)
        d.diffraction = self.diffraction
        d.npass = self.npass
        d.norm = int(norm)
        d.smooth = float(smooth)
        d.reflection = self.reflection
        d.checkpoint = int(checkpoint)
        h = ctypes.c_void_p()
        _lib.check(lib.dtmmx_plan_create(ctypes.byref(h), ctypes.byref(d)))
        self._h = h
        self._lib = lib
        self.workspace_bytes = int(lib.dtmmx_plan_workspace_size(h))

    @property
    def dtype(self):
        torch = _torch()
        return torch.complex64 if self.precision == "c64" else torch.complex128

    def workspace_for(self, stack):
        """Workspace bytes of a transfer of `stack`: the plan's part plus the
        reflected wave of every interface that the 2x2 method keeps between
        passes (npass > 1)."""
        s = stack.c_struct()
        return int(self._lib.dtmmx_transfer_workspace_size(self._h, ctypes.byref(s)))

    def execute(self, stack, field_in, field_out, workspace=None, status=None, stream=None):
        """Run the transfer on device tensors of shape (n_rays, n_wavelengths,
        4, H, W); asynchronous on `stream` (default: torch's current).  With
        npass > 1 `field_in` receives the field at the input side (input plus
        reflection), as the reference writes it into its input array."""
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream(field_in.device)
        for t in (field_in, field_out):
            if t.dtype != self.dtype or not t.is_contiguous() or not t.is_cuda:
                raise ValueError("fields must be contiguous %s device tensors (got %s, "
                                 "contiguous=%s, device=%s)" % (self.dtype, t.dtype,
                                                                t.is_contiguous(), t.device))
            if tuple(t.shape) != (self.n_rays, len(self.ks), 4) + self.shape:
                raise ValueError("field shape %s does not match the plan" % (tuple(t.shape),))
        need = self.workspace_for(stack)
        if need and (workspace is None or workspace.numel() * workspace.element_size() < need):
            workspace = torch.empty(need, dtype=torch.uint8, device=field_in.device)
        s = stack.c_struct()
        _lib.check(self._lib.dtmmx_transfer(
            self._h, ctypes.byref(s), field_in.data_ptr(), field_out.data_ptr(),
            workspace.data_ptr() if workspace is not None else None,
            need, status.data_ptr() if status is not None else None,
            stream.cuda_stream))
        return field_out

    def execute_noalloc(self, stack, field_in, field_out, status=None, stream=None):
        """execute() through ``dtmmx_transfer_noalloc``: the workspace is the
        library's static per-GPU scratch array (at most
        ``NOALLOC_SCRATCH_BYTES``), so the call allocates no device memory.
        Calls that overlap on one GPU share that array and must not run
        concurrently."""
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream(field_in.device)
        for t in (field_in, field_out):
            if t.dtype != self.dtype or not t.is_contiguous() or not t.is_cuda:
                raise ValueError("fields must be contiguous %s device tensors" % self.dtype)
            if tuple(t.shape) != (self.n_rays, len(self.ks), 4) + self.shape:
                raise ValueError("field shape %s does not match the plan" % (tuple(t.shape),))
        s = stack.c_struct()
        _lib.check(self._lib.dtmmx_transfer_noalloc(
            self._h, ctypes.byref(s), field_in.data_ptr(), field_out.data_ptr(),
            status.data_ptr() if status is not None else None, stream.cuda_stream))
        return field_out

    def execute_bulk(self, stack, field_in, field_out, bulk, workspace=None, status=None,
                     stream=None):
        """execute() that also writes the field after every stack layer into
        `bulk`, a contiguous device tensor (n_layers, n_rays, n_wavelengths,
        4, H, W) of the plan's dtype (diffraction 0 or 1; npass 1, or any
        npass with the 2x2 method; dtmm/transfer.py ret_bulk)."""
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream(field_in.device)
        want = (stack.n_layers, self.n_rays, len(self.ks), 4) + self.shape
        if (bulk.dtype != self.dtype or not bulk.is_contiguous() or not bulk.is_cuda or
                tuple(bulk.shape) != want):
            raise ValueError("bulk must be a contiguous %s device tensor of shape %s"
                             % (self.dtype, want))
        need = self.workspace_for(stack)
        if need and (workspace is None or workspace.numel() * workspace.element_size() < need):
            workspace = torch.empty(need, dtype=torch.uint8, device=field_in.device)
        s = stack.c_struct()
        _lib.check(self._lib.dtmmx_transfer_bulk(
            self._h, ctypes.byref(s), field_in.data_ptr(), field_out.data_ptr(), bulk.data_ptr(),
            workspace.data_ptr() if workspace is not None else None, need,
            status.data_ptr() if status is not None else None, stream.cuda_stream))
        return bulk

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                self._lib.dtmmx_plan_destroy(self._h)
                self._h = None
        except Exception:
            pass


def _is_effective_block(e):
    """A custom effective block (d, epsv, epsa) rather than a symmetry or a
    sequence of per-layer symmetries."""
    return (isinstance(e, (tuple, list)) and len(e) == 3 and np.ndim(e[1]) >= 1
            and np.ndim(e[2]) >= 1)


def _resolve_modes(method, npass, diffraction, reflection, eff_data, ret_bulk, norm=DTMM_NORM_FFT):
    """Defaults and reflection auto-selection (dtmm/transfer.py:433-465) and
    the envelope of modes built on the GPU."""
    if npass == -1 or npass == np.inf:
        raise NotImplementedError("npass=-1 (full 3D TMM, dtmm/tmm3d.py) is not supported")
    if method == "4x4" and npass > 1 and diffraction is False:
        warnings.warn("The 4x4 method with diffraction disabled is not yet supported"
                      " for input fields with beta >0. Use 2x2 method insted.")
    if reflection is None:
        reflection = 0 if method == "2x2" else 1
        if method == "4x4" and diffraction == 0:
            reflection = 2
        if npass > 1:
            reflection = 1
            if diffraction > 1 and method == "2x2":
                reflection = 2
    kind = "4x4" if method == "4x4" else "2x2"
    if kind == "4x4" and reflection not in (1, 2, 3, 4):
        raise ValueError("Invalid reflection. The 4x4 method is either reflection mode 1 or 2.")
    if kind == "2x2" and reflection not in (0, 1, 2):
        raise ValueError("Invalid reflection. The 2x2 method supports reflection mode 0,1 or 2.")
    if npass < 1:
        raise ValueError("npass must be >= 1")
    if npass > 1 and kind == "2x2":
        if reflection == 0:
            raise ValueError("Reflection mode `0` not compatible with npass > 1")
        if not ((reflection == 1 and diffraction == 1) or
                (reflection == 2 and diffraction in (0, 1))):
            raise NotImplementedError("npass > 1 with the 2x2 method is supported for "
                                      "reflection mode 1 with diffraction 1 and reflection "
                                      "mode 2 with diffraction 0 or 1")
    if npass > 1 and kind == "4x4" and (norm & DTMM_NORM_REF) and not (0 <= diffraction <= 1):
        raise NotImplementedError("npass > 1 with DTMM_NORM_REF (reference-field "
                                  "normalisation) is supported with diffraction 0 or 1")
    if ret_bulk and (not (0 <= diffraction <= 1) or (npass != 1 and kind == "4x4")):
        raise NotImplementedError("ret_bulk is supported with diffraction 0 or 1, and npass=1 "
                                  "for the 4x4 method")
    if diffraction < 0 or diffraction == np.inf:
        # exact per-mode propagation (dtmm/propagate_4x4.py:515-564)
        if kind != "4x4" or npass != 1 or ret_bulk:
            raise NotImplementedError("diffraction < 0 or inf (exact per-mode propagation) is "
                                      "supported for the 4x4 method with npass=1 and no bulk")
        return kind, reflection
    if diffraction == 0:
        # no diffraction: every pixel is an independent stack (4x4 reflection
        # mode 2, the reference's default for diffraction 0; 2x2 reflection 0)
        if npass != 1 and reflection != 2:
            raise NotImplementedError("diffraction=0 with npass > 1 is supported in reflection "
                                      "mode 2")
        if kind == "4x4" and reflection != 2:
            raise NotImplementedError("4x4 method with diffraction=0: only reflection mode 2 "
                                      "is supported")
        if kind == "2x2" and reflection not in (0, 2):
            raise NotImplementedError("2x2 method with diffraction=0: reflection modes 0 and 2 "
                                      "are supported")
        return kind, reflection
    if kind == "4x4":
        if reflection == 2 and diffraction >= 1 and diffraction != np.inf:
            return kind, reflection
        if reflection != 1:
            raise NotImplementedError("4x4 method: reflection modes 1 and 2 are supported")
        if not (diffraction >= 1) or diffraction == np.inf:
            raise NotImplementedError("4x4 method: only diffraction 0 or finite diffraction >= 1 "
                                      "is supported")
    else:
        if reflection == 2 and diffraction == 1:
            return kind, reflection
        if reflection not in (0, 1):
            raise NotImplementedError("2x2 method: reflection mode 2 is supported with "
                                      "diffraction 0 or 1")
        if diffraction > 1 and (npass != 1 or ret_bulk):
            raise NotImplementedError("2x2 method with diffraction > 1: npass 1 without ret_bulk "
                                      "is supported")
        if not (diffraction >= 1) or diffraction == np.inf:
            raise NotImplementedError("2x2 method: diffraction 0 (reflection 0) or finite "
                                      "diffraction >= 1 is supported")
    return kind, reflection


class _DeviceField:
    """Shape and dtype of a device field (a torch tensor on the GPU) in the
    places where the host path reads them from a NumPy array."""

    def __init__(self, t):
        import torch
        self.t = t
        self.shape = tuple(t.shape)
        self.ndim = t.dim()
        self.dtype = np.complex64 if t.dtype == torch.complex64 else np.complex128

    def host(self):
        return self.t.cpu().numpy()


class _Prepared:
    """A transfer made ready on the device: plan, stack, fields, workspace."""

    def __init__(self, plan, stack, fin, fout, ws, field_in, field_np, npass, out, wavelengths,
                 pixelsize):
        self.plan, self.stack, self.fin, self.fout, self.ws = plan, stack, fin, fout, ws
        self.field_in, self.field_np, self.npass, self.out = field_in, field_np, npass, out
        self.wavelengths, self.pixelsize = wavelengths, pixelsize
        self.on_device = isinstance(field_np, _DeviceField)

    def finish(self):
        if self.on_device:
            # a device field in, a device field out (no host copy)
            res = self.fout.reshape(self.field_np.shape)
            if self.npass > 1:
                self.field_in.copy_(self.fin.reshape(self.field_np.shape))
            if self.out is not None:
                self.out.copy_(res)
                res = self.out
            return res, self.wavelengths, self.pixelsize
        res = self.fout.cpu().numpy().reshape(self.field_np.shape)
        if (self.npass > 1 and isinstance(self.field_in, np.ndarray)
                and self.field_in.flags.writeable):
            # the reference leaves the input-side field (with reflections) in
            # the caller's input array (dtmm/transfer.py:843, :862-866)
            self.field_in[...] = self.fin.cpu().numpy().reshape(self.field_np.shape)
        if self.out is not None:
            self.out[...] = res
            res = self.out
        return res, self.wavelengths, self.pixelsize


def _prepare(field_data, optical_data, beta=None, phi=None, nin=None, nout=None, npass=None,
             nstep=1, diffraction=None, reflection=None, method=None, multiray=False,
             norm=DTMM_NORM_FFT, betamax=None, smooth=SMOOTH, split_layers=False, eff_data=None,
             ret_bulk=False, out=None, precision=None, checkpoint=1):
    """Defaults, validation, plan and device data of one transfer_field call
    (dtmm/transfer.py:433-552); `checkpoint` is the adjoint tape interval."""
    torch = _require_gpu()
    nin = N_COVER if nin is None else nin
    nout = N_COVER if nout is None else nout
    method = _DEFAULT_METHOD if method is None else method
    npass = 1 if npass is None else npass
    diffraction = 1 if diffraction is None else diffraction
    betamax = BETAMAX if betamax is None else betamax
    kind, reflection = _resolve_modes(method, npass, diffraction, reflection, eff_data, ret_bulk,
                                      norm)
    if kind == "4x4" and npass > 1 and smooth <= 0:
        # the notch filter of the pass normalisation divides by its width
        # (dtmm/transfer.py:632-635, dtmm/denoise.py:21-22)
        raise ZeroDivisionError("float division by zero")

    field_in, wavelengths, pixelsize = field_data
    if isinstance(field_in, tuple):
        raise ValueError("a tuple of per-wavelength fields goes through transfer_field")
    if torch.is_tensor(field_in) and field_in.is_cuda:
        # a field that already lives on the device (illumination_data(device=...))
        field_np = _DeviceField(field_in)
    else:
        field_np = np.asarray(field_in.cpu() if torch.is_tensor(field_in) else field_in)
    shape = field_np.shape[-2:]
    wavelengths = np.asarray(wavelengths)
    ks = _k0(wavelengths, pixelsize)
    if field_np.ndim < 4 or field_np.shape[-3] != 4 or field_np.shape[-4] != len(ks):
        raise ValueError("field must have shape (..., n_wavelengths, 4, H, W)")
    if precision is None:
        precision = "c64" if field_np.dtype == np.complex64 else "c128"
    if beta is None and phi is None:
        beta, phi = field2betaphi(field_np.host() if isinstance(field_np, _DeviceField) else
                                  field_np, ks, multiray)
    b, p, per_ray = _validate_betaphi(beta, phi)
    lead = field_np.shape[:-4]
    n_rays = int(np.prod(lead)) if lead else 1
    if per_ray:
        if not lead or lead[0] != len(b):
            raise ValueError("beta/phi of length %d need a leading ray axis of that length" % len(b))
    if n_rays % len(b):
        raise ValueError("rays do not divide into beams")
    rays_per_beam = n_rays // len(b)

    # exact per-mode propagation travels as diffraction -1 through the ABI
    ndiff = -1 if (diffraction < 0 or diffraction == np.inf) else int(diffraction)
    plan = Plan(shape, ks, b, p, n_rays, rays_per_beam, kind, precision, betamax,
                diffraction=ndiff, npass=int(npass), norm=norm, smooth=smooth,
                reflection=reflection, checkpoint=checkpoint)
    if isinstance(optical_data, Stack):
        # a stack already on the device (Stack.from_director): its own layers
        # and sub-steps, this call's input and output media
        stack = optical_data
        if tuple(stack.shape) != tuple(shape):
            raise ValueError("the stack's grid %s does not match the field's %s"
                             % (tuple(stack.shape), tuple(shape)))
        stack.nin, stack.nout = float(nin), float(nout)
    else:
        stack = Stack(optical_data, shape, nstep=nstep, nin=nin, nout=nout,
                      split_layers=split_layers)
    if diffraction > 0 and diffraction != np.inf:
        # no effective medium enters without diffraction or in the exact
        # per-mode propagation
        per_block = (isinstance(optical_data, list) and eff_data is not None
                     and not isinstance(eff_data, (int, str)))
        stack = stack.with_effective(eff_data, per_block)
        if stack.anisotropic_effective and not (
                kind == "4x4" and reflection == 1 and diffraction == 1 and
                not (npass > 1 and norm & DTMM_NORM_REF)):
            raise NotImplementedError("anisotropic effective data is supported with the 4x4 "
                                      "method in reflection mode 1 with diffraction 1 "
                                      "(without DTMM_NORM_REF) and with diffraction 0")
    if isinstance(field_np, _DeviceField):
        fin = field_in.reshape((n_rays, len(ks), 4) + shape).to(dtype=plan.dtype).clone()
    else:
        fin = torch.from_numpy(np.ascontiguousarray(field_np.reshape((n_rays, len(ks), 4) +
                                                                     shape)))
        fin = fin.to(device="cuda", dtype=plan.dtype)
    fout = torch.empty_like(fin)
    ws = None
    need = plan.workspace_for(stack)
    if need:
        ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    return _Prepared(plan, stack, fin, fout, ws, field_in, field_np, npass, out, wavelengths,
                     pixelsize)


def _check_status(status, adjoint=False):
    flags = int(status[0].item())
    if flags & 4 and not flags & 1:
        warnings.warn("the output field holds NaN or infinite values", RuntimeWarning,
                      stacklevel=3)
    if flags & 1:
        if adjoint:
            raise NotImplementedError("biaxial voxels (eps0 != eps1) are not supported by the "
                                      "adjoint transfer")
        raise NotImplementedError("a biaxial voxel (eps0 != eps1) was met in a layer not "
                                  "flagged biaxial (Stack.biaxial); its output is NaN")


def transfer_field(field_data, optical_data, beta=None, phi=None, nin=None, nout=None,
                   npass=None, nstep=1, diffraction=None, reflection=None, method=None,
                   multiray=False, norm=DTMM_NORM_FFT, betamax=None, smooth=SMOOTH,
                   split_rays=False, split_diffraction=False, split_wavelengths=False,
                   split_layers=False, eff_data=None, ret_bulk=False, out=None,
                   precision=None, tune=False):
    """Transfers input field data through optical data (dtmm/transfer.py:322).

    Returns ``(field_out, wavelengths, pixelsize)``.  ``precision`` ("c64" or
    "c128") selects the GPU arithmetic; by default it follows the input
    field dtype (complex64 -> "c64", otherwise "c128").  ``split_rays`` and
    ``split_diffraction`` only bound memory in the reference and do not
    change the result; they are accepted and ignored (the GPU batches all
    planes).  ``tune`` times the kernels' launch shapes on this transfer
    first (dtmm_amd.tune; the winner is cached for later calls of the same
    workload shape).

    Per-wavelength transfers, as in the reference (transfer.py:486-528):
    with ``split_wavelengths`` (and no ``ret_bulk``) each wavelength of the
    (..., wavelengths, 4, H, W) input is transferred on its own and the
    result is put back together; a tuple of per-wavelength fields, each
    (..., 4, H, W), returns a tuple of ``(field, wavelength, pixelsize)``.
    The optical data is evaluated per wavelength, so a dispersive eps
    (``EpsilonCauchy``, ``EpsilonSellmeier`` or any callable of the
    wavelength in nm) is accepted on this route only (transfer.py:535-536).
    """
    kw = dict(beta=beta, phi=phi, nin=nin, nout=nout, npass=npass, nstep=nstep,
              diffraction=diffraction, reflection=reflection, method=method, multiray=multiray,
              norm=norm, betamax=betamax, smooth=smooth, split_layers=split_layers,
              eff_data=eff_data, ret_bulk=ret_bulk, precision=precision, tune=tune)
    field_in, wavelengths, pixelsize = field_data
    if (not isinstance(field_in, tuple) and not (split_wavelengths and not ret_bulk) and
            not isinstance(optical_data, Stack)):
        shape = np.shape(field_in)[-2:]
        if is_optical_data_dispersive(validate_optical_data(optical_data, shape=shape)):
            raise ValueError("You are using dispersive data, so you must use "
                             "`split_wavelengths=True`")
    _require_gpu()
    if split_wavelengths and not isinstance(field_in, tuple) and not ret_bulk:
        arr = field_in if isinstance(field_in, np.ndarray) else np.asarray(field_in)
        res = np.empty_like(arr) if out is None else out
        parts = tuple(arr[..., i, :, :, :] for i in range(len(wavelengths)))
        outs = [res[..., i, :, :, :] for i in range(len(wavelengths))]
        _transfer_wavelengths(parts, wavelengths, pixelsize, optical_data, outs, kw)
        return res, wavelengths, pixelsize
    if isinstance(field_in, tuple):
        return _transfer_wavelengths(field_in, wavelengths, pixelsize, optical_data, out, kw)
    return _transfer_one(field_data, optical_data, out=out, **kw)


def _transfer_wavelengths(fields, wavelengths, pixelsize, optical_data, outs, kw):
    """One transfer per wavelength (dtmm/transfer.py:510-528).  Each field
    (..., 4, H, W) gets a wavelength axis of length 1 as a view, so the
    input-side field of npass > 1 lands in the caller's array.  Optical
    data that does not depend on the wavelength is uploaded once."""
    if len(fields) != len(wavelengths):
        raise ValueError("the tuple of fields needs one wavelength per field")
    outs = [None] * len(fields) if outs is None else list(outs)
    stack = optical_data
    if not isinstance(optical_data, Stack):
        shape = np.shape(fields[0])[-2:]
        if not is_optical_data_dispersive(validate_optical_data(optical_data, shape=shape)):
            stack = Stack(optical_data, shape, nstep=kw["nstep"],
                          nin=N_COVER if kw["nin"] is None else kw["nin"],
                          nout=N_COVER if kw["nout"] is None else kw["nout"],
                          split_layers=kw["split_layers"])
    results = []
    for f, w, o in zip(fields, wavelengths, outs):
        f = f if isinstance(f, np.ndarray) else np.asarray(f)
        opt = stack if isinstance(stack, Stack) else \
            validate_optical_data(optical_data, shape=f.shape[-2:], wavelength=w)
        o1 = None if o is None else np.expand_dims(o, -4)
        r = _transfer_one((np.expand_dims(f, -4), np.asarray([w]), pixelsize), opt, out=o1,
                          **kw)[0]
        results.append((o if o is not None else np.squeeze(r, -4), w, pixelsize))
    return tuple(results)


def _transfer_one(field_data, optical_data, beta=None, phi=None, nin=None, nout=None,
                  npass=None, nstep=1, diffraction=None, reflection=None, method=None,
                  multiray=False, norm=DTMM_NORM_FFT, betamax=None, smooth=SMOOTH,
                  split_layers=False, eff_data=None, ret_bulk=False, out=None, precision=None,
                  tune=False):
    """One transfer of transfer_field on the GPU (dtmm/transfer.py:554-615)."""
    torch = _torch()
    job = _prepare(field_data, optical_data, beta, phi, nin, nout, npass, nstep, diffraction,
                   reflection, method, multiray, norm, betamax, smooth, split_layers, eff_data,
                   ret_bulk, out, precision)
    if tune:
        _tune.tune_plan(job.plan, job.stack, job.fin, job.fout, job.ws)
    status = torch.zeros(4, dtype=torch.int32, device="cuda")
    if ret_bulk:
        # (n_layers + 2, ...): input, the field after every layer, output
        # (dtmm/transfer.py:670-690, :876-877)
        nl = job.stack.n_layers
        bulk = torch.empty((nl,) + tuple(job.fin.shape), dtype=job.plan.dtype, device="cuda")
        job.plan.execute_bulk(job.stack, job.fin, job.fout, bulk, workspace=job.ws, status=status)
        torch.cuda.synchronize()
        _check_status(status)
        shp = job.field_np.shape
        res = np.empty((nl + 2,) + shp, dtype=np.complex64 if job.plan.precision == "c64"
                       else np.complex128)
        # the input side: with npass > 1 its forward part plus the backward
        # field of every backward pass (dtmm/transfer.py:1050, :1080)
        res[0] = job.fin.cpu().numpy().reshape(shp) if job.npass > 1 else \
            (job.field_np.host() if job.on_device else job.field_np)
        res[1:nl + 1] = bulk.cpu().numpy().reshape((nl,) + shp)
        res[nl + 1] = job.fout.cpu().numpy().reshape(shp)
        if out is not None:
            out[...] = res
            res = out
        return res, job.wavelengths, job.pixelsize
    job.plan.execute(job.stack, job.fin, job.fout, workspace=job.ws, status=status)
    torch.cuda.synchronize()
    _check_status(status)
    return job.finish()


def _prepare_samples(samples, kwargs):
    """_prepare for each sample of transfer_batch / TransferGraph."""
    jobs = []
    for smp in samples:
        if isinstance(smp, dict):
            opts = dict(kwargs)
            opts.update({k: v for k, v in smp.items() if k not in ("field_data", "optical_data")})
            fd, od = smp["field_data"], smp["optical_data"]
        else:
            fd, od = smp
            opts = kwargs
        for k in ("split_rays", "split_diffraction", "split_wavelengths"):
            opts = {kk: vv for kk, vv in opts.items() if kk != k}
        jobs.append(_prepare(fd, od, **opts))
    return jobs


def _job_array(jobs):
    """The dtmmx_job descriptors of prepared jobs (and the stack structs they
    point to, which must outlive the array)."""
    structs = [j.stack.c_struct() for j in jobs]
    arr = (_lib.Job * len(jobs))()
    for a, j, st in zip(arr, jobs, structs):
        a.plan = j.plan._h.value
        a.stack = ctypes.pointer(st)
        a.field_in = j.fin.data_ptr()
        a.field_out = j.fout.data_ptr()
        a.workspace = j.ws.data_ptr() if j.ws is not None else None
        a.workspace_bytes = j.ws.numel() if j.ws is not None else 0
    return arr, structs


def transfer_batch(samples, tune=False, **kwargs):
    """Transfers several independent samples of uneven grid, depth and beams
    in one batched engine call (``dtmmx_transfer_batch``).

    ``samples`` is a sequence of ``(field_data, optical_data)`` pairs or of
    dicts ``{"field_data": ..., "optical_data": ..., **per-sample options}``;
    ``kwargs`` are transfer_field options shared by all samples.  Returns
    the list of ``(field_out, wavelengths, pixelsize)``, each equal to what
    transfer_field returns for that sample (the reference users loop
    dtmm.transfer_field over their samples, dtmm/transfer.py:322).  ``tune``
    tunes each sample's launch shapes first (dtmm_amd.tune)."""
    torch = _require_gpu()
    lib = _lib.load()
    jobs = _prepare_samples(samples, kwargs)
    if not jobs:
        return []
    if tune:
        for j in jobs:
            _tune.tune_plan(j.plan, j.stack, j.fin, j.fout, j.ws)
    arr, structs = _job_array(jobs)
    status = torch.zeros(4, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    _lib.check(lib.dtmmx_transfer_batch(arr, len(jobs), status.data_ptr(), stream.cuda_stream))
    torch.cuda.synchronize()
    _check_status(status)
    return [j.finish() for j in jobs]


class TransferGraph:
    """A batch of transfers captured once as a HIP graph and replayed.

    The samples are prepared as by ``transfer_batch`` (plans, stacks, device
    fields, workspaces) and ``dtmmx_transfer_batch`` is captured into one
    ``torch.cuda.CUDAGraph``: a replay launches every sample's kernels with a
    single graph launch and no host work.  Inputs of the same shapes are
    replaced with ``set_inputs``; every replay starts from the stored inputs
    (samples with npass > 1 overwrite their input buffer with the input-side
    field, so the graph first restores it).  Replays are ordered on the
    stream they are launched from; ``results`` waits for them.  ``tune``
    tunes each sample's launch shapes before the capture (dtmm_amd.tune).
    """

    def __init__(self, samples, tune=False, **kwargs):
        torch = _require_gpu()
        self._lib = _lib.load()
        self.jobs = _prepare_samples(samples, kwargs)
        if not self.jobs:
            raise ValueError("no samples")
        if tune:
            for j in self.jobs:
                _tune.tune_plan(j.plan, j.stack, j.fin, j.fout, j.ws)
        self._arr, self._structs = _job_array(self.jobs)
        self._inputs = [j.fin.clone() if j.npass > 1 else None for j in self.jobs]
        self.status = torch.zeros(4, dtype=torch.int32, device="cuda")
        # one eager run: argument checks and lazy initialisation outside capture
        self._run(torch.cuda.current_stream())
        torch.cuda.synchronize()
        _check_status(self.status)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._run(torch.cuda.current_stream())
        torch.cuda.synchronize()

    def _run(self, stream):
        for j, x in zip(self.jobs, self._inputs):
            if x is not None:
                j.fin.copy_(x)
        _lib.check(self._lib.dtmmx_transfer_batch(self._arr, len(self.jobs), self.status.data_ptr(),
                                                  stream.cuda_stream))

    def set_inputs(self, fields):
        """New input fields, one per sample, each of the sample's original
        shape (NumPy arrays or tensors)."""
        torch = _torch()
        if len(fields) != len(self.jobs):
            raise ValueError("expected %d fields" % len(self.jobs))
        for i, (j, f) in enumerate(zip(self.jobs, fields)):
            t = torch.as_tensor(np.asarray(f) if not torch.is_tensor(f) else f)
            if tuple(t.shape) != tuple(j.field_np.shape):
                raise ValueError("field %d has shape %s, expected %s"
                                 % (i, tuple(t.shape), tuple(j.field_np.shape)))
            dst = self._inputs[i] if self._inputs[i] is not None else j.fin
            dst.copy_(t.reshape(dst.shape).to(device=dst.device, dtype=dst.dtype))

    def replay(self):
        self.graph.replay()

    def results(self):
        """The outputs of the last replay, as transfer_batch returns them."""
        torch = _torch()
        torch.cuda.synchronize()
        _check_status(self.status)
        return [j.finish() for j in self.jobs]
