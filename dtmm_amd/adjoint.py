"""Gradients of a loss on the transmitted field with respect to the director
angles of every voxel, by the adjoint of the 4x4 transfer.

The reference (IJSComplexMatter/dtmm) has no gradients.  This module runs
the 4x4 transfer of ``dtmm_amd.transfer_field`` (reflection mode 1,
diffraction 1, npass 1) forward while recording the row-pass inputs, then
the adjoint pass in the engine (``dtmmx_transfer_record`` /
``dtmmx_transfer_adjoint``), which returns dL/dtheta and dL/dphi for the
polar and azimuthal director angles (optical-data angles[..., 1] and
angles[..., 2]) of every voxel.

Convention: for a real loss L of the output field F, ``seed`` is
dL/dconj(F), so that dL = 2 Re <seed, dF>.

The tape keeps the row-pass input of every `checkpoint`-th propagation step
(default 8); the adjoint recomputes the steps between two checkpoints, which
bounds the tape at ceil(steps / checkpoint) fields plus checkpoint - 1
fields of workspace.  ``AdjointTransfer.apply(theta, phi)`` is the same
transfer as a ``torch.autograd.Function`` of device tensors of the director
angles.
"""
import ctypes

import numpy as np

from . import _lib
from .transfer import DTMM_NORM_FFT, SMOOTH, _prepare, _require_gpu, _check_status


class AdjointTransfer:
    """Forward transfer with a tape, then gradients from a seed.

    >>> tr = AdjointTransfer(field_data, optical_data, beta=0., phi=0.)
    >>> out = tr.forward()                 # (n_rays, n_wavelengths, 4, H, W)
    >>> loss, seed = analyser_intensity_loss(out, (1, 1), n=1.5)
    >>> grad = tr.backward(seed)           # (2, n_layers, H, W) float64
    """

    def __init__(self, field_data, optical_data, beta=0., phi=0., nin=None, nout=None, nstep=1,
                 betamax=None, precision=None, split_layers=False, checkpoint=8):
        self.torch = _require_gpu()
        self.lib = _lib.load()
        if int(checkpoint) < 1:
            raise ValueError("checkpoint must be >= 1")
        self.checkpoint = int(checkpoint)
        self.job = _prepare(field_data, optical_data, beta=beta, phi=phi, nin=nin, nout=nout,
                            npass=1, nstep=nstep, diffraction=1, reflection=1, method="4x4",
                            norm=DTMM_NORM_FFT, betamax=betamax, smooth=SMOOTH,
                            split_layers=split_layers, precision=precision,
                            checkpoint=self.checkpoint)
        self._stack = self.job.stack.c_struct()
        tb = int(self.lib.dtmmx_tape_size(self.job.plan._h, ctypes.byref(self._stack)))
        self.tape = self.torch.empty(max(tb, 1), dtype=self.torch.uint8, device="cuda")
        self.tape_bytes = tb
        self.status = self.torch.zeros(4, dtype=self.torch.int32, device="cuda")
        self.ws_bytes = int(self.lib.dtmmx_adjoint_workspace_size(self.job.plan._h))
        self.ws = self.torch.empty(max(self.ws_bytes, 1), dtype=self.torch.uint8, device="cuda")
        self._recorded = False
        self._version = 0  # forward runs; a tape serves the backward of its own run only

    @property
    def n_layers(self):
        return self.job.stack.n_layers

    def set_angles(self, theta, phi):
        """Copies device tensors (n_layers, H, W) of the polar and azimuthal
        director angles into the stack (angles[1], angles[2]); the effective
        media do not depend on them."""
        ang = self.job.stack.angles
        shape = tuple(ang.shape[1:])
        for t in (theta, phi):
            if tuple(t.shape) != shape:
                raise ValueError("angles must have shape %s, got %s" % (shape, tuple(t.shape)))
        ang[1].copy_(theta)
        ang[2].copy_(phi)

    def apply(self, theta, phi):
        """The transfer as a differentiable function of the director angles:
        returns the output field (n_rays, n_wavelengths, 4, H, W) on the
        device; backward gives dL/dtheta and dL/dphi in the dtype of the
        inputs.  Each call runs the transfer (and records its tape); the
        backward of a call must run before the next call."""
        return _DirectorTransfer.apply(theta, phi, self)

    def _record(self, stream):
        self._version += 1
        j = self.job
        _lib.check(self.lib.dtmmx_transfer_record(
            j.plan._h, ctypes.byref(self._stack), j.fin.data_ptr(), j.fout.data_ptr(),
            self.tape.data_ptr(), self.tape_bytes, self.status.data_ptr(), stream.cuda_stream))

    def _adjoint(self, seed, grad, stream):
        j = self.job
        _lib.check(self.lib.dtmmx_transfer_adjoint(
            j.plan._h, ctypes.byref(self._stack), self.tape.data_ptr(), seed.data_ptr(),
            grad.data_ptr(), self.ws.data_ptr(), self.ws_bytes, self.status.data_ptr(),
            stream.cuda_stream))

    def forward(self):
        """Runs the transfer, keeps the tape; returns the output field tensor
        (n_rays, n_wavelengths, 4, H, W) on the device."""
        torch = self.torch
        self._record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        _check_status(self.status, adjoint=True)
        self._recorded = True
        return self.job.fout

    def backward(self, seed, grad=None):
        """dL/dtheta and dL/dphi, (2, n_layers, H, W) float64 on the device,
        from seed = dL/dconj(field_out); added to `grad` when given."""
        torch = self.torch
        if not self._recorded:
            raise RuntimeError("forward() must run before backward()")
        j = self.job
        seed = seed.to(device="cuda", dtype=j.fout.dtype).contiguous()
        if tuple(seed.shape) != tuple(j.fout.shape):
            raise ValueError("seed shape %s does not match the field %s"
                             % (tuple(seed.shape), tuple(j.fout.shape)))
        H, W = j.plan.shape
        if grad is None:
            grad = torch.zeros((2, self.n_layers, H, W), dtype=torch.float64, device="cuda")
        self._adjoint(seed, grad, torch.cuda.current_stream())
        torch.cuda.synchronize()
        _check_status(self.status, adjoint=True)
        return grad

    def gradient_graph(self, loss_fn):
        """Captures forward transfer + loss + adjoint in one HIP graph (the
        per-layer launches are replayed without host work).  loss_fn(field)
        must return (loss, seed) as device tensors without host
        synchronisation.  Returns replay() -> (loss, grad) device tensors,
        overwritten by every replay."""
        torch = self.torch
        H, W = self.job.plan.shape
        grad = torch.zeros((2, self.n_layers, H, W), dtype=torch.float64, device="cuda")
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):  # warm-up outside the capture
            self._record(side)
            _, seed = loss_fn(self.job.fout)
            self._adjoint(seed, grad, side)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        _check_status(self.status, adjoint=True)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            stream = torch.cuda.current_stream()
            self._record(stream)
            loss, seed = loss_fn(self.job.fout)
            grad.zero_()
            self._adjoint(seed, grad, stream)
        self._recorded = True

        def replay():
            graph.replay()
            return loss, grad

        replay.graph = graph
        return replay


def _director_transfer_class():
    import torch

    class DirectorTransfer(torch.autograd.Function):
        """field_out = transfer(theta, phi) with the adjoint as backward.
        torch's gradient of a complex output is 2 dL/dconj(F), so the seed
        is half of it."""

        @staticmethod
        def forward(ctx, theta, phi, tr):
            tr.set_angles(theta.detach(), phi.detach())
            tr._record(torch.cuda.current_stream())
            tr._recorded = True
            ctx.tr = tr
            ctx.version = tr._version
            ctx.dtypes = (theta.dtype, phi.dtype)
            return tr.job.fout.clone()

        @staticmethod
        def backward(ctx, grad_out):
            tr = ctx.tr
            if tr._version != ctx.version:
                raise RuntimeError("the transfer ran again before this backward; the tape "
                                   "holds the later run")
            j = tr.job
            seed = (0.5 * grad_out).to(dtype=j.fout.dtype).contiguous()
            H, W = j.plan.shape
            grad = torch.zeros((2, tr.n_layers, H, W), dtype=torch.float64, device="cuda")
            tr._adjoint(seed, grad, torch.cuda.current_stream())
            _check_status(tr.status, adjoint=True)
            return grad[0].to(ctx.dtypes[0]), grad[1].to(ctx.dtypes[1]), None

    return DirectorTransfer


class _LazyFunction:
    """Defers the torch import to the first use."""
    _cls = None

    def apply(self, *args):
        if _LazyFunction._cls is None:
            _LazyFunction._cls = _director_transfer_class()
        return _LazyFunction._cls.apply(*args)


_DirectorTransfer = _LazyFunction()


def _iso_fmat_normal(n):
    """Normalised field matrix of an isotropic medium at normal incidence
    (columns (Ex, Hy, Ey, Hx) of the +x, -x, +y, -y modes, unit 2-norm;
    dtmm/tmm.py:124-154 with normalize_f, tmm.py:1874)."""
    f = np.zeros((4, 4), np.complex128)
    f[0, 0] = f[0, 1] = 1.
    f[1, 0], f[1, 1] = n, -n
    f[2, 2] = f[2, 3] = 1.
    f[3, 2], f[3, 3] = -n, n
    return f / np.linalg.norm(f, axis=0, keepdims=True)


def polarizer4x4(jones, n=1.):
    """4x4 polarizer of a normalised Jones vector for fields in an isotropic
    medium of index n at normal incidence: F diag-block(P, conj P) F^-1
    with P = j j^H (dtmm/jones4.py:41-101, dtmm/jones.py:236-278, :371-397)."""
    j = np.asarray(jones, np.complex128)
    j = j / np.linalg.norm(j)
    p = np.outer(j, j.conj())
    m = np.zeros((4, 4), np.complex128)
    m[0::2, 0::2] = p
    m[1::2, 1::2] = p.conj()
    f = _iso_fmat_normal(n)
    return f @ m @ np.linalg.inv(f)


# Poynting z of a field sample as a Hermitian form: I(g) = g^H J g
_J = 0.5 * np.array([[0, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, -1], [0, 0, -1, 0]], np.complex128)


def analyser_intensity_loss(field, jones=(1, 1), n=1.5, roi=None, as_tensor=False):
    """Loss = intensity after an analyser (polarizer4x4 of `jones` in a
    medium of index n) summed over the region of interest (y0, y1, x0, x1;
    default: the central quarter in each direction) and over all planes.
    Returns (loss, seed) with seed = dL/dconj(field) (a device tensor of
    the field's shape and dtype); loss is a float, or a 0-d device tensor
    with as_tensor=True (no host synchronisation, for graph capture)."""
    import torch
    H, W = field.shape[-2:]
    if roi is None:
        roi = (H // 2 - H // 8, H // 2 + H // 8, W // 2 - W // 8, W // 2 + W // 8)
    y0, y1, x0, x1 = roi
    m = polarizer4x4(jones, n)
    q = m.conj().T @ _J @ m
    fr = field.to(torch.complex128)
    sub = fr[..., :, y0:y1, x0:x1]

    def apply(a):  # per-pixel 4x4 product, written out
        return torch.stack([sum(complex(a[i, j]) * sub[..., j, :, :] for j in range(4))
                            for i in range(4)], dim=-3)

    g = apply(m)
    inten = (g[..., 0, :, :] * g[..., 1, :, :].conj()).real - \
        (g[..., 2, :, :] * g[..., 3, :, :].conj()).real
    loss = inten.sum()
    if not as_tensor:
        loss = float(loss.item())
    seed = torch.zeros_like(fr)
    seed[..., :, y0:y1, x0:x1] = apply(q)
    return loss, seed.to(field.dtype)


__all__ = ["AdjointTransfer", "polarizer4x4", "analyser_intensity_loss"]
