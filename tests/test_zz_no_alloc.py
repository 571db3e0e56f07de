"""The no-allocation transfer (dtmmx_transfer_noalloc) against dtmmx_transfer.

dtmmx_transfer_noalloc runs the same kernels with their workspace in the
library's static per-GPU scratch array instead of caller memory.  Two calls
in a row, the second on a smaller grid (the array keeps the first call's
contents), must each equal dtmmx_transfer's output bit for bit; a plan whose
workspace exceeds the array is refused before any work.
"""
import ctypes

import numpy as np
import pytest

import dtmm_amd as dm
from dtmm_amd import _lib
from dtmm_amd.transfer import NOALLOC_SCRATCH_BYTES, _prepare
from helpers import droplet_case, flush, random_case


def _both(fd, optical, **kw):
    import torch
    field, wl, ps = fd
    dt = np.complex64 if kw.get("precision") == "c64" else np.complex128
    job = _prepare((field.astype(dt), wl, ps), optical, **kw)
    assert 0 < job.plan.workspace_bytes <= NOALLOC_SCRATCH_BYTES
    ref = torch.empty_like(job.fout)
    job.plan.execute(job.stack, job.fin, ref, workspace=job.ws)
    torch.cuda.synchronize()
    out = torch.full_like(job.fout, float("nan"))
    job.plan.execute_noalloc(job.stack, job.fin, out)
    torch.cuda.synchronize()
    return out.cpu().numpy(), ref.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_matches_transfer_twice_second_smaller():
    flush("case noalloc: 96x96 4x4 c64 with a biaxial-table workspace, then 48x64 2x2 c128")
    fd, optical = random_case(shape=(4, 96, 96), wavelengths=(450., 650.), jones=None)
    d, eps, ang = optical[0]
    eps = eps.copy()
    eps[1, 5, 7, 1] += 0.05       # one biaxial voxel: the layer uses the workspace table
    a, b = _both(fd, [(d, eps, ang)], beta=0., phi=0., method="4x4", precision="c64")
    assert np.isfinite(a).all() and np.array_equal(a, b)
    fd2, optical2 = droplet_case(shape=(5, 48, 64), radius=14, wavelengths=(550.,), jones=None)
    a2, b2 = _both(fd2, optical2, beta=0., phi=0., method="2x2", precision="c128")
    assert np.isfinite(a2).all() and np.array_equal(a2, b2)


@pytest.mark.gpu
def test_noalloc_refuses_a_workspace_beyond_the_scratch():
    import torch
    flush("case noalloc refusal")
    shape = (512, 512)
    ks = dm.k0(np.linspace(400., 700., 8), 100.)
    plan = dm.Plan(shape, ks, np.zeros(1), np.zeros(1), 2, 2, "4x4", "c64", npass=3)
    assert plan.workspace_bytes > NOALLOC_SCRATCH_BYTES
    fd, optical = random_case(shape=(2, 16, 16), wavelengths=(550.,), jones=(1, 0))
    stack = dm.Stack(optical, (16, 16))
    f = torch.zeros((2, 8, 4) + shape, dtype=torch.complex64, device="cuda")
    g = torch.zeros_like(f)
    with pytest.raises(_lib.DtmmxError, match="exceeds the no-allocation scratch"):
        plan.execute_noalloc(stack, f, g)
