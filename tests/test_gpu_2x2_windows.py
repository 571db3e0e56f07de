"""GPU parity: the 2x2 method with the tent windows of diffraction > 1
(dtmm/propagate_2x2.py:212-257), reflection modes 0 and 1, npass 1, against
the complex128 oracle.

Tolerances as in test_gpu_transfer: 1e-9 (c128) and 1e-4 (c64) of
max|F - F_ref| / max|F_ref|."""
import os

import numpy as np
import pytest

import dtmm_amd as dm
import oracle
from helpers import droplet_case, flush, random_case, rel_err

TOL = {"c128": 1e-9, "c64": 1e-4}


def _compare(fd, optical, precision, **kw):
    field, wl, ps = fd
    ref = oracle.transfer_field((field, wl, ps), optical, method="2x2", **kw)[0]
    dt = np.complex64 if precision == "c64" else np.complex128
    flush("  GPU 2x2 windows %s field %s %s" % (precision, field.shape,
                                                  {k: v for k, v in kw.items()
                                                   if k not in ("beta", "phi")}))
    out = dm.transfer_field((field.astype(dt), wl, ps), optical, method="2x2",
                            precision=precision, **kw)[0]
    err = rel_err(out, ref)
    ierr = rel_err(dm.field2intensity(out), oracle.field2intensity(ref))
    flush("  rel field err %.3e  intensity err %.3e" % (err, ierr))
    assert err < TOL[precision], err
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n", [3, 5])
@pytest.mark.parametrize("reflection", [0, 1])
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_2x2_windows(n, reflection, precision):
    flush("case 2x2 windows n=%d reflection %d %s" % (n, reflection, precision))
    fd, optical = random_case(shape=(5, 32, 48), wavelengths=(450., 650.), jones=None,
                              beta=0.1, phi=0.4)
    _compare(fd, optical, precision, beta=0.1, phi=0.4, diffraction=n, reflection=reflection,
             nin=1.2, nout=1.4)


@pytest.mark.gpu
def test_2x2_windows_multi_beam_two_blocks():
    """Two beams, and two blocks with different effective media (the Fresnel
    interface between them in reflection mode 1)."""
    flush("case 2x2 windows multi-beam two blocks")
    beta, phi = [0.05, 0.15], [0.3, -2.0]
    fd, b1 = random_case(shape=(3, 32, 32), wavelengths=(500., 600.), jones=None, beta=beta,
                         phi=phi)
    _, b2 = droplet_case(shape=(3, 32, 32), radius=12, ne=1.75, nhost=1.55)
    _compare(fd, [b1[0], b2[0]], "c128", beta=beta, phi=phi, diffraction=3, reflection=1,
             nin=1.0, nout=1.33)


@pytest.mark.gpu
@pytest.mark.parametrize("reflection", [0, 1])
def test_2x2_windows_substeps(reflection):
    """nstep > 1 runs each window through the sub-steps (one launch per window)."""
    flush("case 2x2 windows nstep 2 reflection %d" % reflection)
    fd, optical = random_case(shape=(3, 24, 32), wavelengths=(550.,), jones=(1, 1j),
                              beta=0.1, phi=0.3)
    _compare(fd, optical, "c128", beta=0.1, phi=0.3, nstep=2, diffraction=3,
             reflection=reflection)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(2, 24, 512), (2, 16, 1024)])
def test_2x2_windows_long_rows(shape):
    """Rows of 512 (two rays per work-group) and 1024 points (one ray, two
    waves per line) in the fused window row kernel."""
    flush("case 2x2 windows rows %s" % (shape,))
    beta, phi = [0.1, 0.1], [0.0, 2.1]
    fd, optical = random_case(shape=shape, wavelengths=(500., 600.), jones=None, beta=beta,
                              phi=phi)
    _compare(fd, optical, "c64", beta=beta, phi=phi, diffraction=3, reflection=1, nin=1.52,
             nout=1.52)


@pytest.mark.gpu
def test_2x2_windows_fused_matches_per_window_launches():
    flush("case 2x2 windows fused vs per-window")
    fd, optical = random_case(shape=(3, 32, 48), wavelengths=(450., 650.), jones=None,
                              beta=0.1, phi=0.7)
    field, wl, ps = fd
    kw = dict(method="2x2", precision="c128", beta=0.1, phi=0.7, diffraction=4, reflection=1)
    out = dm.transfer_field((field.copy(), wl, ps), optical, **kw)[0]
    os.environ["DTMMX_FUSED_WINDOWS"] = "0"
    try:
        ref = dm.transfer_field((field.copy(), wl, ps), optical, **kw)[0]
    finally:
        os.environ.pop("DTMMX_FUSED_WINDOWS", None)
    err = rel_err(out, ref)
    flush("  fused vs per-window rel err %.3e" % err)
    assert err < 1e-12, err
