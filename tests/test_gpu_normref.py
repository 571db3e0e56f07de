"""GPU parity of the reference-field normalisation of the multi-pass 4x4
transfer (norm DTMM_NORM_REF "local" = 2 and DTMM_NORM_FFT | DTMM_NORM_REF
= 3, dtmm/transfer.py:103-138, :631-639, :769-775, :805-835) against the
complex128 oracle, and the cholesteric-droplet example
(examples/3D/cholesteric_droplet.py:41-47) at reduced size.

Tolerances (max|F - F_ref| / max|F_ref|): c128 1e-9, c64 1e-4.
"""
import numpy as np
import pytest

import dtmm_amd as dm
import oracle
from dtmm_amd.samples import random_director_data
from helpers import flush, rel_err

TOL = {"c128": 1e-9, "c64": 1e-4}


def _compare(fd, optical, precision, **kw):
    field, wl, ps = fd
    ref_in = np.array(field, np.complex128)
    ref = oracle.transfer_field((ref_in, wl, ps), optical, method="4x4", **kw)[0]
    gin = np.array(field, np.complex64 if precision == "c64" else np.complex128)
    flush("  GPU 4x4 %s field %s %s" % (kw, field.shape, precision))
    out = dm.transfer_field((gin, wl, ps), optical, method="4x4", precision=precision, **kw)[0]
    err, ein = rel_err(out, ref), rel_err(gin, ref_in)
    flush("  rel output err %.3e input side %.3e" % (err, ein))
    assert np.isfinite(err) and err < TOL[precision] and ein < TOL[precision], (err, ein)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("norm", [2, 3])
@pytest.mark.parametrize("reflection", [1, 2])
def test_random_stack(reflection, norm, precision):
    flush("case normref random refl %d norm %d %s" % (reflection, norm, precision))
    ny, nx = 24, 40
    optical = [random_director_data((3, ny, nx), seed=3, no=1.5, ne=1.7)]
    fd = dm.illumination_data((ny, nx), [450., 650.], pixelsize=100, beta=0.1, phi=0.5,
                              jones=None, window=dm.blackman((ny, nx)))
    _compare(fd, optical, precision, beta=0.1, phi=0.5, nin=1.2, nout=1.4, npass=3, norm=norm,
             reflection=reflection, diffraction=1)


@pytest.mark.gpu
@pytest.mark.parametrize("npass", [2, 4])
def test_no_diffraction_and_substeps(npass):
    flush("case normref diffraction 0 nstep 2 npass %d" % npass)
    ny, nx = 16, 32
    optical = [random_director_data((3, ny, nx), seed=8, no=1.5, ne=1.7)]
    fd = dm.illumination_data((ny, nx), [520.], pixelsize=100, beta=0.12, phi=-0.4, jones=None,
                              window=dm.blackman((ny, nx)))
    _compare(fd, optical, "c128", beta=0.12, phi=-0.4, nin=1.0, nout=1.3, npass=npass, norm=2,
             reflection=2, diffraction=0, nstep=2)


@pytest.mark.gpu
def test_example_cholesteric_droplet():
    """examples/3D/cholesteric_droplet.py:41-47: left-handed cholesteric
    droplet, beam at beta 0.3 focused into the sample from n = 1.5, 4x4,
    diffraction 1, reflection 2, npass 5, norm 2, betamax 0.9 (the
    example's configured default; 40x96x96 -> 12x32x32, 9 -> 3
    wavelengths)."""
    flush("case example cholesteric droplet")
    shape = (12, 32, 32)
    optical = [dm.cholesteric_droplet_data(shape, 8, 7, no=1.5, ne=1.65, nhost=1.5)]
    fd = dm.illumination_data(shape[1:], [380., 580., 780.], jones=None, beta=0.3,
                              pixelsize=50, n=1.5, focus=20, window=None, betamax=0.9)
    _compare(fd, optical, "c128", beta=0.3, phi=0, diffraction=1, smooth=0.1, reflection=2,
             nin=1.5, nout=1.5, npass=5, norm=2, betamax=0.9)
