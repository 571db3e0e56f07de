"""Fused window row kernel (diffraction n > 1, single-step layers): every
tent window of a layer in one launch, the windows summed in real space and
transformed once per line (dtmm/propagate_4x4.py:390-419).

The oracle sums the windows in the reference's order; the fused kernel sums
them in tent-row order and before the forward x-transform, so results agree
to rounding.  The per-window launches (DTMMX_FUSED_WINDOWS=0) are the same
arithmetic one window at a time."""
import contextlib
import os

import numpy as np
import pytest

import dtmm_amd as dm
import oracle
from helpers import droplet_case, flush, random_case, rel_err

TOL = {"c128": 1e-9, "c64": 1e-4}


@contextlib.contextmanager
def _env(**kw):
    old = {k: os.environ.get(k) for k in kw}
    os.environ.update({k: str(v) for k, v in kw.items()})
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _run(fd, optical, precision, **kw):
    field, wl, ps = fd
    dt = np.complex64 if precision == "c64" else np.complex128
    f = np.array(field, dtype=dt)
    return dm.transfer_field((f, wl, ps), optical, method="4x4", precision=precision, **kw)[0], f


@pytest.mark.gpu
@pytest.mark.parametrize("shape,precision", [((3, 32, 48), "c128"), ((2, 24, 512), "c64"),
                                             ((2, 16, 1024), "c64")])
def test_fused_windows_match_per_window_launches(shape, precision):
    flush("case fused windows vs per-window %s %s" % (shape, precision))
    beta, phi = [0.1, 0.1], [0.3, 2.4]
    fd, optical = random_case(shape=shape, wavelengths=(450., 550., 650.), jones=None,
                              beta=beta, phi=phi)
    kw = dict(beta=beta, phi=phi, diffraction=4, nin=1.52, nout=1.52)
    out, _ = _run(fd, optical, precision, **kw)
    with _env(DTMMX_FUSED_WINDOWS=0):
        ref, _ = _run(fd, optical, precision, **kw)
    err = rel_err(out, ref)
    flush("  fused vs per-window rel err %.3e" % err)
    assert err < (1e-12 if precision == "c128" else 1e-5), err


@pytest.mark.gpu
@pytest.mark.parametrize("chunk", [1, 2])
def test_fused_windows_wavelength_chunks_are_bit_identical(chunk):
    """The wavelengths a work-group takes do not change any wavelength's
    arithmetic."""
    flush("case fused windows chunk %d" % chunk)
    fd, optical = random_case(shape=(2, 32, 64), wavelengths=(420., 480., 540., 600., 660.),
                              jones=None, beta=0.1, phi=0.7)
    kw = dict(beta=0.1, phi=0.7, diffraction=3)
    out, _ = _run(fd, optical, "c64", **kw)
    with _env(DTMMX_WIN_C=chunk):
        out_c, _ = _run(fd, optical, "c64", **kw)
    assert np.array_equal(out, out_c)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_fused_windows_droplet_grid_against_oracle(precision):
    """The large-with-reflections physics on a 64 x 128 crop: radial droplet
    between n = 1.52 glass, three beams at beta 0.1 and phi 0/120/240 degrees,
    both polarisations, diffraction 5 and npass 3."""
    flush("case fused windows droplet grid %s" % precision)
    beta = [0.1, 0.1, 0.1]
    phi = [0., 2 * np.pi / 3, 4 * np.pi / 3]
    fd, optical = droplet_case(shape=(4, 64, 128), radius=24, wavelengths=(450., 650.),
                               jones=None, beta=beta, phi=phi)
    kw = dict(beta=beta, phi=phi, diffraction=5, npass=3, nin=1.52, nout=1.52)
    field, wl, ps = fd
    f_ref = np.array(field, dtype=np.complex128)
    ref = oracle.transfer_field((f_ref, wl, ps), optical, method="4x4", **kw)[0]
    out, f = _run(fd, optical, precision, **kw)
    err, rerr = rel_err(out, ref), rel_err(f, f_ref)
    flush("  rel field err %.3e  input-side err %.3e" % (err, rerr))
    assert err < TOL[precision], err
    assert rerr < TOL[precision], rerr
