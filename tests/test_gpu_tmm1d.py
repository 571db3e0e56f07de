"""GPU parity: the non-iterative 1D transfer matrix method
(dtmmx_tmm_transfer4x4, dtmm_amd.tmm) against the oracle's restatement of
dtmm/tmm.py:1112-1673 on the same inputs.

Tolerance: max|f - f_ref| / max|f_ref| < 1e-10 (fp64 throughout; the
eigen-solves differ in rounding only).
"""
import numpy as np
import pytest

import oracle as o
from helpers import flush, rel_err

TOL = 1e-10


def _stack(L, shape, rng, absorbing=False):
    """Per-element layers: isotropic, uniaxial and biaxial in turn."""
    eps = np.empty((L,) + shape + (3,))
    for i in range(L):
        base = rng.uniform(2.0, 2.6, shape + (3,))
        if i % 3 == 0:
            base[..., 1] = base[..., 0] = base[..., 2]
        elif i % 3 == 1:
            base[..., 1] = base[..., 0]
        eps[i] = base
    if absorbing:
        eps = eps + 1j * rng.uniform(0., 0.05, eps.shape)
    ang = rng.uniform(-1.5, 1.5, (L,) + shape + (3,))
    return eps, ang


@pytest.mark.gpu
@pytest.mark.parametrize("absorbing", [False, True])
@pytest.mark.parametrize("with_out", [False, True])
def test_transfer4x4_matches_oracle(absorbing, with_out):
    from dtmm_amd import tmm
    flush("case tmm1d absorbing=%s with_out=%s" % (absorbing, with_out))
    rng = np.random.default_rng(3)
    shape, L = (8, 16), 7
    eps, ang = _stack(L, shape, rng, absorbing)
    kd = rng.uniform(0.3, 3.0, (L,) + shape)
    beta, phi = 0.4, 0.9
    fin = rng.standard_normal(shape + (4,)) + 1j * rng.standard_normal(shape + (4,))
    fout0 = None
    if with_out:
        fout0 = rng.standard_normal(shape + (4,)) + 1j * rng.standard_normal(shape + (4,))
    ref_out, ref_in = o.transfer4x4(fin, kd, eps, ang, beta, phi, nin=1.3, nout=1.6,
                                    fvec_out=fout0)
    f_in = fin.copy()
    f_out = fout0.copy() if with_out else None
    got = tmm.transfer4x4(f_in, kd, eps, ang, beta, phi, nin=1.3, nout=1.6, fvec_out=f_out)
    e1, e2 = rel_err(got, ref_out), rel_err(f_in, ref_in)
    flush("  out err %.3e  in err %.3e" % (e1, e2))
    assert e1 < TOL and e2 < TOL


@pytest.mark.gpu
def test_stack_mat_and_wavelength_broadcast():
    """kd per wavelength (L, 3, 1) against per-pixel layers (L, 1, 10): the
    elements are the broadcast (3, 10)."""
    from dtmm_amd import tmm
    flush("case tmm1d broadcast")
    rng = np.random.default_rng(6)
    L = 5
    eps, ang = _stack(L, (1, 10), rng)
    d = rng.uniform(1., 4., L)
    ks = o.k0([450., 550., 650.], 100.)
    kd = d[:, None, None] * ks[None, :, None]
    ref = o.stack_mat(kd, eps, ang, 0.2, -0.7)
    got = tmm.stack_mat(kd, eps, ang, 0.2, -0.7)
    assert got.shape == ref.shape == (3, 10, 4, 4)
    assert rel_err(got, ref) < TOL
    f0 = o.fvec(o.normalize_f(o.f_iso(0.2, -0.7, 1.0)), jones=(1, 0))
    ref_out, _ = o.transfer4x4(f0, kd, eps, ang, 0.2, -0.7)
    got_out = tmm.transfer4x4(f0, kd, eps, ang, 0.2, -0.7)
    assert got_out.shape == (3, 10, 4)
    assert rel_err(got_out, ref_out) < TOL


@pytest.mark.gpu
def test_full_size_stacks_conserve_energy():
    """512 x 512 independent lossless stacks of 20 layers: R + T = 1 for every
    stack, and a sample of them matches the oracle.  Normal incidence: the
    reference's transfer4x4 builds the stack matrix at beta 0 whatever the
    beam of the media (dtmm/tmm.py:1655), so only there is it one physical
    beam."""
    from dtmm_amd import tmm
    flush("case tmm1d full size")
    rng = np.random.default_rng(12)
    shape, L = (512, 512), 20
    eps = np.broadcast_to(np.asarray([2.25, 2.25, 2.89]), (L,) + shape + (3,))
    ang = np.stack([np.zeros((L,) + shape), rng.uniform(0, np.pi, (L,) + shape),
                    rng.uniform(-np.pi, np.pi, (L,) + shape)], -1)
    kd = np.full(L, 1.3)
    beta, phi = 0., 0.
    f0 = o.fvec(o.normalize_f(o.f_iso(beta, phi, 1.0)), jones=(1, 0))
    f_in = np.broadcast_to(f0, shape + (4,)).copy()
    f_out = tmm.transfer4x4(f_in, kd, eps, ang, beta, phi, nin=1.0, nout=1.5)
    p0 = o.poynting(f0)
    R = -o.poynting(f_in - f0) / p0
    T = o.poynting(f_out) / p0
    err = np.abs(R + T - 1).max()
    flush("  max |R + T - 1| %.3e" % err)
    assert err < 1e-10
    idx = (slice(100, 103), slice(7, 11))
    ref_out, _ = o.transfer4x4(f0, kd, eps[(slice(None),) + idx], ang[(slice(None),) + idx],
                               beta, phi, nin=1.0, nout=1.5)
    assert rel_err(f_out[idx], ref_out) < TOL
