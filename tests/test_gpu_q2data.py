"""GPU parity: optical data of a Q-tensor field built on the device
(dtmmx_q2data, Stack.from_Q; dtmm/data.py:322-373) against the oracle.

The eigenvectors agree up to sign, so the angles are compared through the
rebuilt tensor R diag(eps) R^T (1e-13) and the eigenvalues directly (1e-13);
the transfer through a uniaxial Q stack matches the transfer through the
director stack of the same director (c128, 1e-11).
"""
import numpy as np
import pytest

import dtmm_amd as dm
import oracle as o
from helpers import flush, rel_err


def _q(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape + (3, 3)) * 0.3
    m = (a + np.swapaxes(a, -1, -2)) / 2
    m -= np.trace(m, axis1=-2, axis2=-1)[..., None, None] * np.eye(3) / 3
    q = np.stack([m[..., 0, 0], m[..., 1, 1], m[..., 2, 2], m[..., 0, 1], m[..., 0, 2],
                  m[..., 1, 2]], -1)
    q[0, 0, 0] = 0.                                         # isotropic voxel
    q[0, 0, 1] = (-1. / 3, -1. / 3, 2. / 3, 0., 0., 0.)     # director along z
    q[0, 0, 2] = (2. / 3, -1. / 3, -1. / 3, 0., 0., 0.)     # diagonal, distinct axis x
    # diagonal tensors keep their input order (no ascending sort first)
    q[0, 0, 3] = (0.3, 0.2, -0.5, 0., 0., 0.)
    q[0, 0, 4] = (0.1, 0.05, 0.25, 1e-9, 0., 0.)
    return q


@pytest.mark.gpu
@pytest.mark.parametrize("biaxial", [False, True])
@pytest.mark.parametrize("masked", [False, True])
def test_q2data_matches_oracle(biaxial, masked):
    from dtmm_amd.transfer import Stack
    flush("case q2data biaxial=%s masked=%s" % (biaxial, masked))
    shape = (4, 24, 40)
    q = _q(shape, 5)
    mask = np.random.default_rng(6).random(shape) > 0.3 if masked else None
    kw = dict(no=1.5, ne=1.7, nhost=1.45, scale_factor=0.9)
    _, epsv, epsa = o.Q2data(q, mask=mask, biaxial=biaxial, **kw)
    st = Stack.from_Q(q, mask=mask, biaxial=biaxial, **kw)
    eps_g = np.moveaxis(st.eps_re.cpu().numpy(), 0, -1)
    ang_g = np.moveaxis(st.angles.cpu().numpy(), 0, -1)
    assert np.abs(eps_g - epsv).max() < 1e-13
    if biaxial:
        keep = np.ones(shape, bool) if mask is None else mask.copy()
        # the near-diagonal voxel counts as diagonal: its rebuild is the
        # diagonal part of the tensor (the 1e-9 off-diagonal term is dropped)
        keep[0, 0, 4] = False
        assert np.abs(o.epsva2eps(eps_g, ang_g)[keep] - o.Q2eps(q, 1.5, 1.7, 0.9)[keep]).max() < 1e-13
        if mask is None or mask[0, 0, 4]:
            diag = o.Q2eps(q[0, 0, 4], 1.5, 1.7, 0.9)
            diag[3:] = 0.
            assert np.abs(o.epsva2eps(eps_g[0, 0, 4], ang_g[0, 0, 4]) - diag).max() < 1e-13
    else:
        np.testing.assert_allclose(o.epsva2eps(eps_g, ang_g), o.epsva2eps(epsv, epsa), atol=1e-13)
    if mask is None:
        np.testing.assert_allclose(ang_g[0, 0, 2], [np.pi / 2, np.pi / 2, 0.], atol=1e-15)
        np.testing.assert_array_equal(ang_g[0, 0, 3:5], 0.)
        if biaxial:
            e = o.Q2eps(q[0, 0, 3:5], 1.5, 1.7, 0.9)[..., :3]
            np.testing.assert_allclose(eps_g[0, 0, 3:5], e, rtol=1e-15)
    np.testing.assert_allclose(st.eps_eff_re, epsv.sum(-1).mean() / 3., rtol=1e-13)


@pytest.mark.gpu
def test_transfer_through_q_stack():
    import torch
    from dtmm_amd.transfer import Stack
    flush("case transfer q stack")
    shape = (5, 32, 32)
    rng = np.random.default_rng(7)
    d = rng.standard_normal(shape + (3,))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    kw = dict(no=1.5, ne=1.65)
    fd = dm.illumination_data((32, 32), [500., 620.], pixelsize=100, jones=None,
                              window=dm.blackman((32, 32)))
    ref = dm.transfer_field(fd, Stack.from_director(d, **kw), beta=0., phi=0., method="4x4",
                            precision="c128")[0]
    st = Stack.from_Q(torch.from_numpy(o.director2Q(d)).cuda(), **kw)
    got = dm.transfer_field(fd, st, beta=0., phi=0., method="4x4", precision="c128")[0]
    err = rel_err(got, ref)
    flush("  rel err %.3e" % err)
    assert err < 1e-11
