"""GPU parity: optical data of a director field built on the device
(dtmmx_director2data, Stack.from_director) against the host builder
dtmm_amd.director2data (dtmm/data.py:277-320), and the transfer through the
device-built stack against the transfer of the host-built optical data.

Tolerances: material and angles 1e-14 (device sqrt/atan2 against the host
libm); the effective medium 1e-13; transfer fields (c128) 1e-11.
"""
import numpy as np
import pytest

import dtmm_amd as dm
import oracle
from helpers import flush, rel_err


def _director(shape, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal(shape + (3,))
    d *= rng.uniform(0.2, 1.0, shape + (1,)) / np.linalg.norm(d, axis=-1, keepdims=True)
    d[0, 0, 0] = 0.  # a zero director: isotropic voxel
    d[0, 0, 1] = (0., 0., 1.)  # unit director along z
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("masked", [False, True])
def test_director2data_matches_host_builder(masked):
    import torch
    from dtmm_amd.transfer import Stack
    flush("case director2data masked=%s" % masked)
    shape = (5, 24, 40)
    d = _director(shape, 1)
    mask = None
    if masked:
        mask = np.random.default_rng(2).random(shape) > 0.3
    kw = dict(no=1.5, ne=1.7, nhost=1.45, scale_factor=0.9)
    th, eps, ang = dm.director2data(d, mask=mask, **kw)
    st = Stack.from_director(torch.from_numpy(d).cuda(), mask=mask, **kw)
    eps_g = np.moveaxis(st.eps_re.cpu().numpy(), 0, -1)
    ang_g = np.moveaxis(st.angles.cpu().numpy(), 0, -1)
    assert np.abs(eps_g - eps).max() < 1e-14
    assert np.abs(ang_g - ang).max() < 1e-14
    # closed forms: unit director along z (unmasked) -> (no^2, no^2, (ne / ...)^2 order 1/0.9)
    if not masked or mask[0, 0, 1]:
        o = 1. / 0.9
        m = (2 * 1.5 ** 2 + 1.7 ** 2) / 3.
        dl = 1.7 ** 2 - 1.5 ** 2
        np.testing.assert_allclose(eps_g[0, 0, 1], [m - o * dl / 3, m - o * dl / 3,
                                                     m + 2 * o * dl / 3], rtol=1e-15)
        np.testing.assert_allclose(ang_g[0, 0, 1], [0., 0., 0.], atol=0)
    eff = oracle.uniaxial_order(0., eps).mean(axis=(0, 1, 2))[0]
    np.testing.assert_allclose(st.eps_eff_re, eff, rtol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("split_layers", [False, True])
def test_transfer_through_device_stack(split_layers):
    import torch
    from dtmm_amd.transfer import Stack
    flush("case transfer device stack split=%s" % split_layers)
    shape = (6, 32, 32)
    d = _director(shape, 3)
    mask = np.random.default_rng(4).random(shape) > 0.2
    kw = dict(no=1.5, ne=1.65, nhost=1.5)
    optical = [dm.director2data(d, mask=mask, **kw)]
    fd = dm.illumination_data((32, 32), [500., 620.], pixelsize=100, jones=None,
                              window=dm.blackman((32, 32)))
    ref = dm.transfer_field(fd, optical, beta=0., phi=0., method="4x4", precision="c128",
                            split_layers=split_layers)[0]
    st = Stack.from_director(torch.from_numpy(d).cuda(), mask=torch.from_numpy(mask).cuda(),
                             split_layers=split_layers, **kw)
    got = dm.transfer_field(fd, st, beta=0., phi=0., method="4x4", precision="c128")[0]
    err = rel_err(got, ref)
    flush("  rel err %.3e" % err)
    assert err < 1e-11
