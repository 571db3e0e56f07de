"""GPU parity of the layer-level steps on k-space fields
(dtmm/propagate_4x4.py:376 propagate_4x4_effective_1,
dtmm/propagate_2x2.py:193 propagate_2x2_effective_1) against the oracle's
restatement of the same functions, on random k-space fields through a
random uniaxial layer with an isotropic effective medium."""
import numpy as np
import pytest

import dtmm_amd as dm
import oracle
from helpers import flush, rel_err

TOL = {"c128": 1e-9, "c64": 1e-4}
DT = {"c128": np.complex128, "c64": np.complex64}


def _layer(shape, seed, d=1.3):
    rng = np.random.default_rng(seed)
    eo = rng.uniform(2.2, 2.3, shape)
    eps = np.stack([eo, eo, rng.uniform(2.6, 2.9, shape)], -1)
    ang = np.stack([np.zeros(shape), rng.uniform(0, np.pi, shape),
                    rng.uniform(-np.pi, np.pi, shape)], -1)
    eff = float(np.mean(eps.sum(-1) / 3))
    return (d, eps, ang), (d, np.array([eff] * 3), np.zeros(3))


def _kfield(shape, nc, seed, nlam=2):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((2, nlam, nc) + shape) + 1j * rng.standard_normal((2, nlam, nc) + shape)
    return np.fft.fft2(f)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("nsteps", [1, 2])
def test_propagate_4x4_effective_1(nsteps, precision):
    flush("case layer step 4x4 nsteps=%d %s" % (nsteps, precision))
    shape = (24, 32)
    ks = oracle.k0(np.array([480., 620.]), 100.)
    layer, eff = _layer(shape, 1)
    f = _kfield(shape, 4, 2)
    ref = oracle.propagate_4x4_effective_1(f.copy(), ks, layer, eff, 0.1, 0.4, nsteps, 1, 0.8)
    out = dm.propagate_4x4_effective_1(f.astype(DT[precision]), ks, layer, eff, beta=0.1,
                                       phi=0.4, nsteps=nsteps)
    err = rel_err(out, ref)
    flush("  rel err %.3e" % err)
    assert err < TOL[precision], err


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("reflection,nsteps", [(False, 1), (True, 2)])
def test_propagate_2x2_effective_1(reflection, nsteps, precision):
    flush("case layer step 2x2 reflection=%s nsteps=%d %s" % (reflection, nsteps, precision))
    shape = (32, 24)
    ks = oracle.k0(np.array([500., 650.]), 100.)
    layer, eff = _layer(shape, 3)
    eff_in = (0., np.array([1.2 ** 2] * 3), np.zeros(3))
    f = _kfield(shape, 2, 4)
    ref = oracle.propagate_2x2_effective_1(f.copy(), ks, layer, eff, 0.15, -0.7, nsteps, 1, 0.8,
                                           +1, None, None, 1 if reflection else 0, eff_in)
    out, refl = dm.propagate_2x2_effective_1(f.astype(DT[precision]), ks, None, layer, eff_in,
                                             eff, beta=0.15, phi=-0.7, nsteps=nsteps,
                                             reflection=reflection)
    assert refl is None
    err = rel_err(out, ref)
    flush("  rel err %.3e" % err)
    assert err < TOL[precision], err


@pytest.mark.gpu
def test_layer_step_device_tensor_stays_on_device():
    import torch
    flush("case layer step on a device tensor")
    shape = (16, 16)
    ks = oracle.k0(np.array([550.]), 100.)
    layer, eff = _layer(shape, 5, d=1.)
    f = _kfield(shape, 4, 6, nlam=1)
    t = torch.from_numpy(f).cuda()
    out = dm.propagate_4x4_effective_1(t, ks, layer, eff)
    assert out.is_cuda and out.shape == t.shape
    ref = dm.propagate_4x4_effective_1(f, ks, layer, eff)
    assert np.array_equal(out.cpu().numpy(), ref)
