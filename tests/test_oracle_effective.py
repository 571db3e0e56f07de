"""Pinning of the oracle's anisotropic effective media (eff_data 1, 2,
per-layer symmetries and custom effective blocks; dtmm/transfer.py:926-1005,
dtmm/data.py:1240-1419).  CPU only.

* uniaxial_order / eig_symmetry follow the reference's own doctest vectors
  (dtmm/data.py:1264-1271);
* the effective medium of a laterally uniform layer is the layer itself, so
  the diffraction step of a stack of uniform uniaxial layers is exact: with
  per-layer effective data the 4x4 transfer equals the exact per-mode
  propagation (diffraction -1, itself pinned by Berreman exponentials) for any
  mean beam, while the isotropic effective medium does not;
* rolling the material and the input along x and y rolls the output;
* a custom effective block equal to the isotropic mean reproduces the default.
"""
import numpy as np
import pytest

import oracle as o


def test_uniaxial_order_reference_vectors():
    np.testing.assert_allclose(o.uniaxial_order(0, np.array([1, 2, 3.])), (2, 2, 2))
    np.testing.assert_allclose(o.uniaxial_order(1, np.array([1, 2, 3.])), (1.5, 1.5, 3))
    np.testing.assert_allclose(o.uniaxial_order(-1, np.array([1, 2, 3.])), (1, 2, 3))
    np.testing.assert_allclose(o.uniaxial_order(0.5, np.array([1, 2, 3.])), (1.75, 1.75, 2.5))
    # eig_symmetry: 2 keeps the tensor biaxial
    np.testing.assert_allclose(o.eig_symmetry(2, np.array([1, 2, 3.])), (1, 2, 3))
    np.testing.assert_allclose(o.eig_symmetry([0, 1, 2], np.array([[1, 2, 3.]] * 3)),
                               [(2, 2, 2), (1.5, 1.5, 3), (1, 2, 3)])


def _uniform_stack(H, W):
    vals = [(2.25, 2.25, 2.89), (2.1, 2.1, 2.6), (2.3, 2.3, 2.7)]
    angs = [(0., 0.6, 0.3), (0., 1.1, -0.7), (0., 0.4, 2.0)]
    L = len(vals)
    eps = np.empty((L, H, W, 3))
    ang = np.empty((L, H, W, 3))
    for l in range(L):
        eps[l] = vals[l]
        ang[l] = angs[l]
    return np.array([1.3, 0.8, 2.1]), eps, ang


def _field(H, W, ks, seed, bmax=0.5):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((2, 4, H, W)) + 1j * rng.standard_normal((2, 4, H, W))
    b, _ = o.betaphi((H, W), ks)
    ff = np.fft.fft2(f)[:, None] * (b < bmax)[None, :, None]
    return np.fft.ifft2(ff)


def test_effective_block_of_uniform_layers():
    d, eps, ang = _uniform_stack(4, 5)
    for sym in ([1, 1, 1], [2, 2, 2], ("uniaxial",) * 3):
        de, ev, ea = o.effective_block((d, eps, ang), sym)
        np.testing.assert_allclose(de, d)
        np.testing.assert_allclose(o.epsva2eps(ev, ea), o.epsva2eps(eps[:, 0, 0], ang[:, 0, 0]),
                                   atol=1e-13)
    # one symmetry for the block: the tensor mean of all layers, uniaxial
    de, ev, ea = o.effective_block((d, eps, ang), 1)
    assert ev.shape == (3, 3) and np.all(ev[:, 0] == ev[:, 1])
    mean = o.epsva2eps(eps[:, 0, 0], ang[:, 0, 0]).mean(0)
    w, _ = np.linalg.eigh(np.array([[mean[0], mean[3], mean[4]], [mean[3], mean[1], mean[5]],
                                    [mean[4], mean[5], mean[2]]]))
    # the two closest eigenvalues are averaged, the distinct one is kept
    gaps = [w[1] - w[0], w[2] - w[1]]
    e3 = w[2] if gaps[0] < gaps[1] else w[0]
    np.testing.assert_allclose(ev[0, 2], e3, rtol=1e-12)
    np.testing.assert_allclose(ev[0, 0], (w.sum() - e3) / 2, rtol=1e-12)
    # symmetry 0: the isotropic mean, as the default
    _, ev0, ea0 = o.effective_block((d, eps, ang), 0)
    np.testing.assert_allclose(ev0[0], o.uniaxial_order(0., eps).mean((0, 1, 2)))
    assert np.all(ea0 == 0)
    with pytest.raises(ValueError):
        o.effective_block((d, eps, ang), [1, 1])
    with pytest.raises(ValueError):
        o.effective_block((d, eps, ang), "cubic")


@pytest.mark.parametrize("beta,phi", [(0., 0.), (0.2, 0.5)])
@pytest.mark.parametrize("sym", [[1, 1, 1], [2, 2, 2], [1, 2, 1]])
def test_uniform_uniaxial_layers_are_exact(beta, phi, sym):
    H = W = 16
    d, eps, ang = _uniform_stack(H, W)
    wl, p = np.array([500., 600.]), 100.
    ks = o.k0(wl, p)
    f = _field(H, W, ks, 1)
    kw = dict(beta=beta, phi=phi, method="4x4", nin=1.5, nout=1.5)
    ref, _, _ = o.transfer_field((f.copy(), wl, p), (d, eps, ang), diffraction=-1, **kw)
    got, _, _ = o.transfer_field((f.copy(), wl, p), (d, eps, ang), diffraction=1, reflection=1,
                                 eff_data=sym, **kw)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    iso, _, _ = o.transfer_field((f.copy(), wl, p), (d, eps, ang), diffraction=1, reflection=1,
                                 **kw)
    assert np.abs(iso - ref).max() > 1e-3 * np.abs(ref).max()


def test_uniform_layers_as_blocks():
    """One block per layer with the block symmetry 'uniaxial' is the same
    as per-layer symmetries of one block; a per-block list of identifiers
    too."""
    H = W = 12
    d, eps, ang = _uniform_stack(H, W)
    wl, p = np.array([550.]), 80.
    ks = o.k0(wl, p)
    f = _field(H, W, ks, 2)
    blocks = [(d[i:i + 1], eps[i:i + 1], ang[i:i + 1]) for i in range(3)]
    kw = dict(beta=0.1, phi=-0.3, method="4x4", nin=1.2, nout=1.4, diffraction=1, reflection=1)
    a, _, _ = o.transfer_field((f.copy(), wl, p), (d, eps, ang), eff_data=[1, 1, 1], **kw)
    b, _, _ = o.transfer_field((f.copy(), wl, p), blocks, eff_data="uniaxial", **kw)
    c, _, _ = o.transfer_field((f.copy(), wl, p), blocks, eff_data=[1, "uniaxial", 1], **kw)
    np.testing.assert_allclose(b, a, atol=1e-13 * np.abs(a).max())
    np.testing.assert_allclose(c, a, atol=1e-13 * np.abs(a).max())


def _random_stack(H, W, L, seed):
    rng = np.random.default_rng(seed)
    eps = np.broadcast_to([2.25, 2.25, 2.89], (L, H, W, 3)).copy()
    ang = np.zeros((L, H, W, 3))
    ang[..., 1] = rng.uniform(0, np.pi, (L, H, W))
    ang[..., 2] = rng.uniform(-np.pi, np.pi, (L, H, W))
    return rng.uniform(0.5, 1.5, L), eps, ang


@pytest.mark.parametrize("eff", [1, 2, [1, 2]])
def test_translation_equivariance(eff):
    H, W, L = 12, 10, 2
    d, eps, ang = _random_stack(H, W, L, 3)
    wl, p = np.array([520.]), 90.
    ks = o.k0(wl, p)
    f = _field(H, W, ks, 4)
    kw = dict(beta=0.15, phi=0.7, method="4x4", diffraction=1, reflection=1, eff_data=eff)
    out, _, _ = o.transfer_field((f.copy(), wl, p), (d, eps, ang), **kw)
    dy, dx = 5, 3
    r = lambda a, ax: np.roll(np.roll(a, dy, ax[0]), dx, ax[1])
    out2, _, _ = o.transfer_field((r(f, (-2, -1)), wl, p), (d, r(eps, (1, 2)), r(ang, (1, 2))), **kw)
    np.testing.assert_allclose(out2, r(out, (-2, -1)), atol=1e-13 * np.abs(out).max())


def test_custom_block_equal_to_default():
    H, W, L = 10, 10, 2
    d, eps, ang = _random_stack(H, W, L, 5)
    wl, p = np.array([500.]), 100.
    ks = o.k0(wl, p)
    f = _field(H, W, ks, 6)
    e = o.uniaxial_order(0., eps).mean((0, 1, 2))
    custom = (d, np.broadcast_to(e, (L, 3)), np.zeros((L, 3)))
    kw = dict(method="4x4", diffraction=1, reflection=1, beta=0., phi=0.)
    a, _, _ = o.transfer_field((f.copy(), wl, p), (d, eps, ang), **kw)
    b, _, _ = o.transfer_field((f.copy(), wl, p), (d, eps, ang), eff_data=custom, **kw)
    np.testing.assert_allclose(b, a, atol=1e-14 * np.abs(a).max())
    # a thicker effective layer changes the diffraction step only
    custom2 = (2 * d, custom[1], custom[2])
    c, _, _ = o.transfer_field((f.copy(), wl, p), (d, eps, ang), eff_data=custom2, **kw)
    assert np.abs(c - a).max() > 1e-6 * np.abs(a).max()
