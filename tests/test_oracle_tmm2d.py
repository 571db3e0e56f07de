"""Pinning of the oracle's restatement of dtmm/tmm2d.py (CPU only).

The reference's tests hold no vector for tmm2d, so the 2D method is pinned by
what it reduces to:
* a laterally uniform stack couples no modes: the block matrices are block
  diagonal, each block the 1D characteristic matrix of the layer at that
  mode's beam (pinned in test_oracle_tmm1d by Berreman exponentials), and the
  transfer of every mode is the 1D transmit4x4 at its beam;
* rolling the material and the field along x rolls the output (every
  operator is per pixel or per mode);
* the block products are the products of the flattened 4n x 4n matrices;
* field2modes1 / modes2field1 are inverse on band-limited fields.
"""
import numpy as np
import pytest

import oracle as o


def _fields(rng, W, nlam):
    f = rng.standard_normal((2, nlam, 4, W)) + 1j * rng.standard_normal((2, nlam, 4, W))
    ks = o.k0(np.linspace(450, 650, nlam), 50.)
    mask = o.eigenmask1(W, ks)
    # band-limit to the modes the method keeps
    ff = np.fft.fft(f)
    for i in range(nlam):
        ff[:, i][..., ~mask[i]] = 0.
    return np.fft.ifft(ff), ks


def test_block_products_are_flat_matrix_products():
    rng = np.random.default_rng(1)
    n = 5
    a = rng.standard_normal((n, n, 4, 4)) + 1j * rng.standard_normal((n, n, 4, 4))
    b = rng.standard_normal((n, n, 4, 4)) + 1j * rng.standard_normal((n, n, 4, 4))
    dg = rng.standard_normal((n, 4, 4)) + 1j * rng.standard_normal((n, 4, 4))

    def flat(m):
        return np.moveaxis(m, -2, -3).reshape(4 * n, 4 * n)

    D = np.zeros((4 * n, 4 * n), complex)
    for i in range(n):
        D[4 * i:4 * i + 4, 4 * i:4 * i + 4] = dg[i]
    np.testing.assert_allclose(flat(o.bdotmm(a, b)), flat(a) @ flat(b), atol=1e-12)
    np.testing.assert_allclose(flat(o.bdotmd(a, dg)), flat(a) @ D, atol=1e-12)
    np.testing.assert_allclose(flat(o.bdotdm(dg, a)), D @ flat(a), atol=1e-12)


def test_modes_round_trip():
    rng = np.random.default_rng(2)
    f, ks = _fields(rng, 36, 3)
    mask, modes = o.field2modes1(f, ks)
    np.testing.assert_allclose(o.modes2field1(mask, modes), f, atol=1e-12)


def _uniform_stack(W):
    eps = np.stack([np.broadcast_to([2.3, 2.3, 2.9], (W, 3)),
                    np.broadcast_to([2.1, 2.1, 2.1], (W, 3)),
                    np.broadcast_to([2.2, 2.5, 2.8], (W, 3))])
    ang = np.stack([np.broadcast_to([0., 0.8, 0.4], (W, 3)), np.zeros((W, 3)),
                    np.broadcast_to([0.3, 1.1, -0.6], (W, 3))])
    return np.array([0.7, 1.9, 1.2]), eps.astype(complex), ang


@pytest.mark.parametrize("method", ["4x4", "4x4_1"])
@pytest.mark.parametrize("betay", [0., 0.15])
def test_uniform_stack_is_the_1d_method_per_mode(method, betay):
    """No lateral variation: every mode goes through the 1D stack at its own
    beam (beta_x of the mode, betay) and nothing couples modes."""
    W = 24
    d, eps, ang = _uniform_stack(W)
    rng = np.random.default_rng(3)
    f, ks = _fields(rng, W, 2)
    f0 = f.copy()
    out, _, _ = o.transfer2d((f, np.linspace(450, 650, 2), 50.), (d, eps, ang), betay=betay,
                             nin=1.2, nout=1.5, method=method)
    mask, modes_in = o.field2modes1(f0, ks)
    _, modes_out = o.field2modes1(out, ks)
    _, modes_back = o.field2modes1(f, ks)
    for i, k in enumerate(ks):
        bx = o.betax1(W, k)[mask[i]]
        for j, b in enumerate(bx):
            beta, phi = np.hypot(b, betay), np.arctan2(betay, b)
            cm = o.stack_mat(k * d, eps[:, 0], ang[:, 0], beta, phi, method=method)
            fo, fi = o.transmit4x4(modes_in[i][:, j], cm, o.f_iso(beta, phi, 1.2),
                                   o.f_iso(beta, phi, 1.5))
            np.testing.assert_allclose(modes_out[i][:, j], fo, atol=1e-10 * np.abs(fo).max())
            np.testing.assert_allclose(modes_back[i][:, j], fi, atol=1e-10 * np.abs(fi).max())


def test_transfer2d_is_translation_equivariant():
    W, L = 30, 3
    rng = np.random.default_rng(4)
    e0 = rng.uniform(2.1, 2.4, (L, W))
    eps = np.stack([e0, e0, rng.uniform(2.6, 2.9, (L, W))], -1).astype(complex)
    ang = np.stack([np.zeros((L, W)), rng.uniform(0, np.pi, (L, W)),
                    rng.uniform(-np.pi, np.pi, (L, W))], -1)
    d = rng.uniform(0.5, 2., L)
    f, ks = _fields(rng, W, 2)
    wl = np.linspace(450, 650, 2)
    dx = 7
    out = o.transfer2d((f.copy(), wl, 50.), (d, eps, ang), nin=1.3, nout=1.6)[0]
    out_r = o.transfer2d((np.roll(f, dx, -1), wl, 50.),
                         (d, np.roll(eps, dx, 1), np.roll(ang, dx, 1)), nin=1.3, nout=1.6)[0]
    np.testing.assert_allclose(out_r, np.roll(out, dx, -1), atol=1e-11 * np.abs(out).max())
    out_m = o.transfer2d((np.roll(f, dx, -1), wl, 50.), (d, eps, ang), nin=1.3, nout=1.6)[0]
    assert np.abs(out_m - np.roll(out, dx, -1)).max() > 1e-3 * np.abs(out).max()
