"""Pinning of the oracle's exact per-mode propagation (diffraction < 0 or
inf; dtmm/propagate_4x4.py:515-564, transfer.py:782-803), CPU only.

* Laterally uniform layers: every FFT mode is an eigen-wave of the stack, so
  the result is, per mode with beta < betamax, the ordered product of the
  layers' Berreman exponentials expm(i k0 d Delta(beta)) in the mode's
  plane of incidence (scipy expm, not an eigen-decomposition), and zero for
  the other modes.  A wrong mode angle, a transposed index or a missing
  betamax mask fails it.
* The sub-step thickness: transfer_4x4 hands the layer's t / nstep to the
  propagator, which divides by nstep again, so nstep = 2 propagates as one
  step of t / 2 (the reference's arithmetic, restated).
"""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle as o


def _rz(a):
    return np.array([[np.cos(a), -np.sin(a), 0.], [np.sin(a), np.cos(a), 0.], [0., 0., 1.]])


def _ry(a):
    return np.array([[np.cos(a), 0., np.sin(a)], [0., 1., 0.], [-np.sin(a), 0., np.cos(a)]])


def _mode_matrix(k, beta, phi, d, eps, ang):
    """Ordered product of the layers' Berreman exponentials at (beta, phi)."""
    M = np.eye(4, dtype=complex)
    c, s = np.cos(phi), np.sin(phi)
    T = np.array([[c, 0, -s, 0], [0, c, 0, s], [s, 0, c, 0], [0, -s, 0, c]])
    for j in range(len(d)):
        R = _rz(ang[j, 2] - phi) @ _ry(ang[j, 1]) @ _rz(ang[j, 0])
        et = R @ np.diag(eps[j]) @ R.T
        e6 = np.array([et[0, 0], et[1, 1], et[2, 2], et[0, 1], et[0, 2], et[1, 2]])
        D = o.auxiliary_matrix(beta, e6)
        M = T @ expm(1j * k * d[j] * D) @ np.linalg.inv(T) @ M
    return M


@pytest.mark.parametrize("betamax", [0.8, 10.])
def test_full_uniform_stack_is_per_mode_berreman_exponential(betamax):
    shape = (4, 6)
    wl = np.array([450., 620.])
    ps = 100.
    d = np.array([1.5, 0.7])
    eps = np.array([[2.2, 2.2, 2.9], [2.4, 2.4, 2.4]])
    ang = np.array([[0., 0.6, -0.9], [0., 0., 0.]])
    rng = np.random.default_rng(2)
    field = rng.normal(size=(2, 4) + shape) + 1j * rng.normal(size=(2, 4) + shape)
    block = (d, np.broadcast_to(eps[:, None, None], (2, 1, 1, 3)),
             np.broadcast_to(ang[:, None, None], (2, 1, 1, 3)))
    out = o.transfer_field((field.copy(), wl, ps), [block], beta=0., phi=0., method="4x4",
                           diffraction=-1, betamax=betamax, nin=1.2, nout=1.4)[0]
    inf = o.transfer_field((field.copy(), wl, ps), [block], beta=0., phi=0., method="4x4",
                           diffraction=np.inf, betamax=betamax)[0]
    np.testing.assert_array_equal(out, inf)
    ks = o.k0(wl, ps)
    for i in range(len(wl)):
        b, p = o.betaphi(shape, ks[i])
        ff = np.fft.fft2(field[i])
        res = np.zeros_like(ff)
        for a in range(shape[0]):
            for c in range(shape[1]):
                if b[a, c] < betamax:
                    res[:, a, c] = _mode_matrix(ks[i], b[a, c], p[a, c], d, eps, ang) @ ff[:, a, c]
        np.testing.assert_allclose(out[i], np.fft.ifft2(res), atol=1e-11 * np.abs(field).max())


def test_full_substeps_divide_the_thickness_twice():
    shape = (4, 4)
    wl = np.array([550.])
    rng = np.random.default_rng(3)
    field = rng.normal(size=(1, 4) + shape) + 1j * rng.normal(size=(1, 4) + shape)
    eps = rng.uniform(2.2, 2.9, (1,) + shape + (3,))
    eps[..., 1] = eps[..., 0]
    ang = np.stack([np.zeros((1,) + shape), rng.uniform(0, 3, (1,) + shape),
                    rng.uniform(-3, 3, (1,) + shape)], -1)
    two = o.transfer_field((field, wl, 100.), [(np.array([2.]), eps, ang)], beta=0., phi=0.,
                           method="4x4", diffraction=-1, nstep=2)[0]
    # the same two steps as two one-step layers of t / 4 = 0.5
    quarter = o.transfer_field((field, wl, 100.), [(np.array([0.5, 0.5]),
                                                    np.concatenate([eps, eps]),
                                                    np.concatenate([ang, ang]))],
                               beta=0., phi=0., method="4x4", diffraction=-1, nstep=1)[0]
    np.testing.assert_allclose(two, quarter, atol=1e-12)
    half = o.transfer_field((field, wl, 100.), [(np.array([1.]), eps, ang)], beta=0., phi=0.,
                            method="4x4", diffraction=-1, nstep=1)[0]
    assert np.abs(two - half).max() > 1e-3 * np.abs(half).max()
