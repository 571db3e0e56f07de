"""Pinning of the oracle's eigen-solution and linear algebra (CPU only).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails: the Berreman
eigen-equation Delta.F = F.diag(alpha) holds only for the right alpha and
mode vectors; the Poynting signs fix the +/- interleaving; brute force pins
the batched products; the reference's own continuity tests
(dtmm/test/test_tmm.py:45-62) are restated.
"""
import json
import os

import numpy as np
import pytest

import oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))


def test_betaphi_betaxy_golden_vectors():
    g = json.load(open(os.path.join(HERE, "golden", "wave_vectors.json")))
    shape, ks = tuple(g["shape"]), g["ks"]
    b, p = o.betaphi(shape, ks)
    np.testing.assert_allclose(b, g["beta"], rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(p, g["phi"], rtol=1e-5, atol=1e-8)
    bx, by = o.betaxy(shape, ks)
    np.testing.assert_allclose(bx, g["betax"], rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(by, g["betay"], rtol=1e-5, atol=1e-8)
    b0, _ = o.betaphi(shape, ks[0])
    np.testing.assert_allclose(b0, np.asarray(g["beta"])[0], rtol=1e-5, atol=1e-8)


def _eps6_uniaxial(eo, ee, theta, phi):
    """R diag(eo, eo, ee) R^T written out from the optic axis n:
    eps = eo I + (ee - eo) n n^T."""
    n = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)
    d = (ee - eo)[..., None, None]
    E = eo[..., None, None] * np.eye(3) + d * n[..., :, None] * n[..., None, :]
    return np.stack([E[..., 0, 0], E[..., 1, 1], E[..., 2, 2], E[..., 0, 1], E[..., 0, 2],
                     E[..., 1, 2]], -1)


@pytest.mark.parametrize("beta0", [False, True])
def test_uniaxial_modes_solve_berreman_equation(beta0):
    rng = np.random.default_rng(1)
    n = 400
    beta = np.zeros(n) if beta0 else rng.uniform(0.01, 1.2, n)
    eo = rng.uniform(1.8, 2.6, n) + 0j
    ee = rng.uniform(1.8, 3.2, n) + 0j
    theta = rng.uniform(0, np.pi, n)
    phi = rng.uniform(-np.pi, np.pi, n)
    eps = np.stack([eo, eo, ee], -1)
    ang = np.stack([np.zeros(n), theta, phi], -1)
    alpha, F = o.alphaf(beta, 0., eps, ang)
    D = o.auxiliary_matrix(beta, _eps6_uniaxial(eo, ee, theta, phi))
    np.testing.assert_allclose(D @ F, F * alpha[:, None, :], atol=1e-12)
    # unit 2-norm columns (dtmm/tmm.py:278-288)
    np.testing.assert_allclose(np.linalg.norm(F, axis=-2), 1., atol=1e-12)
    # forward modes in columns 0, 2 and backward in 1, 3 (propagating waves)
    p = o.poynting(np.swapaxes(F, -1, -2))
    prop = np.all(np.abs(alpha.imag) < 1e-12, -1)
    assert np.all(p[prop][:, 0::2] > 0) and np.all(p[prop][:, 1::2] < 0)


def test_isotropic_modes_closed_form_and_equation():
    rng = np.random.default_rng(2)
    n = 200
    beta = rng.uniform(0, 2.0, n)
    e = rng.uniform(1.0, 3.0, n) + 0j
    eps = np.stack([e, e, e], -1)
    alpha, F = o.alphaf(beta, 0., eps, np.zeros(3))
    a = np.sqrt(e - beta ** 2)
    np.testing.assert_allclose(alpha, np.stack([a, -a, a, -a], -1), atol=1e-14)
    e6 = np.stack([e, e, e, 0 * e, 0 * e, 0 * e], -1)
    D = o.auxiliary_matrix(beta, e6)
    np.testing.assert_allclose(D @ F, F * alpha[:, None, :], atol=1e-12)
    # the reference leaves the isotropic columns unnormalised: F[0, 0] = 1/2
    np.testing.assert_allclose(F[:, 0, 0], 0.5)


def test_biaxial_modes_solve_berreman_equation_and_sort():
    rng = np.random.default_rng(3)
    n = 100
    beta = rng.uniform(0, 0.8, n)
    eps = np.stack([rng.uniform(2, 2.2, n), rng.uniform(2.3, 2.5, n),
                    rng.uniform(2.6, 3, n)], -1) + 0j
    ang = np.stack([rng.uniform(0, 3, n), rng.uniform(0, 3, n), rng.uniform(-3, 3, n)], -1)
    alpha, F = o.alphaf(beta, 0., eps, ang)
    # Euler rotation R = Rz(phi) . Ry(theta) . Rz(psi) built from elementary rotations
    def rz(a):
        c, s, z, one = np.cos(a), np.sin(a), np.zeros_like(a), np.ones_like(a)
        return np.stack([np.stack([c, -s, z], -1), np.stack([s, c, z], -1),
                         np.stack([z, z, one], -1)], -2)

    def ry(a):
        c, s, z, one = np.cos(a), np.sin(a), np.zeros_like(a), np.ones_like(a)
        return np.stack([np.stack([c, z, s], -1), np.stack([z, one, z], -1),
                         np.stack([-s, z, c], -1)], -2)

    R = rz(ang[:, 2]) @ ry(ang[:, 1]) @ rz(ang[:, 0])
    E = R @ (eps[:, :, None] * np.swapaxes(R, -1, -2))
    e6 = np.stack([E[:, 0, 0], E[:, 1, 1], E[:, 2, 2], E[:, 0, 1], E[:, 0, 2], E[:, 1, 2]], -1)
    D = o.auxiliary_matrix(beta, e6)
    np.testing.assert_allclose(D @ F, F * alpha[:, None, :], atol=1e-10)
    p = o.poynting(np.swapaxes(F, -1, -2))
    assert np.all(p[:, 0::2] >= 0) and np.all(p[:, 1::2] < 0)


def test_beam_azimuth_rotates_rows():
    """alphaf at beam azimuth phi0 equals the phi0 = 0 solution with the
    material azimuth shifted by -phi0 and (Ex, Ey), (Hy, -Hx) rotated."""
    rng = np.random.default_rng(4)
    n = 50
    beta = rng.uniform(0, 0.9, n)
    phi0 = rng.uniform(-3, 3, n)
    eps = np.stack([np.full(n, 2.25), np.full(n, 2.25), np.full(n, 2.89)], -1) + 0j
    ang = np.stack([np.zeros(n), rng.uniform(0, 3, n), rng.uniform(-3, 3, n)], -1)
    a1, F1 = o.alphaf(beta, phi0, eps, ang)
    ang2 = ang.copy()
    ang2[:, 2] -= phi0
    a0, F0 = o.alphaf(beta, 0., eps, ang2)
    np.testing.assert_allclose(a1, a0, atol=1e-13)
    c, s = np.cos(phi0)[:, None], np.sin(phi0)[:, None]
    np.testing.assert_allclose(F1[:, 0], c * F0[:, 0] - s * F0[:, 2], atol=1e-13)
    np.testing.assert_allclose(F1[:, 2], s * F0[:, 0] + c * F0[:, 2], atol=1e-13)
    np.testing.assert_allclose(F1[:, 1], c * F0[:, 1] + s * F0[:, 3], atol=1e-13)
    np.testing.assert_allclose(F1[:, 3], -s * F0[:, 1] + c * F0[:, 3], atol=1e-13)


def test_alphaf_continuity_reference_cases():
    """Restatement of dtmm/test/test_tmm.py:45-62."""
    for e in (1, 2):
        a0, f0 = o.alphaf(0, 0, (e, e, e), (0, 0, 0))
        a, f = o.alphaf(0.00001, 0.0001, (e, e, e), (0, 0, 0))
        np.testing.assert_allclose(np.round(f, 3), np.round(f0, 3), rtol=1e-5, atol=1e-8)
        a, f = o.alphaf(0., 0, (e, e, e + 0.0001), (0, 0, 0))
        np.testing.assert_allclose(np.round(o.normalize_f(f), 3), np.round(o.normalize_f(f0), 3),
                                   rtol=1e-5, atol=1e-8)
        a, f = o.alphaf(0.0001, 0, (e, e, e + 0.0001), (0, 0, 0))
        np.testing.assert_allclose(np.round(o.normalize_f(f), 3), np.round(o.normalize_f(f0), 3),
                                   rtol=1e-5, atol=1e-8)
        a0, f0 = o.alphaf(0, np.pi / 4, (e, e, e), (0, 0, 0))
        a, f = o.alphaf(0.0001, np.pi / 4, (e, e, e + 0.0001), (0, 0, 0))
        np.testing.assert_allclose(np.round(o.normalize_f(f), 3), np.round(o.normalize_f(f0), 3),
                                   rtol=1e-5, atol=1e-8)


def test_inv_brute_force_and_singular():
    rng = np.random.default_rng(5)
    m = rng.standard_normal((20, 4, 4)) + 1j * rng.standard_normal((20, 4, 4))
    m[3] = 0.
    m[7, :, 1] = 0.
    r = o.inv(m)
    ok = [i for i in range(20) if i not in (3, 7)]
    np.testing.assert_allclose(r[ok], np.linalg.inv(m[ok]), atol=1e-12)
    assert np.all(r[3] == 0) and np.all(r[7] == 0)
    m2 = m[:, :2, :2].copy()
    np.testing.assert_allclose(o.inv(m2)[ok], np.linalg.inv(m2[ok]), atol=1e-12)


def test_dotmf_dotmdmf_brute_force():
    """Restatement of dtmm/test/test_linalg.py:52-65 (multi_dot per pixel)."""
    rng = np.random.default_rng(6)
    ni, nj = 5, 6
    a = rng.standard_normal((ni, nj, 4, 4)) + 1j * rng.standard_normal((ni, nj, 4, 4))
    b = rng.standard_normal((ni, nj, 4, 4)) + 1j * rng.standard_normal((ni, nj, 4, 4))
    d = np.exp(1j * 2.3 * rng.standard_normal((ni, nj, 4)))
    f = rng.standard_normal((4, ni, nj)) + 1j * rng.standard_normal((4, ni, nj))
    out1 = o.dotmf(a, f)
    out2 = o.dotmdmf(a, d, b, f)
    for i in range(ni):
        for j in range(nj):
            np.testing.assert_allclose(out1[:, i, j], a[i, j] @ f[:, i, j], atol=1e-12)
            np.testing.assert_allclose(out2[:, i, j],
                                       np.linalg.multi_dot([a[i, j], np.diag(d[i, j]), b[i, j],
                                                            f[:, i, j]]), atol=1e-12)


def test_isotropic_epn_matrices_are_diagonal_phases():
    """For an isotropic effective medium the 4x4 diffraction chain between
    two identical layers, D1 . D2, is diag(exp(+-i kd (a(k) - a_c))) in the
    (E+x, E-x, E+y, E-y) basis, zero where beta >= betamax."""
    shape = (8, 12)
    ks = o.k0([500., 600.], 100.)
    eps = np.array([2.3, 2.3, 2.3])
    beta, phi, d = np.array(0.2)[None, None, None], np.array(0.4)[None, None, None], 1.0
    d1 = o.first_corrected_Epn_diffraction_matrix(shape, ks, beta, phi, d / 2, eps, np.zeros(3))
    d2 = o.second_corrected_Epn_diffraction_matrix(shape, ks, beta, phi, d / 2, eps, np.zeros(3))
    K = d1 @ d2
    b, _ = o.betaphi(shape, ks)
    ak = np.sqrt(2.3 - b ** 2 + 0j)
    ac = np.sqrt(2.3 - 0.2 ** 2)
    ph = np.exp(1j * ks[:, None, None] * d * (ak - ac))
    ph[b >= o.BETAMAX] = 0.
    expect = np.zeros(K.shape, complex)
    expect[..., 0, 0] = ph
    expect[..., 2, 2] = ph
    expect[..., 1, 1] = np.where(b >= o.BETAMAX, 0., 1. / np.where(b >= o.BETAMAX, 1., ph))
    expect[..., 3, 3] = expect[..., 1, 1]
    np.testing.assert_allclose(K, expect, atol=1e-12)
