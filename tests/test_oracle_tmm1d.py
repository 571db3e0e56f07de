"""Pinning of the oracle's non-iterative 1D transfer matrix method
(dtmm/tmm.py:1112-1673, method "4x4"; CPU only).

The reference's tests hold no vector for it, so it is pinned by:

* layer_mat = expm(-i kd Delta) of the layer's Berreman matrix (scipy expm,
  not the eigen-decomposition), rotated into the laboratory frame;
* a single interface: the Fresnel reflectances R_s, R_p at oblique
  incidence, and R + T = 1;
* an isotropic slab at normal incidence: the Airy transmission
  T = 1 / (1 + F sin^2(delta / 2));
* a quarter-wave stack (HL)^N on a substrate: R = ((n0 - Y) / (n0 + Y))^2 with
  Y = (nH / nL)^(2N) ns;
* energy: R + T = 1 for a random lossless anisotropic stack at oblique
  incidence, for both polarisations.
"""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle as o


def _rz(a):
    return np.array([[np.cos(a), -np.sin(a), 0.], [np.sin(a), np.cos(a), 0.], [0., 0., 1.]])


def _ry(a):
    return np.array([[np.cos(a), 0., np.sin(a)], [0., 1., 0.], [-np.sin(a), 0., np.cos(a)]])


def _berreman_exp(kd, eps, ang, beta, phi):
    R = _rz(ang[2] - phi) @ _ry(ang[1]) @ _rz(ang[0])
    et = R @ np.diag(eps) @ R.T
    e6 = np.array([et[0, 0], et[1, 1], et[2, 2], et[0, 1], et[0, 2], et[1, 2]])
    D = o.auxiliary_matrix(beta, e6)
    c, s = np.cos(phi), np.sin(phi)
    T = np.array([[c, 0, -s, 0], [0, c, 0, s], [s, 0, c, 0], [0, -s, 0, c]])
    return T @ expm(-1j * kd * D) @ np.linalg.inv(T)


@pytest.mark.parametrize("eps,ang", [((2.2, 2.2, 2.9), (0., 0.7, -1.1)),
                                     ((2.1, 2.4, 2.8), (0.4, 1.2, 0.3)),
                                     ((2.5, 2.5, 2.5), (0., 0., 0.))])
def test_layer_mat_is_berreman_exponential(eps, ang):
    beta, phi, kd = 0.3, 0.8, 2.7
    got = o.layer_mat(kd, np.asarray(eps), np.asarray(ang), beta, phi)
    np.testing.assert_allclose(got, _berreman_exp(kd, eps, ang, beta, phi), atol=1e-12)


def _incident(beta, phi, n, pol):
    f = o.f_iso(beta, phi, n)
    return o.fvec(o.normalize_f(f), jones=pol)


def _rt(fin_before, fin_after, fout):
    """Reflectance and transmittance from the Poynting z of the waves."""
    p_in = o.poynting(fin_before)
    refl = fin_after - fin_before
    return -o.poynting(refl) / p_in, o.poynting(fout) / p_in


@pytest.mark.parametrize("n1,n2,beta", [(1.0, 1.5, 0.0), (1.0, 1.5, 0.6), (1.5, 1.0, 0.4),
                                        (1.33, 1.7, 0.9)])
def test_single_interface_is_fresnel(n1, n2, beta):
    phi = 0.
    # a layer of the input medium: the characteristic matrix is the identity
    # up to its phase, so only the interface n1 -> n2 acts
    kd, eps, ang = [0.5], np.asarray([[n1 ** 2] * 3]), np.zeros((1, 3))
    c1 = np.sqrt(1 - (beta / n1) ** 2 + 0j)
    c2 = np.sqrt(1 - (beta / n2) ** 2 + 0j)
    rs = abs((n1 * c1 - n2 * c2) / (n1 * c1 + n2 * c2)) ** 2
    rp = abs((n2 * c1 - n1 * c2) / (n2 * c1 + n1 * c2)) ** 2
    for pol, r_exp in (((1, 0), rp), ((0, 1), rs)):
        f0 = _incident(beta, phi, n1, pol)
        fout, fin = o.transfer4x4(f0, kd, eps, ang, beta, phi, nin=n1, nout=n2)
        R, T = _rt(f0, fin, fout)
        np.testing.assert_allclose(R, r_exp, atol=1e-12)
        np.testing.assert_allclose(R + T, 1., atol=1e-12)


@pytest.mark.parametrize("n,kd", [(1.5, 0.3), (2.0, 1.7), (1.7, 5.0)])
def test_isotropic_slab_is_airy(n, kd):
    f0 = _incident(0., 0., 1.0, (1, 0))
    fout, fin = o.transfer4x4(f0, [kd], np.asarray([[n * n] * 3]), np.zeros((1, 3)),
                              nin=1.0, nout=1.0)
    R, T = _rt(f0, fin, fout)
    r = ((n - 1) / (n + 1)) ** 2
    F = 4 * r / (1 - r) ** 2
    delta = 2 * n * kd
    np.testing.assert_allclose(T, 1 / (1 + F * np.sin(delta / 2) ** 2), atol=1e-12)
    np.testing.assert_allclose(R, 1 - T, atol=1e-12)


@pytest.mark.parametrize("N", [1, 3, 6])
def test_quarter_wave_stack(N):
    nH, nL, ns, n0 = 2.3, 1.38, 1.52, 1.0
    kd = []
    eps = []
    for _ in range(N):
        for nn in (nH, nL):
            kd.append(np.pi / 2 / nn)  # quarter wave: n kd = pi / 2
            eps.append([nn * nn] * 3)
    f0 = _incident(0., 0., n0, (0, 1))
    fout, fin = o.transfer4x4(f0, kd, np.asarray(eps), np.zeros((2 * N, 3)), nin=n0, nout=ns)
    R, T = _rt(f0, fin, fout)
    Y = (nH / nL) ** (2 * N) * ns
    np.testing.assert_allclose(R, ((n0 - Y) / (n0 + Y)) ** 2, atol=1e-12)
    np.testing.assert_allclose(R + T, 1., atol=1e-12)


def test_lossless_anisotropic_stack_conserves_energy():
    rng = np.random.default_rng(2)
    L = 6
    eps = np.stack([rng.uniform(2.0, 2.4, L), rng.uniform(2.0, 2.4, L),
                    rng.uniform(2.5, 3.0, L)], -1)
    eps[::3, 1] = eps[::3, 0]  # some uniaxial layers
    ang = rng.uniform(-1.5, 1.5, (L, 3))
    kd = rng.uniform(0.5, 4.0, L)
    beta, phi = 0.45, 0.7
    for pol in ((1, 0), (0, 1), (1, 1j)):
        f0 = _incident(beta, phi, 1.2, np.asarray(pol) / np.linalg.norm(pol))
        fout, fin = o.transfer4x4(f0, kd, eps, ang, beta, phi, nin=1.2, nout=1.6)
        R, T = _rt(f0, fin, fout)
        assert 0. < T < 1.
        np.testing.assert_allclose(R + T, 1., atol=1e-11)


def test_stack_mat_is_ordered_product():
    rng = np.random.default_rng(8)
    L = 4
    eps = rng.uniform(2.0, 2.8, (L, 3))
    ang = rng.uniform(-1, 1, (L, 3))
    kd = rng.uniform(0.5, 2.0, L)
    got = o.stack_mat(kd, eps, ang, 0.2, 0.4)
    want = np.eye(4)
    for i in range(L):
        want = want @ _berreman_exp(kd[i], eps[i], ang[i], 0.2, 0.4)
    np.testing.assert_allclose(got, want, atol=1e-11)
