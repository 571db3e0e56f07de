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


def _oblique(f0, kd, eps, ang, beta, phi, nin, nout):
    """Transmission through the stack at the beam (beta, phi): the stack
    matrix at that beam and the boundary solve between the isotropic media.
    (transfer4x4 itself builds its stack matrix at normal incidence, as the
    reference does; see test_transfer4x4_builds_the_stack_at_normal_incidence.)"""
    cmat = o.stack_mat(kd, eps, ang, beta, phi)
    return o.transmit4x4(f0, cmat, o.f_iso(beta, phi, nin), o.f_iso(beta, phi, nout))


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
        fout, fin = _oblique(f0, kd, eps, ang, beta, phi, n1, n2)
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
        fout, fin = _oblique(f0, kd, eps, ang, beta, phi, 1.2, 1.6)
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


# ---------------------------------------------------------------------------
# The other 1D methods of dtmm/tmm.py:1112-1748 and transfer's quirk
# ---------------------------------------------------------------------------

def test_transfer4x4_builds_the_stack_at_normal_incidence():
    """dtmm/tmm.py:1655 calls stack_mat(kd, epsv, epsa, method=method) with
    its default beam, while the media use beta and phi."""
    rng = np.random.default_rng(21)
    eps = rng.uniform(2.0, 2.6, (3, 3))
    ang = rng.uniform(-1, 1, (3, 3))
    kd = rng.uniform(0.5, 2.0, 3)
    f0 = _incident(0.3, 0.4, 1.1, (1, 0))
    got = o.transfer4x4(f0, kd, eps, ang, 0.3, 0.4, nin=1.1, nout=1.3)
    want = o.transmit4x4(f0, o.stack_mat(kd, eps, ang), o.f_iso(0.3, 0.4, 1.1),
                         o.f_iso(0.3, 0.4, 1.3))
    np.testing.assert_allclose(got[0], want[0], atol=1e-14)
    np.testing.assert_allclose(got[1], want[1], atol=1e-14)


@pytest.mark.parametrize("n1,n,n2,kd", [(1.0, 1.5, 1.3, 0.7), (1.2, 2.0, 1.0, 2.9)])
def test_4x4_1_slab_is_single_pass(n1, n, n2, kd):
    """"4x4_1" zeroes the backward phases in the layer: no wave travels back
    inside it, so the slab transmits t01 exp(i n kd) t12 (no multiple
    reflections) and reflects r01 (normal incidence)."""
    for pol, comp in (((1, 0), 0), ((0, 1), 2)):
        f0 = o.fvec(o.normalize_f(o.f_iso(0, 0, n1)), jones=pol)
        fout, fin = o.transfer4x4(f0, [kd], np.asarray([[n * n] * 3]), np.zeros((1, 3)),
                                  nin=n1, nout=n2, method="4x4_1")
        t01, t12, r01 = 2 * n1 / (n1 + n), 2 * n / (n + n2), (n1 - n) / (n1 + n)
        np.testing.assert_allclose(fout[comp], f0[comp] * t01 * t12 * np.exp(1j * n * kd),
                                   atol=1e-14)
        np.testing.assert_allclose((fin - f0)[comp], f0[comp] * r01, atol=1e-14)


def test_4x4_2_damping_limits():
    """"4x4_2" adds 2i cfact Re(alpha) to the backward modes: cfact 0 is the
    "4x4" stack, and a strong damping kills the backward waves as "4x4_1"."""
    rng = np.random.default_rng(22)
    eps = rng.uniform(2.0, 2.6, (4, 3))
    ang = rng.uniform(-1, 1, (4, 3))
    kd = rng.uniform(0.5, 2.0, 4)
    s4 = o.stack_mat(kd, eps, ang, 0.2, 0.5)
    np.testing.assert_allclose(o.stack_mat(kd, eps, ang, 0.2, 0.5, cfact=0., method="4x4_2"),
                               s4, atol=1e-14)
    np.testing.assert_allclose(o.stack_mat(kd, eps, ang, 0.2, 0.5, cfact=200., method="4x4_2"),
                               o.stack_mat(kd, eps, ang, 0.2, 0.5, method="4x4_1"), atol=1e-12)


@pytest.mark.parametrize("cfact", [0., 0.05])
def test_4x4_r_block_structure(cfact):
    """layer_mat "4x4_r" is [[F P2 F^-1, 0], [F P F^-1, 0]] (P2 the "4x4_2"
    phases, P the plain ones), so the chained 8x8 matrix is [[A, 0], [B, 0]]
    with A the "4x4_2" stack and B the first layer's plain matrix times the
    "4x4_2" stack of the rest."""
    rng = np.random.default_rng(23)
    eps = rng.uniform(2.0, 2.6, (4, 3))
    ang = rng.uniform(-1, 1, (4, 3))
    kd = rng.uniform(0.5, 2.0, 4)
    c8 = o.stack_mat(kd, eps, ang, 0.3, -0.2, cfact=cfact, method="4x4_r")
    A = o.stack_mat(kd, eps, ang, 0.3, -0.2, cfact=cfact, method="4x4_2")
    B = o.layer_mat(kd[0], eps[0], ang[0], 0.3, -0.2) @ o.stack_mat(
        kd[1:], eps[1:], ang[1:], 0.3, -0.2, cfact=cfact, method="4x4_2")
    np.testing.assert_allclose(c8[0:4, 0:4], A, atol=1e-12)
    np.testing.assert_allclose(c8[4:8, 0:4], B, atol=1e-12)
    assert np.abs(c8[:, 4:8]).max() < 1e-13
    if cfact == 0.:
        np.testing.assert_allclose(B, A, atol=1e-12)


def _rot(phi):
    c, s = np.cos(phi), np.sin(phi)
    return np.array([[c, -s], [s, c]])


def test_2x2_stack_is_ordered_jones_product():
    """"2x2" at normal incidence: each in-plane uniaxial layer is the Jones
    retarder R diag(exp(i ne kd), exp(i no kd)) R^T, and the stack is their
    product with the first layer acting first (M_n ... M_1)."""
    no, ne, L = 1.5, 1.7, 5
    phis = np.linspace(0, np.pi / 2, L)
    kd = np.linspace(0.6, 1.4, L)
    eps = np.broadcast_to([no ** 2, no ** 2, ne ** 2], (L, 3))
    ang = np.zeros((L, 3))
    ang[:, 1] = np.pi / 2
    ang[:, 2] = phis
    J = np.eye(2, dtype=complex)
    for j in range(L):
        J = _rot(phis[j]) @ np.diag([np.exp(1j * kd[j] * ne), np.exp(1j * kd[j] * no)]) \
            @ _rot(phis[j]).T @ J
    np.testing.assert_allclose(o.stack_mat(kd, eps, ang, method="2x2"), J, atol=1e-13)
    # "2x2_1" adds the interface transmissions between the twisted layers
    # (the e and o waves of one layer meet different impedances in the next)
    assert np.abs(o.stack_mat(kd, eps, ang, method="2x2_1") - J).max() > 1e-3


def test_2x2_1_isotropic_layers_are_a_fresnel_chain():
    """"2x2_1" enters every layer after the first from the previous one:
    through isotropic layers at normal incidence the stack is
    prod_j t_(j-1, j) exp(i n_j kd_j) with t = 2 n_(j-1) / (n_(j-1) + n_j);
    "2x2" has the phases only."""
    n = np.array([1.4, 1.9, 1.6, 2.2])
    kd = np.array([0.4, 1.1, 0.8, 1.7])
    eps = np.stack([n ** 2] * 3, -1)
    ang = np.zeros((4, 3))
    expect = np.exp(1j * np.sum(n * kd))
    np.testing.assert_allclose(o.stack_mat(kd, eps, ang, method="2x2"), expect * np.eye(2),
                               atol=1e-13)
    t = np.prod(2 * n[:-1] / (n[:-1] + n[1:]))
    np.testing.assert_allclose(o.stack_mat(kd, eps, ang, method="2x2_1"),
                               t * expect * np.eye(2), atol=1e-13)


@pytest.mark.parametrize("reflect_in,reflect_out", [(False, False), (True, False),
                                                    (False, True), (True, True)])
def test_transfer_2x2_slab_fresnel_flags(reflect_in, reflect_out):
    """transfer with a 2x2 method (dtmm/tmm.py:1697-1748): the forward E of
    the input, the layers' phases, the Fresnel transmissions of the first and
    last interfaces when asked, and H = n_out x E of the output medium."""
    nin, n, nout, kd = 1.0, 1.6, 1.3, 1.9
    f0 = o.fvec(o.normalize_f(o.f_iso(0, 0, nin)), jones=(0.6, 0.8))
    back = o.fvec(o.normalize_f(o.f_iso(0, 0, nin)), jones=(1, 0), mode=-1)
    fout = o.transfer1d(f0 + 0.5 * back, [kd], np.asarray([[n * n] * 3]), np.zeros((1, 3)),
                        nin=nin, nout=nout, method="2x2", reflect_in=reflect_in,
                        reflect_out=reflect_out)
    g = np.exp(1j * n * kd)
    if reflect_in:
        g *= 2 * nin / (nin + n)
    if reflect_out:
        g *= 2 * n / (n + nout)
    np.testing.assert_allclose(fout[0::2], f0[0::2] * g, atol=1e-14)
    np.testing.assert_allclose(fout[1], nout * fout[0], atol=1e-14)
    np.testing.assert_allclose(fout[3], -nout * fout[2], atol=1e-14)


def test_fvec2E_and_E2fvec():
    """fvec2E keeps the forward part of a field vector; E2fvec rebuilds the
    forward wave from its E (oblique, isotropic medium)."""
    f = o.f_iso(0.4, 0.7, 1.5)
    fw = o.fvec(o.normalize_f(f), jones=(0.3, 1j))
    bw = o.fvec(o.normalize_f(f), jones=(1, -0.2), mode=-1)
    e = o.fvec2E(fw + bw, f)
    np.testing.assert_allclose(e, fw[0::2], atol=1e-14)
    np.testing.assert_allclose(o.E2fvec(e, f), fw, atol=1e-14)
    np.testing.assert_allclose(o.fvec2E(bw, f), 0., atol=1e-14)
