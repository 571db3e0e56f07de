"""Pinning of the oracle's reference-field normalisation of the multi-pass
4x4 transfer (norm DTMM_NORM_REF "local" and DTMM_NORM_FFT | DTMM_NORM_REF,
dtmm/transfer.py:103-138, :631-639, :769-775, :805-835) (CPU only).

* The building blocks by what they are: normalised field-matrix columns
  carry unit Poynting flux; a plane wave's mode amplitudes; the Jones
  polarizer and its 4x4 form are projectors that keep a wave of their own
  polarisation and remove the orthogonal one; jones2H gives H = n E x z of
  plane waves (admittance n at normal incidence, eps/a and a obliquely);
  the local projection returns a field whose reference is itself.
* The whole chain on a uniform plane wave through a homogeneous uniaxial
  slab at normal incidence, computed here from normal-incidence formulas
  only: the one-pass 4x4 field (forward/backward split in the effective
  medium for reflection mode 1, in the slab itself for mode 2, phases
  exp(+-i k0 d n), projection onto the forward wave of the output medium),
  the reference (the Jones retarder, H of the output medium, scaled by
  sqrt(nin/nout)) and the normalisation (intensity ratio clipped to 1;
  local: plus the reference's part in the orthogonal polarisation).
* Without reflections (index-matched slab) the normalised passes return
  the one-pass result, and every operator commutes with lateral shifts.
"""
import numpy as np
import pytest

import oracle as o


def _poynting_cols(f):
    return (f[..., 0, :] * np.conj(f[..., 1, :])).real - (f[..., 2, :] * np.conj(f[..., 3, :])).real


@pytest.mark.parametrize("beta", [0., 0.4])
def test_normalized_columns_carry_unit_flux(beta):
    for eps, ang in ((np.full(3, 2.25), np.zeros(3)), (np.array([2.2, 2.2, 2.7]),
                                                       np.array([0., 0.7, 0.3]))):
        f = o.normalize_f(o.alphaf(beta, 0.5, eps, ang)[1])
        np.testing.assert_allclose(np.abs(_poynting_cols(f)), 1., atol=1e-13)


def test_plane_wave_amplitudes():
    n = 1.6
    c, s = 0.6, 0.8j
    fvec = np.array([c, n * c, s, -n * s])  # forward wave (Ex, Hy, Ey, Hx)
    a = o.fvec2avec(fvec, o.f_iso(n=n))
    np.testing.assert_allclose(a[1::2], 0., atol=1e-15)
    # |a|^2 of unit-flux modes is the wave's flux n (|c|^2 + |s|^2)
    np.testing.assert_allclose(np.sum(np.abs(a[0::2]) ** 2), n, atol=1e-13)
    bvec = np.array([c, -n * c, s, n * s])  # backward wave
    np.testing.assert_allclose(o.fvec2avec(bvec, o.f_iso(n=n))[0::2], 0., atol=1e-15)


def test_jones_polarizer_and_4x4_projector():
    rng = np.random.default_rng(1)
    j = o.jonesvec(rng.standard_normal(2) + 1j * rng.standard_normal(2))
    np.testing.assert_allclose(np.vdot(j, j), 1., atol=1e-15)
    J = o.polarizer(j)
    np.testing.assert_allclose(J @ J, J, atol=1e-15)
    np.testing.assert_allclose(J @ j, j, atol=1e-15)
    jp = np.array([-np.conj(j[1]), np.conj(j[0])])  # orthogonal polarisation
    np.testing.assert_allclose(J @ jp, 0., atol=1e-15)
    n = 1.4
    P = o.ray_jonesmat4x4(J, epsv=o.refind2eps([n] * 3))
    np.testing.assert_allclose(P @ P, P, atol=1e-14)
    fw = lambda e: np.array([e[0], n * e[0], e[1], -n * e[1]])  # noqa: E731
    np.testing.assert_allclose(P @ fw(j), fw(j), atol=1e-14)
    np.testing.assert_allclose(P @ fw(jp), 0., atol=1e-14)
    # backward waves: the conjugate polariser
    bw = lambda e: np.array([e[0], -n * e[0], e[1], n * e[1]])  # noqa: E731
    np.testing.assert_allclose(P @ bw(np.conj(j)), bw(np.conj(j)), atol=1e-14)
    with np.errstate(invalid="ignore", divide="ignore"):
        assert np.isnan(o.jonesvec(np.zeros(2))).all()


def test_jones2H_is_plane_wave_admittance():
    shape, wl, ps = (16, 16), np.array([500.]), 100.
    ks = o.k0(wl, ps)
    n = 1.3
    x = np.arange(16)
    # normal incidence, uniform field
    E = np.zeros((1, 2) + shape, complex)
    E[0, 0], E[0, 1] = 0.3 + 0.1j, -0.7j
    H = o.jones2H(E, ks, n=n)
    np.testing.assert_allclose(H[0, 0], n * E[0, 0], atol=1e-14)
    np.testing.assert_allclose(H[0, 1], -n * E[0, 1], atol=1e-14)
    # an oblique wave on the grid (beta = one kx step), p and s admittances
    beta = 500. / (16 * 100.)
    a = np.sqrt(n ** 2 - beta ** 2)
    ramp = np.exp(1j * ks[0] * beta * x)[None, :] * np.ones(shape)
    E = np.zeros((1, 2) + shape, complex)
    E[0, 0] = ramp
    np.testing.assert_allclose(o.jones2H(E, ks, n=n)[0, 0], n ** 2 / a * ramp, atol=1e-13)
    E = np.zeros((1, 2) + shape, complex)
    E[0, 1] = ramp
    np.testing.assert_allclose(o.jones2H(E, ks, n=n)[0, 1], -a * ramp, atol=1e-13)
    # above betamax nothing is kept
    np.testing.assert_allclose(o.jones2H(E, ks, n=n, betamax=0.1), 0., atol=1e-15)


def test_local_projection_of_uniform_forward_waves():
    """On normal-incidence forward waves (uniform fields) the local
    projection keeps a reference of the field's own polarisation scaled by
    its intensity ratio, and passes on the orthogonal polarisation."""
    shape, wl, ps, nout = (12, 16), np.array([480., 600.]), 100., 1.5
    ks = o.k0(wl, ps)
    pout = o.projection_matrix(shape, ks, epsv=o.refind2eps([nout] * 3), mode=+1)

    def wave(e):
        f = np.zeros((2, 4) + shape, complex)
        f[:, 0], f[:, 1], f[:, 2], f[:, 3] = e[0], nout * e[0], e[1], -nout * e[1]
        return f

    j = np.array([0.6, 0.8j])
    jp = np.array([-np.conj(j[1]), np.conj(j[0])])
    fwd = wave(j)
    np.testing.assert_allclose(o.project_normalized_local(fwd, pout, nout, ref=fwd), fwd,
                               atol=1e-13)
    np.testing.assert_allclose(o.project_normalized_local(fwd, pout, nout, ref=0.5 * fwd),
                               0.25 * fwd, atol=1e-13)
    np.testing.assert_allclose(o.project_normalized_local(fwd, pout, nout, ref=wave(jp)),
                               wave(jp), atol=1e-13)
    mixed = wave(0.7 * j + 0.3j * jp)
    np.testing.assert_allclose(o.project_normalized_local(fwd, pout, nout, ref=mixed),
                               0.49 * fwd + wave(0.3j * jp), atol=1e-13)
    np.testing.assert_allclose(o.project_normalized_local(fwd, pout, nout, ref=0 * fwd), 0.,
                               atol=1e-15)


def _slab_pass0(reflection, norm, nin, nout, no, ne, d, k):
    """E of the output after two passes of a uniform 45-degree wave through a
    homogeneous planar slab (optic axis x) at normal incidence, from
    normal-incidence formulas."""
    neff = np.sqrt((2 * no ** 2 + ne ** 2) / 3)
    E0 = np.array([1., 1.]) / np.sqrt(2)
    Et, Eref = np.zeros(2, complex), np.zeros(2, complex)
    for c, n in ((0, ne), (1, no)):
        nsplit = neff if reflection == 1 else n
        ef = (E0[c] + nin * E0[c] / nsplit) / 2  # forward part of (E, nin E)
        eb = (E0[c] - nin * E0[c] / nsplit) / 2
        ph = np.exp(1j * k * d * n)
        e = ef * ph + eb / ph
        h = nsplit * (ef * ph - eb / ph)
        Et[c] = (e + h / nout) / 2  # forward wave of the output medium
        Eref[c] = E0[c] * ph  # the 2x2 reference: Jones retarder
    Eref = Eref * np.sqrt(nin / nout)
    If = nout * np.sum(np.abs(Et) ** 2)
    if norm == 3:
        Iref = nout * np.sum(np.abs(Eref) ** 2)
        return Et * min(1., Iref / If)
    j = Et / np.linalg.norm(Et)
    J = np.outer(j, np.conj(j))
    Iref1 = nout * np.sum(np.abs(J @ Eref) ** 2)
    return Et * min(1., Iref1 / If) + (np.eye(2) - J) @ Eref


@pytest.mark.parametrize("reflection", [1, 2])
@pytest.mark.parametrize("norm", [2, 3])
def test_uniaxial_slab_two_passes_closed_form(reflection, norm):
    nin, nout, no, ne, d = 1.0, 1.6, 1.5, 1.65, 3.3
    shape, wl, ps = (8, 8), np.array([500.]), 100.
    k = o.k0(wl, ps)[0]
    eps = np.broadcast_to([no ** 2, no ** 2, ne ** 2], (1,) + shape + (3,))
    ang = np.broadcast_to([0., np.pi / 2, 0.], (1,) + shape + (3,))
    f = np.zeros((1, 4) + shape, complex)
    f[0, 0], f[0, 1] = 1 / np.sqrt(2), nin / np.sqrt(2)
    f[0, 2], f[0, 3] = 1 / np.sqrt(2), -nin / np.sqrt(2)
    out = o.transfer_4x4((f.copy(), wl, ps), [(np.array([d]), eps, ang)], nin=nin, nout=nout,
                         npass=2,
                         diffraction=1, reflection=reflection, norm=norm, smooth=0.1)[0]
    E = _slab_pass0(reflection, norm, nin, nout, no, ne, d, k)
    np.testing.assert_allclose(out[0, 0], E[0], atol=1e-12)
    np.testing.assert_allclose(out[0, 2], E[1], atol=1e-12)
    np.testing.assert_allclose(out[0, 1], nout * E[0], atol=1e-12)
    np.testing.assert_allclose(out[0, 3], -nout * E[1], atol=1e-12)
    # the normalisation acts: the one-pass field differs
    one = o.transfer_4x4((f.copy(), wl, ps), [(np.array([d]), eps, ang)], nin=nin, nout=nout,
                         npass=1,
                         diffraction=1, reflection=reflection)[0]
    pout = o.projection_matrix(shape, o.k0(wl, ps), epsv=o.refind2eps([nout] * 3), mode=+1)
    assert np.abs(o.diffract(one, pout)[0, 0] - out[0, 0]).max() > 1e-3


@pytest.mark.parametrize("norm,windowed", [(3, True), (2, False)])
def test_index_matched_slab_passes_return_one_pass_result(norm, windowed):
    """No interface reflects: every normalised pass returns the forward
    field it started from, so three passes give the one-pass output (smooth
    0.01: the notch filter is 1 to 1e-30 below betamax).  The local norm
    splits each pixel into normal-incidence modes, which is exact only for
    a plane wave at normal incidence, so it runs on the unwindowed wave."""
    n = 1.5
    shape, wl, ps = (16, 16), np.array([450., 600.]), 100.
    window = o.blackman(shape) if windowed else None
    fd = o.illumination_data(shape, wl, pixelsize=ps, beta=0., phi=0., n=n, window=window,
                             jones=(0.6, 0.8j))
    eps = np.full((2,) + shape + (3,), n ** 2)
    data = [(np.array([2., 3.]), eps, np.zeros_like(eps))]
    one = o.transfer_4x4((fd[0].copy(), wl, ps), data, nin=n, nout=n, npass=1)[0]
    three = o.transfer_4x4((fd[0].copy(), wl, ps), data, nin=n, nout=n, npass=3, norm=norm,
                           smooth=0.01)[0]
    np.testing.assert_allclose(three, one, atol=1e-10 * np.abs(one).max())


@pytest.mark.parametrize("norm,reflection", [(2, 1), (3, 1), (2, 2)])
def test_reference_norm_is_translation_equivariant(norm, reflection):
    rng = np.random.default_rng(21)
    L, shape = 3, (12, 20)
    wl = np.array([520.])
    eps = np.empty((L,) + shape + (3,))
    eps[..., 0] = eps[..., 1] = rng.uniform(2.2, 2.4, (L,) + shape)
    eps[..., 2] = eps[..., 0] + rng.uniform(0.1, 0.3, (L,) + shape)
    ang = np.stack([np.zeros((L,) + shape), rng.uniform(0, np.pi, (L,) + shape),
                    rng.uniform(-np.pi, np.pi, (L,) + shape)], -1)
    f = rng.standard_normal((1, 4) + shape) + 1j * rng.standard_normal((1, 4) + shape)
    dy, dx = 5, 3
    roll = lambda a, ax=(-2, -1): np.roll(np.roll(a, dy, ax[0]), dx, ax[1])  # noqa: E731
    kw = dict(beta=0.1, phi=0.4, nin=1.2, nout=1.7, npass=3, norm=norm, reflection=reflection)
    out = o.transfer_4x4((f.copy(), wl, 100.), [(np.ones(L), eps, ang)], **kw)[0]
    out_r = o.transfer_4x4((roll(f), wl, 100.), [(np.ones(L), roll(eps, (1, 2)),
                                                  roll(ang, (1, 2)))], **kw)[0]
    np.testing.assert_allclose(out_r, roll(out), atol=1e-11 * np.abs(out).max())
