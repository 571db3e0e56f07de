"""The C ABI library: loads, exports every symbol include/dtmmx.h declares,
rejects bad plans before touching the GPU, and its host-side per-pixel
solvers agree with the complex128 oracle (CPU only, no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as o
from dtmm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dtmmx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:long\s+)?[a-z_]+\*?\s+\*?(dtmmx_\w+)\s*\(",
                                 text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert lib.dtmmx_abi_version() == _lib.ABI_VERSION


def _desc(ks, beta, phi, **kw):
    d = _lib.PlanDesc()
    d.height, d.width = kw.get("shape", (32, 32))
    d.n_wavelengths = len(ks)
    d.n_rays = kw.get("n_rays", len(beta))
    d.rays_per_beam = kw.get("rpb", 1)
    d.method = kw.get("method", 4)
    d.precision = kw.get("precision", 0)
    d.betamax = kw.get("betamax", 0.8)
    d.k0 = _lib.dptr(ks)
    d.beta = _lib.dptr(beta)
    d.phi = _lib.dptr(phi)
    d.diffraction = kw.get("diffraction", 1)
    d.npass = kw.get("npass", 1)
    d.norm = kw.get("norm", 1)
    d.smooth = 0.1
    d.reflection = kw.get("reflection", 1 if d.method == 4 else 0)
    return d


def _plan_rc(**kw):
    lib = _lib.load()
    ks = np.array([1.0])
    b = np.array([0.0])
    d = _lib.PlanDesc()
    d.height, d.width = kw.get("shape", (32, 32))
    d.n_wavelengths = kw.get("nlam", 1)
    d.n_rays = kw.get("n_rays", 1)
    d.rays_per_beam = kw.get("rpb", 1)
    d.method = kw.get("method", 4)
    d.precision = kw.get("precision", 0)
    d.betamax = 0.8
    d.k0 = _lib.dptr(ks)
    d.beta = _lib.dptr(b)
    d.phi = _lib.dptr(b)
    d.diffraction = kw.get("diffraction", 1)
    d.npass = kw.get("npass", 1)
    d.norm = kw.get("norm", 1)
    d.smooth = 0.1
    d.reflection = kw.get("reflection", 1 if d.method == 4 else 0)
    h = ctypes.c_void_p()
    return lib.dtmmx_plan_create(ctypes.byref(h), ctypes.byref(d)), lib.dtmmx_last_error()


@pytest.mark.parametrize("kw,code", [
    (dict(shape=(0, 32)), _lib.DTMMX_E_ARG),
    (dict(shape=(7, 32)), _lib.DTMMX_E_UNSUPPORTED),
    (dict(shape=(32, 1000)), _lib.DTMMX_E_UNSUPPORTED),
    (dict(nlam=0), _lib.DTMMX_E_ARG),
    (dict(nlam=65), _lib.DTMMX_E_ARG),
    (dict(n_rays=3, rpb=2), _lib.DTMMX_E_ARG),
    (dict(method=3), _lib.DTMMX_E_ARG),
    (dict(precision=2), _lib.DTMMX_E_ARG),
    (dict(diffraction=0), _lib.DTMMX_E_UNSUPPORTED),
    (dict(method=2, diffraction=3, reflection=1, npass=2), _lib.DTMMX_E_UNSUPPORTED),
    (dict(npass=0), _lib.DTMMX_E_ARG),
    (dict(method=2, npass=3), _lib.DTMMX_E_UNSUPPORTED),
    (dict(npass=3, norm=2, diffraction=3), _lib.DTMMX_E_UNSUPPORTED),
    (dict(npass=3, norm=4), _lib.DTMMX_E_ARG),
    (dict(method=2, reflection=0, npass=2), _lib.DTMMX_E_UNSUPPORTED),
    (dict(method=2, reflection=1, diffraction=0, npass=2), _lib.DTMMX_E_UNSUPPORTED),
    (dict(method=2, reflection=2, diffraction=3, npass=2), _lib.DTMMX_E_UNSUPPORTED),
    (dict(reflection=3), _lib.DTMMX_E_UNSUPPORTED),
    (dict(method=2, reflection=2, diffraction=3), _lib.DTMMX_E_UNSUPPORTED),
])
def test_plan_create_rejects_bad_descriptions(kw, code):
    rc, msg = _plan_rc(**kw)
    assert rc == code
    assert msg


def test_plan_tiles_reject_bad_arguments():
    lib = _lib.load()
    assert lib.dtmmx_plan_set_tiles(None, 0, 0) == _lib.DTMMX_E_ARG
    c = ctypes.c_int()
    assert lib.dtmmx_plan_get_tiles(None, ctypes.byref(c), ctypes.byref(c)) == _lib.DTMMX_E_ARG
    assert lib.dtmmx_plan_set_row_split(None, 0) == _lib.DTMMX_E_ARG
    assert lib.dtmmx_plan_get_row_split(None, ctypes.byref(c)) == _lib.DTMMX_E_ARG


def _voxels(n, seed, biaxial=False):
    rng = np.random.default_rng(seed)
    beta = rng.uniform(0, 0.9, n)
    beta[: n // 4] = 0.
    phi = rng.uniform(-3, 3, n)
    eo = rng.uniform(1.5, 2.5, n)
    ee = rng.uniform(1.5, 3.0, n)
    eps = np.stack([eo, eo, ee], -1)
    eps[3 * n // 4:] = eo[3 * n // 4:, None]  # isotropic voxels
    ang = np.stack([np.zeros(n), rng.uniform(0, np.pi, n), rng.uniform(-np.pi, np.pi, n)], -1)
    if biaxial:  # eps0 != eps1 in the first half, with a third Euler angle psi
        h = n // 2
        eps[:h, 1] = rng.uniform(1.5, 2.5, h)
        ang[:h, 0] = rng.uniform(-np.pi, np.pi, h)
    return beta, phi, np.ascontiguousarray(eps), np.ascontiguousarray(ang)


def test_host_alphaf_matches_oracle():
    lib = _lib.load()
    beta, phi, eps, ang = _voxels(1000, 1)
    n = len(beta)
    al = np.zeros((n, 4, 2))
    fm = np.zeros((n, 4, 4, 2))
    _lib.check(lib.dtmmx_host_alphaf(n, _lib.dptr(beta), _lib.dptr(phi), _lib.dptr(eps), None,
                                     _lib.dptr(ang), _lib.dptr(al), _lib.dptr(fm)))
    ao, fo = o.alphaf(beta, phi, eps, ang)
    np.testing.assert_allclose(al[..., 0] + 1j * al[..., 1], ao, atol=1e-13)
    np.testing.assert_allclose(fm[..., 0] + 1j * fm[..., 1], fo, atol=1e-13)


def _pair_match(a, ao):
    """Index map taking the oracle's modes to the engine's: the two modes
    of a direction (columns 0, 2 forward, 1, 3 backward) may come in either
    order from two eigen-solvers."""
    n = len(a)
    perm = np.tile(np.arange(4), (n, 1))
    for c0, c1 in ((0, 2), (1, 3)):
        swap = (np.abs(a[:, c0] - ao[:, c1]) + np.abs(a[:, c1] - ao[:, c0]) <
                np.abs(a[:, c0] - ao[:, c0]) + np.abs(a[:, c1] - ao[:, c1]))
        perm[swap, c0], perm[swap, c1] = c1, c0
    return perm


def test_host_alphaf_biaxial_matches_oracle():
    """Biaxial voxels (numeric eigensolve: Hessenberg + shifted QR in the
    engine, LAPACK zgeev through numpy.linalg.eig in the oracle): the same
    eigenvalues and unit eigenvectors per direction, after matching the
    order of the two modes within a direction."""
    lib = _lib.load()
    beta, phi, eps, ang = _voxels(2000, 7, biaxial=True)
    n = len(beta)
    al = np.zeros((n, 4, 2))
    fm = np.zeros((n, 4, 4, 2))
    _lib.check(lib.dtmmx_host_alphaf(n, _lib.dptr(beta), _lib.dptr(phi), _lib.dptr(eps), None,
                                     _lib.dptr(ang), _lib.dptr(al), _lib.dptr(fm)))
    a = al[..., 0] + 1j * al[..., 1]
    f = fm[..., 0] + 1j * fm[..., 1]
    ao, fo = o.alphaf(beta, phi, eps, ang)
    bia = eps[:, 0] != eps[:, 1]
    assert bia.sum() == n // 2
    perm = _pair_match(a, ao)
    ao = np.take_along_axis(ao, perm, 1)
    fo = np.take_along_axis(fo, perm[:, None, :], 2)
    np.testing.assert_allclose(a, ao, atol=1e-12)
    # unit columns with the largest component real: equal up to the rotation
    # by the beam azimuth, which preserves the column norm but not which
    # component is largest, so compare up to one phase per column
    ph = np.sum(np.conj(fo) * f, axis=1)
    ph = ph / np.abs(ph)
    np.testing.assert_allclose(np.abs(ph[bia]), 1.)
    np.testing.assert_allclose(f[bia], fo[bia] * ph[bia][:, None, :], atol=1e-10)
    np.testing.assert_allclose(f[~bia], fo[~bia], atol=1e-13)


@pytest.mark.parametrize("absorbing", [False, True])
@pytest.mark.parametrize("biaxial", [False, True])
def test_host_layer_operator_matches_oracle(absorbing, biaxial):
    """E diag(p) E^-1 from the engine's per-pixel solve equals the oracle's
    (dtmm/propagate_4x4.py:88-97 with E = E_mat(F) and its 4x4 inverse);
    biaxial voxels through the numeric eigensolve."""
    lib = _lib.load()
    beta, phi, eps, ang = _voxels(600, 2, biaxial=biaxial)
    n = len(beta)
    eps_im = np.full(eps.shape, 0.02) if absorbing else None
    al = np.zeros((n, 4, 2))
    pr = np.zeros((n, 2, 2, 2, 2))
    ok = np.zeros((n, 2), np.int32)
    _lib.check(lib.dtmmx_host_rop(n, _lib.dptr(beta), _lib.dptr(phi), _lib.dptr(eps),
                                  _lib.dptr(eps_im), _lib.dptr(ang), _lib.dptr(al), _lib.dptr(pr),
                                  ok.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    a = al[..., 0] + 1j * al[..., 1]
    P = pr[..., 0] + 1j * pr[..., 1]
    epsc = eps + (1j * eps_im if absorbing else 0.)
    ao, fo = o.alphaf(beta, phi, epsc, ang)
    ao_m = np.take_along_axis(ao, _pair_match(a, ao), 1)
    np.testing.assert_allclose(a, ao_m, atol=1e-12 if biaxial else 1e-13)
    assert np.all(ok == 1)
    kd = 1.7
    p = np.exp(1j * kd * a)
    # engine operator in the (E+x, E-x, E+y, E-y) basis
    M = np.zeros((n, 4, 4), complex)
    I2 = np.eye(2)
    Mp = p[:, 2, None, None] * I2 + (p[:, 0] - p[:, 2])[:, None, None] * P[:, 0]
    Mn = p[:, 3, None, None] * I2 + (p[:, 1] - p[:, 3])[:, None, None] * P[:, 1]
    M[:, 0::2, 0::2] = Mp
    M[:, 1::2, 1::2] = Mn
    e = o.E_mat(fo)
    ref = o.dotmdm(e, np.exp(1j * kd * ao), o.inv(e))
    np.testing.assert_allclose(M, ref, atol=1e-11)


def test_host_rop_real_path_matches_complex_path():
    """Lossless isotropic / uniaxial voxels with propagating modes take a
    real-arithmetic solve; a 1e-300 imaginary eps sends the same voxels
    through the complex closed forms, and both must agree.  Voxels with
    evanescent modes (beta^2 > eps) must stay on the complex solve: their
    ordinary-mode alpha is imaginary."""
    lib = _lib.load()
    beta, phi, eps, ang = _voxels(800, 3)
    n = len(beta)

    def solve(b, eps_im):
        al = np.zeros((n, 4, 2))
        pr = np.zeros((n, 2, 2, 2, 2))
        ok = np.zeros((n, 2), np.int32)
        _lib.check(lib.dtmmx_host_rop(n, _lib.dptr(b), _lib.dptr(phi), _lib.dptr(eps),
                                      _lib.dptr(eps_im), _lib.dptr(ang), _lib.dptr(al),
                                      _lib.dptr(pr), ok.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return al[..., 0] + 1j * al[..., 1], pr, ok

    a_r, p_r, ok_r = solve(beta, None)
    a_c, p_c, ok_c = solve(beta, np.full(eps.shape, 1e-300))
    assert np.array_equal(ok_r, ok_c)
    assert np.all(a_r.imag == 0.)
    np.testing.assert_allclose(a_r, a_c, rtol=0, atol=1e-13)
    np.testing.assert_allclose(p_r, p_c, rtol=0, atol=1e-11 * np.abs(p_c).max())
    # evanescent: beta above sqrt(eps_o)
    b_ev = np.sqrt(eps[:, 0]) * np.random.default_rng(4).uniform(1.05, 1.3, n)
    a_e, _, _ = solve(b_ev, None)
    np.testing.assert_allclose(np.abs(a_e[:, 2].imag), np.sqrt(b_ev ** 2 - eps[:, 0]), rtol=1e-12)


@pytest.mark.parametrize("n", [2, 3, 5])
@pytest.mark.parametrize("case", [
    dict(shape=(32, 48), wl=(450., 650.), beta=(0.,), phi=(0.,)),
    dict(shape=(24, 40), wl=(550.,), beta=(0.12,), phi=(0.4,)),
    dict(shape=(40, 32), wl=(500., 600., 700.), beta=(0., 0.2, 0.3), phi=(0., 1.1, -2.4)),
])
def test_host_windows_match_oracle(n, case):
    """The plan's window list and mean beams (dtmm/mode.py fft_mask with the
    beam offsets of propagate_4x4.py:401-402) against the oracle."""
    lib = _lib.load()
    ks = o.k0(np.asarray(case["wl"]), 100.)
    beta = np.asarray(case["beta"], float)
    phi = np.asarray(case["phi"], float)
    nb = len(beta)
    d = _desc(ks, beta, phi, shape=case["shape"], diffraction=n)
    cnt = ctypes.c_int()
    wi = np.zeros(n * n, np.int32)
    wj = np.zeros(n * n, np.int32)
    wb = np.zeros(n * n * nb)
    wp = np.zeros(n * n * nb)
    rc = lib.dtmmx_host_windows(ctypes.byref(d), n * n, ctypes.byref(cnt), _lib.iptr(wi),
                                _lib.iptr(wj), _lib.dptr(wb), _lib.dptr(wp))
    assert rc == 0, lib.dtmmx_last_error()
    m = cnt.value
    # oracle: per beam offset, as the reference broadcasts beta (rays, 1, 1, 1)
    off_b = beta[:, None, None]
    off_p = phi[:, None, None]
    ks_b = np.asarray(ks)
    windows, (betas, phis) = o.fft_mask(case["shape"], ks_b, n, off_b * np.cos(off_p),
                                              off_b * np.sin(off_p))
    full = o.fft_windows(*o.betaxy(case["shape"], ks_b[..., None]), n,
                              off_b * np.cos(off_p), off_b * np.sin(off_p))
    keep = [i for i in range(n * n) if not np.all(full[i] == 0.)]
    assert m == len(keep)
    assert [int(a) * n + int(b) for a, b in zip(wi[:m], wj[:m])] == keep
    betas = betas.reshape(m, nb)
    phis = phis.reshape(m, nb)
    got_b = wb[:m * nb].reshape(m, nb)
    got_p = wp[:m * nb].reshape(m, nb)
    np.testing.assert_allclose(got_b * np.cos(got_p), betas * np.cos(phis), atol=1e-13)
    np.testing.assert_allclose(got_b * np.sin(got_p), betas * np.sin(phis), atol=1e-13)


def test_host_rop_grad_matches_finite_differences():
    """The forward-mode derivatives of the per-pixel operator data (used by
    the adjoint row kernel) against central differences of dtmmx_host_rop,
    and their values against dtmmx_host_rop itself."""
    lib = _lib.load()
    beta, phi, eps_re, ang = _voxels(64, 7)
    n = len(beta)
    args = [np.ascontiguousarray(a) for a in (beta, phi, eps_re)]
    eim = None
    a0 = np.zeros((n, 4, 2))
    p0 = np.zeros((n, 2, 2, 2, 2))
    da = np.zeros((n, 2, 4, 2))
    dp = np.zeros((n, 2, 2, 2, 2, 2))
    rc = lib.dtmmx_host_rop_grad(n, *[_lib.dptr(a) for a in args], _lib.dptr(eim),
                                 _lib.dptr(np.ascontiguousarray(ang)), _lib.dptr(a0), _lib.dptr(p0),
                                 _lib.dptr(da), _lib.dptr(dp))
    assert rc == 0, lib.dtmmx_last_error()

    def rop(angles):
        a = np.zeros((n, 4, 2))
        p = np.zeros((n, 2, 2, 2, 2))
        ok = np.zeros((n, 2), np.int32)
        rc = lib.dtmmx_host_rop(n, *[_lib.dptr(x) for x in args], _lib.dptr(eim),
                                _lib.dptr(np.ascontiguousarray(angles)), _lib.dptr(a), _lib.dptr(p),
                                ok.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        assert rc == 0
        return a, p

    av, pv = rop(ang)
    np.testing.assert_allclose(a0, av, atol=1e-13)
    np.testing.assert_allclose(p0, pv, atol=1e-12)
    h = 1e-6
    for k, col in ((0, 1), (1, 2)):
        ap = ang.copy()
        am = ang.copy()
        ap[:, col] += h
        am[:, col] -= h
        (a_p, p_p), (a_m, p_m) = rop(ap), rop(am)
        np.testing.assert_allclose(da[:, k], (a_p - a_m) / (2 * h), atol=2e-7)
        np.testing.assert_allclose(dp[:, k], (p_p - p_m) / (2 * h), atol=2e-7)


def test_mode_resolution_no_diffraction():
    """diffraction=0 selects reflection mode 2 for the 4x4 method
    (dtmm/transfer.py:460-461) and keeps mode 0 for an explicit 2x2 method;
    reflection mode 2 is also built with finite diffraction >= 1 and with
    npass > 1; other combinations are refused before any GPU call."""
    from dtmm_amd.transfer import _resolve_modes
    assert _resolve_modes("4x4", 1, 0, None, None, False) == ("4x4", 2)
    assert _resolve_modes("2x2", 1, 0, None, None, False) == ("2x2", 0)
    assert _resolve_modes("4x4", 1, False, None, None, False) == ("4x4", 2)
    with pytest.raises(NotImplementedError):
        _resolve_modes('"2x2"', 1, 0, None, None, False)   # configured default: reflection 1
    with pytest.raises(NotImplementedError):
        _resolve_modes("4x4", 1, 0, 1, None, False)
    assert _resolve_modes("4x4", 1, 1, 2, None, False) == ("4x4", 2)
    assert _resolve_modes("4x4", 1, 3, 2, None, False) == ("4x4", 2)
    # diffraction inf or < 0: exact per-mode propagation, 4x4 with one pass
    assert _resolve_modes("4x4", 1, np.inf, 2, None, False) == ("4x4", 2)
    assert _resolve_modes("4x4", 1, -1, None, None, False) == ("4x4", 1)
    with pytest.raises(NotImplementedError):
        _resolve_modes("4x4", 3, np.inf, None, None, False)
    with pytest.raises(NotImplementedError):
        _resolve_modes("2x2", 1, -1, None, None, False)
    assert _resolve_modes("4x4", 2, 1, 2, None, False) == ("4x4", 2)
    assert _resolve_modes("4x4", 3, 0, 2, None, False) == ("4x4", 2)
    with pytest.raises(NotImplementedError):
        _resolve_modes("4x4", 3, 0, None, None, False)  # npass > 1 selects reflection 1
    with pytest.raises(ValueError):
        _resolve_modes("2x2", 2, 0, 0, None, False)  # as dtmm/transfer.py:1077-1078
    assert _resolve_modes("2x2", 1, 0, 2, None, False) == ("2x2", 2)
    # 2x2 with npass > 1: reflection 1 (the default, dtmm/transfer.py:459-463)
    # with diffraction 1, reflection 2 with diffraction 0 or 1
    assert _resolve_modes("2x2", 3, 1, None, None, False) == ("2x2", 1)
    assert _resolve_modes("2x2", 3, 1, 2, None, False) == ("2x2", 2)
    assert _resolve_modes("2x2", 4, 0, 2, None, True) == ("2x2", 2)
    with pytest.raises(NotImplementedError):
        _resolve_modes("2x2", 3, 0, 1, None, False)
    with pytest.raises(NotImplementedError):
        _resolve_modes("2x2", 3, 3, None, None, False)  # selects reflection 2, windows
    assert _resolve_modes("2x2", 1, 1, 2, None, False) == ("2x2", 2)
    with pytest.raises(NotImplementedError):
        _resolve_modes("2x2", 1, 3, 2, None, False)
    # 2x2 with tent windows (diffraction > 1): reflection 0 or 1, npass 1, no bulk
    assert _resolve_modes("2x2", 1, 3, None, None, False) == ("2x2", 0)
    assert _resolve_modes('"2x2"', 1, 5, None, None, False) == ('2x2', 1)
    with pytest.raises(NotImplementedError):
        _resolve_modes("2x2", 1, 3, None, None, True)
    # DTMM_NORM_REF with npass > 1: diffraction 0 or 1
    assert _resolve_modes("4x4", 3, 1, None, None, False, 2) == ("4x4", 1)
    assert _resolve_modes("4x4", 3, 1, None, None, False, 3) == ("4x4", 1)
    with pytest.raises(NotImplementedError):
        _resolve_modes("4x4", 3, 3, None, None, False, 2)
    # ret_bulk: diffraction 0 or 1; the 4x4 method with npass 1
    assert _resolve_modes("4x4", 1, 1, None, None, True) == ("4x4", 1)
    assert _resolve_modes("2x2", 1, 1, None, None, True) == ("2x2", 0)
    with pytest.raises(NotImplementedError):
        _resolve_modes("4x4", 3, 1, None, None, True)
    with pytest.raises(NotImplementedError):
        _resolve_modes("4x4", 1, 3, None, None, True)


def _pom_desc(**kw):
    d = _lib.PomDesc()
    ks = np.asarray([1.0, 1.2])
    d.height, d.width, d.n_wavelengths = 32, 48, 2
    d.k0 = _lib.dptr(ks)
    d.n_fields, d.n_pol, d.diffraction, d.mode = 2, 1, 1, 0
    d.eps_re, d.eps_im, d.betamax, d.d = 1.0, 0.0, 0.8, 0.0
    d.precision = _lib.DTMMX_C64
    for k, v in kw.items():
        setattr(d, k, v)
    return d, ks


def test_pom_workspace_size_and_rejects_bad_descriptions():
    """Host-side sizing and validation of dtmmx_pom_specter (no GPU call)."""
    lib = _lib.load()
    d, ks = _pom_desc()
    n = lib.dtmmx_pom_workspace_size(ctypes.byref(d))
    r = lambda b: (b + 255) // 256 * 256  # noqa: E731
    assert n == r(2 * 2 * 4 * 32 * 48 * 8) + r(48 * 8) + r(32 * 8) + r(2 * 32 * 4)
    d0, ks0 = _pom_desc(diffraction=0)
    assert lib.dtmmx_pom_workspace_size(ctypes.byref(d0)) == 0
    bad = [dict(n_pol=3), dict(mode=2), dict(n_fields=0), dict(n_wavelengths=0),
           dict(n_wavelengths=_lib.MAX_WAVELENGTHS + 1), dict(precision=7)]
    for kw in bad:
        db, ksb = _pom_desc(**kw)
        assert lib.dtmmx_pom_workspace_size(ctypes.byref(db)) == 0, kw
        rc = lib.dtmmx_pom_specter(ctypes.byref(db), None, None, None, 0, None)
        assert rc == _lib.DTMMX_E_ARG, kw
    du, ksu = _pom_desc(height=30)
    assert lib.dtmmx_pom_specter(ctypes.byref(du), None, None, None, 0, None) == \
        _lib.DTMMX_E_UNSUPPORTED
    kneg = np.asarray([1.0, -1.0])
    dn, _ = _pom_desc()
    dn.k0 = _lib.dptr(kneg)
    assert lib.dtmmx_pom_specter(ctypes.byref(dn), None, None, None, 0, None) == _lib.DTMMX_E_ARG
    # valid description, missing buffers
    assert lib.dtmmx_pom_specter(ctypes.byref(d), None, None, None, 0, None) == _lib.DTMMX_E_ARG


def test_specter2color_rejects_bad_arguments():
    lib = _lib.load()
    cmf = np.ones((3, 3))
    assert lib.dtmmx_specter2color_workspace_size(1000, 3) == 256 + 4 * 8
    p = ctypes.c_void_p(8)  # never dereferenced: every call below fails validation
    for args in [(0, 3, 0, 0), (10, 0, 0, 0), (10, 3, 3, 0), (10, 3, 0, 3)]:
        npix, nlam, nm, gm = args
        rc = lib.dtmmx_specter2color(p, p, npix, nlam, _lib.dptr(cmf), nm, 0., gm, 0., 0,
                                     _lib.DTMMX_C64, p, 1 << 20, None)
        assert rc == _lib.DTMMX_E_ARG, args
    rc = lib.dtmmx_specter2color(p, p, 10, 3, _lib.dptr(cmf), 0, 0., 0, 0., 0, _lib.DTMMX_C64,
                                 p, 8, None)
    assert rc == _lib.DTMMX_E_ARG


def test_director2data_rejects_bad_arguments():
    """Host-side validation of dtmmx_director2data (no GPU call)."""
    lib = _lib.load()
    p = ctypes.c_void_p(8)  # never dereferenced: every call below fails validation
    assert lib.dtmmx_director2data_blocks(1000) == 4
    assert lib.dtmmx_director2data_blocks(0) == 0
    bad = [
        (None, None, 10, 1, 1.5, 1.6, 1.5, 1.0),        # null director
        (p, None, 0, 1, 1.5, 1.6, 1.5, 1.0),            # empty layer
        (p, None, 10, 0, 1.5, 1.6, 1.5, 1.0),           # no layers
        (p, None, 10, 1, 1.5, 1.6, 1.5, 0.0),           # scale factor 0
        (p, p, 10, 1, 1.5, 1.6, float("nan"), 1.0),     # mask without nhost
    ]
    for d, m, ls, nl, no, ne, nh, sf in bad:
        rc = lib.dtmmx_director2data(d, m, ls, nl, no, ne, nh, sf, p, p, p, None)
        assert rc == _lib.DTMMX_E_ARG, (ls, nl, sf)


def test_float_sincos_of_the_layer_kernels():
    """sincos_cw (the complex64 kernels' phase factors) against the exact
    sine and cosine of the same float argument, over phases up to 1e6 rad
    (thick layers): error below 1e-7."""
    lib = _lib.load()
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.uniform(-10, 10, 2000), rng.uniform(-1e6, 1e6, 4000),
                        np.array([0., 5e4, -5e4, 6.5e4, 1e6, np.pi / 4, 3 * np.pi / 4])])
    x = x.astype(np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    fp = ctypes.POINTER(ctypes.c_float)
    _lib.check(lib.dtmmx_host_sincos(len(x), x.ctypes.data_as(fp), s.ctypes.data_as(fp),
                                     c.ctypes.data_as(fp)))
    xd = x.astype(np.float64)
    assert np.abs(s - np.sin(xd)).max() < 1e-7
    assert np.abs(c - np.cos(xd)).max() < 1e-7


def test_stale_library_is_refused(tmp_path):
    """The build records the hash of the sources it compiled next to the
    library; a record that does not match the sources present, or no record,
    refuses the library before it is loaded."""
    lib = tmp_path / "libx.so"
    lib.write_bytes(b"not loaded")
    with pytest.raises(RuntimeError, match="no build record"):
        _lib._check_fresh(str(lib))
    (tmp_path / "libx.so.srchash").write_text("0" * 64 + "\n")
    with pytest.raises(RuntimeError, match="other sources"):
        _lib._check_fresh(str(lib))
    (tmp_path / "libx.so.srchash").write_text(_lib.source_hash() + " \n")
    _lib._check_fresh(str(lib))
    # the library in the tree passes (it was built from these sources)
    _lib._check_fresh(_lib.LIB_PATH)
