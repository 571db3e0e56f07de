"""GPU parity of anisotropic effective media (eff_data 1, 2, per-layer
symmetries, custom effective blocks; dtmm/transfer.py:926-1005,
dtmm/data.py:1364-1419) against the oracle: the 4x4 method in reflection
mode 1 with diffraction 1, whose layer boundaries then apply the dense
diffraction operators of the effective media (dtmm/propagate_4x4.py:376-388).

Tolerances (max|F - F_ref| / max|F_ref|): c128 1e-9, c64 1e-4.
"""
import numpy as np
import pytest

import dtmm_amd as dm
import oracle
from helpers import flush, random_case, rel_err

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-9, "c64": 1e-4}


def _compare(fd, optical, precision, eff_data, beta=0., phi=0., **kw):
    field, wl, ps = fd
    ref = oracle.transfer_field((field.copy(), wl, ps), optical, beta=beta, phi=phi,
                                method="4x4", eff_data=eff_data, **kw)[0]
    dt = np.complex64 if precision == "c64" else np.complex128
    flush("  GPU eff_data=%r %s field %s %s" % (eff_data if not isinstance(eff_data, tuple)
                                                 else "custom", precision, field.shape, kw))
    out = dm.transfer_field((field.astype(dt), wl, ps), optical, beta=beta, phi=phi,
                            method="4x4", eff_data=eff_data, precision=precision, **kw)[0]
    err = rel_err(out, ref)
    flush("  rel field err %.3e" % err)
    assert err < TOL[precision], err
    return out, ref


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("eff", [1, 2, "uniaxial", [1, 2, 1, 2]])
def test_random_director(eff, precision):
    flush("case eff random_director %r %s" % (eff, precision))
    fd, optical = random_case(shape=(4, 24, 32), wavelengths=(480., 620.), jones=None)
    # a list of blocks takes one entry per block: the block itself for
    # per-layer symmetries
    _compare(fd, optical[0] if isinstance(eff, list) else optical, precision, eff)


def test_multi_beam_and_substeps():
    flush("case eff multi_beam nstep")
    beta = [0., 0.2, 0.15]
    phi = [0., 1.0, -2.0]
    fd, optical = random_case(shape=(3, 32, 32), wavelengths=(500., 600.), jones=None,
                              beta=beta, phi=phi)
    _compare(fd, optical, "c128", 1, beta=beta, phi=phi, nstep=2)


def test_uniform_uniaxial_layers_exact():
    """Laterally uniform uniaxial layers with per-layer effective media: the
    diffraction step is exact, so the output is the exact per-mode
    propagation (the oracle's diffraction -1)."""
    flush("case eff uniform layers exact")
    H = W = 32
    vals = [(2.25, 2.25, 2.89), (2.1, 2.1, 2.6), (2.3, 2.3, 2.7)]
    angs = [(0., 0.6, 0.3), (0., 1.1, -0.7), (0., 0.4, 2.0)]
    eps = np.empty((3, H, W, 3))
    ang = np.empty((3, H, W, 3))
    for l in range(3):
        eps[l] = vals[l]
        ang[l] = angs[l]
    optical = (np.array([1.3, 0.8, 2.1]), eps, ang)
    fd = dm.illumination_data((H, W), [500., 600.], pixelsize=100, beta=0.1, phi=0.4,
                              window=dm.blackman((H, W)))
    field, wl, ps = fd
    exact = oracle.transfer_field((field.copy(), wl, ps), optical, beta=0.1, phi=0.4,
                                  method="4x4", diffraction=-1)[0]
    out = dm.transfer_field((field, wl, ps), optical, beta=0.1, phi=0.4, method="4x4",
                            eff_data=[1, 1, 1], precision="c128")[0]
    assert rel_err(out, exact) < 1e-9


def test_npass3():
    flush("case eff npass 3")
    fd, optical = random_case(shape=(3, 24, 24), wavelengths=(550.,), jones=(1, 0))
    _compare(fd, optical, "c128", 1, npass=3, nin=1.5, nout=1.5)


def test_ret_bulk():
    flush("case eff ret_bulk")
    fd, optical = random_case(shape=(3, 24, 24), wavelengths=(550.,), jones=(1, 0))
    field, wl, ps = fd
    ref = oracle.transfer_field((field.copy(), wl, ps), optical, beta=0., phi=0., method="4x4",
                                eff_data=2, ret_bulk=True)[0]
    out = dm.transfer_field((field, wl, ps), optical, beta=0., phi=0., method="4x4",
                            eff_data=2, ret_bulk=True, precision="c128")[0]
    assert out.shape == ref.shape
    assert rel_err(out, ref) < 1e-9


def test_custom_block_and_per_block_list():
    """A custom anisotropic effective block with its own thickness, and a
    list of blocks with one effective identifier each (isotropic for one)."""
    flush("case eff custom block")
    fd, b1 = random_case(shape=(2, 24, 24), wavelengths=(520.,), jones=None, seed=5)
    _, b2 = random_case(shape=(3, 24, 24), wavelengths=(520.,), jones=None, seed=6)
    d1 = np.asarray(b1[0][0])
    custom = (1.5 * d1, np.array([[2.2, 2.2, 2.7], [2.3, 2.3, 2.6]]),
              np.array([[0., 0.5, 0.2], [0., 1.0, -1.1]]))
    _compare(fd, b1[0], "c128", custom)
    _compare(fd, b1 + b2, "c128", ["uniaxial", "isotropic"])
    _compare(fd, b1 + b2, "c128", [2, 1])


def test_isotropic_equivalents_take_the_isotropic_path():
    """Per-layer symmetry 0 and a custom isotropic block with the layer
    thickness are isotropic effective media: same result as the default up
    to the rounding of the mean."""
    flush("case eff isotropic equivalents")
    fd, optical = random_case(shape=(3, 24, 24), wavelengths=(550.,), jones=(1, 0))
    field, wl, ps = fd
    base = dm.transfer_field((field, wl, ps), optical, beta=0., phi=0., method="4x4",
                             precision="c128")[0]
    d, epsv, epsa = optical[0]
    e = dm.uniaxial_order(0., np.asarray(epsv)).reshape(-1, 3).mean(0)
    custom = (np.asarray(d), np.tile(e, (len(d), 1)), np.zeros((len(d), 3)))
    for eff in ([0, 0, 0], custom):
        out = dm.transfer_field((field, wl, ps), optical[0], beta=0., phi=0., method="4x4",
                                eff_data=eff, precision="c128")[0]
        assert rel_err(out, base) < 1e-9


def test_no_diffraction_ignores_effective_data():
    flush("case eff diffraction 0")
    fd, optical = random_case(shape=(3, 16, 16), wavelengths=(550.,), jones=(1, 0))
    field, wl, ps = fd
    a = dm.transfer_field((field, wl, ps), optical, beta=0., phi=0., method="4x4",
                          diffraction=0, precision="c128")[0]
    b = dm.transfer_field((field, wl, ps), optical, beta=0., phi=0., method="4x4",
                          diffraction=0, eff_data=1, precision="c128")[0]
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("kw", [dict(method="2x2"), dict(method="4x4", reflection=2),
                                dict(method="4x4", diffraction=3)])
def test_unsupported_modes_raise(kw):
    fd, optical = random_case(shape=(2, 16, 16), wavelengths=(550.,), jones=(1, 0))
    field, wl, ps = fd
    with pytest.raises(NotImplementedError):
        dm.transfer_field((field, wl, ps), optical, beta=0., phi=0., eff_data=1,
                          precision="c128", **kw)
