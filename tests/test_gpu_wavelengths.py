"""GPU parity of the per-wavelength routes of transfer_field: dispersive
materials with split_wavelengths, tuple field data, and the beam inferred
from the input field when beta and phi are not given
(dtmm/transfer.py:486-536, :662-664; examples/hello_dispersion.py)."""
import numpy as np
import pytest

import dtmm_amd as dm
import oracle
from dtmm_amd.data import EpsilonCauchy
from helpers import flush, random_case, rel_err

TOL = {"c128": 1e-9, "c64": 1e-4}
DT = {"c128": np.complex128, "c64": np.complex64}


def _dispersive_droplet(shape=(12, 32, 32)):
    """The dispersive droplet of examples/hello_dispersion.py:17-33 at a
    reduced size: Cauchy n = a + b / w^2 with a = sqrt(eps) and b = 0.005 a."""
    d, epsv, epsa = dm.nematic_droplet_data(shape, 10, profile="r", no=1.5, ne=1.6, nhost=1.5)
    epsc = EpsilonCauchy(shape=shape, n=2)
    epsc.coefficients[..., 0] = epsv ** 0.5
    epsc.coefficients[..., 0:2, 1] = 0.005 * epsv[..., 0:2] ** 0.5
    epsc.coefficients[..., 2, 1] = 0.005 * epsv[..., 2] ** 0.5
    coef = epsc.coefficients.copy()
    return (d, epsc, epsa), (d, lambda w: oracle.cauchy2eps(coef, w), epsa)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_hello_dispersion_pattern(precision):
    """Unpolarised eigen-wave illumination, the configured default method,
    no beta/phi (inferred from the field), split_wavelengths=True."""
    flush("case hello dispersion %s" % precision)
    block, oblock = _dispersive_droplet()
    wl = np.linspace(380, 780, 5)
    fd = dm.illumination_data((32, 32), wl, pixelsize=200)
    ref = oracle.transfer_field(fd, [oblock], split_wavelengths=True)[0]
    field = fd[0].astype(DT[precision])
    out, w_out, ps = dm.transfer_field((field, fd[1], fd[2]), [block], split_wavelengths=True,
                                       precision=precision)
    assert out.shape == ref.shape and ps == fd[2]
    err = rel_err(out, ref)
    flush("  rel field err %.3e" % err)
    assert err < TOL[precision], err
    # the dispersion matters here: the same droplet without it differs
    plain = oracle.transfer_field(fd, [(block[0], oblock[1](550.), block[2])])[0]
    assert rel_err(plain, ref) > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_tuple_field_data(method):
    """A tuple of per-wavelength fields returns a tuple of (field,
    wavelength, pixelsize), each the oracle's."""
    flush("case tuple field data %s" % method)
    fd, optical = random_case(shape=(5, 24, 32), wavelengths=(450., 550., 650.), jones=None,
                              beta=0.1, phi=0.4)
    field, wl, ps = fd
    parts = tuple(field[..., i, :, :, :] for i in range(len(wl)))
    ref = oracle.transfer_field((parts, wl, ps), optical, beta=0.1, phi=0.4, method=method)
    out = dm.transfer_field((parts, wl, ps), optical, beta=0.1, phi=0.4, method=method,
                            precision="c128")
    assert isinstance(out, tuple) and len(out) == len(wl)
    for (a, wa, pa), (b, wb, pb) in zip(out, ref):
        assert wa == wb and pa == pb and a.shape == b.shape
        assert rel_err(a, b) < TOL["c128"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_split_wavelengths_npass_writes_input_side(precision):
    """split_wavelengths with npass 3: per-wavelength transfers whose
    input-side fields land in the caller's array, as the joint call's."""
    flush("case split wavelengths npass 3 %s" % precision)
    fd, optical = random_case(shape=(4, 24, 32), wavelengths=(480., 640.), jones=None)
    field, wl, ps = fd
    f_ref = np.array(field, np.complex128)
    ref = oracle.transfer_field((f_ref, wl, ps), optical, beta=0., phi=0., method="4x4",
                                npass=3, nin=1.0, nout=1.4)[0]
    f_gpu = np.array(field, DT[precision])
    out = dm.transfer_field((f_gpu, wl, ps), optical, beta=0., phi=0., method="4x4", npass=3,
                            nin=1.0, nout=1.4, split_wavelengths=True, precision=precision)[0]
    assert rel_err(out, ref) < TOL[precision]
    assert rel_err(f_gpu, f_ref) < TOL[precision]


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_inferred_beam(method):
    """No beta/phi: the beam of an oblique windowed plane wave is inferred
    from the field (multiray off) on the GPU path as in the oracle."""
    flush("case inferred beam %s" % method)
    fd, optical = random_case(shape=(4, 32, 32), wavelengths=(550.,), jones=(1, 0), beta=0.15,
                              phi=0.7)
    ref = oracle.transfer_field(fd, optical, method=method)[0]
    out = dm.transfer_field(fd, optical, method=method, precision="c128")[0]
    assert rel_err(out, ref) < TOL["c128"]
