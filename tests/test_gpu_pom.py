"""GPU parity: the polarizing-microscope specter (dtmmx_pom_specter) and the
colour conversion (dtmmx_specter2color) against the oracle's restatement of
dtmm/field_viewer.py, dtmm/field.py and dtmm/color.py on the same inputs.

Tolerances (max|S - S_ref| / max|S_ref|):
  c128 (float64 specter): 1e-9;  c64 (float32 specter): 1e-4.
Colour: float64 1e-12 absolute; float32 1e-5 absolute (values in [0, 1]).
"""
import numpy as np
import pytest

import oracle
from helpers import droplet_case, flush, rel_err

TOL = {"c128": 1e-9, "c64": 1e-4}


def _field(lead, nlam, shape, seed):
    rng = np.random.default_rng(seed)
    sh = lead + (nlam, 4) + shape
    return rng.standard_normal(sh) + 1j * rng.standard_normal(sh)


CASES = [
    # (lead, jvec, pmat kind, diffraction, mode, focus, n)
    ((2,), None, None, False, None, 0., 1.0),
    ((2,), None, "per_lambda", True, None, 5.5, 1.5),
    ((3, 2), (0.6, 0.8j), "shared", True, None, -3.0, 1.0),
    ((1, 2), (1., 0.), "per_lambda", True, +1, 2.0, 1.5),
    ((2,), None, None, True, -1, 4.0, 1.2),
    ((2, 2), (0.3 - 0.1j, 0.9), "shared", False, None, 0., 1.0),
    # a medium with eps below betamax^2: evanescent modes pass the mask, so
    # the unfused k-space operator runs instead of the fused column pass
    ((2,), None, None, True, None, 2.0, 0.75),
]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_pom_specter_matches_oracle(case, precision):
    from dtmm_amd import pom
    lead, jvec, pk, diffraction, mode, focus, n = CASES[case]
    shape, nlam = (32, 48), 3
    flush("case pom_specter %d %s" % (case, precision))
    f = _field(lead, nlam, shape, 100 + case)
    ks = oracle.k0([450., 550., 650.], 100.)
    rng = np.random.default_rng(7)
    pmat = None
    if pk == "per_lambda":
        pmat = rng.standard_normal((nlam, 4, 4)) + 1j * rng.standard_normal((nlam, 4, 4))
    elif pk == "shared":
        th = 0.7
        pmat = pom.jonesmat4x4(pom.polarizer([np.cos(th), np.sin(th)]), n)
    kw = dict(jvec=jvec, pmat=pmat, focus=focus, n=n, mode=mode, betamax=0.8,
              diffraction=diffraction)
    ref = oracle.pom_specter(f, ks, **kw)
    dt = np.complex64 if precision == "c64" else np.complex128
    got = pom.calculate_pom_specter(f.astype(dt), ks, precision=precision, **kw)
    assert got.shape == ref.shape == shape + (nlam,)
    err = rel_err(got, ref)
    flush("  rel specter err %.3e" % err)
    assert err < TOL[precision], err


@pytest.mark.gpu
def test_pom_specter_of_a_transfer_output():
    """The viewer step on what the transfer produced: crossed polarizers
    around a droplet, refocused by 3 pixels in a medium of index 1.5."""
    import dtmm_amd as dm
    from dtmm_amd import pom
    flush("case pom_after_transfer")
    fd, optical = droplet_case(shape=(12, 32, 32), wavelengths=(500., 600.), jones=None)
    field, wl, ps = fd
    out = dm.transfer_field((field, wl, ps), optical, precision="c128")[0]
    ks = oracle.k0(wl, ps)
    pmat = pom.jonesmat4x4(pom.polarizer([0., 1.]), 1.5)
    kw = dict(jvec=(1., 0.), pmat=pmat, focus=3., n=1.5, betamax=0.8)
    ref = oracle.pom_specter(out, ks, **kw)
    got = pom.calculate_pom_specter(out, ks, **kw)
    err = rel_err(got, ref)
    flush("  rel specter err %.3e" % err)
    assert err < 1e-9


@pytest.mark.gpu
def test_field2specter_keeps_leading_axes():
    from dtmm_amd import pom
    flush("case field2specter")
    f = _field((2, 3), 4, (16, 24), 5)
    ref = oracle.field2specter(f)
    got = pom.field2specter(f)
    assert got.shape == ref.shape == (2, 3, 16, 24, 4)
    assert rel_err(got, ref) < 1e-12
    got_sum = pom.field2spectersum(f[0])
    assert rel_err(got_sum, ref[0].sum(0)) < 1e-12


@pytest.mark.gpu
def test_pom_specter_full_size_properties():
    """512 x 512, 9 wavelengths, 2 polarisations (the transfer benchmark's
    output shape), c64: with focus 0 and both directions the diffraction
    matrix is the identity on a band-limited field, and the Jones
    combination (1, 0) selects the x-polarised field."""
    import torch
    from dtmm_amd import pom
    flush("case pom_full_size")
    H = W = 512
    nlam = 9
    ks = oracle.k0(np.linspace(400., 700., nlam), 100.)
    g = torch.Generator(device="cuda").manual_seed(3)
    f = torch.randn((2, nlam, 4, H, W), dtype=torch.complex64, device="cuda", generator=g)
    # band-limit on the device with the engine's own diffraction (mode both, d 0)
    beta = torch.from_numpy(oracle.betaphi((H, W), ks)[0]).to("cuda")
    F = torch.fft.fft2(f)
    F = F * (beta < 0.8)[None, :, None].to(F.dtype)
    f = torch.fft.ifft2(F).contiguous()
    s_diff = pom.calculate_pom_specter(f[None], ks, jvec=(1., 0.), focus=0., n=1.0,
                                       betamax=0.8)
    s_plain = pom.calculate_pom_specter(f[0], ks, diffraction=False)
    torch.cuda.synchronize()
    err = (s_diff - s_plain).abs().max().item() / s_plain.abs().max().item()
    flush("  identity rel err %.3e" % err)
    assert err < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("norm,gamma,gray", [(False, True, False), (True, False, False),
                                             (2.5, 2.2, False), (True, True, True),
                                             (False, False, True)])
def test_specter2color_matches_oracle(dtype, norm, gamma, gray):
    from dtmm_amd import pom
    flush("case specter2color %s %s %s %s" % (dtype, norm, gamma, gray))
    rng = np.random.default_rng(9)
    spec = (rng.random((17, 23, 5)) * 0.4).astype(dtype)
    cmf = rng.random((5, 3)) * 0.5
    ref = oracle.specter2color(spec.astype(np.float64), cmf, norm=norm, gamma=gamma, gray=gray)
    got = pom.specter2color(spec, cmf, norm=norm, gamma=gamma, gray=gray)
    assert got.shape == ref.shape == (17, 23, 3)
    assert got.dtype == spec.dtype
    err = np.abs(got - ref).max()
    flush("  abs err %.3e" % err)
    assert err < (1e-12 if dtype == "float64" else 1e-5)


@pytest.mark.gpu
def test_specter2color_long_cmf():
    """More wavelengths than the kernel arguments carry: the cmf is uploaded."""
    from dtmm_amd import pom
    flush("case specter2color long cmf")
    rng = np.random.default_rng(4)
    spec = rng.random((9, 11, 150)) * 0.01
    cmf = rng.random((150, 3)) * 0.1
    ref = oracle.specter2color(spec, cmf, norm=True)
    got = pom.specter2color(spec, cmf, norm=True)
    assert np.abs(got - ref).max() < 1e-12


@pytest.mark.gpu
def test_pom_specter_repeated_calls_reuse_tables():
    """Calls alternating between two grids and two betamax values (kept
    device tables) give the oracle's result every time."""
    from dtmm_amd import pom
    flush("case pom repeated calls")
    ks = oracle.k0([500., 600.], 100.)
    for rep in range(2):
        for shape, bm in (((32, 48), 0.8), ((64, 40), 0.6), ((32, 48), 0.6)):
            f = _field((1,), 2, shape, 40 + rep)
            ref = oracle.pom_specter(f, ks, focus=2., n=1.3, betamax=bm)
            got = pom.calculate_pom_specter(f, ks, focus=2., n=1.3, betamax=bm)
            assert rel_err(got, ref) < 1e-9, (rep, shape, bm)
