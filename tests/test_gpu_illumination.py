"""GPU parity: illumination_data built on the device (dtmmx_illumination)
against the host builder, which the CPU tests pin against the oracle's
restatement of dtmm/field.py:586-713 (eigen-wave snapping, windows, focus,
power normalisation).  Tolerance 1e-12 of the largest field value (c128),
1e-6 (c64)."""
import numpy as np
import pytest

import dtmm_amd as dm
from helpers import flush, rel_err

CASES = [
    dict(beta=0., phi=0., jones=None, window=None),
    dict(beta=0.23, phi=0.7, jones=(1, 1j), window=None),
    dict(beta=0.1, phi=-2.0, jones=None, window="blackman"),
    dict(beta=[0., 0.2, 0.2], phi=[0., 1.0, 2.1], jones=(1, 0), window="blackman",
         intensity=1.3),
    dict(beta=0.15, phi=0.4, jones=(0.3, 1), window="blackman", focus=4.5),
    dict(beta=0.15, phi=0.4, jones=None, window="blackman", backdir=True, focus=-2.),
    dict(beta=0.3, phi=1.3, jones=(1, 0), window=None, diffraction=False),
    dict(beta=[0.05, 0.1], phi=[0.3, -0.4], jones=None, window=None, n=1.0, betamax=0.6),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_device_illumination_matches_host(case, precision):
    kw = dict(CASES[case])
    shape = (24, 40)
    if kw.get("window") == "blackman":
        kw["window"] = dm.blackman(shape)
    flush("case illumination %d %s" % (case, precision))
    wl = [450., 550., 650.]
    host = dm.illumination_data(shape, wl, pixelsize=100, **kw)[0]
    dev, w, p = dm.illumination_data(shape, wl, pixelsize=100, device="cuda", precision=precision,
                                     **kw)
    assert tuple(dev.shape) == host.shape
    err = rel_err(dev.cpu().numpy(), host)
    flush("  rel err %.3e" % err)
    assert err < (1e-12 if precision == "c128" else 1e-6), err


@pytest.mark.gpu
def test_device_field_through_transfer_field():
    """A device illumination goes into transfer_field and comes out as a
    device field equal to the host route's."""
    import torch
    flush("case illumination device transfer")
    shape = (32, 32)
    optical = [dm.nematic_droplet_data((6,) + shape, 12, profile="r", no=1.5, ne=1.6,
                                       nhost=1.5)]
    kw = dict(beta=0.1, phi=0.3, jones=None, window=dm.blackman(shape))
    host = dm.illumination_data(shape, [500., 600.], pixelsize=100, **kw)
    dev = dm.illumination_data(shape, [500., 600.], pixelsize=100, device="cuda", **kw)
    out_h = dm.transfer_field(host, optical, beta=0.1, phi=0.3, method="4x4")[0]
    out_d = dm.transfer_field(dev, optical, beta=0.1, phi=0.3, method="4x4")[0]
    assert torch.is_tensor(out_d) and out_d.is_cuda
    err = rel_err(out_d.cpu().numpy(), out_h)
    flush("  rel err %.3e" % err)
    assert err < 1e-12


@pytest.mark.gpu
def test_device_field_npass_writes_input_side():
    """npass > 1 leaves the input-side field in the caller's device tensor, as
    the host route does in its array."""
    import torch
    flush("case illumination device npass")
    shape = (32, 32)
    optical = [dm.nematic_droplet_data((4,) + shape, 12, profile="r", no=1.5, ne=1.6,
                                       nhost=1.5)]
    kw = dict(beta=0., phi=0., jones=(1, 0), window=dm.blackman(shape))
    fh = dm.illumination_data(shape, [550.], pixelsize=100, **kw)
    fd = dm.illumination_data(shape, [550.], pixelsize=100, device="cuda", **kw)
    tkw = dict(beta=0., phi=0., method="4x4", npass=3, nin=1.0, nout=1.2)
    out_h = dm.transfer_field(fh, optical, **tkw)[0]
    out_d = dm.transfer_field(fd, optical, **tkw)[0]
    assert rel_err(out_d.cpu().numpy(), out_h) < 1e-12
    assert rel_err(fd[0].cpu().numpy(), fh[0]) < 1e-12
