"""Failure reporting of the GPU transfer: a non-finite output is flagged in
the status word (bit 2) and reported as a RuntimeWarning; the reference has
no such check and returns the values as they are."""
import warnings

import numpy as np
import pytest

import dtmm_amd as dm
from helpers import flush, random_case


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_non_finite_output_warns(method):
    flush("case non-finite output %s" % method)
    fd, optical = random_case(shape=(2, 16, 32), wavelengths=(500.,))
    field, wl, ps = fd
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        good = dm.transfer_field((field.copy(), wl, ps), optical, method=method)[0]
    assert np.isfinite(good).all()
    bad = field.copy()
    bad[..., 3, 5] = np.nan
    with pytest.warns(RuntimeWarning, match="NaN or infinite"):
        out = dm.transfer_field((bad, wl, ps), optical, method=method)[0]
    assert not np.isfinite(out).all()
