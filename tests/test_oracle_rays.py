"""Pinning of the illumination rays (dtmm/field.py:342-404, window.py:107-156;
CPU only) and of the product's host builders against the oracle.

Closed forms: without smoothing a 5-pixel aperture is the disk of pixel
centres with x^2 + y^2 < 6.25 (21 rays: the 25 centres minus the 4
corners); beta = |r| * 2 NA / (D - 1) with the centre ray at 0 and the
outermost on-axis ray at NA; the ray set is point symmetric (the mean
transverse wave vector vanishes); intensities sum to the number of rays;
the Tukey edge is 1/2 half way through the edge and continuous; an annular
aperture removes the inner disk.
"""
import numpy as np
import pytest

import oracle as o


def test_hard_aperture_rays():
    b, p, i = o.illumination_rays(0.2, 5, smooth=0.)
    assert len(b) == 21
    np.testing.assert_allclose(i, 1.)
    assert np.isclose(b.min(), 0.)
    np.testing.assert_allclose(np.sort(b)[-8:], np.sqrt(5) * 0.1)  # (+-2, +-1), (+-1, +-2)
    assert np.any(np.isclose(b, 0.2))                                  # (+-2, 0): beta = NA
    np.testing.assert_allclose((b * np.cos(p)).sum(), 0., atol=1e-12)
    np.testing.assert_allclose((b * np.sin(p)).sum(), 0., atol=1e-12)


def test_tukey_edge():
    r = np.linspace(0, 1.2, 241)
    t = o.tukey(r, 0.2)
    np.testing.assert_allclose(t[r <= 0.8], 1.)
    np.testing.assert_allclose(t[r >= 1.], 0.)
    np.testing.assert_allclose(o.tukey(np.asarray([0.9]), 0.2), 0.5)
    assert np.abs(np.diff(t)).max() < 0.05


def test_smooth_and_annular_apertures():
    b, p, i = o.illumination_rays(0.3, 9, smooth=0.3)
    np.testing.assert_allclose(i.sum(), len(i))
    assert i.max() > i.min()
    bi, _, _ = o.illumination_rays(0.3, (4, 9), smooth=0.)
    assert bi.min() > 0.3 * 2 / 8 * 1.4  # the inner disk is dark
    with pytest.raises(ValueError):
        o.illumination_aperture((5, 5))


@pytest.mark.parametrize("args", [(0.2, 5, 0.1), (0.35, 11, 0.), (0.1, (3, 8), 0.2),
                                  (0.25, 6, 0.5)])
def test_product_rays_match_oracle(args):
    import dtmm_amd as dm
    NA, D, sm = args
    for got, want in zip(dm.illumination_rays(NA, D, sm), o.illumination_rays(NA, D, sm)):
        np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(dm.aperture((12, 10), (0.3, 0.9), 0.2),
                               o.aperture((12, 10), (0.3, 0.9), 0.2), atol=1e-15)
