"""The .dtms / .dtmf file formats (dtmm/data.py:1530-1627,
dtmm/field.py:903-967), host only: round trips, the byte layout (magic,
version, numpy.save records), version-1 stacks, and refusal of pickled
records."""
import io

import numpy as np
import pytest

import dtmm_amd as dm
from dtmm_amd import io as dio


def _blocks():
    rng = np.random.default_rng(1)
    b1 = (np.ones(3), rng.uniform(2, 3, (3, 4, 5, 3)), rng.uniform(-1, 1, (3, 4, 5, 3)))
    b2 = (np.asarray([0.5, 2.]), rng.uniform(2, 3, (2, 4, 5, 3)) + 0.01j,
          rng.uniform(-1, 1, (2, 4, 5, 3)))
    return [b1, b2]


def test_stack_round_trip_and_layout(tmp_path):
    blocks = _blocks()
    p = tmp_path / "sample"
    dio.save_stack(str(p), blocks)
    raw = (tmp_path / "sample.dtms").read_bytes()
    assert raw[:5] == b"dtms\x02"
    # the body is plain numpy.save records: d, epsv, epsa per block
    f = io.BytesIO(raw[5:])
    for d, e, a in blocks:
        np.testing.assert_array_equal(np.load(f), d)
        np.testing.assert_array_equal(np.load(f), e)
        np.testing.assert_array_equal(np.load(f), a)
    assert f.read() == b""
    got = dio.load_stack(str(tmp_path / "sample.dtms"))
    assert len(got) == 2
    for (d, e, a), (d2, e2, a2) in zip(blocks, got):
        np.testing.assert_array_equal(d, d2)
        np.testing.assert_array_equal(e, e2)
        np.testing.assert_array_equal(a, a2)


def test_version1_stack_single_block_without_epsa():
    f = io.BytesIO()
    f.write(b"dtms\x01")
    np.save(f, np.ones(2))
    np.save(f, np.full((2, 3, 3, 3), 2.25))
    f.seek(0)
    d, e, a = dio.load_stack(f)
    assert a is None and e.shape == (2, 3, 3, 3)


def test_pickled_record_is_refused():
    f = io.BytesIO()
    f.write(b"dtms\x02")
    np.save(f, np.ones(1))
    np.save(f, np.asarray(("cauchy", [1.5, 0.01]), dtype=object), allow_pickle=True)
    np.save(f, np.zeros((1, 3)))
    f.seek(0)
    with pytest.raises(NotImplementedError):
        dio.load_stack(f)


def test_bad_magic_and_newer_version():
    with pytest.raises(OSError):
        dio.load_stack(io.BytesIO(b"xxxx\x02"))
    with pytest.raises(OSError):
        dio.load_stack(io.BytesIO(b"dtms\x03"))
    with pytest.raises(OSError):
        dio.load_field(io.BytesIO(b"dtmf\x01"))


def test_field_round_trip(tmp_path):
    fd = dm.illumination_data((8, 12), [500., 600.], pixelsize=50., jones=None,
                              window=dm.blackman((8, 12)))
    dio.save_field(str(tmp_path / "f"), fd)
    raw = (tmp_path / "f.dtmf").read_bytes()
    assert raw[:5] == b"dtmf\x00"
    field, wl, ps = dio.load_field(str(tmp_path / "f.dtmf"))
    np.testing.assert_array_equal(field, fd[0])
    np.testing.assert_array_equal(wl, fd[1])
    assert ps == 50.


def test_field_is_saved_in_double_precision(tmp_path):
    """save_field casts like validate_field_data (dtmm/field.py:891-894): a
    complex64 field and integer wavelengths are stored as complex128 and
    float64, so the file is the reference's byte for byte."""
    f = (np.arange(2 * 4 * 3 * 5).reshape(2, 4, 3, 5) * (1 + 0.5j)).astype(np.complex64)
    path = str(tmp_path / "f.dtmf")
    dio.save_field(path, (f, [500, 600], 100))
    g, wl, ps = dio.load_field(path)
    assert g.dtype == np.complex128 and wl.dtype == np.float64
    np.testing.assert_array_equal(g, f.astype(np.complex128))
    np.testing.assert_array_equal(wl, [500., 600.])


def test_cholesteric_droplet_builders():
    """sphere_mask counts the voxel centres within the radius (the z offset
    enters with the reference's opposite sign), and the cholesteric
    director turns by 2 pi per pitch, left-handed clockwise."""
    import dtmm_amd as dm
    shape = (6, 9, 10)
    c = [np.arange(n) - (n - 1) / 2. for n in shape]
    zz, yy, xx = np.meshgrid(*c, indexing="ij")
    for off in ((0, 0, 0), (1., -0.5, 1.5)):
        want = (xx - off[0]) ** 2 + (yy - off[1]) ** 2 + (zz + off[2]) ** 2 <= 3.2 ** 2
        np.testing.assert_array_equal(dm.sphere_mask(shape, 3.2, off), want)
    d = dm.cholesteric_director(shape, 4.)
    np.testing.assert_allclose(np.arctan2(d[1, 0, 0, 1], d[1, 0, 0, 0]), -np.pi / 2, atol=1e-15)
    np.testing.assert_allclose(d[4], d[0], atol=1e-12)
    np.testing.assert_allclose(dm.cholesteric_director(shape, 4., "right")[1, 0, 0, 1], 1.,
                               atol=1e-15)
    t, eps, ang = dm.cholesteric_droplet_data(shape, 3.2, 4., no=1.5, ne=1.7, nhost=1.4)
    outside = ~dm.sphere_mask(shape, 3.2)
    np.testing.assert_allclose(eps[outside], 1.4 ** 2)
