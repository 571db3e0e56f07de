"""The reference's file formats for optical data (``.dtms``) and field data
(``.dtmf``), byte for byte.

``.dtms`` (dtmm/data.py:1530-1627): the magic ``b"dtms"``, a version byte
(``\\x02``; ``\\x01`` files hold a single block whose epsa may be missing) and,
per block, three ``numpy.save`` records: thickness, epsv, epsa.
``.dtmf`` (dtmm/field.py:903-967): ``b"dtmf"``, version ``\\x00``, then the
field, wavelengths and pixel size as ``numpy.save`` records.

Loading never unpickles: the reference stores dispersive epsv models
(EpsilonCauchy / EpsilonSellmeier) as pickled object arrays, and such a
record is refused with an error instead of being executed.  The GPU
transfer takes numeric epsv only.
"""
import os

import numpy as np

from .data import validate_optical_data

STACK_MAGIC = b"dtms"
STACK_VERSION = b"\x02"
_STACK_VERSION_1 = b"\x01"
FIELD_MAGIC = b"dtmf"
FIELD_VERSION = b"\x00"

__all__ = ["save_stack", "load_stack", "save_field", "load_field"]


def _open(file, mode, ext):
    if isinstance(file, (str, os.PathLike)):
        name = os.fspath(file)
        if "w" in mode and not name.endswith(ext):
            name = name + ext
        return open(name, mode), True
    return file, False


def _at_end(f):
    if hasattr(f, "peek"):
        return f.peek(1) == b""
    pos = f.tell()
    end = f.seek(0, 2)
    f.seek(pos)
    return pos >= end


def _load(f):
    try:
        return np.load(f, allow_pickle=False)
    except ValueError as e:
        if "allow_pickle" in str(e):
            raise NotImplementedError(
                "the file holds a pickled (dispersive epsv) record; it is not loaded, since "
                "loading it would execute code from the file") from e
        raise


def save_stack(file, optical_data):
    """Save optical data in ``.dtms`` format (dtmm/data.py:1536-1580)."""
    optical_data = validate_optical_data(optical_data, shape=(1, 1))
    f, own = _open(file, "wb", ".dtms")
    try:
        f.write(STACK_MAGIC)
        f.write(STACK_VERSION)
        for d, epsv, epsa in optical_data:
            epsv = np.asarray(epsv)
            if epsv.dtype == object:
                raise NotImplementedError("dispersive epsv models are not saved here")
            np.save(f, np.asarray(d), allow_pickle=False)
            np.save(f, epsv, allow_pickle=False)
            np.save(f, np.asarray(epsa), allow_pickle=False)
    finally:
        if own:
            f.close()


def load_stack(file):
    """Load optical data from a ``.dtms`` file (dtmm/data.py:1582-1627).
    Version 1 files return their single block as ``[d, epsv, epsa]``, as in
    the reference; version 2 files return the list of block tuples."""
    f, own = _open(file, "rb", ".dtms")
    try:
        magic = f.read(len(STACK_MAGIC))
        if magic != STACK_MAGIC:
            raise OSError("Failed to interpret file {}".format(file))
        version = f.read(1)
        if ord(version) > ord(STACK_VERSION):
            raise OSError("This file was created with a more recent version of dtmm. "
                          "Please upgrade your dtmm package!")
        if version == _STACK_VERSION_1:
            d = _load(f)
            epsv = _load(f)
            epsa = None if _at_end(f) else _load(f)
            return [d, epsv, epsa]
        out = []
        while not _at_end(f):
            d = _load(f)
            epsv = _load(f)
            epsa = _load(f)
            out.append((d, epsv, epsa))
        return out
    finally:
        if own:
            f.close()


def save_field(file, field_data):
    """Save field data in ``.dtmf`` format (dtmm/field.py:906-935)."""
    field, wavelengths, pixelsize = field_data
    # validate_field_data casts to complex128 / float64 before saving
    # (dtmm/field.py:891-894), so the stored arrays always have those dtypes
    field = np.asarray(field, np.complex128)
    wavelengths = np.asarray(wavelengths, np.float64)
    if field.ndim < 4 or field.shape[-3] != 4:
        raise ValueError("field must have shape (..., n_wavelengths, 4, H, W)")
    if wavelengths.ndim != 1 or wavelengths.shape[0] != field.shape[-4]:
        raise ValueError("wavelengths do not match the field")
    f, own = _open(file, "wb", ".dtmf")
    try:
        f.write(FIELD_MAGIC)
        f.write(FIELD_VERSION)
        np.save(f, field, allow_pickle=False)
        np.save(f, wavelengths, allow_pickle=False)
        np.save(f, np.asarray(float(pixelsize)), allow_pickle=False)
    finally:
        if own:
            f.close()


def load_field(file):
    """Load ``(field, wavelengths, pixelsize)`` from a ``.dtmf`` file
    (dtmm/field.py:937-965)."""
    f, own = _open(file, "rb", ".dtmf")
    try:
        if f.read(len(FIELD_MAGIC)) != FIELD_MAGIC:
            raise OSError("Failed to interpret file {}".format(file))
        if f.read(1) != FIELD_VERSION:
            raise OSError("This file was created with a more recent version of dtmm. "
                          "Please upgrade your dtmm package!")
        field = _load(f)
        wavelengths = _load(f)
        pixelsize = float(_load(f))
        return field, wavelengths, pixelsize
    finally:
        if own:
            f.close()
