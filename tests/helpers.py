"""Shared builders for the parity tests (inputs only; no computation)."""
import numpy as np

import dtmm_amd as dm
from dtmm_amd.samples import random_director_data


def rel_err(a, b):
    """max|a - b| / max|b| (the field-error measure of the benchmark)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


def droplet_case(shape=(12, 32, 32), radius=None, wavelengths=(550.,), jones=(1, 0),
                 beta=0., phi=0., no=1.5, ne=1.6, nhost=1.5):
    nz, ny, nx = shape
    radius = radius if radius is not None else 0.4 * min(shape)
    optical = [dm.nematic_droplet_data(shape, radius, profile="r", no=no, ne=ne, nhost=nhost)]
    window = dm.blackman((ny, nx))
    fd = dm.illumination_data((ny, nx), list(wavelengths), pixelsize=100, beta=beta, phi=phi,
                              jones=jones, window=window)
    return fd, optical


def random_case(shape=(10, 24, 40), seed=3, wavelengths=(450., 650.), jones=None, beta=0.,
                phi=0., no=1.5, ne=1.7):
    nz, ny, nx = shape
    optical = [random_director_data(shape, seed=seed, no=no, ne=ne)]
    window = dm.blackman((ny, nx))
    fd = dm.illumination_data((ny, nx), list(wavelengths), pixelsize=100, beta=beta, phi=phi,
                              jones=jones, window=window)
    return fd, optical


def flush(msg):
    print(msg, flush=True)
