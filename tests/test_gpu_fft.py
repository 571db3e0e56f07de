"""GPU parity of the LDS Stockham FFT against numpy.fft (the reference's
fft2/ifft2 are pinned to numpy.fft by dtmm/test/test_fft.py)."""
import ctypes

import numpy as np
import pytest

from dtmm_amd import _lib
from helpers import flush


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c64", "c128"])
@pytest.mark.parametrize("shape", [(96, 96), (24, 40), (512, 64), (128, 1024), (1024, 32), (20, 12), (384, 192)])
def test_fft2(shape, precision):
    import torch
    flush("case fft2 %s %s" % (shape, precision))
    rng = np.random.default_rng(0)
    batch = 3
    x = rng.standard_normal((batch,) + shape) + 1j * rng.standard_normal((batch,) + shape)
    dt = torch.complex64 if precision == "c64" else torch.complex128
    lib = _lib.load()
    for inverse in (0, 1):
        t = torch.from_numpy(x).to("cuda", dt).contiguous()
        s = torch.cuda.current_stream()
        _lib.check(lib.dtmmx_fft2(ctypes.c_void_p(t.data_ptr()), batch, shape[0], shape[1], inverse,
                                  _lib.DTMMX_C64 if precision == "c64" else _lib.DTMMX_C128,
                                  ctypes.c_void_p(s.cuda_stream)))
        torch.cuda.synchronize()
        ref = np.fft.fft2(x) if not inverse else np.fft.ifft2(x) * (shape[0] * shape[1])
        err = np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max()
        tol = 1e-6 if precision == "c64" else 1e-13
        assert err < tol, (inverse, err)
