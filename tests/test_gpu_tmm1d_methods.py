"""GPU parity: the 1D methods "4x4_1", "4x4_2", "4x4_r", "2x2" and "2x2_1"
(dtmmx_tmm_transfer, dtmm_amd.tmm) against the oracle's restatement of
dtmm/tmm.py:1112-1748 on the same inputs.

Tolerance: max|f - f_ref| / max|f_ref| < 1e-10 (fp64 throughout)."""
import numpy as np
import pytest

import oracle as o
from helpers import flush, rel_err

TOL = 1e-10


def _stack(L, shape, rng, absorbing=False):
    eps = np.empty((L,) + shape + (3,))
    for i in range(L):
        base = rng.uniform(2.0, 2.6, shape + (3,))
        if i % 3 == 0:
            base[..., 1] = base[..., 0] = base[..., 2]
        elif i % 3 == 1:
            base[..., 1] = base[..., 0]
        eps[i] = base
    if absorbing:
        eps = eps + 1j * rng.uniform(0., 0.05, eps.shape)
    ang = rng.uniform(-1.5, 1.5, (L,) + shape + (3,))
    return eps, ang


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "4x4_1", "4x4_2", "4x4_r", "2x2", "2x2_1"])
@pytest.mark.parametrize("cfact", [0.01, 0.3])
def test_stack_and_layer_mat(method, cfact):
    from dtmm_amd import tmm
    flush("case tmm1d stack_mat %s cfact %g" % (method, cfact))
    rng = np.random.default_rng(31)
    shape, L = (4, 6), 5
    eps, ang = _stack(L, shape, rng, absorbing=True)
    kd = rng.uniform(0.3, 3.0, (L,) + shape)
    ref = o.stack_mat(kd, eps, ang, 0.3, 0.8, cfact=cfact, method=method)
    got = tmm.stack_mat(kd, eps, ang, 0.3, 0.8, cfact=cfact, method=method)
    assert got.shape == ref.shape
    assert rel_err(got, ref) < TOL
    ref1 = o.layer_mat(kd[0], eps[0], ang[0], 0.3, 0.8, cfact=cfact, method=method)
    got1 = tmm.layer_mat(kd[0], eps[0], ang[0], 0.3, 0.8, cfact=cfact, method=method)
    assert rel_err(got1, ref1) < TOL


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4_1", "4x4_2", "4x4_r"])
@pytest.mark.parametrize("with_out", [False, True])
def test_transfer4x4_methods(method, with_out):
    from dtmm_amd import tmm
    flush("case tmm1d transfer4x4 %s with_out=%s" % (method, with_out))
    rng = np.random.default_rng(32)
    shape, L = (8, 16), 6
    eps, ang = _stack(L, shape, rng)
    kd = rng.uniform(0.3, 3.0, (L,) + shape)
    fin = rng.standard_normal(shape + (4,)) + 1j * rng.standard_normal(shape + (4,))
    fout0 = None
    if with_out:
        fout0 = rng.standard_normal(shape + (4,)) + 1j * rng.standard_normal(shape + (4,))
    ref_out, ref_in = o.transfer4x4(fin, kd, eps, ang, 0.35, -0.6, nin=1.3, nout=1.6,
                                    fvec_out=fout0, method=method)
    f_in = fin.copy()
    f_out = fout0.copy() if with_out else None
    got = tmm.transfer4x4(f_in, kd, eps, ang, 0.35, -0.6, nin=1.3, nout=1.6, method=method,
                          fvec_out=f_out)
    e1, e2 = rel_err(got, ref_out), rel_err(f_in, ref_in)
    flush("  out err %.3e  in err %.3e" % (e1, e2))
    assert e1 < TOL and e2 < TOL


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["2x2", "2x2_1"])
@pytest.mark.parametrize("reflect_in,reflect_out", [(False, False), (True, True), (True, False)])
def test_transfer_2x2_methods(method, reflect_in, reflect_out):
    from dtmm_amd import tmm
    flush("case tmm1d transfer %s in=%s out=%s" % (method, reflect_in, reflect_out))
    rng = np.random.default_rng(33)
    shape, L = (5, 7), 6
    eps, ang = _stack(L, shape, rng, absorbing=True)
    kd = rng.uniform(0.3, 3.0, (L,) + shape)
    fin = rng.standard_normal(shape + (4,)) + 1j * rng.standard_normal(shape + (4,))
    ref = o.transfer1d(fin, kd, eps, ang, 0.25, 1.1, nin=1.2, nout=1.45, method=method,
                       reflect_in=reflect_in, reflect_out=reflect_out)
    f_in = fin.copy()
    got = tmm.transfer(f_in, kd, eps, ang, 0.25, 1.1, nin=1.2, nout=1.45, method=method,
                       reflect_in=reflect_in, reflect_out=reflect_out)
    assert np.array_equal(f_in, fin)  # the 2x2 methods leave the input alone
    err = rel_err(got, ref)
    flush("  out err %.3e" % err)
    assert err < TOL
    ev = rng.standard_normal(shape + (2,)) + 1j * rng.standard_normal(shape + (2,))
    nin = 1.2 if reflect_in else None
    nout = 1.45 if reflect_out else None
    ref2 = o.transfer2x2(ev, kd, eps, ang, 0.25, 1.1, nin=nin, nout=nout, method=method)
    got2 = tmm.transfer2x2(ev, kd, eps, ang, 0.25, 1.1, nin=nin, nout=nout, method=method)
    err2 = rel_err(got2, ref2)
    flush("  transfer2x2 err %.3e" % err2)
    assert err2 < TOL


@pytest.mark.gpu
def test_transfer_default_method_is_2x2():
    """dtmm/tmm.py:1697: transfer's method defaults to "2x2"."""
    from dtmm_amd import tmm
    flush("case tmm1d transfer default method")
    rng = np.random.default_rng(34)
    eps, ang = _stack(4, (3,), rng)
    kd = rng.uniform(0.3, 3.0, (4, 3))
    f0 = o.fvec(o.normalize_f(o.f_iso(0, 0, 1.0)), jones=(1, 0))
    ref = o.transfer1d(f0, kd, eps, ang)
    got = tmm.transfer(f0, kd, eps, ang)
    assert rel_err(got, ref) < TOL
