"""Host logic of the anisotropic effective media (dtmm_amd.data, CPU only)
against the oracle's restatement of dtmm/data.py:489-532, :1240-1419 and
dtmm/linalg.py:364-424."""
import numpy as np
import pytest

import dtmm_amd.data as dd
import oracle as o


def _tensor(e):
    return np.array([[e[0], e[3], e[4]], [e[3], e[1], e[5]], [e[4], e[5], e[2]]])


def test_tensor_eig_order_and_frame():
    rng = np.random.default_rng(0)
    t = rng.standard_normal((200, 6))
    t[:, :3] += 3.
    # diagonal tensors in every order, including ties and equal spacing
    t[:6, 3:] = 0.
    t[:6, :3] = [[1, 2, 3], [3, 1, 2], [2, 3, 1], [1, 1, 2], [2, 1, 1], [1.5, 1.5, 1.5]]
    w, v = dd.tensor_eig(t)
    wo, _ = o.tensor_eig(t)
    np.testing.assert_allclose(w, wo, rtol=1e-13, atol=1e-13)
    for i in range(len(t)):
        # v diag(w) v^T rebuilds the tensor; v is a right-handed rotation
        np.testing.assert_allclose(v[i] @ np.diag(w[i]) @ v[i].T, _tensor(t[i]), atol=1e-12)
        np.testing.assert_allclose(np.linalg.det(v[i]), 1., atol=1e-12)


def test_angles_rebuild_tensor():
    rng = np.random.default_rng(1)
    t = rng.standard_normal((100, 6))
    t[:, :3] += 3.
    w, a = dd.eps2epsva(t)
    np.testing.assert_allclose(dd.epsva2eps(w, a), t, atol=1e-12)
    np.testing.assert_allclose(dd.epsva2eps(w, a), o.epsva2eps(w, a), atol=1e-13)
    with pytest.raises(NotImplementedError):
        dd.eps2epsva(t + 0.1j)


def test_eig_symmetry_reference_vectors():
    np.testing.assert_allclose(dd.eig_symmetry(0, np.array([1, 2, 3.])), (2, 2, 2))
    np.testing.assert_allclose(dd.eig_symmetry(1, np.array([1, 2, 3.])), (1.5, 1.5, 3))
    np.testing.assert_allclose(dd.eig_symmetry(2, np.array([1, 2, 3.])), (1, 2, 3))
    np.testing.assert_allclose(dd.uniaxial_order(0.5, np.array([1, 2, 3.])), (1.75, 1.75, 2.5))
    assert dd.parse_symmetry("biaxial") == 2
    assert dd.parse_symmetry(["uniaxial", 0, 2]) == (1, 0, 2)
    with pytest.raises(ValueError):
        dd.parse_symmetry(3)


@pytest.mark.parametrize("sym", [0, 1, 2, "uniaxial", [1, 2, 0, 1]])
def test_effective_block_matches_oracle(sym):
    rng = np.random.default_rng(2)
    L, H, W = 4, 6, 7
    eps = np.broadcast_to([2.25, 2.25, 2.89], (L, H, W, 3)).copy()
    eps[1] = [2.1, 2.3, 2.8]  # a biaxial layer
    ang = rng.uniform(-1, 1, (L, H, W, 3))
    d = rng.uniform(0.5, 1.5, L)
    de, ev, ea = dd.effective_block((d, eps, ang), sym)
    do, evo, eao = o.effective_block((d, eps, ang), sym)
    np.testing.assert_allclose(de, do)
    # the same effective tensors and eigenvalues (frames may differ in sign)
    np.testing.assert_allclose(ev, np.asarray(evo).reshape(L, 3), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(dd.epsva2eps(ev, ea),
                               o.epsva2eps(np.asarray(evo).reshape(L, 3), np.asarray(eao).reshape(L, 3)),
                               atol=1e-12)


def test_effective_from_layer_tensors():
    """The per-layer mean tensors (what the device computes) give the same
    effective layers as the whole block."""
    rng = np.random.default_rng(3)
    L, H, W = 3, 5, 5
    eps = np.broadcast_to([2.25, 2.25, 2.89], (L, H, W, 3))
    ang = rng.uniform(-1, 1, (L, H, W, 3))
    t = dd.epsva2eps(eps, ang).reshape(L, -1, 6).mean(1)
    for sym in (1, 2, [2, 1, 1]):
        w, a = dd.effective_from_tensors(t, sym)
        _, ev, ea = dd.effective_block((np.ones(L), eps, ang), sym)
        np.testing.assert_allclose(w, ev, rtol=1e-13)
        np.testing.assert_allclose(dd.epsva2eps(w, a), dd.epsva2eps(ev, ea), atol=1e-13)
