"""Pinning of the oracle's Q-tensor optical data (dtmm/data.py:322-373 Q2data,
Q2eps, eps2epsva; dtmm/linalg.py:172-222 eigenvalue order;
dtmm/rotation.py:154-171 Euler angles; CPU only).

The reference's tests hold no vector for it, so it is pinned by:

* the rebuild: R(angles) diag(eigenvalues) R^T equals the eps tensor of Q
  (a wrong angle formula, sign or transposed frame fails it);
* the order rule on diagonal tensors, by hand: a diagonal tensor keeps its
  input order and the identity frame, then the most distinct eigenvalue
  moves to index 2 by a cyclic permutation only when it is not there
  already: (3, 1, 1) -> (1, 1, 3) with angles (pi/2, pi/2, 0);
  (1, 3, 3.5) -> (3, 3.5, 1); (3.5, 3, 1) and (2.3, 2.2, 2.6) stay, with
  angles 0 (an ascending sort first would give (3, 3.5, 1) and
  (2.2, 2.3, 2.6));
* the diagonal threshold: off-diagonal elements 1e-7 of the largest one
  (squared norm 2e-14) count as diagonal, 1e-5 (2e-10) do not;
* a uniaxial Q of a unit director n: eigenvalues (no^2, no^2, ne^2) and the
  last eigenvector +-n.
"""
import numpy as np
import pytest

import oracle as o


def _random_q(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape + (3, 3)) * 0.3
    m = (a + np.swapaxes(a, -1, -2)) / 2
    m -= np.trace(m, axis1=-2, axis2=-1)[..., None, None] * np.eye(3) / 3
    return np.stack([m[..., 0, 0], m[..., 1, 1], m[..., 2, 2], m[..., 0, 1], m[..., 0, 2],
                     m[..., 1, 2]], -1)


def test_rebuild_is_the_eps_tensor():
    q = _random_q((3, 5, 7), 1)
    _, epsv, epsa = o.Q2data(q, no=1.5, ne=1.7, scale_factor=0.8, biaxial=True)
    np.testing.assert_allclose(o.epsva2eps(epsv, epsa), o.Q2eps(q, 1.5, 1.7, 0.8), atol=1e-13)
    assert np.all(np.abs(np.diff(epsv[..., :2], axis=-1)) <= np.abs(epsv[..., 2:] - epsv[..., :2]).min(-1, keepdims=True) + 1e-12)


@pytest.mark.parametrize("t6,w_exp,ang_exp", [
    ((3., 1., 1., 0., 0., 0.), (1., 1., 3.), (np.pi / 2, np.pi / 2, 0.)),
    ((1., 3., 3.5, 0., 0., 0.), (3., 3.5, 1.), None),
    ((3.5, 3., 1., 0., 0., 0.), (3.5, 3., 1.), (0., 0., 0.)),
    ((2.3, 2.2, 2.6, 0., 0., 0.), (2.3, 2.2, 2.6), (0., 0., 0.))])
def test_order_rule_on_diagonal_tensors(t6, w_exp, ang_exp):
    w, v = o.tensor_eig(np.asarray(t6))
    np.testing.assert_allclose(w, w_exp, atol=1e-14)
    np.testing.assert_allclose(np.linalg.det(v), 1., atol=1e-14)
    if ang_exp is not None:
        np.testing.assert_allclose(np.abs(o.rotation_angles(v)), ang_exp, atol=1e-14)


def test_diagonal_threshold():
    near = o.tensor_eig(np.asarray([3.5, 3., 1., 3.5e-7, 0., 0.]))
    assert tuple(near[0]) == (3.5, 3., 1.)
    np.testing.assert_array_equal(near[1], np.eye(3))
    off = o.tensor_eig(np.asarray([3.5, 3., 1., 3.5e-5, 0., 0.]))
    np.testing.assert_allclose(np.sort(off[0]), [1., 3., 3.5], atol=1e-8)
    assert not np.array_equal(off[1], np.eye(3))


def test_uniaxial_q_of_a_director():
    rng = np.random.default_rng(4)
    n = rng.standard_normal((2, 4, 6, 3))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    _, epsv, epsa = o.Q2data(o.director2Q(n), no=1.5, ne=1.65)
    np.testing.assert_allclose(epsv, np.broadcast_to([2.25, 2.25, 1.65 ** 2], epsv.shape),
                               atol=1e-13)
    th, ph = epsa[..., 1], epsa[..., 2]
    axis = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    np.testing.assert_allclose(np.abs(np.sum(axis * n, -1)), 1., atol=1e-13)
