"""Pins the oracle's exp_su3_luscher (src/su3/utils.cpp:145-193) and
proj_lie_su3 (:130-134) against scipy.linalg.expm golden vectors, closed
forms and algebraic identities."""
import json
import os

import numpy as np
import pytest

import fields
import oracle

rng = np.random.default_rng(7)


def load_golden(golden_dir):
    with open(os.path.join(golden_dir, "expm_golden.json")) as f:
        cases = json.load(f)["cases"]
    Z = np.array([np.array(c["Z"]).reshape(3, 3, 2) for c in cases])
    E = np.array([np.array(c["expm"]).reshape(3, 3, 2) for c in cases])
    co = np.array([c["coeff"] for c in cases])
    return Z[..., 0] + 1j * Z[..., 1], co, E[..., 0] + 1j * E[..., 1]


def test_exp_matches_scipy_expm_golden(golden_dir):
    Z, co, E = load_golden(golden_dir)
    for k in range(len(co)):
        got = oracle.exp_su3(Z[k][None], co[k])[0]
        # N = 18 Horner truncation is far below 1e-14 for |X| <= sqrt(2)
        assert np.abs(got - E[k]).max() < 2e-14, k


def test_exp_zero_is_identity():
    assert np.array_equal(oracle.exp_su3(np.zeros((1, 3, 3)))[0], np.eye(3))


def test_exp_diagonal_closed_form():
    for a, b in [(0.3, -0.1), (1e-3, 2e-3), (0.7, 0.2)]:
        X = np.diag([1j * a, 1j * b, -1j * (a + b)])
        got = oracle.exp_su3(X[None])[0]
        want = np.diag(np.exp([1j * a, 1j * b, -1j * (a + b)]))
        assert np.abs(got - want).max() < 1e-15


def test_exp_su2_embedding_closed_form():
    sig = [np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]])]
    for theta in [0.01, 0.4, 0.9]:
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        ns = sum(n[k] * sig[k] for k in range(3))
        for (i, j) in [(0, 1), (1, 2), (0, 2)]:
            X = np.zeros((3, 3), complex)
            X[np.ix_([i, j], [i, j])] = 1j * theta * ns
            want = np.eye(3, dtype=complex)
            want[np.ix_([i, j], [i, j])] = np.cos(theta) * np.eye(2) + 1j * np.sin(theta) * ns
            got = oracle.exp_su3(X[None])[0]
            assert np.abs(got - want).max() < 1e-15


def test_exp_group_properties():
    X = fields.random_algebra(rng, 32)
    X *= (0.9 / np.linalg.norm(X, axis=(1, 2)))[:, None, None]
    Ep = oracle.exp_su3(X)
    Em = oracle.exp_su3(-X)
    I = np.eye(3)
    assert np.abs(Ep @ Em - I).max() < 1e-14
    assert np.abs(Ep @ np.conj(np.swapaxes(Ep, 1, 2)) - I).max() < 1e-14
    assert np.abs(np.linalg.det(Ep) - 1).max() < 1e-14


def test_exp_coeff_scales_argument():
    X = fields.random_algebra(rng, 8)
    a = oracle.exp_su3(X, 0.005)
    b = oracle.exp_su3(0.005 * X, 1.0)
    assert np.abs(a - b).max() < 1e-16


def test_exp_warns_when_t_exceeds_one():
    before = oracle.exp_warnings()
    X = fields.random_algebra(rng)
    X *= 1.3 / np.linalg.norm(X)  # t = |X|^2 / 2 = 0.845
    oracle.exp_su3(X[None])
    assert oracle.exp_warnings() == before
    X *= 2.0  # t = 3.38
    oracle.exp_su3(X[None])
    assert oracle.exp_warnings() == before + 1


@pytest.mark.parametrize("seed", [0, 1])
def test_proj_lie_su3_properties(seed):
    r = np.random.default_rng(seed)
    A = r.normal(size=(16, 3, 3)) + 1j * r.normal(size=(16, 3, 3))
    P = oracle.proj_lie_su3(A)
    assert np.abs(P + np.conj(np.swapaxes(P, 1, 2))).max() < 1e-15     # anti-hermitian
    assert np.abs(np.trace(P, axis1=1, axis2=2)).max() < 1e-15        # traceless
    assert np.abs(oracle.proj_lie_su3(P) - P).max() < 1e-15            # idempotent
    # hermitian part and trace are dropped entirely
    Hm = A + np.conj(np.swapaxes(A, 1, 2))
    assert np.abs(oracle.proj_lie_su3(Hm)).max() < 1e-15
    assert np.abs(oracle.proj_lie_su3(1j * np.eye(3)[None])).max() < 1e-16
    # closed form: (A - A^dag)/2 - tr/3
    B = 0.5 * (A - np.conj(np.swapaxes(A, 1, 2)))
    want = B - (np.trace(B, axis1=1, axis2=2) / 3)[:, None, None] * np.eye(3)
    assert np.abs(P - want).max() < 1e-15
