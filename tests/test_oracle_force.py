"""Pins the oracle's staple sum (src/gauge/GaugeField.cpp:74-97) and force
F = P_TA(-U Sigma) (src/flow/gradient_flow.cpp:17-37) without retyping them:
a finite-difference identity against the plaquette action, gauge covariance,
and zero force on constant abelian fields."""
import numpy as np
import pytest

import fields
import oracle

LATTICES = [(4, 4, 4, 4), (3, 4, 5, 4)]


def total_re_tr(f, L):
    """sum over all plaquettes of Re tr U_p (from the oracle's plaquette)."""
    return oracle.plaquette(f, L) * 6 * 3 * int(np.prod(L))


@pytest.mark.parametrize("L", LATTICES)
def test_force_is_minus_gradient_of_plaquette_sum(L):
    """Re tr(X F(x,mu)) = -d/ds sum_p Re tr U_p(exp(sX) U_mu(x)) at s = 0 for X in su(3).
    This fixes the sign, the normalisation and both staple orientations."""
    rng = np.random.default_rng(3)
    f = oracle.hot_start(L, 11)
    oracle.heatbath(f, L, 5.0, 2, 12)  # rough but not random: O(1) forces
    F = fields.to_cplx(oracle.compute_force(f, L))
    V = int(np.prod(L))
    for trial in range(6):
        s = int(rng.integers(V))
        mu = int(rng.integers(4))
        X = fields.random_algebra(rng)
        X /= np.linalg.norm(X)
        U0 = fields.to_cplx(f[s, mu]).copy()

        def S(h):
            g = f.copy()
            E = oracle.exp_su3(X[None], h)[0]
            g[s, mu] = fields.to_real(E @ U0)
            return total_re_tr(g, L)

        def D(h):
            return (S(h) - S(-h)) / (2 * h)

        h = 2e-3
        deriv = (4 * D(h / 2) - D(h)) / 3  # Richardson, O(h^4)
        lhs = np.real(np.trace(X @ F[s, mu]))
        assert abs(lhs + deriv) < 1e-8 * max(1.0, abs(deriv)), (trial, lhs, deriv)


@pytest.mark.parametrize("L", LATTICES)
def test_force_gauge_covariance(L):
    f = oracle.hot_start(L, 5)
    F = fields.to_cplx(oracle.compute_force(f, L))
    g_field = f.copy()
    G = oracle.gauge_transform(g_field, L, 77)
    Fg = fields.to_cplx(oracle.compute_force(g_field, L))
    Gd = np.conj(np.swapaxes(G, 1, 2))
    want = G[:, None] @ F @ Gd[:, None]
    assert np.abs(Fg - want).max() < 1e-12


def test_force_traceless_antihermitian():
    L = (4, 4, 4, 4)
    F = fields.to_cplx(oracle.compute_force(oracle.hot_start(L, 2), L))
    assert np.abs(F + np.conj(np.swapaxes(F, -1, -2))).max() < 1e-14
    assert np.abs(np.trace(F, axis1=-2, axis2=-1)).max() < 1e-14


def test_force_vanishes_on_identity_and_abelian_fields():
    L = (4, 4, 6, 4)
    f = oracle.cold_start(L)
    assert np.abs(oracle.compute_force(f, L)).max() == 0.0
    fa, _ = fields.abelian_flux(L, [1, 2, -1, 1, 3, 2])
    # forward and backward staples give exp(+i phi H) + exp(-i phi H): hermitian
    assert np.abs(oracle.compute_force(fa, L)).max() < 1e-14


def test_staple_contains_all_six_plaquettes():
    """Re tr(U_mu(x) Sigma_mu(x)) = sum of the six plaquettes through (x, mu),
    i.e. the change of the total plaquette sum when U_mu(x) is removed."""
    L = (4, 4, 4, 4)
    f = oracle.hot_start(L, 9)
    s, mu = 37, 2
    st = oracle.staple(f, L, s, mu)
    U = fields.to_cplx(f[s, mu])
    lhs = np.real(np.trace(U @ st))
    g = f.copy()
    g[s, mu] = 0.0  # a zero link kills exactly the plaquettes through it
    rhs = total_re_tr(f, L) - total_re_tr(g, L)
    assert abs(lhs - rhs) < 1e-11
