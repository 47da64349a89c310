"""Pins the oracle's plaquette (src/observables/observables_mpi.cpp:11-48),
clover (:66-90) and Q/E (:93-156) by closed forms on constant abelian fields
and by symmetries."""
import numpy as np
import pytest

import fields
import oracle


@pytest.mark.parametrize("L,n", [
    ((4, 4, 4, 4), [1, 0, 0, 0, 0, 1]),
    ((4, 6, 4, 8), [1, 2, -1, 1, 3, 2]),
    ((6, 4, 5, 4), [2, -1, 1, 3, 1, -2]),
])
def test_abelian_closed_forms(L, n):
    f, phis = fields.abelian_flux(L, n)
    P, E, Q = fields.abelian_expected(L, phis)
    assert abs(oracle.plaquette(f, L) - P) < 1e-14
    q, e = oracle.q_e_clover(f, L)
    assert abs(e - E) < 1e-13 * max(1.0, E)
    assert abs(q - Q) < 1e-11 * max(1.0, abs(Q))


def test_identity_field():
    L = (4, 4, 4, 4)
    f = oracle.cold_start(L)
    assert oracle.plaquette(f, L) == 1.0
    assert oracle.q_e_clover(f, L) == (0.0, 0.0)


def test_clover_antisymmetry():
    """F_numu = -F_munu: the (nu,mu) leaves are the inverses of the (mu,nu) leaves."""
    L = (4, 4, 4, 4)
    f = oracle.hot_start(L, 3)
    for s in (0, 77, 255):
        for mu, nu in fields.PLANES:
            a = oracle.clover(f, L, s, mu, nu)
            b = oracle.clover(f, L, s, nu, mu)
            assert np.abs(a + b).max() < 1e-15


def test_observables_gauge_invariant():
    L = (4, 4, 4, 4)
    f = oracle.heatbath(oracle.hot_start(L, 8), L, 6.0, 3, 9)
    oracle.rk3(f, L, 0.02, 5)  # smooth a little so Q is O(1)
    P0 = oracle.plaquette(f, L)
    Q0, E0 = oracle.q_e_clover(f, L)
    g = f.copy()
    oracle.gauge_transform(g, L, 123)
    assert abs(oracle.plaquette(g, L) - P0) < 1e-13
    Q1, E1 = oracle.q_e_clover(g, L)
    assert abs(E1 - E0) < 1e-13 * E0 and abs(Q1 - Q0) < 1e-12


def test_observables_translation_invariant():
    L = (4, 5, 4, 6)
    f = oracle.hot_start(L, 4)
    oracle.rk3(f, L, 0.02, 3)
    P0 = oracle.plaquette(f, L)
    Q0, E0 = oracle.q_e_clover(f, L)
    g = fields.translate(f, L, (1, 2, 3, 5))
    assert abs(oracle.plaquette(g, L) - P0) < 1e-13
    Q1, E1 = oracle.q_e_clover(g, L)
    assert abs(E1 - E0) < 1e-13 * E0 and abs(Q1 - Q0) < 1e-12


@pytest.mark.parametrize("d", [0, 1, 2, 3])
def test_reflection_flips_q(d):
    L = (4, 4, 4, 4)
    f = oracle.heatbath(oracle.hot_start(L, 12), L, 6.0, 2, 13)
    oracle.rk3(f, L, 0.02, 8)
    Q0, E0 = oracle.q_e_clover(f, L)
    g = fields.reflect(f, L, d)
    Q1, E1 = oracle.q_e_clover(g, L)
    assert abs(Q0) > 1e-3
    assert abs(Q1 + Q0) < 1e-12 and abs(E1 - E0) < 1e-13 * E0
    assert abs(oracle.plaquette(g, L) - oracle.plaquette(f, L)) < 1e-14


def test_heatbath_generator_thermalises():
    """Test-input generator only: plaquette near its beta = 6.0 value and links in SU(3)."""
    L = (6, 6, 6, 6)
    f = oracle.thermalised(L, 6.0, 1234, 40)
    P = oracle.plaquette(f, L)
    assert 0.585 < P < 0.605, P
    U = fields.to_cplx(f).reshape(-1, 3, 3)
    assert np.abs(U @ np.conj(np.swapaxes(U, 1, 2)) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-13


def test_local_q_e_closed_form_and_sums():
    """local_q_e_clover per site: on a constant abelian field every site carries
    q = Q/V and e = E (closed forms); on a random field the site values sum to
    the lattice totals of topo_q_e_clover."""
    L = (4, 6, 4, 8)
    f, phis = fields.abelian_flux(L, [1, 2, -1, 1, 3, 2])
    P, E, Q = fields.abelian_expected(L, phis)
    V = int(np.prod(L))
    for site in (0, 7, V // 2 + 3, V - 1):
        q, e = oracle.local_q_e(f, L, site)
        assert abs(q - Q / V) < 1e-14 * max(1.0, abs(Q / V)) and abs(e - E) < 1e-13 * E
    L = (4, 4, 4, 4)
    g = oracle.hot_start(L, 3)
    Qt, Et = oracle.q_e_clover(g, L)
    qe = np.array([oracle.local_q_e(g, L, s) for s in range(256)])
    assert abs(qe[:, 0].sum() - Qt) < 1e-12 and abs(qe[:, 1].mean() - Et) < 1e-12 * Et
