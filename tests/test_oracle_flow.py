"""Pins the oracle's RK3 step (src/flow/gradient_flow.cpp:40-135) and the
topo_charge_flowed cadence (src/observables/observables_mpi.cpp:159-186)."""
import json
import os

import numpy as np

import fields
import oracle


def thermal(L, seed, sweeps=5, beta=6.0):
    f = oracle.hot_start(L, seed)
    return oracle.heatbath(f, L, beta, sweeps, seed + 1)


def test_rk3_is_third_order():
    """Global error at fixed flow time scales as eps^3: a wrong RK3 coefficient
    (e.g. 17/36 -> 1/2 or 3/4 -> 1) drops the order and fails the ratio."""
    L = (4, 4, 4, 4)
    f0 = thermal(L, 21)
    T = 0.16

    def flow(eps):
        f = f0.copy()
        oracle.rk3(f, L, eps, int(round(T / eps)))
        return f

    ref = flow(0.00125)
    e = [np.abs(flow(eps) - ref).max() for eps in (0.04, 0.02, 0.01)]
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 6.5 < r1 < 11.0 and 6.5 < r2 < 11.0, (e, r1, r2)


def test_flow_is_gauge_covariant():
    L = (4, 3, 4, 5)
    f = thermal(L, 31, sweeps=3)
    g = f.copy()
    G = oracle.gauge_transform(g, L, 5)
    oracle.rk3(f, L, 0.02, 3)
    oracle.rk3(g, L, 0.02, 3)
    want = fields.gauge_apply(f, L, G)
    assert np.abs(g - want).max() < 1e-12


def test_flow_keeps_links_in_su3():
    L = (4, 4, 4, 4)
    f = oracle.hot_start(L, 41)
    oracle.rk3(f, L, 0.02, 10)
    U = fields.to_cplx(f).reshape(-1, 3, 3)
    I = np.eye(3)
    assert np.abs(U @ np.conj(np.swapaxes(U, 1, 2)) - I).max() < 1e-13
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-13


def test_flow_fixed_points():
    L = (4, 4, 4, 4)
    f = oracle.cold_start(L)
    oracle.rk3(f, L, 0.02, 2)
    assert np.abs(f - oracle.cold_start(L)).max() < 1e-15
    fa, _ = fields.abelian_flux(L, [1, 0, 2, 1, 0, 1])
    g = fa.copy()
    oracle.rk3(g, L, 0.02, 3)
    assert np.abs(g - fa).max() < 1e-13


def test_flow_smooths_energy_monotonically():
    """dE/dt <= 0 along the Wilson flow (the flow is the gradient flow of the
    Wilson action): E_ref decreases at every step of a thermalised config."""
    L = (4, 4, 4, 4)
    f = thermal(L, 51)
    tqe, plaq, _ = oracle.flow_measure(f, L, 0.02, 15, 1)
    assert np.all(np.diff(tqe[:, 2]) < 0)
    assert np.all(np.diff(plaq) > 0)


def test_flow_measure_cadence():
    """N_steps_gf measurements, (N_steps_gf - 1) * N_rk_steps RK3 steps; t is the
    running sum of eps; the first point is the unflowed field; input unchanged."""
    L = (4, 4, 4, 4)
    f = thermal(L, 61, sweeps=2)
    keep = f.copy()
    tqe, plaq, out = oracle.flow_measure(f, L, 0.02, 4, 3, want_field=True)
    assert np.array_equal(f, keep)
    t = 0.0
    ts = [0.0]
    for _ in range(3):
        for _ in range(3):
            t += 0.02
        ts.append(t)
    assert np.array_equal(tqe[:, 0], np.array(ts))
    Q0, E0 = oracle.q_e_clover(f, L)
    assert tqe[0, 1] == Q0 and tqe[0, 2] == E0
    g = f.copy()
    oracle.rk3(g, L, 0.02, 9)
    assert np.array_equal(g, out)
    assert oracle.plaquette(g, L) == plaq[-1]


def test_golden_trajectory_regression(golden_dir):
    """The oracle reproduces the committed 4^4 trajectory (from the committed
    ILDG fixture) bit for bit."""
    import limefile

    with open(os.path.join(golden_dir, "flow_4x4x4x4.json")) as fp:
        g = json.load(fp)
    links = limefile.read_ildg(os.path.join(golden_dir, "conf_4x4x4x4.ildg"))
    L = tuple(g["lattice"])
    f = np.ascontiguousarray(links.reshape(-1, 4, 3, 3, 2))
    tqe, plaq, _ = oracle.flow_measure(f, L, g["eps"], g["N_steps_gf"], g["N_rk_steps"])
    for k, key in enumerate(["t", "Q", "E"]):
        want = np.array([float(v) for v in g[key]])
        assert np.allclose(tqe[:, k], want, rtol=1e-14, atol=1e-15), key
    assert np.allclose(plaq, np.array([float(v) for v in g["P"]]), rtol=1e-14)
