"""FP32-mixed against FP64 at flow time t = 2 on the 48^4 precision workload
and the 64^4 large-lattice workload.

Gates (FP32-mixed vs the FP64 GPU path on the same input, same cadence):
|dE/E| < 1e-6 and |dQ| < 1e-4 at t = 2.00.  On 48^4 also the flow times
where t^2 E reaches 0.15 and 0.3 (E in the reference's normalisation,
(1/2) sum |F|^2 / V, so 0.15 is the customary t0) agree to 1e-5.  The inputs
are thermalised on the device: Haar hot start, then checkerboard heatbath
sweeps."""
import numpy as np
import pytest

import qcbflow

pytestmark = pytest.mark.gpu

TOL_DE_REL = 1e-6
TOL_DQ_ABS = 1e-4
TOL_T0 = 1e-5


def say(msg):
    print(msg, flush=True)


def t_at(tqe, target):
    """Flow time where t^2 E(t) first reaches target (linear interpolation)."""
    t, E = tqe[:, 0], tqe[:, 2]
    y = t * t * E
    k = np.nonzero(y >= target)[0]
    if len(k) == 0 or k[0] == 0:
        return None
    k = k[0]
    return t[k - 1] + (target - y[k - 1]) * (t[k] - t[k - 1]) / (y[k] - y[k - 1])


def flows(L, beta, seed, sweeps, n_steps, every):
    """(tQE, plaquette) of the FP64 and of the FP32-mixed flow of one
    thermalised input, n_steps RK3 steps measured every `every` steps."""
    with qcbflow.GradientFlow(L, 0.02) as g64, \
            qcbflow.GradientFlow(L, 0.02, qcbflow.FP32_MIXED) as g32:
        g64.hot_start(seed)
        g64.heatbath(beta, sweeps, seed)
        g32.copy_from(g64)
        say(f"  thermalised: plaquette {g64.plaquette():.6f}")
        r32 = g32.flow_measure(n_steps // every + 1, every)
        du, dd = g32.unitarity()
        r64 = g64.flow_measure(n_steps // every + 1, every)
        assert g64.exp_warnings() == 0 and g32.exp_warnings() == 0
    say(f"  FP32-mixed field after the flow: unitarity {du:.1e}, det {dd:.1e}")
    assert du < 1e-5 and dd < 1e-5
    return r64, r32


def gate(r64, r32, t_gate=2.0):
    (t64, p64), (t32, p32) = r64, r32
    assert np.array_equal(t64[:, 0], t32[:, 0])
    k = int(np.argmin(np.abs(t64[:, 0] - t_gate)))
    assert t64[k, 0] == pytest.approx(t_gate, abs=1e-12)
    dE = np.abs(t32[:, 2] - t64[:, 2]) / np.abs(t64[:, 2])
    dQ = np.abs(t32[:, 1] - t64[:, 1])
    dP = np.abs(p32 - p64) / np.abs(p64)
    say(f"  t={t64[k, 0]:.2f}: E={t64[k, 2]:.10g} dE/E={dE[k]:.2e}  Q={t64[k, 1]:.6f} "
        f"dQ={dQ[k]:.2e}  dP/P={dP[k]:.2e}; max over the flow: dE/E={dE.max():.2e} "
        f"dQ={dQ.max():.2e}")
    assert dE[k] < TOL_DE_REL and dQ[k] < TOL_DQ_ABS
    return dE, dQ, dP


def test_precision_gates_48x48x48x48():
    """48^4, beta = 6.2, seed 11, 100 heatbath sweeps; 250 RK3 steps measured
    after every step (t = 0 .. 5)."""
    L = (48, 48, 48, 48)
    say("48^4 beta 6.2 seed 11: FP64 and FP32-mixed, 250 steps")
    r64, r32 = flows(L, 6.2, 11, 100, 250, 1)
    gate(r64, r32)
    for thr in (0.15, 0.3):
        a, b = t_at(r64[0], thr), t_at(r32[0], thr)
        say(f"  t^2 E = {thr}: t = {a} (FP64), {b} (FP32-mixed)")
        assert a is not None and b is not None
        assert abs(a - b) < TOL_T0


def test_precision_gates_64x64x64x64():
    """64^4, beta = 6.4, seed 2026, 50 heatbath sweeps; 100 RK3 steps measured
    every 10 steps (t = 0, 0.2, .., 2.0)."""
    L = (64, 64, 64, 64)
    say("64^4 beta 6.4 seed 2026: FP64 and FP32-mixed, 100 steps measured every 10")
    r64, r32 = flows(L, 6.4, 2026, 50, 100, 10)
    gate(r64, r32)
