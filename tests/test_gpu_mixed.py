"""FP32-mixed precision form against the FP64 path.

Links are stored as two FP32 rows, flow arithmetic is FP32, the
Cayley-Hamilton coefficients, the row re-unitarisation and the sums are FP64.
Gates at flow time t = 2.0 against the FP64 path: |dE/E| < 1e-6 and
|dQ| < 1e-4; the plaquette and t0 (t^2 E = 0.3, with the reference's E)
deviations are printed."""
import numpy as np
import pytest

import fields
import oracle
import qcbflow

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def thermal(L, seed, sweeps, beta=6.0):
    f = oracle.hot_start(L, seed)
    return oracle.heatbath(f, L, beta, sweeps, seed + 0x5EED)


def t0_of(tqe, target=0.3):
    """Linear interpolation of t^2 E(t) = target (None if not reached)."""
    t, E = tqe[:, 0], tqe[:, 2]
    y = t * t * E
    k = np.nonzero(y >= target)[0]
    if len(k) == 0 or k[0] == 0:
        return None
    k = k[0]
    return t[k - 1] + (target - y[k - 1]) * (t[k] - t[k - 1]) / (y[k] - y[k - 1])


def test_mixed_roundtrip_is_su3_to_fp32():
    L = (8, 8, 8, 8)
    f = oracle.hot_start(L, 3)
    with qcbflow.GradientFlow(L, precision=qcbflow.FP32_MIXED) as gf:
        gf.copy(f)
        g = gf.field()
        du, dd = gf.unitarity()
    assert np.abs(g - f).max() < 2e-7
    assert du < 1e-6 and dd < 1e-6


def test_mixed_upload_reunitarises_scaled_links():
    """The FP32-mixed upload re-unitarises the two stored rows (FP64 row
    Gram-Schmidt with a Newton-refined reciprocal square root): links scaled
    by 1.5 and 0.25 (squared row norms 2.25 and 0.0625, far from 1) come back
    as the SU(3) links they were scaled from, to FP32 rounding."""
    L = (4, 4, 4, 8)
    f = oracle.hot_start(L, 11)
    for scale in (1.5, 0.25):
        with qcbflow.GradientFlow(L, precision=qcbflow.FP32_MIXED) as gf:
            gf.copy(f * scale)
            g = gf.field()
            du, dd = gf.unitarity()
        say(f"scale {scale}: max |g - f| {np.abs(g - f).max():.2e}, unitarity {du:.2e} {dd:.2e}")
        assert np.abs(g - f).max() < 2e-7
        assert du < 1e-6 and dd < 1e-6


def test_mixed_measure_closed_form():
    L = (4, 6, 4, 8)
    f, phis = fields.abelian_flux(L, [1, 2, -1, 1, 3, 2])
    P, E, Q = fields.abelian_expected(L, phis)
    with qcbflow.GradientFlow(L, precision=qcbflow.FP32_MIXED) as gf:
        gf.copy(f)
        p, e, q = gf.measure()
    assert abs(p - P) < 1e-6 and abs(e - E) < 1e-6 * E and abs(q - Q) < 1e-5 * max(1, abs(Q))


def test_mixed_unfused_matches_fused():
    L = (4, 4, 6, 4)
    f = thermal(L, 3, 3)
    with qcbflow.GradientFlow(L, 0.02, qcbflow.FP32_MIXED) as a, \
            qcbflow.GradientFlow(L, 0.02, qcbflow.FP32_MIXED) as b:
        a.copy(f)
        b.copy(f)
        a.rk3_step(1)
        b.compute_force(0)
        b.compute_w1()
        b.compute_force(1)
        b.compute_w2()
        b.replace_force_0()
        b.compute_force(1)
        b.compute_w3()
        assert np.abs(a.field() - b.field()).max() < 2e-6


def test_mixed_accuracy_gates_at_t2_12x12x12x12():
    """12^4, beta = 6.0, 30 heatbath sweeps from seed 11; 100 RK3 steps to
    t = 2.00, measuring every step, FP32-mixed vs FP64."""
    L = (12, 12, 12, 12)
    say("mixed vs fp64: thermalising 12^4 on the CPU")
    f = thermal(L, 11, 30)
    with qcbflow.GradientFlow(L, 0.02) as g64:
        g64.copy(f)
        ref, refp = g64.flow_measure(101, 1)
    with qcbflow.GradientFlow(L, 0.02, qcbflow.FP32_MIXED) as g32:
        g32.copy(f)
        got, gotp = g32.flow_measure(101, 1)
        du, _ = g32.unitarity()
    assert got[-1, 0] == pytest.approx(2.0, abs=1e-12)
    dE = np.abs(got[:, 2] - ref[:, 2]) / ref[:, 2]
    dQ = np.abs(got[:, 1] - ref[:, 1])
    dP = np.abs(gotp - refp) / refp
    t0a, t0b = t0_of(ref), t0_of(got)
    say(f"  t=2: E64={ref[-1, 2]:.10g} dE/E={dE[-1]:.2e} Q64={ref[-1, 1]:.6f} dQ={dQ[-1]:.2e} "
        f"dP/P={dP[-1]:.2e}; max over t: dE/E={dE.max():.2e} dQ={dQ.max():.2e}; "
        f"t0 {t0a} vs {t0b}; unitarity {du:.1e}")
    assert dE[-1] < 1e-6 and dQ[-1] < 1e-4
    if t0a is not None and t0b is not None:
        assert abs(t0a - t0b) < 1e-5


def test_mixed_ildg_load(golden_dir):
    import os
    import limefile

    path = os.path.join(golden_dir, "conf_4x4x4x4.ildg")
    want = limefile.read_ildg(path).reshape(-1, 4, 3, 3, 2)
    with qcbflow.GradientFlow((4, 4, 4, 4), precision=qcbflow.FP32_MIXED) as gf:
        gf.load_ildg(path)
        got = gf.field()
    assert np.abs(got - want).max() < 2e-7


def test_device_copies_between_handles_match_host_copies():
    """qcbf_copy_field converts on the device through the same kernels as a
    host round trip: FP64 -> FP64 is exact, FP64 -> FP32-mixed and back give
    what copying the host field gives."""
    L = (8, 6, 4, 10)
    f = thermal(L, 21, 2)
    with qcbflow.GradientFlow(L) as a, qcbflow.GradientFlow(L) as b, \
            qcbflow.GradientFlow(L, precision=qcbflow.FP32_MIXED) as m, \
            qcbflow.GradientFlow(L, precision=qcbflow.FP32_MIXED) as m_host, \
            qcbflow.GradientFlow(L) as back_host:
        a.copy(f)
        b.copy_from(a)
        assert np.array_equal(b.field(), f)
        m.copy_from(a)
        m_host.copy(f)
        assert np.array_equal(m.field(), m_host.field())
        b.copy_from(m)
        back_host.copy(m.field())
        assert np.array_equal(b.field(), back_host.field())
        assert np.array_equal(m.flow_measure(3, 1)[0], m_host.flow_measure(3, 1)[0])
        with qcbflow.GradientFlow((8, 6, 4, 8)) as other:
            with pytest.raises(qcbflow.QcbfError):
                other.copy_from(a)
