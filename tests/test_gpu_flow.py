"""GPU parity of the FP64 flow path against the oracle, through the C ABI.

Tolerances (FP64 GPU vs FP64 CPU oracle): plaquette and E within 1e-12
relative, Q within 1e-10 absolute, at every measured flow time."""
import json
import os

import numpy as np
import pytest

import fields
import oracle
import qcbflow

pytestmark = pytest.mark.gpu

TOL_REL_P = 1e-12
TOL_REL_E = 1e-12
TOL_ABS_Q = 1e-10


def say(msg):
    print(msg, flush=True)


def thermal(L, seed, sweeps, beta=6.0):
    f = oracle.hot_start(L, seed)
    return oracle.heatbath(f, L, beta, sweeps, seed + 0x5EED)


def check_traj(tqe, plaq, ref_tqe, ref_plaq):
    assert np.array_equal(tqe[:, 0], ref_tqe[:, 0])
    dP = np.max(np.abs(plaq - ref_plaq) / np.abs(ref_plaq))
    dE = np.max(np.abs(tqe[:, 2] - ref_tqe[:, 2]) / np.abs(ref_tqe[:, 2]))
    dQ = np.max(np.abs(tqe[:, 1] - ref_tqe[:, 1]))
    say(f"  max rel dP={dP:.2e} rel dE={dE:.2e} abs dQ={dQ:.2e}")
    assert dP < TOL_REL_P and dE < TOL_REL_E and dQ < TOL_ABS_Q


def test_device_present():
    name, cus = qcbflow.device_info()
    say(f"device: {name}, {cus} CUs")
    assert cus > 0


def test_exp_matches_scipy_golden_and_oracle(golden_dir):
    say("exp: 64 golden cases")
    with open(os.path.join(golden_dir, "expm_golden.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        Z = np.array(c["Z"]).reshape(3, 3, 2)
        Z = Z[..., 0] + 1j * Z[..., 1]
        E = np.array(c["expm"]).reshape(3, 3, 2)
        E = E[..., 0] + 1j * E[..., 1]
        got, flagged = qcbflow.su3_exp(Z[None], c["coeff"])
        assert not flagged
        assert np.abs(got[0] - E).max() < 2e-14
        assert np.abs(got[0] - oracle.exp_su3(Z[None], c["coeff"])[0]).max() < 1e-15


def test_exp_flags_large_argument():
    rng = np.random.default_rng(1)
    X = fields.random_algebra(rng, 4)
    X *= (2.0 / np.linalg.norm(X, axis=(1, 2)))[:, None, None]  # t = 2
    _, flagged = qcbf_flag(X)
    assert flagged


def qcbf_flag(X):
    return qcbflow.su3_exp(X, 1.0)


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (3, 5, 4, 6)])
def test_force_parity(L):
    say(f"force parity {L}")
    f = oracle.hot_start(L, 17)
    ref = oracle.compute_force(f, L)
    with qcbflow.GradientFlow(L) as gf:
        gf.copy(f)
        gf.compute_force(0)
        got = gf.force(0)
    assert np.abs(got - ref).max() < 1e-13


@pytest.mark.parametrize("L,n", [((4, 6, 4, 8), [1, 2, -1, 1, 3, 2]), ((5, 4, 6, 4), [1, 0, 2, 0, 1, 1])])
def test_measure_closed_forms(L, n):
    say(f"abelian closed form {L}")
    f, phis = fields.abelian_flux(L, n)
    P, E, Q = fields.abelian_expected(L, phis)
    with qcbflow.GradientFlow(L) as gf:
        gf.copy(f)
        p, e, q = gf.measure()
        assert abs(gf.plaquette() - P) < 1e-14
    assert abs(p - P) < 1e-14 and abs(e - E) < 1e-13 * max(1, E) and abs(q - Q) < 1e-11 * max(1, abs(Q))


def test_unfused_stages_equal_fused_step():
    """compute_force/w1/w2/replace_force_0/w3 in the reference's order give the
    same field as the fused rk3_step."""
    L = (4, 4, 6, 4)
    say(f"unfused vs fused {L}")
    f = thermal(L, 3, 3)
    with qcbflow.GradientFlow(L, 0.02) as a, qcbflow.GradientFlow(L, 0.02) as b:
        a.copy(f)
        b.copy(f)
        for _ in range(2):
            a.rk3_step()
            b.compute_force(0)
            b.compute_w1()
            b.compute_force(1)
            b.compute_w2()
            b.replace_force_0()
            b.compute_force(1)
            b.compute_w3()
        fa, fb = a.field(), b.field()
    ref = f.copy()
    oracle.rk3(ref, L, 0.02, 2)
    assert np.abs(fa - fb).max() < 1e-14
    assert np.abs(fa - ref).max() < 1e-13


def test_flow_parity_8x8x8x8_thermalised():
    """Config: 8^4, Haar hot start seed 1234, 20 heatbath sweeps at beta = 6.0;
    N_steps_gf = 100, N_rk_steps = 1, eps = 0.02 (99 RK3 steps, last t = 1.98)."""
    L = (8, 8, 8, 8)
    say("flow parity 8^4 x 99 RK3 steps")
    f = thermal(L, 1234, 20)
    ref_tqe, ref_plaq, ref_f = oracle.flow_measure(f, L, 0.02, 100, 1, want_field=True)
    with qcbflow.GradientFlow(L, 0.02) as gf:
        gf.copy(f)
        tqe, plaq = gf.flow_measure(100, 1)
        out = gf.field()
        assert gf.exp_warnings() == 0
    check_traj(tqe, plaq, ref_tqe, ref_plaq)
    say(f"  max |dU| = {np.abs(out - ref_f).max():.2e}")
    assert np.abs(out - ref_f).max() < 1e-12


# x-extents 2 and 4 divide 64: those lattices take the FP64 x-neighbour links
# from lane shuffles (when the volume fills whole 64-site tiles), the others
# load them.  Extent 2: the neighbours x + mu and x - mu are the same site.
@pytest.mark.parametrize("L", [(6, 4, 5, 8), (3, 5, 4, 7), (12, 4, 4, 4), (2, 4, 8, 4), (4, 6, 2, 4),
                               (2, 2, 2, 2), (2, 2, 2, 8)])
def test_flow_parity_anisotropic(L):
    say(f"flow parity {L}")
    f = thermal(L, 77, 5)
    ref_tqe, ref_plaq, _ = oracle.flow_measure(f, L, 0.02, 11, 2)
    with qcbflow.GradientFlow(L, 0.02) as gf:
        gf.copy(f)
        tqe, plaq = gf.flow_measure(11, 2)
    check_traj(tqe, plaq, ref_tqe, ref_plaq)


@pytest.mark.parametrize("n_gf,n_rk", [(1, 1), (4, 0), (5, 3), (0, 1)])
def test_flow_measure_cadences(n_gf, n_rk):
    """topo_charge_flowed's cadence (observables_mpi.cpp:159-186) at its
    edges: one measurement and no step, repeated measurements of an unflowed
    field, several RK3 steps between measurements (the measurement rides on
    the first one's stage 1), and an empty series."""
    L = (4, 4, 6, 4)
    say(f"cadence N_steps_gf={n_gf} N_rk_steps={n_rk}")
    f = thermal(L, 21, 2)
    ref_tqe, ref_plaq, ref_f = oracle.flow_measure(f, L, 0.02, n_gf, n_rk, want_field=True)
    with qcbflow.GradientFlow(L, 0.02) as gf:
        gf.copy(f)
        tqe, plaq = gf.flow_measure(n_gf, n_rk)
        out = gf.field()
    assert tqe.shape == (n_gf, 3) and plaq.shape == (n_gf,)
    if n_gf:
        check_traj(tqe, plaq, ref_tqe, ref_plaq)
    assert np.abs(out - ref_f).max() < 1e-12


def test_flow_is_deterministic():
    L = (8, 8, 8, 8)
    say("determinism 8^4")
    with qcbflow.GradientFlow(L, 0.02) as gf:
        gf.hot_start(5)
        f = gf.field()
        a, pa = gf.flow_measure(6, 1)
        gf.copy(f)
        b, pb = gf.flow_measure(6, 1)
    assert np.array_equal(a, b) and np.array_equal(pa, pb)


def test_hot_start_is_su3_and_haar_like():
    L = (8, 8, 8, 8)
    with qcbflow.GradientFlow(L) as gf:
        gf.hot_start(42)
        du, dd = gf.unitarity()
        f = gf.field()
        p = gf.plaquette()
    assert du < 1e-13 and dd < 1e-13
    # Haar: <Re tr U> = 0, <|tr U|^2> = 1; plaquette of random links ~ 0
    tr = np.trace(fields.to_cplx(f).reshape(-1, 3, 3), axis1=1, axis2=2)
    assert abs(np.mean(tr.real)) < 0.02 and abs(np.mean(np.abs(tr) ** 2) - 1) < 0.03
    assert abs(p) < 0.01


def test_exp_warning_flag_on_flow():
    L = (4, 4, 4, 4)
    with qcbflow.GradientFlow(L, 0.5) as gf:
        gf.hot_start(1)
        gf.rk3_step(1)
        assert gf.exp_warnings() == 1
        assert gf.exp_warnings() == 0  # reading resets
        # a warning not yet read survives other calls on the handle
        gf.hot_start(1)
        gf.rk3_step(1)
        gf.heatbath(6.0, 1, 3)
        gf.measure()
        assert gf.exp_warnings() == 1


def test_large_lattice_properties_32x32x32x32():
    """Size-independent properties at the full 32^4 workload: one RK3 step
    against the oracle, unitarity after 10 steps, translation and gauge
    invariance of P, E, Q."""
    L = (32, 32, 32, 32)
    say("32^4: hot start, 1 RK3 step vs oracle")
    with qcbflow.GradientFlow(L, 0.02) as gf:
        gf.hot_start(7)
        f = gf.field()
        gf.rk3_step(1)
        g1 = gf.field()
        p1, e1, q1 = gf.measure()
        ref = f.copy()
        oracle.rk3(ref, L, 0.02, 1)
        assert np.abs(g1 - ref).max() < 1e-13
        P, (Q, E) = oracle.plaquette(ref, L), oracle.q_e_clover(ref, L)
        assert abs(p1 - P) < 1e-12 * abs(P) and abs(e1 - E) < 1e-12 * E and abs(q1 - Q) < 1e-10
        say("32^4: 9 more steps, unitarity")
        gf.rk3_step(9)
        du, dd = gf.unitarity()
        assert du < 1e-13 and dd < 1e-13
        p0, e0, q0 = gf.measure()
        g = gf.field()
        say("32^4: translation")
        gf.copy(fields.translate(g, L, (3, 1, 4, 1)))
        p2, e2, q2 = gf.measure()
        assert abs(p2 - p0) < 1e-13 and abs(e2 - e0) < 1e-12 * e0 and abs(q2 - q0) < 1e-10
        say("32^4: gauge transformation")
        G = fields.random_su3(np.random.default_rng(3), int(np.prod(L)))
        gf.copy(fields.gauge_apply(g, L, G))
        p3, e3, q3 = gf.measure()
        assert abs(p3 - p0) < 1e-13 and abs(e3 - e0) < 1e-12 * e0 and abs(q3 - q0) < 1e-10


def test_lattice_properties_64x64x64x64():
    """The 64^4 configuration (16.8M sites, 23.6 GB per FP64 handle), on the
    device only: unitarity after two RK3 steps, P, E, Q unchanged by a
    periodic shift, the flow commuting with the shift bit for bit (every link
    is formed by the same code wherever it sits, across the periodic
    boundary too), and the FP32-mixed form within its gates."""
    L = (64, 64, 64, 64)
    s = (5, 17, 33, 63)
    say("64^4: hot start, 2 RK3 steps, unitarity")
    with qcbflow.GradientFlow(L, 0.02) as a, qcbflow.GradientFlow(L, 0.02) as b:
        a.hot_start(2026)
        b.copy_from(a)
        b.shift(s)
        a.rk3_step(2)
        du, dd = a.unitarity()
        assert du < 1e-13 and dd < 1e-13
        p0, e0, q0 = a.measure()
        say("64^4: flow of the shifted field")
        b.rk3_step(2)
        p1, e1, q1 = b.measure()
        assert abs(p1 - p0) < 1e-13 and abs(e1 - e0) < 1e-12 * e0 and abs(q1 - q0) < 1e-10
        say("64^4: shift of the flowed field, compared link by link")
        a.shift(s)
        assert np.array_equal(a.field(), b.field())
    say("64^4: FP32-mixed against FP64, 2 RK3 steps")
    with qcbflow.GradientFlow(L, 0.02) as a, qcbflow.GradientFlow(L, 0.02, qcbflow.FP32_MIXED) as m:
        a.hot_start(2026)
        m.copy_from(a)
        a.rk3_step(2)
        m.rk3_step(2)
        pa, ea, qa = a.measure()
        pm, em, qm = m.measure()
        say(f"  dE/E {abs(em - ea) / ea:.2e}, dQ {abs(qm - qa):.2e}, dP/P {abs(pm - pa) / pa:.2e}")
        assert abs(em - ea) < 1e-6 * ea and abs(qm - qa) < 1e-4


def test_lattice_of_120_gb_96x96x96x96():
    """A 96^4 lattice (84.9M sites, 2.7e9 link words past 2^31) on one GPU: an
    FP64 handle keeps 120 GB resident, an FP32-mixed one 43.5 GB more.  On the
    device only (a host copy of the field would be 49 GB): unitarity after two
    RK3 steps, P, E, Q unchanged by a periodic shift, the flow of the shifted
    field measuring as the flow of the field, and the FP32-mixed gates."""
    L = (96, 96, 96, 96)
    s = (95, 7, 48, 33)
    with qcbflow.GradientFlow(L, 0.02) as a:
        say("96^4: hot start, P, E, Q before and after a periodic shift")
        a.hot_start(96)
        m0 = a.measure()
        a.shift(s)
        m0s = a.measure()
        say(f"  P {m0[0]:.15f} E {m0[1]:.12g} Q {m0[2]:.6g}; shifted: dP {abs(m0s[0] - m0[0]):.1e} "
            f"dE {abs(m0s[1] - m0[1]):.1e} dQ {abs(m0s[2] - m0[2]):.1e}")
        # a shift regroups the site sums: rounding only (Q sums 85M rough terms)
        assert abs(m0s[0] - m0[0]) < 1e-13 and abs(m0s[1] - m0[1]) < 1e-12 * m0[1]
        assert abs(m0s[2] - m0[2]) < 1e-8
        say("96^4: 2 RK3 steps, unitarity")
        a.hot_start(96)
        a.rk3_step(2)
        du, dd = a.unitarity()
        assert du < 1e-13 and dd < 1e-13
        m1 = a.measure()
        say("96^4: 2 RK3 steps of the shifted field")
        a.hot_start(96)
        a.shift(s)
        a.rk3_step(2)
        m2 = a.measure()
        say(f"  P {m1[0]:.15f} E {m1[1]:.12g} Q {m1[2]:.6g}; shifted: dP {abs(m2[0] - m1[0]):.1e} "
            f"dE {abs(m2[1] - m1[1]):.1e} dQ {abs(m2[2] - m1[2]):.1e}")
        assert abs(m2[0] - m1[0]) < 1e-13 and abs(m2[1] - m1[1]) < 1e-12 * m1[1]
        assert abs(m2[2] - m1[2]) < 1e-8
        say("96^4: FP32-mixed against FP64, 2 RK3 steps")
        with qcbflow.GradientFlow(L, 0.02, qcbflow.FP32_MIXED) as m:
            a.hot_start(96)
            m.copy_from(a)
            a.rk3_step(2)
            m.rk3_step(2)
            pa, ea, qa = a.measure()
            pm, em, qm = m.measure()
            say(f"  dE/E {abs(em - ea) / ea:.2e}, dQ {abs(qm - qa):.2e}, dP/P {abs(pm - pa) / pa:.2e}")
            assert abs(em - ea) < 1e-6 * ea and abs(qm - qa) < 1e-4


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
def test_flow_commutes_with_shift_48x48x48x48(prec):
    """The 48^4 configuration (x-extent 48: the FP64 kernels load the
    x-neighbours), on the device: the flow of a shifted field is the shifted
    flow, link for link, in both storage forms, and P, E, Q are unchanged."""
    L = (48, 48, 48, 48)
    s = (47, 1, 24, 13)
    say(f"48^4 precision {prec}: flow and periodic shift")
    with qcbflow.GradientFlow(L, 0.02) as src, qcbflow.GradientFlow(L, 0.02, prec) as a, \
            qcbflow.GradientFlow(L, 0.02, prec) as b:
        src.hot_start(11)
        src.heatbath(6.2, 2, 11)
        a.copy_from(src)
        b.copy_from(src)
        b.shift(s)
        ta, pa = a.flow_measure(4, 1)
        tb, pb = b.flow_measure(4, 1)
        assert np.max(np.abs(pa - pb)) < 1e-13
        assert np.max(np.abs(ta[:, 2] - tb[:, 2]) / ta[:, 2]) < 1e-12
        assert np.max(np.abs(ta[:, 1] - tb[:, 1])) < 1e-10
        a.shift(s)
        assert np.array_equal(a.field(), b.field())


def _su3_defects(f):
    """max |U U^dag - 1| and max |det U - 1| over all links of a host field."""
    U = fields.to_cplx(f).reshape(-1, 3, 3)
    uu = np.einsum("nij,nkj->nik", U, U.conj()) - np.eye(3)
    return np.abs(uu).max(), np.abs(np.linalg.det(U) - 1).max()


def test_flow_output_leaves_device_whole():
    """The FP64 flow stores two rows per link; every path that hands the field
    out (get_links, ILDG save, device copies, ensembles) rebuilds the third,
    so what leaves the device is SU(3) to rounding and agrees with the oracle,
    and a field that was never flowed leaves bit-for-bit as it came in."""
    L = (6, 4, 4, 8)
    f = thermal(L, 77, 2)
    ref = f.copy()
    oracle.rk3(ref, L, 0.02, 3)
    with qcbflow.GradientFlow(L) as gf, qcbflow.GradientFlow(L) as other:
        gf.copy(f)
        assert np.array_equal(gf.field(), f)
        gf.rk3_step(3)
        out = gf.field()
        du, dd = _su3_defects(out)
        assert du < 1e-14 and dd < 1e-14
        assert np.abs(out - ref).max() < 1e-13
        gf.rk3_step(2)
        other.copy_from(gf)  # source completed on the device before the copy
        g = gf.field()
        assert np.array_equal(other.field(), g)
        assert max(_su3_defects(g)) < 1e-14
    with qcbflow.Ensemble([L, (4, 4, 4, 4)]) as ens:
        ens.copy(0, f)
        ens.rk3_step(3)
        e0 = ens.field(0)
        assert np.array_equal(e0, out)


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
@pytest.mark.parametrize("L", [(8, 8, 8, 8), (3, 5, 4, 7)])
def test_fused_measurement_is_bit_identical(prec, L):
    """qcbf_flow_measure measures inside the next RK3 stage 1; the series is
    bit-for-bit the one of separate measurement kernels and of measuring step
    by step (a lattice whose volume is not a multiple of 64 included)."""
    say(f"fused measurement {L} precision {prec}")
    f = thermal(L, 5, 2)
    with qcbflow.GradientFlow(L, 0.02, prec) as a, qcbflow.GradientFlow(L, 0.02, prec) as b, \
            qcbflow.GradientFlow(L, 0.02, prec) as c:
        for h in (a, b, c):
            h.copy(f)
        assert a.get_tune(qcbflow.TUNE_FUSED_MEASURE) == 1
        b.tune(qcbflow.TUNE_FUSED_MEASURE, 0)
        ta, pa = a.flow_measure(6, 1)
        tb, pb = b.flow_measure(6, 1)
        assert np.array_equal(ta, tb) and np.array_equal(pa, pb)
        assert np.array_equal(a.field(), b.field())
        for k in range(6):
            p, e, q = c.measure()
            assert (p, e, q) == (pa[k], ta[k, 2], ta[k, 1]), k
            if k < 5:
                c.rk3_step(1)
        assert np.array_equal(a.field(), c.field())
        ta2, pa2 = a.flow_measure(4, 2)  # measurement every second step
        tb2, pb2 = b.flow_measure(4, 2)
        assert np.array_equal(ta2, tb2) and np.array_equal(pa2, pb2)


def test_timing_samples_every_nth_launch():
    """qcbf_timing_enable(h, n): the first, (n+1)-th, ... launch of each kernel
    class is timed; n = 1 times all of them; a negative n is refused."""
    L = (8, 8, 8, 8)
    with qcbflow.GradientFlow(L) as gf:
        gf.hot_start(3)
        for n, want in ((1, 7), (3, 3), (4, 2), (8, 1)):
            gf.timing(n)
            gf.rk3_step(7)
            ms, launches = gf.timing_read()
            assert launches[:3] == [want] * 3, (n, launches)
            assert all(m > 0 for m in ms[:3])
        gf.timing(2)
        gf.flow_measure(5, 1)  # 4 fused stage-1 launches, stages 2 and 3, 1 measurement
        ms, launches = gf.timing_read()
        assert launches == [0, 2, 2, 1, 2], launches
        gf.timing(False)
        with pytest.raises(qcbflow.QcbfError):
            gf.timing(-1)
