"""clover_site and local_q_e_clover of one site on the device
(qcbf_clover_site, qcbf_local_q_e) against the oracle, in both storage
forms, and their closed forms on a constant abelian field."""
import numpy as np
import pytest

import fields
import oracle
import qcbflow

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (3, 5, 4, 6), (8, 6, 4, 4)])
def test_clover_site_and_local_q_e_match_oracle(L):
    print(f"clover_site / local_q_e on {L}", flush=True)
    f = oracle.hot_start(L, 41)
    f = oracle.rk3(f, L, 0.02, 3)  # a smoother, non-trivial field
    V = int(np.prod(L))
    sites = [0, 1, V // 3, V - 1]
    with qcbflow.GradientFlow(L) as gf, \
            qcbflow.GradientFlow(L, precision=qcbflow.FP32_MIXED) as gm:
        gf.copy(f)
        gm.copy(f)
        for s in sites:
            for mu in range(4):
                for nu in range(4):
                    ref = oracle.clover(f, L, s, mu, nu) if mu != nu else np.zeros((3, 3))
                    assert np.abs(gf.clover_site(s, mu, nu) - ref).max() < 1e-14
                    assert np.abs(gm.clover_site(s, mu, nu) - ref).max() < 2e-6
            q, e = oracle.local_q_e(f, L, s)
            gq, ge = gf.local_q_e(s)
            assert abs(gq - q) < 1e-15 and abs(ge - e) < 1e-13 * e
            mq, me = gm.local_q_e(s)
            assert abs(mq - q) < 1e-6 and abs(me - e) < 1e-5 * e
        with pytest.raises(qcbflow.QcbfError):
            gf.clover_site(V, 0, 1)
        with pytest.raises(qcbflow.QcbfError):
            gf.local_q_e(-1)


def test_local_q_e_sums_to_the_lattice_totals():
    L = (4, 4, 6, 4)
    f = oracle.hot_start(L, 5)
    V = int(np.prod(L))
    with qcbflow.GradientFlow(L) as gf:
        gf.copy(f)
        Q, E = gf.q_e_clover()
        qe = np.array([gf.local_q_e(s) for s in range(V)])
    assert abs(qe[:, 0].sum() - Q) < 1e-12 and abs(qe[:, 1].mean() - E) < 1e-12 * E


def test_local_q_e_closed_form():
    L = (4, 6, 4, 8)
    f, phis = fields.abelian_flux(L, [1, 2, -1, 1, 3, 2])
    P, E, Q = fields.abelian_expected(L, phis)
    V = int(np.prod(L))
    with qcbflow.GradientFlow(L) as gf:
        gf.copy(f)
        for s in (0, V // 2, V - 1):
            q, e = gf.local_q_e(s)
            assert abs(q - Q / V) < 1e-14 * max(1.0, abs(Q / V)) and abs(e - E) < 1e-13 * E
