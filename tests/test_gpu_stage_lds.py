"""The RK3 stage kernel with the tile's shared links staged in LDS
(QCBF_TUNE_LDS_STAGE = 1) against the plain stage kernel, in both storage
forms: the same products on the same words in the same order, so fields and
measured series must be bit-identical.  Lattices with whole x-rows per
wavefront (x-neighbours read from LDS) and without (Lx not dividing 64: the
same-site links are still staged), and one whose volume is not a multiple of
64 (the plain kernel runs)."""
import numpy as np
import pytest

import oracle
import qcbflow

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
@pytest.mark.parametrize("L", [(8, 8, 8, 8), (16, 8, 4, 8), (32, 4, 4, 4), (64, 2, 2, 4),
                               (12, 8, 8, 8), (6, 4, 8, 4), (3, 5, 4, 7)])
def test_lds_stage_bit_identical(L, prec):
    say(f"LDS stage vs plain stage, lattice {L}, precision {prec}")
    f = oracle.hot_start(L, 99)
    out = {}
    for lds in (0, 1):
        with qcbflow.GradientFlow(L, 0.02, prec) as gf:
            gf.tune(qcbflow.TUNE_LDS_STAGE, lds)
            assert gf.get_tune(qcbflow.TUNE_LDS_STAGE) == lds
            gf.copy(f)
            gf.rk3_step(3)
            fld = gf.field()
            tqe, plaq = gf.flow_measure(4, 2)
            out[lds] = (fld, tqe, plaq, gf.field())
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_lds_stage_matches_oracle():
    L = (8, 8, 8, 8)
    f = oracle.hot_start(L, 5)
    ref_tqe, ref_plaq, _ = oracle.flow_measure(f.copy(), L, 0.02, 6, 2)
    with qcbflow.GradientFlow(L, 0.02) as gf:
        gf.tune(qcbflow.TUNE_LDS_STAGE, 1)
        gf.copy(f)
        tqe, plaq = gf.flow_measure(6, 2)
    assert np.max(np.abs(plaq - ref_plaq) / ref_plaq) < 1e-12
    assert np.max(np.abs(tqe[:, 2] - ref_tqe[:, 2]) / ref_tqe[:, 2]) < 1e-12
    assert np.max(np.abs(tqe[:, 1] - ref_tqe[:, 1])) < 1e-10


def test_lds_stage_tune_rejects_bad_value():
    with qcbflow.GradientFlow((4, 4, 4, 4)) as gf:
        with pytest.raises(qcbflow.QcbfError):
            gf.tune(qcbflow.TUNE_LDS_STAGE, 2)
