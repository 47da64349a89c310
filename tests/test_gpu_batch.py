"""Ensembles of uneven volumes flowed together (persistent stage kernels,
hipGraph-replayed steps, autotuned launch shape) against the single-lattice
path and the oracle."""
import numpy as np
import pytest

import oracle
import qcbflow

pytestmark = pytest.mark.gpu

DIMS = [(4, 4, 4, 4), (6, 4, 5, 8), (8, 8, 8, 8), (3, 5, 4, 7), (4, 4, 4, 4)]


def say(msg):
    print(msg, flush=True)


def fields_for(dims, sweeps=3):
    out = []
    for k, L in enumerate(dims):
        f = oracle.hot_start(L, 500 + k)
        out.append(oracle.heatbath(f, L, 6.0, sweeps, 900 + k))
    return out


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_single_lattice_path(prec, graph):
    """Each member of the ensemble gets bit-for-bit the series the
    single-lattice path gives (same per-link code, same reduction order),
    whatever the tile shape of the members it divides."""
    say(f"ensemble of {len(DIMS)} uneven lattices, precision {prec}, graph {graph}")
    fs = fields_for(DIMS)
    with qcbflow.Ensemble(DIMS, 0.02, prec) as ens:
        ens.use_graph(graph)
        ens.tune(qcbflow.TUNE_TILE_SHAPE, qcbflow.tile_code(4, 4, 2, 2) if graph else -1)
        for c, f in enumerate(fs):
            ens.copy(c, f)
        tqe, plaq = ens.flow_measure(9, 1)
        outs = [ens.field(c) for c in range(len(DIMS))]
    for c, (L, f) in enumerate(zip(DIMS, fs)):
        with qcbflow.GradientFlow(L, 0.02, prec) as gf:
            gf.copy(f)
            t1, p1 = gf.flow_measure(9, 1)
            assert np.array_equal(tqe[c], t1), c
            assert np.array_equal(plaq[c], p1), c
            assert np.array_equal(outs[c], gf.field()), c


def test_batch_fp64_against_oracle():
    fs = fields_for(DIMS)
    with qcbflow.Ensemble(DIMS, 0.02) as ens:
        ens.use_graph(True)
        for c, f in enumerate(fs):
            ens.copy(c, f)
        tqe, plaq = ens.flow_measure(11, 2)
    for c, (L, f) in enumerate(zip(DIMS, fs)):
        ref, refp, _ = oracle.flow_measure(f, L, 0.02, 11, 2)
        assert np.array_equal(tqe[c][:, 0], ref[:, 0])
        assert np.max(np.abs(tqe[c][:, 2] - ref[:, 2]) / ref[:, 2]) < 1e-12
        assert np.max(np.abs(tqe[c][:, 1] - ref[:, 1])) < 1e-10
        assert np.max(np.abs(plaq[c] - refp) / refp) < 1e-12


def test_graph_replays_match_eager_over_many_steps():
    """Graph replays interleaved with eager work and synchronisations give the
    eager results at every step."""
    dims = [(8, 8, 8, 8), (12, 12, 12, 12), (4, 6, 8, 10)]
    with qcbflow.Ensemble(dims) as a, qcbflow.Ensemble(dims) as b:
        for c in range(len(dims)):
            a.hot_start(c, 40 + c)
            b.hot_start(c, 40 + c)
        b.use_graph(True)
        for rep in range(3):
            ta, pa = a.flow_measure(6, 1)
            tb, pb = b.flow_measure(6, 1)
            assert np.array_equal(ta, tb) and np.array_equal(pa, pb), rep
            a.measure()
            b.measure()
            b.synchronize()


def test_autotune_leaves_the_ensemble_unchanged():
    dims = [(8, 8, 8, 8), (12, 12, 12, 12)]
    with qcbflow.Ensemble(dims) as ens:
        for c in range(len(dims)):
            ens.hot_start(c, 7 + c)
        before = [ens.field(c) for c in range(len(dims))]
        m, w, bpc = ens.autotune(trial_steps=1)
        say(f"  autotune picked mode {m} win {w} blocks/CU {bpc}")
        assert ens.get_tune(qcbflow.TUNE_MAP_MODE) == m
        for c in range(len(dims)):
            assert np.array_equal(ens.field(c), before[c])
        t1, _ = ens.flow_measure(3, 1)
    with qcbflow.Ensemble(dims) as ref:
        for c in range(len(dims)):
            ref.hot_start(c, 7 + c)
        t0, _ = ref.flow_measure(3, 1)
    assert np.array_equal(t0, t1)


def test_autotune_keeps_the_exp_warning_state():
    """The autotuners' trial steps may raise the exp warning; the caller's
    flag is what it was before tuning, set or clear."""
    L = (8, 8, 8, 8)
    with qcbflow.GradientFlow(L, 0.5) as gf, qcbflow.Ensemble([L], 0.5) as ens:
        gf.hot_start(1)
        ens.copy_from(0, gf)
        gf.autotune(trial_steps=1)       # eps = 0.5 on a hot start warns in the trials
        ens.autotune(trial_steps=1)
        assert gf.exp_warnings() == 0 and ens.exp_warnings() == 0
        gf.rk3_step(1)
        ens.rk3_step(1)
        gf.autotune(trial_steps=1)
        ens.autotune(trial_steps=1)
        assert gf.exp_warnings() == 1 and ens.exp_warnings() == 1


def test_single_lattice_autotune_and_tile_maps_agree():
    L = (16, 8, 8, 12)
    with qcbflow.GradientFlow(L) as gf:
        gf.hot_start(3)
        f = gf.field()
        m, w = gf.autotune(trial_steps=1)
        assert np.array_equal(gf.field(), f)
        results = []
        shapes = [-1, qcbflow.tile_code(4, 4, 2, 2), qcbflow.tile_code(8, 4, 2, 1),
                  qcbflow.tile_code(8, 8, 1, 1), qcbflow.tile_code(2, 2, 4, 4)]
        for mode, win, shape in [(0, 64, -1), (1, 64, shapes[1]), (2, 64, shapes[2]),
                                 (2, 1024, shapes[3]), (0, 64, shapes[4]), (3, 2, -1),
                                 (3, 1, shapes[1]), (3, 3, shapes[2]), (2, 20, -1)]:
            gf.tune(qcbflow.TUNE_MAP_MODE, mode)
            gf.tune(qcbflow.TUNE_MAP_WIN, win)
            gf.tune(qcbflow.TUNE_TILE_SHAPE, shape)
            assert gf.get_tune(qcbflow.TUNE_TILE_SHAPE) == shape
            gf.copy(f)
            results.append(gf.flow_measure(4, 1))
        with pytest.raises(qcbflow.QcbfError):  # 32 sites along x do not tile a 16-site extent
            gf.tune(qcbflow.TUNE_TILE_SHAPE, qcbflow.tile_code(32, 2, 1, 1))
    for r in results[1:]:
        assert np.array_equal(r[0], results[0][0]) and np.array_equal(r[1], results[0][1])


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
def test_batch_copy_from_single_lattice_handles(prec):
    """Ensemble.copy_from (device to device, FP64 source) equals a host copy."""
    dims = [(4, 4, 4, 4), (6, 4, 8, 4)]
    fs = fields_for(dims, sweeps=1)
    with qcbflow.Ensemble(dims, 0.02, prec) as dev, qcbflow.Ensemble(dims, 0.02, prec) as host:
        for c, (L, f) in enumerate(zip(dims, fs)):
            with qcbflow.GradientFlow(L) as gf:
                gf.copy(f)
                dev.copy_from(c, gf)
            host.copy(c, f)
            assert np.array_equal(dev.field(c), host.field(c)), c
        with qcbflow.GradientFlow(dims[0]) as gf:
            with pytest.raises(qcbflow.QcbfError):
                dev.copy_from(1, gf)


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
def test_batch_fused_measurement_is_bit_identical(prec):
    """Measurements riding on the next stage 1 (graphs on and off) give the
    series of separate measurement kernels, on the persistent grid and with
    one workgroup per tile."""
    dims = [(4, 4, 4, 4), (3, 5, 4, 7), (8, 8, 8, 8)]
    fs = fields_for(dims, sweeps=1)
    out = []
    for fused, graph, bpc in [(1, True, 2), (0, True, 2), (1, False, 2), (0, False, 2),
                              (1, True, 0), (0, False, 0), (1, False, 3)]:
        with qcbflow.Ensemble(dims, 0.02, prec) as ens:
            ens.tune(qcbflow.TUNE_FUSED_MEASURE, fused)
            ens.tune(qcbflow.TUNE_BLOCKS_PER_CU, bpc)
            ens.use_graph(graph)
            for c, f in enumerate(fs):
                ens.copy(c, f)
            out.append(ens.flow_measure(7, 1) + tuple(ens.field(c) for c in range(len(dims))))
    for r in out[1:]:
        for x, y in zip(out[0], r):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
def test_batch_lds_stage_is_bit_identical(prec):
    """The ensemble stage kernel with each tile's shared links staged in LDS
    (QCBF_TUNE_LDS_STAGE = 1, the default where every member is whole 64-site
    tiles on the one-workgroup-per-tile grid) against the plain one, graphs on
    and off, and against the single-lattice path; an ensemble with a member
    that is not whole tiles runs the plain kernel."""
    for dims in ([(4, 4, 4, 4), (8, 8, 8, 8), (12, 4, 4, 4), (6, 4, 8, 4)],
                 [(4, 4, 4, 4), (3, 5, 4, 7)]):
        fs = fields_for(dims, sweeps=1)
        out = []
        for lds, graph in [(1, True), (0, True), (1, False), (0, False)]:
            with qcbflow.Ensemble(dims, 0.02, prec) as ens:
                ens.tune(qcbflow.TUNE_BLOCKS_PER_CU, 0)
                ens.tune(qcbflow.TUNE_LDS_STAGE, lds)
                assert ens.get_tune(qcbflow.TUNE_LDS_STAGE) == lds
                ens.use_graph(graph)
                for c, f in enumerate(fs):
                    ens.copy(c, f)
                ens.rk3_step(2)
                out.append(ens.flow_measure(5, 1) + tuple(ens.field(c) for c in range(len(dims))))
        for r in out[1:]:
            for x, y in zip(out[0], r):
                assert np.array_equal(x, y)
        for c, (L, f) in enumerate(zip(dims, fs)):
            with qcbflow.GradientFlow(L, 0.02, prec) as gf:
                gf.copy(f)
                gf.rk3_step(2)
                tqe, plaq = gf.flow_measure(5, 1)
            assert np.array_equal(tqe, out[0][0][c]) and np.array_equal(plaq, out[0][1][c])
    with qcbflow.Ensemble([(4, 4, 4, 4)], 0.02, prec) as ens:
        with pytest.raises(qcbflow.QcbfError):
            ens.tune(qcbflow.TUNE_LDS_STAGE, 2)


def test_batch_errors():
    with pytest.raises(qcbflow.QcbfError) as ei:
        qcbflow.Ensemble([(4, 4, 4, 4), (4, 1, 4, 4)])
    assert ei.value.code == qcbflow.QCBF_ERR_ARG
    with qcbflow.Ensemble([(4, 4, 4, 4)]) as ens:
        with pytest.raises(qcbflow.QcbfError):
            ens.hot_start(3, 1)
        with pytest.raises(qcbflow.QcbfError):
            ens.tune(qcbflow.TUNE_BLOCKS_PER_CU, 17)


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
def test_full_ensemble_configuration_equals_single_lattices(prec):
    """The 64-configuration ensemble (16 each of 8^4, 12^4, 16^4, 24^4; 6.75M
    sites), device hot starts and two heatbath sweeps each, flowed together
    with graph-replayed steps and measured every step: every member's series
    equals its single-lattice series bit for bit (the 8^4 and 16^4 members
    take the FP64 x-row lane shuffles there, the ensemble kernels load)."""
    say(f"64-configuration ensemble, precision {prec}")
    dims = [((8, 12, 16, 24)[c // 16],) * 4 for c in range(64)]
    singles = []
    with qcbflow.Ensemble(dims, 0.02, prec) as ens:
        for c, L in enumerate(dims):
            src = qcbflow.GradientFlow(L, 0.02)
            src.hot_start(100 + c)
            src.heatbath((5.8, 6.0, 6.2)[c % 3], 2, 300 + c)
            ens.copy_from(c, src)
            gf = qcbflow.GradientFlow(L, 0.02, prec)
            gf.copy_from(src)
            src.close()
            singles.append(gf)
        ens.use_graph(True)
        tqe, plaq = ens.flow_measure(4, 1)
    try:
        for c, gf in enumerate(singles):
            t1, p1 = gf.flow_measure(4, 1)
            assert np.array_equal(tqe[c], t1), c
            assert np.array_equal(plaq[c], p1), c
    finally:
        for gf in singles:
            gf.close()


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("n_gf,n_rk", [(1, 1), (4, 0), (5, 3)])
def test_batch_cadences_equal_single_lattices(graph, n_gf, n_rk):
    """The ensemble's flow_measure at the cadence edges (one measurement, no
    steps between measurements, three between), eager and graph-replayed,
    against each member's single-lattice series."""
    dims = [(4, 4, 4, 4), (3, 5, 4, 7), (8, 8, 8, 8)]
    say(f"ensemble cadence N_steps_gf={n_gf} N_rk_steps={n_rk} graph {graph}")
    fs = fields_for(dims, sweeps=1)
    with qcbflow.Ensemble(dims, 0.02) as ens:
        ens.use_graph(graph)
        for c, f in enumerate(fs):
            ens.copy(c, f)
        tqe, plaq = ens.flow_measure(n_gf, n_rk)
        outs = [ens.field(c) for c in range(len(dims))]
    for c, (L, f) in enumerate(zip(dims, fs)):
        with qcbflow.GradientFlow(L, 0.02) as gf:
            gf.copy(f)
            t1, p1 = gf.flow_measure(n_gf, n_rk)
            assert np.array_equal(tqe[c], t1), c
            assert np.array_equal(plaq[c], p1), c
            assert np.array_equal(outs[c], gf.field()), c


def test_tile_shape_change_waits_for_queued_steps():
    """qcbf_batch_tune(TILE_SHAPE) right after asynchronous RK3 steps: the
    configuration descriptors are rewritten only after the queued kernels have
    read them, so the result equals the synchronised sequence bit for bit."""
    dims = [(8, 8, 8, 8), (4, 6, 8, 10), (16, 8, 8, 8)]
    fs = fields_for(dims, sweeps=1)
    outs = []
    for sync in (True, False):
        say(f"tile shape change after queued steps, synchronised first: {sync}")
        with qcbflow.Ensemble(dims, 0.02) as ens:
            for c, f in enumerate(fs):
                ens.copy(c, f)
            ens.rk3_step(6)
            if sync:
                ens.synchronize()
            ens.tune(qcbflow.TUNE_TILE_SHAPE, qcbflow.tile_code(4, 2, 4, 2))
            ens.rk3_step(2)
            outs.append([ens.field(c) for c in range(len(dims))])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
