"""Threads per site of the RK3 stage kernels (QCBF_TUNE_THREADS_PER_SITE: 4,
2 or 1).  With fewer than four, one wavefront updates two or four directions
of its sites in turn and a workgroup takes two or four tiles; which wavefront
updates a link must not change the link.  Flowed fields and measured series
are compared bit for bit with the default four threads per site, on lattices
whose tile count is not a multiple of the tiles per workgroup, with a partial
last tile, with x-rows inside a wavefront (FP64 lane shuffles), with box tiles
and with slab sweeps whose t-slices do not split into whole tile groups."""
import numpy as np
import pytest

import oracle
import qcbflow

pytestmark = pytest.mark.gpu

T = qcbflow.TUNE_THREADS_PER_SITE


def say(msg):
    print(msg, flush=True)


def field_for(L, seed):
    return oracle.heatbath(oracle.hot_start(L, seed), L, 6.0, 2, seed + 400)


# (extents, tile shape code or -1, (map mode, map window) or None for the default)
SINGLE_CASES = [
    ((8, 8, 8, 8), -1, None),             # x-rows in a wavefront; slab sweep
    ((6, 4, 5, 8), -1, None),             # 15 tiles: a last group of 3 (or 1) tiles
    ((3, 5, 4, 7), -1, (2, 1024)),        # 420 sites: partial last tile
    ((8, 8, 3, 4), -1, (3, 1)),           # 3 tiles per t-slice: groups straddle t
    ((8, 8, 2, 5), -1, (0, 64)),          # lane shuffles, 10 tiles: a last group of 2
    ((16, 8, 8, 12), qcbflow.tile_code(4, 4, 2, 2), (3, 2)),  # box tiles, slab sweep
    ((16, 8, 8, 12), qcbflow.tile_code(8, 4, 2, 1), (1, 64)),
]


def run_single(L, prec, shape, mapping, tps, f):
    with qcbflow.GradientFlow(L, 0.02, prec) as gf:
        gf.tune(qcbflow.TUNE_TILE_SHAPE, shape)
        if mapping is not None:
            gf.tune(qcbflow.TUNE_MAP_MODE, mapping[0])
            gf.tune(qcbflow.TUNE_MAP_WIN, mapping[1])
        gf.tune(T, tps)
        assert gf.get_tune(T) == tps
        gf.copy(f)
        tqe, plaq = gf.flow_measure(4, 2)  # fused stage 1 + stages 2, 3 at tps
        gf.rk3_step(2)                     # plain stage 1 at tps as well
        return tqe, plaq, gf.field()


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
@pytest.mark.parametrize("case", range(len(SINGLE_CASES)))
def test_single_lattice_threads_per_site_bit_identical(prec, case):
    L, shape, mapping = SINGLE_CASES[case]
    say(f"single lattice {L}, precision {prec}, tile shape {shape}, map {mapping}")
    f = field_for(L, 31 + case)
    ref = run_single(L, prec, shape, mapping, 4, f)
    for tps in (2, 1):
        say(f"  threads per site {tps}")
        got = run_single(L, prec, shape, mapping, tps, f)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), tps


ENS_DIMS = [(4, 4, 4, 4), (6, 4, 5, 8), (3, 5, 4, 7), (8, 8, 8, 8), (4, 6, 2, 4)]


def run_ensemble(prec, tps, bpc, graph, shape, fs):
    with qcbflow.Ensemble(ENS_DIMS, 0.02, prec) as ens:
        ens.use_graph(graph)
        ens.tune(qcbflow.TUNE_BLOCKS_PER_CU, bpc)
        ens.tune(qcbflow.TUNE_TILE_SHAPE, shape)
        ens.tune(T, tps)
        assert ens.get_tune(T) == tps
        for c, f in enumerate(fs):
            ens.copy(c, f)
        tqe, plaq = ens.flow_measure(4, 2)
        ens.rk3_step(1)
        return tqe, plaq, [ens.field(c) for c in range(len(ENS_DIMS))]


@pytest.mark.parametrize("prec", [qcbflow.FP64, qcbflow.FP32_MIXED])
@pytest.mark.parametrize("bpc,graph,shape", [(0, True, -1), (2, False, -1),
                                             (0, False, qcbflow.tile_code(4, 4, 2, 2))])
def test_ensemble_threads_per_site_bit_identical(prec, bpc, graph, shape):
    """Flat and persistent grids; a tile group may span two members; a box
    tile shape taken by the members it divides (4^4, 8^4), not by the others."""
    say(f"ensemble of {len(ENS_DIMS)}, precision {prec}, blocks/CU {bpc}, graph {graph}, "
        f"tile shape {shape}")
    fs = [field_for(L, 70 + k) for k, L in enumerate(ENS_DIMS)]
    ref = run_ensemble(prec, 4, bpc, graph, shape, fs)
    for tps in (2, 1):
        say(f"  threads per site {tps}")
        got = run_ensemble(prec, tps, bpc, graph, shape, fs)
        assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1]), tps
        for a, b in zip(ref[2], got[2]):
            assert np.array_equal(a, b), tps


def test_threads_per_site_arguments_and_autotune():
    L = (8, 8, 8, 8)
    with qcbflow.GradientFlow(L) as gf, qcbflow.Ensemble([L, (4, 4, 4, 4)]) as ens:
        for h in (gf, ens):
            assert h.get_tune(T) == 4  # the default
            for bad in (0, 3, 8, -1):
                with pytest.raises(qcbflow.QcbfError):
                    h.tune(T, bad)
            assert h.get_tune(T) == 4
        gf.hot_start(5)
        f = gf.field()
        gf.autotune(trial_steps=1)
        say(f"  single-lattice autotune: threads per site {gf.get_tune(T)}")
        assert gf.get_tune(T) in (1, 2, 4)
        assert np.array_equal(gf.field(), f)
        ens.copy(0, f)
        ens.hot_start(1, 6)
        ens.autotune(trial_steps=1)
        say(f"  ensemble autotune: threads per site {ens.get_tune(T)}")
        assert ens.get_tune(T) in (1, 2, 4)
        assert np.array_equal(ens.field(0), f)
