"""ILDG load/save to HBM (read_ildg_clime / save_ildg_clime,
src/io/ildg.cpp:34-237) and the error paths of the C ABI."""
import json
import os

import numpy as np
import pytest

import limefile
import oracle
import qcbflow

pytestmark = pytest.mark.gpu

L4 = (4, 4, 4, 4)


def say(msg):
    print(msg, flush=True)


def fixture(golden_dir):
    return os.path.join(golden_dir, "conf_4x4x4x4.ildg")


def test_load_fixture_exact(golden_dir):
    say("load 4^4 ILDG fixture")
    want = limefile.read_ildg(fixture(golden_dir)).reshape(-1, 4, 3, 3, 2)
    with qcbflow.GradientFlow(L4) as gf:
        gf.load_ildg(fixture(golden_dir))
        got = gf.field()
    assert np.array_equal(got, want)


def test_golden_trajectory_from_ildg(golden_dir):
    """ILDG -> HBM -> 20 RK3 steps with 21 measurements, against the committed
    oracle trajectory (t exact; P, E 1e-12 relative; Q 1e-10)."""
    say("golden 4^4 trajectory")
    with open(os.path.join(golden_dir, "flow_4x4x4x4.json")) as fp:
        g = json.load(fp)
    with qcbflow.GradientFlow(L4, g["eps"]) as gf:
        gf.load_ildg(fixture(golden_dir))
        tqe, plaq = gf.flow_measure(g["N_steps_gf"], g["N_rk_steps"])
    t = np.array([float(v) for v in g["t"]])
    Q = np.array([float(v) for v in g["Q"]])
    E = np.array([float(v) for v in g["E"]])
    P = np.array([float(v) for v in g["P"]])
    assert np.array_equal(tqe[:, 0], t)
    assert np.max(np.abs(tqe[:, 2] - E) / E) < 1e-12
    assert np.max(np.abs(tqe[:, 1] - Q)) < 1e-10
    assert np.max(np.abs(plaq - P) / P) < 1e-12


def test_save_is_byte_identical_to_reference_framing(golden_dir, tmp_path):
    say("save 4^4 ILDG")
    out = str(tmp_path / "saved.ildg")
    with qcbflow.GradientFlow(L4) as gf:
        gf.load_ildg(fixture(golden_dir))
        gf.save_ildg(out)
    assert open(out, "rb").read() == open(fixture(golden_dir), "rb").read()
    recs = limefile.records(open(out, "rb").read())
    assert [(r[0], r[1], r[2]) for r in recs] == [("ildg-format", True, True),
                                                 ("ildg-binary-data", False, True)]


def test_roundtrip_anisotropic_and_standard_framing(tmp_path):
    L = (6, 4, 5, 3)
    say(f"ILDG round trip {L}")
    f = oracle.hot_start(L, 5)
    p = str(tmp_path / "std.ildg")
    limefile.write_ildg(p, f, L, framing="standard")
    with qcbflow.GradientFlow(L) as gf:
        gf.load_ildg(p)
        assert np.array_equal(gf.field(), f)
        q = str(tmp_path / "again.ildg")
        gf.save_ildg(q)
    assert np.array_equal(limefile.read_ildg(q).reshape(f.shape), f)


def test_roundtrip_across_transfer_chunks(tmp_path):
    """80640 sites: more than one 65536-site transfer chunk, the last one
    partial, through the double-buffered load, the chunked save, and
    set_links / get_links."""
    L = (20, 18, 16, 14)
    say(f"ILDG round trip {L} across chunks")
    f = oracle.hot_start(L, 8)
    p = str(tmp_path / "big.ildg")
    limefile.write_ildg(p, f, L, framing="standard")
    with qcbflow.GradientFlow(L) as gf:
        gf.load_ildg(p)
        assert np.array_equal(gf.field(), f)
        q = str(tmp_path / "big_again.ildg")
        gf.save_ildg(q)
        g = oracle.hot_start(L, 9)
        gf.copy(g)
        assert np.array_equal(gf.field(), g)
    assert np.array_equal(limefile.read_ildg(q).reshape(f.shape), f)
    # the FP32-mixed form converts per chunk on the device
    with qcbflow.GradientFlow(L, 0.02, qcbflow.FP32_MIXED) as gm:
        gm.load_ildg(p)
        assert np.abs(gm.field() - f).max() < 1e-6


def test_force_readback_across_transfer_chunks():
    """qcbf_get_force streams through the 65536-site staging chunk: 80640
    sites, the last chunk partial, both force slots, against the oracle's
    force (and force_1 equal to force_0 of the same field)."""
    L = (20, 18, 16, 14)
    say(f"force read-back {L} across chunks")
    f = oracle.hot_start(L, 31)
    ref = oracle.compute_force(f.copy(), L)
    with qcbflow.GradientFlow(L) as gf:
        gf.copy(f)
        gf.compute_force(0)
        gf.compute_force(1)
        f0, f1 = gf.force(0), gf.force(1)
    assert np.abs(f0 - ref).max() < 1e-13
    assert np.array_equal(f0, f1)


def test_load_single_precision_payload(tmp_path):
    L = (4, 4, 4, 4)
    f = oracle.hot_start(L, 6)
    p = str(tmp_path / "f32.ildg")
    limefile.write_ildg(p, f, L, precision=32)
    with qcbflow.GradientFlow(L) as gf:
        gf.load_ildg(p)
        got = gf.field()
    assert np.array_equal(got, f.astype(np.float32).astype(np.float64))


def _expect(code, fn):
    with pytest.raises(qcbflow.QcbfError) as ei:
        fn()
    assert ei.value.code == code, str(ei.value)


def test_ildg_error_paths(golden_dir, tmp_path):
    good = open(fixture(golden_dir), "rb").read()
    with qcbflow.GradientFlow(L4) as gf:
        _expect(qcbflow.QCBF_ERR_IO, lambda: gf.load_ildg(str(tmp_path / "missing")))
        p = tmp_path / "trunc.ildg"
        p.write_bytes(good[:-1000])
        _expect(qcbflow.QCBF_ERR_FORMAT, lambda: gf.load_ildg(str(p)))
        p = tmp_path / "magic.ildg"
        p.write_bytes(b"\0\0\0\0" + good[4:])
        _expect(qcbflow.QCBF_ERR_FORMAT, lambda: gf.load_ildg(str(p)))
        # no binary record: keep only the ildg-format record
        recs = limefile.records(good)
        end = recs[1][3] - 144
        p = tmp_path / "nobin.ildg"
        p.write_bytes(good[:end])
        _expect(qcbflow.QCBF_ERR_FORMAT, lambda: gf.load_ildg(str(p)))
        # binary size that does not match V*576
        f = oracle.hot_start((4, 4, 4, 2), 1)
        p = tmp_path / "small.ildg"
        limefile.write_ildg(str(p), f, (4, 4, 4, 4))
        _expect(qcbflow.QCBF_ERR_FORMAT, lambda: gf.load_ildg(str(p)))
        # extents in ildg-format different from the handle's
        f = oracle.hot_start((4, 4, 2, 8), 1)
        p = tmp_path / "dims.ildg"
        limefile.write_ildg(str(p), f, (4, 4, 2, 8))
        _expect(qcbflow.QCBF_ERR_FORMAT, lambda: gf.load_ildg(str(p)))
        # the handle still works after the failures
        gf.load_ildg(fixture(golden_dir))
        assert gf.plaquette() > 0


def test_abi_argument_errors():
    _expect(qcbflow.QCBF_ERR_ARG, lambda: qcbflow.GradientFlow((4, 4, 4, 1)))
    _expect(qcbflow.QCBF_ERR_ARG, lambda: qcbflow.GradientFlow((4, 4, 4, 4), precision=7))
    with qcbflow.GradientFlow(L4) as gf:
        _expect(qcbflow.QCBF_ERR_STATE, lambda: gf.compute_w2())
        _expect(qcbflow.QCBF_ERR_STATE, lambda: gf.force(1))
        _expect(qcbflow.QCBF_ERR_ARG, lambda: gf.compute_force(2))
        with pytest.raises(ValueError):
            gf.copy(np.zeros(10))
    lib = qcbflow.load()
    assert lib.qcbf_rk3_step(None, 0.02, 1) == qcbflow.QCBF_ERR_ARG
    assert b"null handle" in lib.qcbf_last_error(None)


def test_oversized_allocation_reports_oom():
    # 256^3 x 15 sites: 145 GB per FP64 link field; two of them plus the
    # RK3 accumulator exceed the 288 GB of HBM
    _expect(qcbflow.QCBF_ERR_OOM, lambda: qcbflow.GradientFlow((256, 256, 256, 15)))
