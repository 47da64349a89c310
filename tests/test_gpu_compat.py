"""The C++ drop-in (the reference's own GradientFlow / mpi::observables /
read_ildg_clime signatures over libqcbflow) driven by a caller shaped like the
reference's src/bin/test.cpp, against the committed oracle trajectory."""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "ecmc-qcb_amd", "compat", "flow_measure")
STAGES = os.path.join(ROOT, "ecmc-qcb_amd", "compat", "flow_stages")


def test_compat_caller_matches_golden(golden_dir, tmp_path):
    print("compat flow_measure on the 4^4 fixture", flush=True)
    d = tmp_path / "conf"
    d.mkdir()
    shutil.copy(os.path.join(golden_dir, "conf_4x4x4x4.ildg"), d / "conf")
    out = subprocess.run([EXE, str(tmp_path), "conf", "4", "21", "1"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    with open(os.path.join(golden_dir, "flow_4x4x4x4.json")) as fp:
        g = json.load(fp)
    rows = [l.split()[1:] for l in out.stdout.splitlines() if l.startswith("tQE ")]
    tqe = np.array(rows, dtype=float)
    assert tqe.shape == (21, 3)
    assert np.array_equal(tqe[:, 0], np.array([float(v) for v in g["t"]]))
    E = np.array([float(v) for v in g["E"]])
    assert np.max(np.abs(tqe[:, 2] - E) / E) < 1e-12
    assert np.max(np.abs(tqe[:, 1] - np.array([float(v) for v in g["Q"]]))) < 1e-10
    P = float([l for l in out.stdout.splitlines() if l.startswith("P ")][0].split()[1])
    assert abs(P - float(g["P"][0])) < 1e-12 * abs(P)
    # the reference's progress lines and its "t Q E" file (fixed 10 decimals)
    assert "n = 20, t = 0.40, Q = " in out.stdout
    topo = (d / "conf_topo.txt").read_text().splitlines()
    assert len(topo) == 21 and topo[1].split()[0] == "0.0200000000"


def test_compat_missing_file_exits_with_status_1(tmp_path):
    d = tmp_path / "conf"
    d.mkdir()
    out = subprocess.run([EXE, str(tmp_path), "missing", "4", "2", "1"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 1 and "missing" in out.stderr


def test_compat_stage_members_match_rk3_step_and_oracle(golden_dir, tmp_path):
    """GradientFlow's members called one by one in rk3_step's order
    (src/flow/gradient_flow.cpp:124-135) through the drop-in, with the host
    mirrors force_0 and field_c, against rk3_step and against the oracle."""
    import limefile
    import oracle

    print("compat flow_stages on the 4^4 fixture", flush=True)
    conf = os.path.join(golden_dir, "conf_4x4x4x4.ildg")
    d = tmp_path / "conf"
    d.mkdir()
    shutil.copy(conf, d / "conf")
    out = subprocess.run([STAGES, str(tmp_path), "conf", "4", str(tmp_path)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    dmax = float(re.search(r"stage sequence\| = (\S+)", out.stdout).group(1))
    assert dmax < 1e-13  # fused stage kernels against the unfused members
    assert "differing from their periodic image: 0" in out.stdout
    L = (4, 4, 4, 4)
    f = np.ascontiguousarray(limefile.read_ildg(conf))
    # the observables' local and global forms (observables_mpi.cpp:11-48, 115-156)
    vals = {l.split()[0]: [float(v) for v in l.split()[1:]] for l in out.stdout.splitlines()
            if l.split() and l.split()[0] in ("P_local", "QE_local", "QE_global")}
    V = 4 ** 4
    P, (Q, E) = oracle.plaquette(f, L), oracle.q_e_clover(f, L)
    assert abs(vals["P_local"][0] - 6 * V * P) < 1e-12 * 6 * V * abs(P)
    assert abs(vals["QE_global"][0] - Q) < 1e-10 and abs(vals["QE_global"][1] - E) < 1e-12 * E
    assert abs(vals["QE_local"][0] - Q) < 1e-10 and abs(vals["QE_local"][1] - V * E) < 1e-12 * V * E
    force0 = np.fromfile(tmp_path / "force0.bin")
    assert np.max(np.abs(force0 - oracle.compute_force(f.copy(), L))) < 1e-13
    rk3 = np.fromfile(tmp_path / "rk3.bin")
    assert np.max(np.abs(rk3 - oracle.rk3(f.copy(), L, 0.02, 1))) < 1e-13
    stepwise = np.fromfile(tmp_path / "stepwise.bin")
    assert np.max(np.abs(stepwise - oracle.rk3(f.copy(), L, 0.02, 1))) < 1e-13
    # save_ildg_clime of the flowed field_c: the reference's framing, the field exactly
    saved = (tmp_path / "flowed" / "flowed").read_bytes()
    recs = limefile.records(saved)
    assert [(r[0], r[1], r[2]) for r in recs] == [("ildg-format", True, True),
                                                 ("ildg-binary-data", False, True)]
    assert np.array_equal(limefile.read_ildg(str(tmp_path / "flowed" / "flowed")), rk3)


SAMPLING = os.path.join(ROOT, "ecmc-qcb_amd", "compat", "hb_sampling")


def test_sampling_loop_caller_matches_oracle_on_evolving_configurations(tmp_path):
    """The sampling loop of the reference's heatbath driver
    (src/bin/gauge_hb_cb.cpp:150-197) through the drop-in: random shift and
    mpi::heatbathcb::samples on the device, plaquette, and topo_charge_flowed
    through ONE GradientFlow object on the evolving configuration.  Every
    measured series equals the oracle's flow of the configuration the caller
    saved just before measuring (P and E to 1e-12 relative, Q to 1e-10), and
    the measured configurations differ from one another."""
    import limefile
    import oracle

    L = (8, 8, 8, 8)
    print("compat hb_sampling: 8^4, 3 thermalisation + 6 sampling shifts", flush=True)
    args = [SAMPLING, str(tmp_path), "run", "8", "6.0", "1234", "3", "6", "2", "1", "1", "2",
            "11", "2"]
    out = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    rows = [l.split()[1:] for l in out.stdout.splitlines() if l.startswith("tQE ")]
    got = {}
    for k, t, q, e in rows:
        got.setdefault(int(k), []).append((float(t), float(q), float(e)))
    assert sorted(got) == [0, 1, 2]
    fields_seen = []
    for k in range(3):
        tqe = np.array(got[k])
        f = limefile.read_ildg(str(tmp_path / f"run_m{k}" / f"run_m{k}")).reshape(-1, 4, 3, 3, 2)
        fields_seen.append(f)
        ref, _, _ = oracle.flow_measure(f.copy(), L, 0.02, 11, 2)
        assert tqe.shape == (11, 3)
        assert np.array_equal(tqe[:, 0], ref[:, 0])
        assert np.max(np.abs(tqe[:, 2] - ref[:, 2]) / ref[:, 2]) < 1e-12
        assert np.max(np.abs(tqe[:, 1] - ref[:, 1])) < 1e-10
    assert not np.array_equal(fields_seen[0], fields_seen[1])
    assert not np.array_equal(fields_seen[1], fields_seen[2])
    plaq = [float(l.split()[2]) for l in out.stdout.splitlines() if l.startswith("P ")]
    assert len(plaq) == 6 and all(0.5 < p < 0.7 for p in plaq)
    topo = (tmp_path / "run" / "run_topo.txt").read_text().splitlines()
    assert len(topo) == 33


def test_sampling_loop_with_sub_domains_runs_them_as_one_lattice(tmp_path):
    """The reference's input files ask for n_core_dims = 2 (16 sub-domains of
    L_core^4): on one GPU the drop-in runs them as one lattice of 2 L_core
    per dimension, so the loop with L_core = 4, n_core_dims = 2 prints exactly
    what it prints for one 8^4 lattice."""
    print("compat hb_sampling: 8^4 as L_core 8 and as 2^4 sub-domains of 4^4", flush=True)
    outs = []
    for lc, n in ((8, 1), (4, 2)):
        d = tmp_path / f"L{lc}n{n}"
        d.mkdir()
        args = [SAMPLING, str(d), "run", str(lc), "6.0", "77", "2", "3", "2", "1", "1", "2",
                "6", "2", str(n)]
        out = subprocess.run(args, capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr[-2000:]
        outs.append([l for l in out.stdout.splitlines() if l.startswith(("tQE ", "P "))])
        if n == 2:
            assert "n_core_dim = 2 on one GPU" in out.stderr
    assert len(outs[0]) == 3 + 2 * 6 and outs[0] == outs[1]


def test_compat_observables_per_site_and_levi_civita(tmp_path):
    """clover_site, local_q_e_clover and levi_civita of the drop-in
    (src/observables/observables_mpi.h:8-19) against the oracle, through a
    C++ caller that only calls observables (no GradientFlow), and exits
    cleanly."""
    import limefile
    import oracle

    src = os.path.join(ROOT, "tests", "compat_observables_check.cpp")
    exe = str(tmp_path / "obs_check")
    compat = os.path.join(ROOT, "ecmc-qcb_amd", "compat")
    pkg = os.path.join(ROOT, "ecmc-qcb_amd")
    subprocess.run(["g++", "-O1", "-std=c++20", "-I", compat, "-I", os.path.join(ROOT, "include"),
                    src, os.path.join(compat, "libqcbcompat.a"), "-L", pkg, "-lqcbflow",
                    f"-Wl,-rpath,{pkg}", "-o", exe], check=True, timeout=120)
    d = tmp_path / "conf"
    d.mkdir()
    L = (6, 6, 6, 6)
    f = oracle.rk3(oracle.hot_start(L, 77), L, 0.02, 2)
    limefile.write_ildg(str(d / "conf"), f, L)
    out = subprocess.run([exe, str(tmp_path), "conf", "6"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = out.stdout.splitlines()
    for l in lines:
        w = l.split()
        if w[0] == "F":  # F <x> <y> <z> <t> <mu> <nu> 18 doubles (extended-box coordinates)
            x, y, z, t, mu, nu = map(int, w[1:7])
            site = (((t - 1) % 6 * 6 + (z - 1) % 6) * 6 + (y - 1) % 6) * 6 + (x - 1) % 6
            v = np.array(w[7:], dtype=float)
            got = v[0::2].reshape(3, 3) + 1j * v[1::2].reshape(3, 3)
            ref = oracle.clover(f, L, site, mu, nu) if mu != nu else np.zeros((3, 3))
            assert np.abs(got - ref).max() < 1e-14
        elif w[0] == "QE":
            x, y, z, t = map(int, w[1:5])
            site = (((t - 1) % 6 * 6 + (z - 1) % 6) * 6 + (y - 1) % 6) * 6 + (x - 1) % 6
            q, e = oracle.local_q_e(f, L, site)
            assert abs(float(w[5]) - q) < 1e-15 and abs(float(w[6]) - e) < 1e-13 * e
        elif w[0] == "EPS":
            assert w[1] == "ok"
    assert sum(1 for l in lines if l.startswith("F ")) >= 8
    assert sum(1 for l in lines if l.startswith("QE ")) >= 3
    assert "EPS ok" in out.stdout
    assert "mu = nu => G(site, mu, nu) = 0" in out.stderr
