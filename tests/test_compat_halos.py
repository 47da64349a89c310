"""The compat halos and neighbour table on a host GaugeField (no GPU): the
extended box holds periodic images, faces, edges and corners included, and
one- and two-hop neighbours (x +- mu, x + mu - nu, x - mu - nu) land on the
right links."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMPAT = os.path.join(ROOT, "ecmc-qcb_amd", "compat")
CXX = shutil.which("g++")


@pytest.mark.skipif(CXX is None, reason="g++ not installed")
def test_compat_halos_are_periodic_images(tmp_path):
    exe = tmp_path / "compat_halo_check"
    srcs = [os.path.join(COMPAT, p) for p in ("geometry/GeometryCB.cpp", "gauge/GaugeField.cpp")]
    subprocess.run([CXX, "-O1", "-std=c++20", "-I", COMPAT, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "compat_halo_check.cpp"), *srcs, "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "wrong: 0" in r.stdout


@pytest.mark.skipif(CXX is None, reason="g++ not installed")
def test_compat_topology_folds_sub_domains_into_one_lattice(tmp_path):
    """MpiTopology(n_core_dim = 2) then GeometryCB(L_core = 4): one 8^4 lattice
    on the one rank (the reference's 16 sub-domains of 4^4), as the reference's
    own input files ask for."""
    exe = tmp_path / "compat_fold_check"
    srcs = [os.path.join(COMPAT, "geometry/GeometryCB.cpp")]
    subprocess.run([CXX, "-O1", "-std=c++20", "-I", COMPAT,
                    os.path.join(ROOT, "tests", "compat_fold_check.cpp"), *srcs, "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "wrong: 0" in r.stdout, r.stdout
    assert "n_core_dim = 2 on one GPU" in r.stderr
