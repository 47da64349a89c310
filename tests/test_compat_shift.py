"""The compat mpi::shift on a host GaugeField (no GPU): periodic translation,
pack / unpack, and random_shift's use of the caller's std::mt19937_64."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMPAT = os.path.join(ROOT, "ecmc-qcb_amd", "compat")
CXX = shutil.which("g++")


@pytest.mark.skipif(CXX is None, reason="g++ not installed")
def test_compat_shift_on_host(tmp_path):
    exe = tmp_path / "compat_shift_check"
    srcs = [os.path.join(COMPAT, p) for p in
            ("geometry/GeometryCB.cpp", "gauge/GaugeField.cpp", "mpi/Shift.cpp")]
    subprocess.run([CXX, "-O1", "-std=c++20", "-I", COMPAT, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "compat_shift_check.cpp"), *srcs, "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "Shift (" in r.stdout
