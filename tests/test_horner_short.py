"""The exponential's shortened Cayley-Hamilton coefficients (su3.hpp,
horner_exp), built for the host.

Below t = -tr(X^2)/2 = 0.01 the kernels evaluate the degree-10 Taylor
polynomial instead of the reference's 18 terms (exp_su3_luscher,
src/su3/utils.cpp:162-186).  Pinned here on random su(3) elements with t
spread over [1e-8, 0.01) and beyond: the degree-10 exponential differs from
the 18-term one by at most 2^-55 per entry (the remainder bound is 1.2e-17
plus rounding), both match scipy's expm to 1e-15, horner_exp takes exactly
the degree-10 coefficients below the threshold and the 18-term ones above it,
and degree 8 would not meet the bound near the threshold (so the check can
fail).  For an FP32 result (the FP32-mixed flow) horner_exp takes the
degree-7 polynomial below t = 0.005, equal to the 18-term one within 2^-41
(1/2^17 of an FP32 ulp of 1); degree 5 would not meet that bound there."""
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg

from conftest import ROOT


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("horner") / "horner_check")
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-host-only", "-O2", "-std=c++17",
                    "-I", os.path.join(ROOT, "ecmc-qcb_amd", "csrc"),
                    os.path.join(ROOT, "tests", "horner_check.cpp"), "-o", exe],
                   check=True, timeout=300, capture_output=True)
    return exe


def su3_algebra(rng, n, t_target):
    """n random traceless anti-hermitian X scaled to t = -tr(X^2)/2 = t_target."""
    A = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    X = 0.5 * (A - np.conj(np.transpose(A, (0, 2, 1))))
    X -= np.trace(X, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    t = -0.5 * np.trace(X @ X, axis1=1, axis2=2).real
    return X * np.sqrt(t_target / t)[:, None, None]


def coefficients(exe, X):
    t = -0.5 * np.trace(X @ X, axis1=1, axis2=2).real
    d = (1j * np.linalg.det(X)).real
    r = subprocess.run([exe], input=np.stack([t, d], axis=1).astype(np.float64).tobytes(),
                       capture_output=True, timeout=60, check=True)
    q = np.frombuffer(r.stdout, dtype=np.float64).reshape(-1, 5, 3, 2)
    # [case, variant (10, 18, exp, 7, exp for an FP32 result), q0..q2]
    return t, q[..., 0] + 1j * q[..., 1]


def exp_from(q, X):
    return q[:, 0, None, None] * np.eye(3) + q[:, 1, None, None] * X + q[:, 2, None, None] * (X @ X)


def test_degree_10_matches_18_terms_below_threshold(checker):
    rng = np.random.default_rng(5)
    X = np.concatenate([su3_algebra(rng, 400, tt) for tt in
                        (1e-8, 1e-6, 1e-4, 1e-3, 4e-3, 8e-3, 0.0099999)])
    t, q = coefficients(checker, X)
    assert np.all(t < 0.01)
    e10, e18, eh = (exp_from(q[:, k], X) for k in range(3))
    assert np.max(np.abs(e10 - e18)) <= 2.0 ** -55
    ref = np.stack([scipy.linalg.expm(x) for x in X])
    assert np.max(np.abs(e18 - ref)) < 1e-15 and np.max(np.abs(e10 - ref)) < 1e-15
    assert np.array_equal(q[:, 2], q[:, 0])  # horner_exp took degree 10


def test_18_terms_at_and_above_threshold(checker):
    rng = np.random.default_rng(6)
    X = np.concatenate([su3_algebra(rng, 200, tt) for tt in (0.01, 0.05, 0.5)])
    t, q = coefficients(checker, X)
    assert np.all(t >= 0.01 * (1 - 1e-12))
    sel = t >= 0.01
    assert np.array_equal(q[sel, 2], q[sel, 1])  # horner_exp took all 18 terms


def test_degree_8_would_fail_the_bound():
    """The degree-8 polynomial misses exp(X) by more than 2^-55 near t = 0.01:
    the threshold is not loose enough to hide a degree off by two."""
    rng = np.random.default_rng(7)
    X = su3_algebra(rng, 50, 0.0099)
    p8 = sum(np.linalg.matrix_power(X, n) / float(np.prod(np.arange(1, n + 1))) for n in range(9))
    ref = np.stack([scipy.linalg.expm(x) for x in X])
    # about 7e-15, against 2^-55 = 2.8e-17 for degree 10 (and 2.2e-16 of
    # rounding in this numpy power sum)
    assert np.max(np.abs(p8 - ref)) > 2e-15


def test_degree_7_for_fp32_results_below_its_threshold(checker):
    rng = np.random.default_rng(8)
    X = np.concatenate([su3_algebra(rng, 400, tt) for tt in
                        (1e-8, 1e-6, 1e-4, 1e-3, 2e-3, 4e-3, 0.0049999)])
    t, q = coefficients(checker, X)
    assert np.all(t < 0.005)
    e7, e18 = exp_from(q[:, 3], X), exp_from(q[:, 1], X)
    assert np.max(np.abs(e7 - e18)) <= 2.0 ** -41
    ref = np.stack([scipy.linalg.expm(x) for x in X])
    assert np.max(np.abs(e7 - ref)) < 2.0 ** -41
    assert np.array_equal(q[:, 4], q[:, 3])  # the FP32-result choice took degree 7


def test_fp32_result_choice_above_the_degree_7_threshold(checker):
    """From t = 0.005 the FP32-result choice is the FP64 one: degree 10 below
    t = 0.01, the 18 terms from there on."""
    rng = np.random.default_rng(9)
    X = np.concatenate([su3_algebra(rng, 200, tt) for tt in (0.005, 0.008, 0.02, 0.5)])
    t, q = coefficients(checker, X)
    assert np.all(t >= 0.005 * (1 - 1e-12))
    sel = t >= 0.005
    assert np.array_equal(q[sel, 4], q[sel, 2])


def test_degree_5_would_fail_the_fp32_bound():
    """The degree-5 polynomial misses exp(X) by more than 2^-41 near
    t = 0.005: the degree-7 threshold is not loose enough to hide a degree
    off by two."""
    rng = np.random.default_rng(10)
    X = su3_algebra(rng, 50, 0.0049)
    p5 = sum(np.linalg.matrix_power(X, n) / float(np.prod(np.arange(1, n + 1))) for n in range(6))
    ref = np.stack([scipy.linalg.expm(x) for x in X])
    assert np.max(np.abs(p5 - ref)) > 2.0 ** -38
