"""Pair-level checks of the hand-written adjoint (shamrock_amd/csrc/adjoint_pair.hpp), on the CPU.

The header is compiled for the host (tests/adjoint_pair_check.cpp) and its reverse sweep is
compared with central finite differences of its own forward pair value.  The kernel partial
derivatives are pinned to finite differences and W, dW/dr, dW/dh to the oracle's kernels.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as orc

HERE = os.path.dirname(os.path.abspath(__file__))
KIDS = {"M4": 0, "M6": 1, "M8": 2, "C2": 3}
HFACT = {"M4": 1.2, "M6": 1.0, "M8": 1.0, "C2": 1.4}
RKERN = {"M4": 2.0, "M6": 3.0, "M8": 4.0, "C2": 2.0}


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("adj") / "adj_check.so")
    subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-O1", "-std=c++20", "-fPIC",
                    "-shared", os.path.join(HERE, "adjoint_pair_check.cpp"), "-o", out],
                   check=True, capture_output=True)
    lib = C.CDLL(out)
    dp = C.POINTER(C.c_double)
    lib.pair_eval.argtypes = [C.c_int, dp, dp, dp, dp, dp]
    lib.kern_deriv.argtypes = [C.c_int, C.c_double, C.c_double, dp]
    return lib


def _p(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))


def kd(lib, k, r, h):
    out = np.zeros(6)
    assert lib.kern_deriv(KIDS[k], r, h, _p(out)) == 0
    return out


@pytest.mark.parametrize("k", list(KIDS))
def test_kernel_partials(harness, k):
    R = RKERN[k]
    rng = np.random.default_rng(1)
    for _ in range(40):
        h = rng.uniform(0.5, 2.0)
        r = rng.uniform(0.01, R * h * 0.999)
        # skip points next to the polynomial breakpoints (q = 1, 2, 3)
        if np.min(np.abs(r / h - np.arange(1, 5))) < 1e-3:
            continue
        W, Wr, Wh, Wrr, Wrh, Whh = kd(harness, k, r, h)
        e = 1e-6
        f = lambda rr, hh: kd(harness, k, rr, hh)  # noqa: E731
        np.testing.assert_allclose(Wr, (f(r + e, h)[0] - f(r - e, h)[0]) / (2 * e), rtol=1e-6,
                                   atol=1e-9)
        np.testing.assert_allclose(Wh, (f(r, h + e)[0] - f(r, h - e)[0]) / (2 * e), rtol=1e-6,
                                   atol=1e-9)
        np.testing.assert_allclose(Wrr, (f(r + e, h)[1] - f(r - e, h)[1]) / (2 * e), rtol=1e-5,
                                   atol=1e-8)
        np.testing.assert_allclose(Wrh, (f(r, h + e)[1] - f(r, h - e)[1]) / (2 * e), rtol=1e-5,
                                   atol=1e-8)
        np.testing.assert_allclose(Wrh, (f(r + e, h)[2] - f(r - e, h)[2]) / (2 * e), rtol=1e-5,
                                   atol=1e-8)
        np.testing.assert_allclose(Whh, (f(r, h + e)[2] - f(r, h - e)[2]) / (2 * e), rtol=1e-5,
                                   atol=1e-8)
        o = orc.kernel_eval(k, [r], [h])
        np.testing.assert_allclose([W, Wr, Wh], [o["W"][0], o["dW"][0], o["dhW"][0]],
                                   rtol=1e-13, atol=1e-15)


def _pair(lib, k, par, x, seeds, grad=True):
    loss = np.zeros(1)
    g = np.zeros(17)
    assert lib.pair_eval(KIDS[k], _p(par), _p(x), _p(seeds), _p(loss),
                         _p(g) if grad else None) == 0
    return loss[0], g


@pytest.mark.parametrize("k", ["M4", "M6"])
@pytest.mark.parametrize("eos", [0, 1, 2])
def test_force_pair_reverse_sweep_matches_finite_differences(harness, k, eos):
    """d(seeds . (a_ab, du_ab)) / d(r, v_a, v_b, u_a, u_b, h_a, h_b, Omega_a, Omega_b) and, for
    LP07 (eos 2), d / d(c_s,a, c_s,b) of the sides' own sound speeds."""
    rng = np.random.default_rng(3 + eos)
    par = np.array([1e-3, HFACT[k], 5 / 3, 1.0, 1.0, 2.0, eos, 0.7])
    checked = 0
    for trial in range(60):
        h_a, h_b = rng.uniform(0.08, 0.12, 2)
        d = rng.uniform(0.05, RKERN[k] * min(h_a, h_b) * 0.95)
        n = rng.normal(size=3)
        r = d * n / np.linalg.norm(n)
        va, vb = rng.normal(0, 0.3, 3), rng.normal(0, 0.3, 3)
        ua, ub = rng.uniform(0.5, 2.0, 2)
        om_a, om_b = rng.uniform(0.8, 1.2, 2)
        cs_a, cs_b = rng.uniform(0.3, 1.5, 2)
        x = np.concatenate([r, va, vb, [ua, ub, h_a, h_b, om_a, om_b, cs_a, cs_b]])
        seeds = rng.normal(size=4)
        _, g = _pair(harness, k, par, x, seeds)
        fd = np.zeros(17)
        for i in range(17):
            e = 1e-6 * max(abs(x[i]), 1e-2)
            xp, xm = x.copy(), x.copy()
            xp[i] += e
            xm[i] -= e
            fd[i] = (_pair(harness, k, par, xp, seeds, False)[0] -
                     _pair(harness, k, par, xm, seeds, False)[0]) / (2 * e)
        # the AV switch max(., 0) and |v.r| are kinks: skip draws within the FD step of them
        w = np.dot(va - vb, r / d)
        if abs(w) < 1e-4:
            continue
        scale = np.abs(fd).max()
        np.testing.assert_allclose(g, fd, rtol=2e-5, atol=2e-6 * scale,
                                   err_msg=f"trial {trial}")
        checked += 1
    assert checked > 40
