"""The reference's Sod-tube CI workload on the HIP library: HCP lattices at resolution 128
(82944 particles), M6, Cullen & Dehnen viscosity, run to t = 0.245 with the last step clipped,
then the L2 errors of rho, v and P against the exact Riemann solution compared with the
reference's expected values (examples/tests_ci/sod_tube_sph.py:114-121, relative 1e-11).
"""
import json
import os
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from shamrock_amd import analysis, setups  # noqa: E402
from shamrock_amd.solver import GpuSolver  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_sod_tube_ci_l2():
    fx = json.load(open(os.path.join(HERE, "golden", "sod_tube_ci.json")))
    cfg, f = setups.sod_tube(fx["resol"])
    g = GpuSolver(cfg, f)
    target = fx["t_target"]
    steps = 0
    while g.time()["t"] < target:
        tm = g.time()
        if tm["t"] + tm["dt"] > target:
            g.set_time(dt=target - tm["t"])
        g.evolve_once()
        steps += 1
        if steps % 100 == 0:
            say(f"[sod] step {steps} t {g.time()['t']:.5f}")
        assert steps < 5000
    out = {k: g.get(k) for k in ("x", "y", "z", "vx", "vy", "vz", "h", "u")}
    g.close()
    sod = analysis.SodTube(1.4, 1, 1, 0.125, 0.1)
    rho, v, P = analysis.sod_l2(sod, out, cfg["gpart_mass"], 1.0, 1.4, target)
    got = dict(rho=rho, vx=v[0], vy=v[1], vz=v[2], P=P)
    say(f"[sod] {steps} steps, t = {target}")
    for k, val in got.items():
        e = fx["expect_" + k]
        say(f"[sod] {k}: {val!r} expected {e!r} rel {(val - e) / e:.3e}")
    for k, val in got.items():
        e = fx["expect_" + k]
        assert abs(val - e) <= fx["tol"] * abs(e), (k, val, e)
