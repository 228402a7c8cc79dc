"""Runs the Sod-tube CI workload (examples/tests_ci/sod_tube_sph.py) on the oracle and prints its
L2 errors next to the reference's expected values (tests/golden/sod_tube_ci.json).  Several
minutes on 8 cores; the GPU test test_gpu_sod.py runs the same workload on the HIP library."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as orc  # noqa: E402
from shamrock_amd import analysis, setups  # noqa: E402


def main():
    fx = json.load(open(os.path.join(HERE, "sod_tube_ci.json")))
    cfg, f = setups.sod_tube(fx["resol"])
    s = orc.Solver(cfg, f)
    t0 = time.time()
    steps = 0
    target = fx["t_target"]
    while s.time()["t"] < target:
        tm = s.time()
        if tm["t"] + tm["dt"] > target:
            s.set_time(dt=target - tm["t"])
        s.evolve_once()
        steps += 1
        if steps % 50 == 0:
            print(f"step {steps} t {s.time()['t']:.5f} ({time.time() - t0:.0f} s)", flush=True)
    out = {k: s.get(k) for k in ("x", "y", "z", "vx", "vy", "vz", "h", "u")}
    sod = analysis.SodTube(1.4, 1, 1, 0.125, 0.1)
    rho, v, P = analysis.sod_l2(sod, out, cfg["gpart_mass"], 1.0, 1.4, target)
    got = dict(rho=rho, vx=v[0], vy=v[1], vz=v[2], P=P)
    print(f"steps {steps}")
    for k, g in got.items():
        e = fx["expect_" + k]
        print(f"{k}: got {g!r} expected {e!r} rel {(g - e) / e:.3e}")


if __name__ == "__main__":
    main()
