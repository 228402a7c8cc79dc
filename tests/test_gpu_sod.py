"""The reference's Sod-tube CI run (examples/tests_ci/sod_tube_sph.py) through the public Python
surface only, `shamrock_amd` imported as `shamrock`, with the script's own sequence of calls:
config (M6, Cullen & Dehnen viscosity, periodic, adiabatic gamma = 1.4), get_box_dim_fcc_3d,
resize_simulation_box, two HCP generators joined by make_combiner_add, set_value_in_a_box,
total_mass_to_part_mass, CFL 0.1, evolve_until(0.245), then make_analysis_sodtube and
compute_L2_dist, whose values must match the script's expected ones (:114-121, relative 1e-11).
82,944 particles, 646 steps on the HIP library.
"""
import json
import os
import sys

import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import shamrock_amd as shamrock  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_sod_tube_ci_script_calls():
    fx = json.load(open(os.path.join(HERE, "golden", "sod_tube_ci.json")))
    gamma = 1.4
    rho_g, rho_d = 1, 0.125
    fact = (rho_g / rho_d) ** (1.0 / 3.0)
    P_g, P_d = 1, 0.1
    u_g = P_g / ((gamma - 1) * rho_g)
    u_d = P_d / ((gamma - 1) * rho_d)
    resol = fx["resol"]

    ctx = shamrock.Context()
    ctx.pdata_layout_new()
    model = shamrock.get_Model_SPH(context=ctx, vector_type="f64_3", sph_kernel="M6")
    cfg = model.gen_default_config()
    cfg.set_artif_viscosity_VaryingCD10(alpha_min=0.0, alpha_max=1, sigma_decay=0.1,
                                        alpha_u=1, beta_AV=2)
    cfg.set_boundary_periodic()
    cfg.set_eos_adiabatic(gamma)
    model.set_solver_config(cfg)
    model.init_scheduler(int(1e8), 1)

    xs, ys, zs = model.get_box_dim_fcc_3d(1, resol, 24, 24)
    dr = 1 / xs
    xs, ys, zs = model.get_box_dim_fcc_3d(dr, resol, 24, 24)
    model.resize_simulation_box((-xs, -ys / 2, -zs / 2), (xs, ys / 2, zs / 2))
    setup = model.get_setup()
    gen1 = setup.make_generator_lattice_hcp(dr, (-xs, -ys / 2, -zs / 2), (0, ys / 2, zs / 2))
    gen2 = setup.make_generator_lattice_hcp(dr * fact, (0, -ys / 2, -zs / 2),
                                            (xs, ys / 2, zs / 2))
    comb = setup.make_combiner_add(gen1, gen2)
    setup.apply_setup(comb)
    model.set_value_in_a_box("uint", "f64", u_g, (-xs, -ys / 2, -zs / 2), (0, ys / 2, zs / 2))
    model.set_value_in_a_box("uint", "f64", u_d, (0, -ys / 2, -zs / 2), (xs, ys / 2, zs / 2))
    vol_b = xs * ys * zs
    totmass = (rho_d * vol_b) + (rho_g * vol_b)
    pmass = model.total_mass_to_part_mass(totmass)
    model.set_particle_mass(pmass)
    model.set_cfl_cour(0.1)
    model.set_cfl_force(0.1)
    assert model.get_solver_tex() and model.get_solver_dot_graph()
    say(f"[sod] {model.get_total_part_count()} particles, evolve_until {fx['t_target']}")

    t_target = fx["t_target"]
    res = model.evolve_until(t_target)
    assert res.reach_target_time
    say(f"[sod] {res.iter_count} steps, t = {model.get_time()}")

    sod = shamrock.phys.SodTube(gamma=gamma, rho_1=1, P_1=1, rho_5=0.125, P_5=0.1)
    sodanalysis = model.make_analysis_sodtube(sod, (1, 0, 0), t_target, 0.0, -0.5, 0.5)
    rho, v, P = sodanalysis.compute_L2_dist()
    vx, vy, vz = v
    got = dict(rho=rho, vx=vx, vy=vy, vz=vz, P=P)
    for k, val in got.items():
        e = fx["expect_" + k]
        say(f"[sod] {k}: {val!r} expected {e!r} rel {(val - e) / e:.3e}")
    for k, val in got.items():
        e = fx["expect_" + k]
        assert abs(val - e) < fx["tol"] * e, (k, val, e)
