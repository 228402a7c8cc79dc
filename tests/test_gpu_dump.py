"""A run continued from a restart dump (Model::dump / load_from_dump, Model.hpp:906-991) is the
same run: a gas disc with a sink and the Cullen & Dehnen viscosity (whose alpha and saved sound
speed are state carried from step to step) is stepped, dumped, reloaded into a fresh model, and
both models take the same further steps; every field, the sink and the time edges must agree
bit for bit."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import shamrock_amd as shamrock  # noqa: E402
from shamrock_amd import setups  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _model(cfg, f):
    """The disc of setups.disc built through the public surface only: config setters, box,
    push_particle, a velocity field set from the positions (Keplerian about the sink), particle
    mass, the sink."""
    ctx = shamrock.Context()
    ctx.pdata_layout_new()
    m = shamrock.get_Model_SPH(context=ctx, vector_type="f64_3", sph_kernel=cfg["kernel"])
    c = m.gen_default_config()
    c.set_artif_viscosity_VaryingCD10(alpha_min=0.0, alpha_max=1.0, sigma_decay=0.1,
                                      alpha_u=1.0, beta_AV=2.0)
    c.set_boundary_free()
    c.set_eos_locally_isothermalLP07(cfg["eos_cs0"], cfg["eos_q"], cfg["eos_r0"])
    # code units with G = 1: a mass unit of 1 / G_SI kg
    c.set_units(shamrock.UnitSystem(unit_mass=1.0 / 6.6743015e-11))
    c.set_sink_softening(cfg["sink_softening"])
    m.set_solver_config(c)
    m.resize_simulation_box(cfg["box_min"], cfg["box_max"])
    pos = np.stack([f["x"], f["y"], f["z"]], 1)
    m.push_particle(pos, f["h"], f["u"])
    G = c.get_constant_G()

    def kepler(p):
        r = np.hypot(p[0], p[1])
        vk = np.sqrt(G * 1.0 / r)
        return (-vk * p[1] / r, vk * p[0] / r, 0.0)

    m.set_field_value_lambda_f64_3("vxyz", kepler)
    m.set_particle_mass(cfg["gpart_mass"])
    for s in cfg.get("sinks", ()):
        m.add_sink(s["mass"], s["pos"], s["velocity"], s["accretion_radius"])
    return ctx, m


def test_restart_continues_the_run(tmp_path):
    cfg, f = setups.disc(20000, seed=3)
    ctx, m = _model(cfg, f)
    assert abs(m.get_current_config().get_constant_G() - 1.0) < 1e-15
    for i in range(3):
        say(f"[dump] step {i}")
        m.evolve_once()
    path = str(tmp_path / "disc.shamx")
    m.dump(path)
    ctx2 = shamrock.Context()
    ctx2.pdata_layout_new()
    m2 = shamrock.get_Model_SPH(context=ctx2, vector_type="f64_3", sph_kernel=cfg["kernel"])
    m2.load_from_dump(path)
    for i in range(2):
        say(f"[dump] continued step {i}")
        m.evolve_once()
        m2.evolve_once()
    d1, d2 = ctx.collect_data(), ctx2.collect_data()
    assert "alpha_AV" in d1 and np.count_nonzero(d1["alpha_AV"]) > 0
    for k in d1:
        np.testing.assert_array_equal(d1[k], d2[k], err_msg=k)
    assert (m.get_time(), m.get_dt()) == (m2.get_time(), m2.get_dt())
    assert m.get_sinks() == m2.get_sinks()
