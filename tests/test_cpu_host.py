"""CPU tests of the host side: the C-ABI library loads and exports every symbol the header
declares (no compute call without a GPU), setups, and the Python model surface."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shamx_sph.h")


def header_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(shamx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    import ctypes

    from shamrock_amd import _lib

    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the binding declares a signature for each of them
    assert set(syms) == set(_lib.SIGNATURES)


def test_binding_loads_and_reports_errors_without_gpu():
    import ctypes

    from shamrock_amd import _lib

    L = _lib.lib()
    # argument validation happens before any device call
    h = ctypes.c_void_p()
    cfg = _lib.ShamxConfig()
    cfg.kernel = 9
    rc = L.shamx_solver_create(ctypes.byref(cfg), 10, ctypes.byref(h))
    assert rc == 1
    assert b"kernel" in L.shamx_last_error()


def test_struct_layouts_match_header(tmp_path):
    """Size and every field offset of shamx_config and shamx_step_log as the C compiler lays
    them out from include/shamx_sph.h, against the ctypes mirrors."""
    import ctypes
    import os
    import subprocess

    from shamrock_amd import _lib

    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "shamx_sph.h"',
             "int main(void) {"]
    for cname, ct in (("shamx_config", _lib.ShamxConfig), ("shamx_step_log", _lib.ShamxStepLog)):
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in ct._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    c = {}
    for line in out.splitlines():
        name, field, v = line.split()
        c[(name, field)] = int(v)
    for cname, ct in (("shamx_config", _lib.ShamxConfig), ("shamx_step_log", _lib.ShamxStepLog)):
        assert c[(cname, "size")] == ctypes.sizeof(ct), cname
        for f, _ in ct._fields_:
            assert c[(cname, f)] == getattr(ct, f).offset, (cname, f)


def ctypes_size(t):
    import ctypes

    return ctypes.sizeof(t)


def test_lattice_cubic_counts_and_bounds():
    from shamrock_amd import setups

    p = setups.lattice_cubic(1 / 8, (0, 0, 0), (1, 1, 1))
    assert p.shape == (512, 3)
    assert p.min() == 0.0 and p.max() == 7 / 8
    p = setups.lattice_cubic(1 / 16, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    assert p.shape == (4096, 3)
    assert np.all(p >= -0.5) and np.all(p < 0.5)


def test_sedov_energy_budget():
    from shamrock_amd import setups

    cfg, f = setups.sedov(16)
    n = f["x"].size
    m = cfg["gpart_mass"]
    assert n == 4096 and m == 1 / 4096
    E = m * f["u"].sum()
    assert abs(E - (1.0 + n * m * 1e-5)) < 1e-12
    r = np.sqrt(f["x"] ** 2 + f["y"] ** 2 + f["z"] ** 2)
    # energy only inside the kernel support 2 dr
    hot = f["u"] > 1e-5
    assert r[hot].max() < 2 / 16


def test_sound_wave_setup():
    from shamrock_amd import setups

    cfg, f = setups.sound_wave(8)
    assert f["x"].size == 512
    np.testing.assert_allclose(f["vx"], 1e-4 * np.sin(2 * np.pi * f["x"]))
    assert cfg["gpart_mass"] == 1 / 512 and np.all(f["h"] == 1 / 8)


def test_get_model_sph_surface():
    import shamrock_amd as shamrock

    ctx = shamrock.Context()
    ctx.pdata_layout_new()
    with pytest.raises(ValueError, match="unknown combination"):
        shamrock.get_Model_SPH(context=ctx, vector_type="f32_3", sph_kernel="M4")
    with pytest.raises(ValueError):
        shamrock.get_Model_SPH(context=ctx, vector_type="f64_3", sph_kernel="M5")
    model = shamrock.get_Model_SPH(context=ctx, vector_type="f64_3", sph_kernel="M6")
    assert model.name == "SPHModel_f64_3_M6"
    cfg = model.gen_default_config()
    cfg.set_artif_viscosity_Constant(alpha_u=1.0, alpha_AV=1.0, beta_AV=2.0)
    cfg.set_boundary_periodic()
    cfg.set_eos_adiabatic(5 / 3)
    cfg.set_particle_mass(1e-3)
    model.set_solver_config(cfg)
    model.resize_simulation_box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    setup = model.get_setup()
    model.get_setup().apply_setup(setup.make_generator_lattice_cubic(0.1, (-0.5,) * 3, (0.5,) * 3))
    assert model.get_total_part_count() == 1000
    model.set_value_in_a_box("uint", "f64", 2.0, (-0.5,) * 3, (0.5,) * 3)
    model.add_kernel_value("uint", "f64", 1.0, (0, 0, 0), 0.1)
    d = ctx.collect_data()
    assert d["xyz"].shape == (1000, 3) and d["hpart"].shape == (1000,)
    assert np.all(d["uint"] >= 2.0) and d["uint"].max() > 2.0
    assert abs(model.rho_h(0.1) - 1.0) < 1e-12  # m (hfact / h)^3 with hfact 1.0 for M6
    assert model.get_hfact() == 1.0
    with pytest.raises(RuntimeError):
        model2 = shamrock.get_Model_SPH(context=shamrock.Context(), vector_type="f64_3",
                                        sph_kernel="M4")
        model2.evolve_once()  # no particles: fails before touching the GPU
