"""Python surface of the SPH model, mirroring the reference bindings for the hot path.

Reference: shammodels/sph/src/pySPHModel.cpp (get_Model_SPH :1725-1754, SPHModel_f64_3_<K>
:761-1510, SolverConfig setters :69-400, TimestepLog :1763-1770, EvolveUntilResults
:1692-1705) and shampylib/src/pyShamrockCtx.cpp (Context, pdata_layout_new, collect_data
:71-128).  Particles live on the host until the first step; from then on the fields are owned
by the GPU solver (libshamx_sph.so) and copied back on demand.
"""
from __future__ import annotations

import dataclasses
import json
import math
import time
from typing import Callable

import numpy as np

from .analysis import AnalysisSodTube
from .phys import SodTube
from .setups import (KERNEL_HFACT, KERNEL_RKERN, base_config, fcc_cube, hcp_box_dim,
                     hcp_lattice, kernel_W, lattice_cubic)
from .solver import FIELDS, G_SI, GpuSolver, sink_dict

_VEC_FIELDS = {"xyz": ("x", "y", "z"), "vxyz": ("vx", "vy", "vz"), "axyz": ("ax", "ay", "az"),
               "axyz_ext": ("axe", "aye", "aze"), "curlv": ("curlx", "curly", "curlz")}
_SCAL_FIELDS = {"hpart": "h", "uint": "u", "duint": "du", "alpha_AV": "alpha_AV",
                "divv": "divv", "dtdivv": "dtdivv", "soundspeed": "soundspeed"}
# fields of the layout (SolverConfig.cpp:24-121) for the two viscosities
_LAYOUT_CONSTANT = ("xyz", "vxyz", "axyz", "axyz_ext", "hpart", "uint", "duint")
_LAYOUT_CD10 = _LAYOUT_CONSTANT + ("alpha_AV", "divv", "dtdivv", "curlv", "soundspeed")


class UnitSystem:
    """shamunits::UnitSystem (UnitSystem.hpp:148-156): SI value of each code unit."""

    def __init__(self, unit_time=1.0, unit_length=1.0, unit_mass=1.0, unit_current=1.0,
                 unit_temperature=1.0, unit_qte=1.0, unit_lumint=1.0):
        self.unit_time, self.unit_length, self.unit_mass = unit_time, unit_length, unit_mass
        self.unit_current, self.unit_temperature = unit_current, unit_temperature
        self.unit_qte, self.unit_lumint = unit_qte, unit_lumint


class Constants:
    """shamunits::Constants: physical constants expressed in a unit system."""

    def __init__(self, units: UnitSystem | None = None):
        self.units = units or UnitSystem()

    def G(self):
        """6.6743015e-11 m^3 kg^-1 s^-2 (Constants.hpp:130) in code units."""
        u = self.units
        return G_SI * u.unit_mass * u.unit_time ** 2 / u.unit_length ** 3


@dataclasses.dataclass
class TimestepLog:
    """Solver.hpp:43-54"""
    rank: int = 0
    rate: float = 0.0
    npart: int = 0
    tcompute: float = 0.0

    def rate_sum(self):
        return self.rate

    def npart_sum(self):
        return self.npart


@dataclasses.dataclass
class EvolveUntilResults:
    """Solver.hpp:56-62"""
    reach_target_time: bool = False
    reach_niter_max: bool = False
    reach_max_walltime: bool = False
    iter_count: int = 0


class Context:
    """ShamrockCtx subset: one patch, one GPU."""

    def __init__(self):
        self.layout_ready = False
        self.model = None

    def pdata_layout_new(self):
        self.layout_ready = True

    def init_sched(self, split_crit=0, merge_crit=0):
        self.sched_ready = True

    def collect_data(self) -> dict:
        """pyShamrockCtx.cpp:97-128: dict of numpy arrays keyed by field name."""
        if self.model is None:
            return {}
        return self.model.collect_data()


class SolverConfig:
    """Hot-path subset of SolverConfig (SolverConfig.hpp:432-1400)."""

    def __init__(self, kernel: str):
        self.kernel = kernel
        self.d = base_config(kernel, (0, 0, 0), (1, 1, 1))
        self.two_stage_search = True

    # artificial viscosity (AVConfig.hpp:49-55)
    def set_artif_viscosity_Constant(self, *, alpha_u, alpha_AV, beta_AV):
        self.d.update(av_kind=0, alpha_u=float(alpha_u), alpha_AV=float(alpha_AV),
                      beta_AV=float(beta_AV))

    def set_artif_viscosity_VaryingCD10(self, *, alpha_min, alpha_max, sigma_decay, alpha_u,
                                        beta_AV):
        """Cullen & Dehnen 2010 (pySPHModel.cpp:170-186, AVConfig VaryingCD10)."""
        self.d.update(av_kind=1, alpha_min=float(alpha_min), alpha_max=float(alpha_max),
                      sigma_decay=float(sigma_decay), alpha_u=float(alpha_u),
                      beta_AV=float(beta_AV))

    # units: G of the sink forces (SolverConfig.hpp:469-477)
    def set_units(self, units: UnitSystem):
        self.units = units
        self.d["constant_G"] = Constants(units).G()

    def get_constant_G(self):
        return self.d["constant_G"]

    def set_sink_softening(self, eps):
        """Plummer softening of the sink-gas force: an extension of this library (the
        reference's sink force is unsoftened); 0 restores the reference's force."""
        self.d["sink_softening"] = float(eps)

    def set_boundary_periodic(self):
        self.d["periodic"] = 1

    def set_boundary_free(self):
        self.d["periodic"] = 0

    def set_eos_adiabatic(self, gamma):
        self.d.update(eos_kind=0, gamma=float(gamma))

    def set_eos_isothermal(self, cs):
        self.d.update(eos_kind=1, eos_cs=float(cs))

    def set_eos_locally_isothermalLP07(self, cs0, q, r0):
        self.d.update(eos_kind=2, eos_cs0=float(cs0), eos_q=float(q), eos_r0=float(r0))

    def set_tree_reduction_level(self, level):
        self.d["tree_reduction_level"] = int(level)

    def set_two_stage_search(self, enable):
        if not enable:
            raise NotImplementedError("only the two-stage neighbour search is implemented")
        self.two_stage_search = True

    def set_smoothing_length_density_based(self):
        pass

    def set_cfl_cour(self, v):
        self.d["cfl_cour"] = float(v)

    def set_cfl_force(self, v):
        self.d["cfl_force"] = float(v)

    def set_cfl_mult_stiffness(self, v):
        self.d["cfl_multiplier_stiffness"] = float(v)

    def set_particle_mass(self, m):
        self.d["gpart_mass"] = float(m)

    def set_mixed_precision(self, enable: bool):
        """Build extension: FP32 pair sums with Kahan compensation on FP64 r_ab."""
        self.d["mixed_precision"] = int(bool(enable))

    def print_status(self):
        print(self.d)


class SetupNode:
    """A node of the setup graph (ISPHSetupNode, modules/setup/ISPHSetupNode.hpp): a lattice
    generator or a combiner of two nodes.  `positions()` returns the particles the node emits
    (positions and smoothing lengths) in emission order."""

    def __init__(self, kind, dr=0.0, box_min=None, box_max=None, parents=()):
        self.kind, self.dr = kind, float(dr)
        self.box_min = None if box_min is None else tuple(map(float, box_min))
        self.box_max = None if box_max is None else tuple(map(float, box_max))
        self.parents = tuple(parents)

    def positions(self):
        if self.kind == "combiner_add":
            parts = [p.positions() for p in self.parents]
            return (np.concatenate([q for q, _ in parts]), np.concatenate([h for _, h in parts]))
        gen = hcp_lattice if self.kind == "lattice_hcp" else lattice_cubic
        pos = gen(self.dr, self.box_min, self.box_max)
        return pos, np.full(pos.shape[0], self.dr)

    def get_dot(self):
        """Graphviz text of the node graph (ISPHSetupNode::get_dot)."""
        lines, ids = [], {}

        def visit(node):
            if id(node) in ids:
                return ids[id(node)]
            ids[id(node)] = k = len(ids)
            lines.append(f'  n{k} [label="{node.kind}"];')
            for p in node.parents:
                lines.append(f"  n{visit(p)} -> n{k};")
            return k

        visit(self)
        return "digraph setup {\n" + "\n".join(lines) + "\n}"


class _Setup:
    """SPHSetup subset (SPHSetup.cpp, pySPHModel.cpp:449-640): lattice generators, the add
    combiner and apply_setup.  Generators give h = their spacing and u = 0
    (GeneratorLatticeHCP.hpp:83-130, GeneratorLatticeCubic.hpp); the combiner emits its first
    parent's particles, then its second's."""

    def __init__(self, model):
        self.model = model

    def make_generator_lattice_cubic(self, dr, box_min, box_max):
        return SetupNode("lattice_cubic", dr, box_min, box_max)

    def make_generator_lattice_hcp(self, dr, box_min, box_max, discontinuous=True):
        """pySPHModel.cpp:452-462.  `discontinuous` selects the reference's emission order
        (an interleaved walk of the index ranges); the point set is the same either way and
        only particle ids differ, so one order is used here."""
        return SetupNode("lattice_hcp", dr, box_min, box_max)

    def make_combiner_add(self, parent1, parent2):
        """pySPHModel.cpp:624-630"""
        return SetupNode("combiner_add", parents=(parent1, parent2))

    def apply_setup(self, gen, **kw):
        """SPHSetup::apply_setup: insert the particles of the node graph (the keyword
        arguments of the reference tune its MPI insertion and are accepted and ignored)."""
        pos, h = gen.positions()
        self.model.push_particle(pos, h, np.zeros(pos.shape[0]))


class SPHModel:
    """SPHModel_f64_3_<K> (pySPHModel.cpp:761-1510), hot-path subset."""

    def __init__(self, context: Context, kernel: str):
        self.ctx = context
        self.kernel = kernel
        self.name = f"SPHModel_f64_3_{kernel}"
        self.cfg = SolverConfig(kernel)
        self.host = {k: np.zeros(0) for k in FIELDS}
        self.sinks = np.zeros((0, 17))
        # storage order of the particles (caller ids per slot) to restore when the solver is
        # built again, so that a rebuilt or reloaded solver sums in the same order
        self.storage_order = None
        self.gpu: GpuSolver | None = None
        self.t, self.dt, self.cfl_mult = 0.0, 0.0, 1e-2
        self.logs: list[dict] = []
        self.step_count = 0
        self.callbacks: list[tuple[Callable | None, Callable | None]] = []
        context.model = self

    # --- configuration -------------------------------------------------------------------
    def gen_default_config(self) -> SolverConfig:
        return SolverConfig(self.kernel)

    def set_solver_config(self, cfg: SolverConfig):
        self.cfg = cfg
        self._invalidate()

    def get_current_config(self) -> SolverConfig:
        return self.cfg

    def init_scheduler(self, split_crit=0, merge_crit=0):
        pass

    def init(self):
        pass

    def resize_simulation_box(self, box_min, box_max):
        self.cfg.d["box_min"] = [float(v) for v in box_min]
        self.cfg.d["box_max"] = [float(v) for v in box_max]
        self._invalidate()

    def get_setup(self):
        return _Setup(self)

    def get_box_dim_fcc_3d(self, dr, xcnt, ycnt, zcnt):
        """pySPHModel.cpp:802-806 (Model::get_box_dim_fcc_3d, generators.hpp:31-51): extent of
        xcnt x ycnt x zcnt cells of the HCP lattice of spacing dr."""
        return hcp_box_dim(float(dr), int(xcnt), int(ycnt), int(zcnt))

    def get_ideal_hcp_box(self, dr, box_min, box_max):
        """pySPHModel.cpp:818-827 (deprecated there in favour of shamrock.math)."""
        from .math import get_ideal_hcp_box

        return get_ideal_hcp_box(dr, box_min, box_max)

    def add_cube_fcc_3d(self, dr, box_min, box_max):
        """pySPHModel.cpp:849-853 (Model::add_cube_fcc_3d, Model.cpp:1093-1190): the lattice of
        add_particles_fcc (generators.hpp:82-130), anchored at box_min, kept inside the
        half-open box; h = dr, u = 0."""
        pos = fcc_cube(float(dr), box_min, box_max)
        self.push_particle(pos, np.full(pos.shape[0], float(dr)), np.zeros(pos.shape[0]))

    def make_analysis_sodtube(self, sod, direction, time_val, x_ref, x_min, x_max):
        """pySPHModel.cpp:1466-1485: an analysis whose compute_L2_dist() returns the mean
        squared deviations (rho, (vx, vy, vz), P) from the exact solution `sod`."""
        return AnalysisSodTube(self, sod, direction, time_val, x_ref, x_min, x_max)

    def get_closest_part_to(self, pos):
        """Model::get_closest_part_to: position of the particle nearest to pos."""
        p = self._field_get("xyz")
        if p.shape[0] == 0:
            raise RuntimeError("no particles")
        d = ((p - np.asarray(pos, float)) ** 2).sum(1)
        return tuple(float(v) for v in p[int(np.argmin(d))])

    def get_solver_tex(self):
        """A text description of the step this model runs (the reference renders its solver
        graph; this library's step is the fixed sequence below)."""
        av = "Cullen \\& Dehnen varying" if self.cfg.d.get("av_kind", 0) else "constant"
        return ("SPH step (" + self.kernel + " kernel, " + av + " viscosity): leapfrog "
                "predictor; periodic images; Morton codes, radix sort, leaf reduction, Karras "
                "tree, cell boxes; two-stage neighbour cache; smoothing-length Newton "
                "iteration; $\\Omega$; equation of state; forces and $du/dt$; leapfrog "
                "corrector with rerun on $\\epsilon_v > 10^{-2}$; CFL time step.")

    def get_solver_dot_graph(self):
        """Graphviz edges of the step's phases, in execution order."""
        phases = ["predictor", "ghosts", "tree", "neighbour_cache", "h_iteration", "omega",
                  "eos", "force", "corrector", "cfl"]
        return "\n".join(f"{a} -> {b};" for a, b in zip(phases, phases[1:]))

    # --- particles -----------------------------------------------------------------------
    def _sync_from_gpu(self):
        if self.gpu is not None:
            for k in self.host:
                self.host[k] = self.gpu.get(k)
            self.sinks = self.gpu.sink_records()
            self.storage_order = self.gpu.storage_order()
            tm = self.gpu.time()
            self.t, self.dt, self.cfl_mult = tm["t"], tm["dt"], tm["cfl_mult"]

    def _invalidate(self):
        self._sync_from_gpu()
        if self.gpu is not None:
            self.gpu.close()
        self.gpu = None

    def push_particle(self, pos, hpart, upart):
        self._invalidate()
        pos = np.asarray(pos, float).reshape(-1, 3)
        n = pos.shape[0]
        add = dict(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], h=np.asarray(hpart, float),
                   u=np.asarray(upart, float))
        for k in self.host:
            self.host[k] = np.concatenate([self.host[k], add.get(k, np.zeros(n))])
        self.storage_order = None

    def get_total_part_count(self):
        return int(self.host["x"].size) if self.gpu is None else self.gpu.n

    def _field_get(self, name):
        self._sync_from_gpu()
        if name in _VEC_FIELDS:
            return np.stack([self.host[k] for k in _VEC_FIELDS[name]], axis=1)
        return self.host[_SCAL_FIELDS[name]].copy()

    def _field_set(self, name, val):
        self._invalidate()
        if name in _VEC_FIELDS:
            val = np.asarray(val, float).reshape(-1, 3)
            for i, k in enumerate(_VEC_FIELDS[name]):
                self.host[k] = val[:, i].copy()
        else:
            self.host[_SCAL_FIELDS[name]] = np.asarray(val, float).copy()

    def set_value_in_a_box(self, field_name, field_type, val, box_min, box_max):
        pos = self._field_get("xyz")
        bmin, bmax = np.asarray(box_min, float), np.asarray(box_max, float)
        inside = np.all((pos >= bmin) & (pos < bmax), axis=1)
        f = self._field_get(field_name)
        f[inside] = val
        self._field_set(field_name, f)

    def set_field_value_lambda_f64(self, field_name, pos_to_val, offset=0):
        pos = self._field_get("xyz")
        f = self._field_get(field_name)
        f[offset:] = [pos_to_val(tuple(p)) for p in pos[offset:]]
        self._field_set(field_name, f)

    def set_field_value_lambda_f64_3(self, field_name, pos_to_val, offset=0):
        pos = self._field_get("xyz")
        f = self._field_get(field_name)
        f[offset:] = np.array([pos_to_val(tuple(p)) for p in pos[offset:]], float).reshape(-1, 3)
        self._field_set(field_name, f)

    def add_kernel_value(self, field_name, field_type, val, center, h_ker):
        """Model.hpp:739-768: f += val * W_3d(|r - center|, h_ker)."""
        pos = self._field_get("xyz")
        r = np.sqrt(((pos - np.asarray(center, float)) ** 2).sum(axis=1))
        f = self._field_get(field_name)
        f = f + val * kernel_W(self.kernel, r, h_ker)
        self._field_set(field_name, f)

    def get_sum(self, field_name, field_type):
        f = self._field_get(field_name)
        return f.sum(axis=0)

    def collect_data(self) -> dict:
        """pyShamrockCtx.cpp:97-128: every field of the layout, (N, 3) or (N,) float64."""
        names = _LAYOUT_CD10 if self.cfg.d.get("av_kind", 0) == 1 else _LAYOUT_CONSTANT
        return {k: self._field_get(k) for k in names}

    # --- restart dumps ---------------------------------------------------------------------
    def dump(self, fname):
        """Model::dump (Model.hpp:959-991): the whole state needed to continue the run, i.e.
        the solver config, the time edges (t, dt, CFL multiplier), the sinks and every field of
        the particles, in one file.  The layout is this library's own (see write_dump), not
        the reference's MPI patch serialisation."""
        self._sync_from_gpu()
        write_dump(fname, self.kernel, self.cfg.d, dict(t=self.t, dt=self.dt,
                                                         cfl_mult=self.cfl_mult),
                   self.sinks, self.host, self.step_count, self.storage_order)

    def load_from_dump(self, fname):
        """Model::load_from_dump (Model.hpp:906-957) for files written by dump."""
        kernel, cfg, tm, sinks, fields, steps, order = read_dump(fname)
        if kernel != self.kernel:
            raise ValueError(f"dump holds a {kernel} model, this model is {self.kernel}")
        self._invalidate()
        self.cfg = SolverConfig(kernel)
        self.cfg.d.update(cfg)
        self.t, self.dt, self.cfl_mult = tm["t"], tm["dt"], tm["cfl_mult"]
        self.sinks = sinks
        self.host = {k: fields.get(k, np.zeros(fields["x"].size)) for k in FIELDS}
        self.storage_order = order
        self.step_count = steps

    # --- sinks ---------------------------------------------------------------------------
    def add_sink(self, mass, pos, velocity, accretion_radius):
        """Model::add_sink (Model.hpp:183-200)."""
        self._sync_from_gpu()
        rec = np.zeros(17)
        rec[0:3], rec[3:6], rec[12], rec[16] = pos, velocity, mass, accretion_radius
        self.sinks = np.concatenate([self.sinks, rec[None]])
        if self.gpu is not None:
            self.gpu.set_sink_records(self.sinks)

    def get_sinks(self) -> list:
        """pySPHModel.cpp:1047-1065"""
        self._sync_from_gpu()
        return [sink_dict(r) for r in self.sinks]

    # --- mass / density ------------------------------------------------------------------
    def set_particle_mass(self, m):
        self.cfg.set_particle_mass(m)
        self._invalidate()

    def get_particle_mass(self):
        return self.cfg.d["gpart_mass"]

    def total_mass_to_part_mass(self, total_mass):
        return total_mass / self.get_total_part_count()

    def get_hfact(self):
        return KERNEL_HFACT[self.kernel]

    def rho_h(self, h):
        """density.hpp:22-25"""
        hf = self.get_hfact()
        m = self.get_particle_mass()
        return m * (hf / h) * (hf / h) * (hf / h)

    # --- time ----------------------------------------------------------------------------
    def _time(self):
        if self.gpu is not None:
            tm = self.gpu.time()
            self.t, self.dt, self.cfl_mult = tm["t"], tm["dt"], tm["cfl_mult"]

    def get_time(self):
        self._time()
        return self.t

    def get_dt(self):
        self._time()
        return self.dt

    def set_time(self, t):
        self._time()
        self.t = float(t)
        if self.gpu is not None:
            self.gpu.set_time(t=self.t)

    def set_next_dt(self, dt):
        self._time()
        self.dt = float(dt)
        if self.gpu is not None:
            self.gpu.set_time(dt=self.dt)

    def set_cfl_multipler(self, m):
        self._time()
        self.cfl_mult = float(m)
        if self.gpu is not None:
            self.gpu.set_time(cfl_mult=self.cfl_mult)

    def set_cfl_cour(self, v):
        self.cfg.set_cfl_cour(v)
        self._invalidate()

    def set_cfl_force(self, v):
        self.cfg.set_cfl_force(v)
        self._invalidate()

    def change_htolerances(self, coarse, fine):
        self.cfg.d.update(htol_up_coarse=float(coarse), htol_up_fine=float(fine))
        self._invalidate()

    # --- stepping ------------------------------------------------------------------------
    def _ensure_gpu(self):
        if self.gpu is None:
            n = self.host["x"].size
            if n == 0:
                raise RuntimeError("no particles: add particles before stepping")
            if self.cfg.d["gpart_mass"] == 0:
                raise RuntimeError("invalid gpart_mass 0, this configuration can not converge.\n"
                                   "Please set it using either model.set_particle_mass(pmass) "
                                   "or cfg.set_particle_mass(pmass)")
            self.gpu = GpuSolver(self.cfg.d, self.host)
            if self.storage_order is not None and self.storage_order.size == self.gpu.n:
                self.gpu.set_storage_order(self.storage_order)
            self.gpu.set_sink_records(self.sinks)
            self.gpu.set_time(self.t, self.dt, self.cfl_mult)
        return self.gpu

    def add_timestep_callback(self, step_begin=None, step_end=None):
        self.callbacks.append((step_begin, step_end))

    def evolve_once(self):
        """Solver::evolve_once (Solver.cpp:1942-3272)"""
        g = self._ensure_gpu()
        for b, _ in self.callbacks:
            if b:
                b()
        log = g.evolve_once()
        for _, e in reversed(self.callbacks):
            if e:
                e()
        self.logs.append(log)
        self.step_count += 1
        self._time()
        return log

    def evolve_once_override_time(self, t_curr, dt_input):
        """Solver::evolve_once_time_expl (Solver.hpp:298-303)"""
        self.set_time(t_curr)
        self.set_next_dt(dt_input)
        self.evolve_once()
        return self.get_dt()

    def timestep(self) -> TimestepLog:
        log = self.evolve_once()
        return TimestepLog(rank=0, rate=log["rate"], npart=log["n_local"],
                           tcompute=log["t_total"])

    def evolve_until(self, target_time, niter_max=-1, max_walltime=-1):
        """Solver::evolve_until (Solver.hpp:305-433)"""
        start = time.time()
        it = 0
        while self.get_time() < target_time:
            t, dt = self.get_time(), self.get_dt()
            if t > target_time:
                raise ValueError("the target time is higher than the current time")
            if t + dt > target_time:
                self.set_next_dt(target_time - t)
            self.evolve_once()
            it += 1
            if niter_max >= 0 and it >= niter_max:
                return EvolveUntilResults(False, True, False, it)
            if max_walltime >= 0 and time.time() - start >= max_walltime:
                return EvolveUntilResults(False, False, True, it)
        return EvolveUntilResults(True, False, False, it)

    # --- logs ----------------------------------------------------------------------------
    def solver_logs_last_rate(self):
        return self.logs[-1]["rate"] if self.logs else math.nan

    def solver_logs_last_obj_count(self):
        return self.logs[-1]["n_local"] if self.logs else 0

    def solver_logs_step_count(self):
        return self.step_count

    def solver_logs_reset_step_count(self):
        self.step_count = 0


_DUMP_MAGIC = b"SHAMXDMP"
_DUMP_VERSION = 2


def write_dump(fname, kernel, cfg, time_state, sinks, fields, step_count=0,
               storage_order=None):
    """Restart file: 8-byte magic, u64 version, u64 header length, a JSON header (kernel,
    config, time edges, step count, sink records, field names, particle count and whether a
    storage order follows), then each field as n little-endian float64 values in header order,
    then (if present) the storage order as n little-endian u32 particle ids."""
    n = int(np.asarray(fields["x"]).size)
    names = [k for k in FIELDS if k in fields]
    cfg = {k: (list(v) if isinstance(v, (list, tuple, np.ndarray)) else v)
           for k, v in cfg.items() if k != "sinks"}
    has_order = storage_order is not None and np.asarray(storage_order).size == n
    head = dict(kernel=kernel, config=cfg, time=dict(time_state), step_count=int(step_count),
                sinks=np.asarray(sinks, float).reshape(-1, 17).tolist(), n=n, fields=names,
                storage_order=bool(has_order))
    hb = json.dumps(head, sort_keys=True).encode()
    with open(fname, "wb") as f:
        f.write(_DUMP_MAGIC)
        f.write(np.array([_DUMP_VERSION, len(hb)], "<u8").tobytes())
        f.write(hb)
        for k in names:
            a = np.ascontiguousarray(fields[k], "<f8")
            if a.size != n:
                raise ValueError(f"field {k} has {a.size} values, expected {n}")
            f.write(a.tobytes())
        if has_order:
            f.write(np.ascontiguousarray(storage_order, "<u4").tobytes())


def read_dump(fname):
    """Inverse of write_dump: (kernel, config, time edges, sink records, fields, step count,
    storage order or None).  Version-1 files (no storage order) are read too."""
    with open(fname, "rb") as f:
        if f.read(8) != _DUMP_MAGIC:
            raise ValueError(f"{fname}: not a dump written by this library")
        version, hlen = np.frombuffer(f.read(16), "<u8")
        if int(version) not in (1, _DUMP_VERSION):
            raise ValueError(f"{fname}: dump version {int(version)} is not supported")
        head = json.loads(f.read(int(hlen)).decode())
        n = int(head["n"])
        fields = {}
        for k in head["fields"]:
            a = np.frombuffer(f.read(8 * n), "<f8")
            if a.size != n:
                raise ValueError(f"{fname}: truncated field {k}")
            fields[k] = a.astype(np.float64)
        order = None
        if head.get("storage_order", False):
            order = np.frombuffer(f.read(4 * n), "<u4")
            if order.size != n:
                raise ValueError(f"{fname}: truncated storage order")
            order = order.astype(np.uint32)
            if not np.array_equal(np.sort(order), np.arange(n, dtype=np.uint32)):
                raise ValueError(f"{fname}: the storage order is not a permutation")
    sinks = np.asarray(head["sinks"], float).reshape(-1, 17)
    return (head["kernel"], head["config"], head["time"], sinks, fields,
            int(head.get("step_count", 0)), order)


_KERNELS = ("M4", "M6", "M8", "C2", "C4", "C6")


def get_Model_SPH(*, context: Context, vector_type: str, sph_kernel: str) -> SPHModel:
    """pySPHModel.cpp:1725-1754"""
    if vector_type != "f64_3" or sph_kernel not in _KERNELS:
        raise ValueError("unknown combination of representation and kernel")
    return SPHModel(context, sph_kernel)
