"""Python handle on the C ABI Solver (shamx_solver_*): one patch, fields on the GPU.

Mirrors shammodels::sph::Solver<f64_3, Kern> (Solver.hpp:70-434) for the hot path: field
set/get, the time-state edges (t, dt, cfl multiplier) and evolve_once.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

# gravitational constant with no unit system set (shamunits Constants.hpp:130)
G_SI = 6.6743015e-11

FIELDS = ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "h", "u", "du", "alpha_AV",
          "soundspeed", "divv", "curlx", "curly", "curlz", "dtdivv", "axe", "aye", "aze"]


def sink_dict(r) -> dict:
    """One 17-double sink record with the keys of get_sinks (pySPHModel.cpp:1047-1065)."""
    r = np.asarray(r, float)
    return dict(pos=tuple(r[0:3]), velocity=tuple(r[3:6]), sph_acceleration=tuple(r[6:9]),
                ext_acceleration=tuple(r[9:12]), mass=float(r[12]),
                angular_momentum=tuple(r[13:16]), accretion_radius=float(r[16]))


def make_config(cfg: dict) -> _lib.ShamxConfig:
    c = _lib.ShamxConfig()
    kernel = cfg["kernel"]
    c.kernel = _lib.KERNEL_IDS[kernel] if isinstance(kernel, str) else int(kernel)
    for k in ("periodic", "eos_kind", "tree_reduction_level", "h_iter_per_subcycles",
              "h_max_subcycles", "mixed_precision"):
        setattr(c, k, int(cfg[k]))
    c.av_kind = int(cfg.get("av_kind", 0))
    c.keep_storage_order = int(cfg.get("keep_storage_order", 0))
    c.adjoint_fp32 = int(cfg.get("adjoint_fp32", 0))
    for k, d in (("alpha_min", 0.0), ("alpha_max", 1.0), ("sigma_decay", 0.1),
                 ("constant_G", G_SI), ("sink_softening", 0.0)):
        setattr(c, k, float(cfg.get(k, d)))
    for i in range(3):
        c.box_min[i] = float(cfg["box_min"][i])
        c.box_max[i] = float(cfg["box_max"][i])
    for k in ("gpart_mass", "gamma", "eos_cs", "eos_cs0", "eos_q", "eos_r0", "alpha_u",
              "alpha_AV", "beta_AV", "htol_up_coarse", "htol_up_fine", "epsilon_h", "cfl_cour",
              "cfl_force", "cfl_multiplier_stiffness"):
        setattr(c, k, float(cfg[k]))
    return c


class GpuSolver:
    """Owns a shamx_solver; fields are kept in HBM between steps."""

    def __init__(self, cfg: dict, fields: dict):
        self.cfg = dict(cfg)
        self.n = int(np.asarray(fields["x"]).size)
        self._c = make_config(cfg)
        h = C.c_void_p()
        check(lib().shamx_solver_create(C.byref(self._c), self.n, C.byref(h)))
        self.h = h
        for k, v in fields.items():
            self.set(k, v)
        for s in cfg.get("sinks", ()):
            self.add_sink(s["mass"], s["pos"], s["velocity"], s["accretion_radius"])

    def set(self, name, v):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (self.n,)))
        check(lib().shamx_solver_set_field(self.h, FIELDS.index(name), a.ctypes.data, 0))

    def set_device(self, name, tensor):
        check(lib().shamx_solver_set_field(self.h, FIELDS.index(name), tensor.data_ptr(), 1))

    def get(self, name):
        a = np.zeros(self.n)
        check(lib().shamx_solver_get_field(self.h, FIELDS.index(name), a.ctypes.data, 0))
        return a

    def storage_order(self) -> np.ndarray:
        """Caller particle id of every storage slot, in the order the next step runs in."""
        out = np.zeros(self.n, np.uint32)
        check(lib().shamx_solver_get_storage_order(self.h, out.ctypes.data))
        return out

    def set_storage_order(self, order):
        """Store the particles in `order` (a permutation of the caller's ids)."""
        a = np.ascontiguousarray(order, np.uint32)
        if a.size != self.n:
            raise ValueError(f"storage order has {a.size} ids, the solver {self.n} particles")
        check(lib().shamx_solver_set_storage_order(self.h, a.ctypes.data))

    def products_save(self, slot):
        """Keep what the adjoint of the last step reads in `slot` (the adjoint's product store)."""
        check(lib().shamx_solver_products_save(self.h, int(slot)))

    def products_load(self, slot):
        """Put the products of `slot` back (the fields then hold that step's end positions)."""
        check(lib().shamx_solver_products_load(self.h, int(slot)))

    def products_free(self):
        check(lib().shamx_solver_products_free(self.h))

    def get_device(self, name, out=None):
        """Device copy of a field as a torch tensor (float64, cuda)."""
        import torch

        if out is None:
            out = torch.empty(self.n, dtype=torch.float64, device="cuda")
        check(lib().shamx_solver_get_field(self.h, FIELDS.index(name), out.data_ptr(), 1))
        return out

    def snapshot(self) -> dict:
        """Device copy of the whole state (fields and time) for an exact replay."""
        st = {k: self.get_device(k) for k in FIELDS}
        st["_time"] = self.time()
        st["_sinks"] = self.sink_records()
        return st

    def restore(self, st: dict):
        for k in FIELDS:
            self.set_device(k, st[k])
        if "_sinks" in st:
            self.set_sink_records(st["_sinks"])
        tm = st["_time"]
        self.set_time(tm["t"], tm["dt"], tm["cfl_mult"])

    def adjoint_step(self, seeds):
        """Seeds (11 cuda float64 tensors: x y z vx vy vz u ax ay az du) at the end of the last
        step become those at its start (in place)."""
        arr = (C.c_void_p * 11)(*[t.data_ptr() for t in seeds])
        check(lib().shamx_solver_adjoint_step(self.h, arr))

    def adjoint_density(self, rho_bar, xbar, ybar, zbar):
        check(lib().shamx_solver_adjoint_density(self.h, rho_bar.data_ptr(), xbar.data_ptr(),
                                                 ybar.data_ptr(), zbar.data_ptr()))

    def time(self):
        a = np.zeros(3)
        check(lib().shamx_solver_get_time(self.h, _lib.dptr(a)))
        return dict(t=a[0], dt=a[1], cfl_mult=a[2])

    def set_time(self, t=None, dt=None, cfl_mult=None):
        cur = self.time()
        a = np.array([cur["t"] if t is None else t, cur["dt"] if dt is None else dt,
                      cur["cfl_mult"] if cfl_mult is None else cfl_mult])
        check(lib().shamx_solver_set_time(self.h, _lib.dptr(a)))

    def add_sink(self, mass, pos, velocity, accretion_radius):
        """Model::add_sink (Model.hpp:183-200)."""
        p = np.ascontiguousarray(pos, np.float64)
        v = np.ascontiguousarray(velocity, np.float64)
        check(lib().shamx_solver_add_sink(self.h, float(mass), _lib.dptr(p), _lib.dptr(v),
                                          float(accretion_radius)))

    def sink_records(self) -> np.ndarray:
        n = C.c_uint32(0)
        check(lib().shamx_solver_get_sinks(self.h, C.byref(n), None))
        out = np.zeros((n.value, 17))
        if n.value:
            check(lib().shamx_solver_get_sinks(self.h, C.byref(n), out.ctypes.data))
        return out

    def set_sink_records(self, rec):
        rec = np.ascontiguousarray(rec, np.float64).reshape(-1, 17)
        check(lib().shamx_solver_set_sinks(self.h, rec.shape[0],
                                           rec.ctypes.data if rec.size else None))

    def get_sinks(self) -> list:
        """get_sinks (pySPHModel.cpp:1047-1065): one dict per sink."""
        return [sink_dict(r) for r in self.sink_records()]

    def evolve_once(self) -> dict:
        log = _lib.ShamxStepLog()
        check(lib().shamx_solver_evolve_once(self.h, C.byref(log)))
        return log.as_dict()

    def density_force(self) -> dict:
        """One density and force pass without time integration: tree, neighbour cache, one
        smoothing-length subcycle (unconverged h kept clamped), Omega, EOS, forces and the CFL
        step.  Constant viscosity only."""
        log = _lib.ShamxStepLog()
        check(lib().shamx_solver_density_force(self.h, C.byref(log)))
        return log.as_dict()

    def stream(self) -> int:
        s = C.c_void_p()
        check(lib().shamx_solver_stream(self.h, C.byref(s)))
        return s.value or 0

    def sizes(self) -> dict:
        s = np.zeros(6, np.uint64)
        check(lib().shamx_solver_sizes(self.h, s.ctypes.data_as(C.POINTER(C.c_uint64))))
        return dict(Nm=int(s[0]), G=int(s[1]), L=int(s[2]), I=int(s[3]), P=int(s[4]),
                    PL=int(s[5]))

    def get_f64(self, what, n):
        out = np.zeros(max(n, 1))
        check(lib().shamx_solver_get_f64(self.h, what, out.ctypes.data))
        return out[:n]

    def get_u32(self, what, n):
        out = np.zeros(max(n, 1), np.uint32)
        check(lib().shamx_solver_get_u32(self.h, what, out.ctypes.data))
        return out[:n]

    def get_u8(self, what, n):
        out = np.zeros(max(n, 1), np.uint8)
        check(lib().shamx_solver_get_u8(self.h, what, out.ctypes.data))
        return out[:n]

    def transient(self) -> dict:
        """Intermediate products of the last step (same keys as the oracle's)."""
        s = self.sizes()
        Nm, G, L, I, P, PL = s["Nm"], s["G"], s["L"], s["I"], s["P"], s["PL"]
        N = self.n
        d = dict(sizes=s)
        for i, k in enumerate(["mx", "my", "mz", "mh", "mvx", "mvy", "mvz", "mu", "momega", "P",
                               "cs"]):
            d[k] = self.get_f64(i, Nm)
        d["omega"] = self.get_f64(11, N)
        d["vsig"] = self.get_f64(12, N)
        d["tree_bbox"] = self.get_f64(13, 6)
        d["rint_tree"] = self.get_f64(14, I + L)
        d["aabb_min"] = self.get_f64(15, 3 * (I + L)).reshape(-1, 3)
        d["aabb_max"] = self.get_f64(16, 3 * (I + L)).reshape(-1, 3)
        for i, k, n in [(0, "ghost_src", G), (1, "sorted_codes", Nm), (2, "sort_index", Nm),
                        (3, "reduc_map", L + 2), (4, "owner", N), (5, "part_cnt", N),
                        (6, "part_offset", N), (7, "part_ids", P), (8, "leaf_cnt", L),
                        (9, "leaf_offset", L), (10, "leaf_ids", PL), (11, "lchild", I),
                        (12, "rchild", I), (13, "endrange", I), (14, "codes", Nm)]:
            d[k] = self.get_u32(i, n)
        d["lflag"] = self.get_u8(0, I)
        d["rflag"] = self.get_u8(1, I)
        return d

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().shamx_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
