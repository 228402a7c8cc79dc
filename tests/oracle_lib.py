"""ctypes binding of the CPU oracle (oracle/_build/libsph_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this module.  The oracle restates the reference SPH step on the CPU (see oracle/sph_oracle.hpp).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "libsph_oracle.so")
# a prebuilt oracle to load instead (the sanitizer build of test_oracle_sanitized.py)
LIB_OVERRIDE = os.environ.get("SHAMX_ORACLE_LIB")

KERNELS = {"M4": 0, "M6": 1, "M8": 2, "C2": 3, "C4": 4, "C6": 5}

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OrcConfig(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32), ("periodic", C.c_int32), ("eos_kind", C.c_int32),
        ("tree_reduction_level", C.c_int32), ("h_iter_per_subcycles", C.c_int32),
        ("h_max_subcycles", C.c_int32),
        ("box_min", C.c_double * 3), ("box_max", C.c_double * 3),
        ("gpart_mass", C.c_double), ("gamma", C.c_double), ("eos_cs", C.c_double),
        ("eos_cs0", C.c_double), ("eos_q", C.c_double), ("eos_r0", C.c_double),
        ("alpha_u", C.c_double), ("alpha_AV", C.c_double), ("beta_AV", C.c_double),
        ("htol_up_coarse", C.c_double), ("htol_up_fine", C.c_double), ("epsilon_h", C.c_double),
        ("cfl_cour", C.c_double), ("cfl_force", C.c_double),
        ("cfl_multiplier_stiffness", C.c_double),
        ("av_kind", C.c_int32), ("alpha_min", C.c_double), ("alpha_max", C.c_double),
        ("sigma_decay", C.c_double), ("constant_G", C.c_double), ("sink_softening", C.c_double),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("t_total", C.c_double), ("t_ghost", C.c_double), ("t_tree", C.c_double),
        ("t_neigh", C.c_double), ("t_hiter", C.c_double), ("t_omega", C.c_double),
        ("t_eos", C.c_double), ("t_force", C.c_double), ("t_corr", C.c_double),
        ("t_cfl", C.c_double),
        ("n_merged", C.c_uint32), ("n_leaf", C.c_uint32), ("subcycles", C.c_uint32),
        ("h_iterations", C.c_uint32), ("corrector_iterations", C.c_uint32),
        ("n_pairs", C.c_uint64),
        ("energy", C.c_double), ("eps_v", C.c_double), ("next_dt", C.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if LIB_OVERRIDE:
            L = C.CDLL(LIB_OVERRIDE)
        else:
            build()
            L = C.CDLL(LIB_PATH)
        L.orc_last_error.restype = C.c_char_p
        L.orc_kernel_eval.argtypes = [C.c_int, C.c_uint32, _f64p, _f64p, _f64p, _f64p, _f64p,
                                      _f64p, _f64p, _f64p]
        L.orc_bbox.argtypes = [C.c_uint32, _f64p, _f64p, _f64p, _f64p]
        L.orc_morton_u32.argtypes = [C.c_uint32, _f64p, _f64p, _f64p, _f64p, C.c_uint32, _u32p]
        L.orc_morton_u64.argtypes = [C.c_uint32, _f64p, _f64p, _f64p, _f64p, C.c_uint32, _u64p]
        L.orc_bitonic_u32.argtypes = [_u32p, _u32p, C.c_uint32]
        L.orc_bitonic_u64.argtypes = [_u64p, _u32p, C.c_uint32]
        L.orc_reduce_u32.argtypes = [_u32p, C.c_uint32, C.c_uint32, _u32p, C.POINTER(C.c_uint32)]
        L.orc_reduce_u64.argtypes = [_u64p, C.c_uint32, C.c_uint32, _u32p, C.POINTER(C.c_uint32)]
        L.orc_karras_u32.argtypes = [_u32p, C.c_uint32, _u32p, _u32p, _u8p, _u8p, _u32p]
        L.orc_karras_u64.argtypes = [_u64p, C.c_uint32, _u32p, _u32p, _u8p, _u8p, _u32p]
        L.orc_tree_build.restype = C.c_void_p
        L.orc_tree_build.argtypes = [C.c_int, C.c_uint32, _f64p, _f64p, _f64p, _f64p, C.c_uint32,
                                     C.c_uint32]
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_sizes.argtypes = [C.c_void_p, _u32p]
        L.orc_tree_get_u32.argtypes = [C.c_void_p, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p,
                                       _u32p, _u8p, _u8p, _u32p, _f64p, _f64p]
        L.orc_tree_get_u64.argtypes = [C.c_void_p, _u64p, _u64p, _u32p, _u32p, _u64p, _u32p,
                                       _u32p, _u8p, _u8p, _u32p, _f64p, _f64p]
        L.orc_tree_field_max.argtypes = [C.c_void_p, _f64p, _f64p]
        L.orc_tree_box_query.argtypes = [C.c_void_p, C.c_uint32, _f64p, _f64p, _f64p, C.c_double,
                                         _u32p, _u32p]
        L.orc_iterate_h.argtypes = [C.c_int, C.c_uint32, _u32p, _u32p, _u32p, C.c_uint32, _f64p,
                                    _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                    C.c_double]
        L.orc_cd10_operators.argtypes = ([C.c_int, C.c_uint32, _u32p, _u32p, _u32p, C.c_uint32] +
                                         [_f64p] * 11 + [C.c_double] + [_f64p] * 5)
        L.orc_cd10_alpha.argtypes = ([C.c_uint32, C.c_double] + [_f64p] * 8 +
                                     [C.c_double] * 3 + [_f64p])
        L.orc_sink_gas.argtypes = [C.c_uint32, _f64p, _f64p, _f64p, C.c_uint32, _f64p, C.c_double,
                                   C.c_double, C.c_double, _f64p]
        L.orc_periodic_ghosts.argtypes = [C.c_uint32, _f64p, _f64p, _f64p, _f64p, C.c_double,
                                          C.c_uint32, _u32p, _f64p, C.POINTER(C.c_uint32)]
        L.orc_neigh_build.restype = C.c_void_p
        L.orc_neigh_build.argtypes = [C.c_uint32, _f64p, _f64p, _f64p, _f64p, C.c_uint32,
                                      C.c_uint32, C.c_double, C.c_double, _u64p]
        L.orc_neigh_get.argtypes = [C.c_void_p, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p]
        L.orc_neigh_free.argtypes = [C.c_void_p]
        L.orc_solver_new.restype = C.c_void_p
        L.orc_solver_new.argtypes = [C.POINTER(OrcConfig), C.c_uint32]
        L.orc_solver_free.argtypes = [C.c_void_p]
        L.orc_solver_field.argtypes = [C.c_void_p, C.c_int, C.c_int, _f64p, C.c_int]
        L.orc_solver_time.argtypes = [C.c_void_p, C.c_int, _f64p, C.c_int]
        L.orc_solver_add_sink.argtypes = [C.c_void_p, C.c_int, _f64p]
        L.orc_solver_get_sinks.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_uint32), C.c_void_p]
        L.orc_solver_evolve_once.argtypes = [C.c_void_p, C.c_int, C.POINTER(StepStats)]
        L.orc_solver_density_force.argtypes = [C.c_void_p, C.c_int, C.POINTER(StepStats)]
        L.orc_solver_sizes.argtypes = [C.c_void_p, C.c_int, _u64p]
        L.orc_solver_get_f64.argtypes = [C.c_void_p, C.c_int, C.c_int, _f64p]
        L.orc_solver_get_u32.argtypes = [C.c_void_p, C.c_int, C.c_int, _u32p]
        L.orc_solver_get_u8.argtypes = [C.c_void_p, C.c_int, C.c_int, _u8p]
        L.orc_solver_get_aabb.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p]
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def soa(pos):
    p = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    return _f(p[:, 0]), _f(p[:, 1]), _f(p[:, 2])


# ---------------------------------------------------------------- single phases
def kernel_eval(name, r, h):
    r, h = _f(r), _f(h)
    n = r.size
    out = [np.zeros(n) for _ in range(5)]
    consts = np.zeros(3)
    _chk(lib().orc_kernel_eval(KERNELS[name], n, r, h, *out, consts))
    W, dW, dhW, f, df = out
    return dict(W=W, dW=dW, dhW=dhW, f=f, df=df, Rkern=consts[0], hfactd=consts[1],
                norm_3d=consts[2])


def bbox(pos):
    x, y, z = soa(pos)
    out = np.zeros(6)
    _chk(lib().orc_bbox(x.size, x, y, z, out))
    return out


def morton(pos, bb6, morton_count, bits=32):
    x, y, z = soa(pos)
    bb6 = _f(bb6)
    if bits == 32:
        out = np.zeros(morton_count, np.uint32)
        _chk(lib().orc_morton_u32(x.size, x, y, z, bb6, morton_count, out))
    else:
        out = np.zeros(morton_count, np.uint64)
        _chk(lib().orc_morton_u64(x.size, x, y, z, bb6, morton_count, out))
    return out


def bitonic(keys, vals=None):
    keys = np.array(keys, copy=True)
    if vals is None:
        vals = np.arange(keys.size, dtype=np.uint32)
    vals = np.ascontiguousarray(vals, dtype=np.uint32).copy()
    if keys.dtype == np.uint32:
        _chk(lib().orc_bitonic_u32(keys, vals, keys.size))
    else:
        keys = keys.astype(np.uint64)
        _chk(lib().orc_bitonic_u64(keys, vals, keys.size))
    return keys, vals


def reduce(sorted_codes, cnt, level):
    out = np.zeros(cnt + 2, np.uint32)
    L = C.c_uint32(0)
    if sorted_codes.dtype == np.uint32:
        _chk(lib().orc_reduce_u32(np.ascontiguousarray(sorted_codes), cnt, level, out, C.byref(L)))
    else:
        _chk(lib().orc_reduce_u64(np.ascontiguousarray(sorted_codes, dtype=np.uint64), cnt, level,
                                  out, C.byref(L)))
    return out[: L.value + 2]


def karras(codes):
    L = codes.size
    I = max(L - 1, 0)
    lc, rc, er = (np.zeros(max(I, 1), np.uint32) for _ in range(3))
    lf, rf = (np.zeros(max(I, 1), np.uint8) for _ in range(2))
    f = lib().orc_karras_u32 if codes.dtype == np.uint32 else lib().orc_karras_u64
    _chk(f(np.ascontiguousarray(codes), L, lc, rc, lf, rf, er))
    return dict(lchild=lc[:I], rchild=rc[:I], lflag=lf[:I], rflag=rf[:I], endrange=er[:I])


class Tree:
    """CLBVH built by the oracle (CompressedLeafBVH::rebuild_from_positions)."""

    def __init__(self, pos, bb6, level, bits=32, morton_count=0):
        self.x, self.y, self.z = soa(pos)
        self.bits = bits
        self.h = lib().orc_tree_build(bits, self.x.size, self.x, self.y, self.z, _f(bb6), level,
                                     morton_count)
        if not self.h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        s = np.zeros(3, np.uint32)
        lib().orc_tree_sizes(self.h, s)
        self.morton_count, self.L, self.I = (int(v) for v in s)
        mt = np.uint32 if bits == 32 else np.uint64
        n = self.x.size
        self.codes = np.zeros(self.morton_count, mt)
        self.sorted_codes = np.zeros(self.morton_count, mt)
        self.sort_index = np.zeros(self.morton_count, np.uint32)
        self.reduc_map = np.zeros(self.L + 2, np.uint32)
        self.leaf_codes = np.zeros(self.L, mt)
        k = max(self.I, 1)
        self.lchild, self.rchild, self.endrange = (np.zeros(k, np.uint32) for _ in range(3))
        self.lflag, self.rflag = (np.zeros(k, np.uint8) for _ in range(2))
        self.aabb_min = np.zeros(3 * (self.I + self.L))
        self.aabb_max = np.zeros(3 * (self.I + self.L))
        g = lib().orc_tree_get_u32 if bits == 32 else lib().orc_tree_get_u64
        g(self.h, self.codes, self.sorted_codes, self.sort_index, self.reduc_map, self.leaf_codes,
          self.lchild, self.rchild, self.lflag, self.rflag, self.endrange, self.aabb_min,
          self.aabb_max)
        for a in ("lchild", "rchild", "endrange", "lflag", "rflag"):
            setattr(self, a, getattr(self, a)[: self.I])
        self.aabb_min = self.aabb_min.reshape(-1, 3)
        self.aabb_max = self.aabb_max.reshape(-1, 3)

    def field_max(self, field):
        out = np.zeros(self.I + self.L)
        _chk(lib().orc_tree_field_max(self.h, _f(field), out))
        return out

    def box_query(self, s):
        n = self.x.size
        counts = np.zeros(n, np.uint32)
        neigh = np.zeros(n * n, np.uint32)
        _chk(lib().orc_tree_box_query(self.h, n, self.x, self.y, self.z, s, counts, neigh))
        return counts, neigh[: int(counts.sum())]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None


def iterate_h(kernel, cnt, offset, ids, pos, h_old, h_new, eps, mass, h_evol_max,
              h_evol_iter_max):
    x, y, z = soa(pos)
    _chk(lib().orc_iterate_h(KERNELS[kernel], cnt.size, np.ascontiguousarray(cnt, np.uint32),
                             np.ascontiguousarray(offset, np.uint32),
                             np.ascontiguousarray(ids, np.uint32), ids.size, x, y, z, _f(h_old),
                             h_new, eps, mass, h_evol_max, h_evol_iter_max))


def cd10_operators(kernel, cache, pos, vel, acc, h, omega, mass):
    """div v, curl v and d(div v)/dt of the cache rows (DiffOperator.cpp:24-268,
    DiffOperatorDtDivv.cpp:89-175)."""
    x, y, z = soa(pos)
    vx, vy, vz = soa(vel)
    ax, ay, az = soa(acc)
    n = cache["cnt"].size
    out = [np.zeros(n) for _ in range(5)]
    ids = np.ascontiguousarray(cache["ids"], np.uint32)
    _chk(lib().orc_cd10_operators(KERNELS[kernel], n, np.ascontiguousarray(cache["cnt"], np.uint32),
                                  np.ascontiguousarray(cache["offset"], np.uint32), ids, ids.size,
                                  x, y, z, vx, vy, vz, ax, ay, az, _f(h), _f(omega), mass, *out))
    return dict(zip(("divv", "curlx", "curly", "curlz", "dtdivv"), out))


def cd10_alpha(dt, cs, h, alpha, divv, curl, dtdivv, alpha_min, alpha_max, sigma_decay):
    """Cullen & Dehnen alpha update (UpdateViscosity.cpp:120-219)."""
    cs = _f(cs)
    out = np.zeros(cs.size)
    _chk(lib().orc_cd10_alpha(cs.size, dt, cs, _f(h), _f(alpha), _f(divv), _f(curl[0]),
                              _f(curl[1]), _f(curl[2]), _f(dtdivv), alpha_min, alpha_max,
                              sigma_decay, out))
    return out


def sink_gas(pos, sinks, G, gpart_mass, soft=0.0):
    """Sink-gas gravity (SinkParticlesUpdate.cpp:69-150): the particles' external accelerations
    (n, 3) and each sink's acceleration from the gas (ns, 3)."""
    x, y, z = soa(pos)
    rec = np.zeros((len(sinks), 17))
    for i, s in enumerate(sinks):
        rec[i, 0:3], rec[i, 12], rec[i, 16] = s["pos"], s["mass"], s["accretion_radius"]
    rec = np.ascontiguousarray(rec)
    ext = np.zeros(3 * x.size)
    _chk(lib().orc_sink_gas(x.size, x, y, z, len(sinks), rec.reshape(-1), G, gpart_mass, soft,
                            ext))
    return ext.reshape(3, -1).T.copy(), rec[:, 6:9].copy()


def periodic_ghosts(pos, box6, int_range):
    x, y, z = soa(pos)
    cap = 27 * x.size
    src = np.zeros(cap, np.uint32)
    off = np.zeros(3 * cap)
    cnt = C.c_uint32(0)
    _chk(lib().orc_periodic_ghosts(x.size, x, y, z, _f(box6), int_range, cap, src, off,
                                   C.byref(cnt)))
    n = cnt.value
    return src[:n], off[: 3 * n].reshape(-1, 3)


def neigh_cache(pos, h, n_local, level=3, htol=1.1, Rkern=2.0):
    """Two-stage neighbour cache (NeighbourCache.cpp:222-604) on arbitrary positions."""
    x, y, z = soa(pos)
    h = _f(h)
    sz = np.zeros(4, np.uint64)
    hd = lib().orc_neigh_build(x.size, x, y, z, h, n_local, level, htol, Rkern, sz)
    if not hd:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())
    L, I, P, PL = (int(v) for v in sz)
    out = dict(cnt=np.zeros(max(n_local, 1), np.uint32), offset=np.zeros(max(n_local, 1), np.uint32),
               ids=np.zeros(max(P, 1), np.uint32), owner=np.zeros(max(n_local, 1), np.uint32),
               leaf_cnt=np.zeros(max(L, 1), np.uint32), leaf_ids=np.zeros(max(PL, 1), np.uint32))
    lib().orc_neigh_get(hd, out["cnt"], out["offset"], out["ids"], out["owner"], out["leaf_cnt"],
                        out["leaf_ids"])
    lib().orc_neigh_free(hd)
    for k, n in (("cnt", n_local), ("offset", n_local), ("owner", n_local), ("ids", P),
                 ("leaf_cnt", L), ("leaf_ids", PL)):
        out[k] = out[k][:n]
    out.update(L=L, I=I, P=P, PL=PL)
    return out


# ---------------------------------------------------------------- full solver
FIELDS = ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "h", "u", "du", "alpha_AV",
          "soundspeed", "divv", "curlx", "curly", "curlz", "dtdivv", "axe", "aye", "aze"]


def sink_dict(r):
    """One sink record (17 doubles) as the keys of get_sinks (pySPHModel.cpp:1047-1065)."""
    r = np.asarray(r, float)
    return dict(pos=r[0:3].copy(), velocity=r[3:6].copy(), sph_acceleration=r[6:9].copy(),
                ext_acceleration=r[9:12].copy(), mass=float(r[12]),
                angular_momentum=r[13:16].copy(), accretion_radius=float(r[16]))


class Solver:
    """Oracle Solver<f64_3, Kern>: evolve_once on one periodic or free patch."""

    def __init__(self, cfg: dict, fields: dict):
        c = OrcConfig()
        self.kernel = KERNELS[cfg["kernel"]]
        c.kernel = self.kernel
        c.periodic = int(cfg.get("periodic", 1))
        c.eos_kind = int(cfg.get("eos_kind", 0))
        c.tree_reduction_level = int(cfg.get("tree_reduction_level", 3))
        c.h_iter_per_subcycles = int(cfg.get("h_iter_per_subcycles", 50))
        c.h_max_subcycles = int(cfg.get("h_max_subcycles", 100))
        for i in range(3):
            c.box_min[i] = cfg["box_min"][i]
            c.box_max[i] = cfg["box_max"][i]
        c.gpart_mass = cfg["gpart_mass"]
        c.gamma = cfg.get("gamma", 5.0 / 3.0)
        c.eos_cs = cfg.get("eos_cs", 1.0)
        c.eos_cs0 = cfg.get("eos_cs0", 1.0)
        c.eos_q = cfg.get("eos_q", 0.5)
        c.eos_r0 = cfg.get("eos_r0", 1.0)
        c.alpha_u = cfg.get("alpha_u", 1.0)
        c.alpha_AV = cfg.get("alpha_AV", 1.0)
        c.beta_AV = cfg.get("beta_AV", 2.0)
        c.htol_up_coarse = cfg.get("htol_up_coarse", 1.1)
        c.htol_up_fine = cfg.get("htol_up_fine", 1.1)
        c.epsilon_h = cfg.get("epsilon_h", 1e-6)
        c.cfl_cour = cfg.get("cfl_cour", 0.3)
        c.cfl_force = cfg.get("cfl_force", 0.3)
        c.cfl_multiplier_stiffness = cfg.get("cfl_multiplier_stiffness", 2.0)
        c.av_kind = int(cfg.get("av_kind", 0))
        c.alpha_min = cfg.get("alpha_min", 0.0)
        c.alpha_max = cfg.get("alpha_max", 1.0)
        c.sigma_decay = cfg.get("sigma_decay", 0.1)
        c.constant_G = cfg.get("constant_G", 6.6743015e-11)
        c.sink_softening = cfg.get("sink_softening", 0.0)
        self.cfg = c
        self.n = int(np.asarray(fields["x"]).size)
        self.h = lib().orc_solver_new(C.byref(c), self.n)
        if not self.h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        for k, v in fields.items():
            self.set(k, v)
        for s in cfg.get("sinks", ()):
            self.add_sink(s["mass"], s["pos"], s["velocity"], s["accretion_radius"])

    def set(self, name, v):
        a = _f(np.broadcast_to(np.asarray(v, dtype=np.float64), (self.n,)))
        _chk(lib().orc_solver_field(self.h, self.kernel, FIELDS.index(name), a.copy(), 1))

    def get(self, name):
        a = np.zeros(self.n)
        _chk(lib().orc_solver_field(self.h, self.kernel, FIELDS.index(name), a, 0))
        return a

    def time(self):
        a = np.zeros(3)
        _chk(lib().orc_solver_time(self.h, self.kernel, a, 0))
        return dict(t=a[0], dt=a[1], cfl_mult=a[2])

    def set_time(self, t=None, dt=None, cfl_mult=None):
        cur = self.time()
        a = np.array([cur["t"] if t is None else t, cur["dt"] if dt is None else dt,
                      cur["cfl_mult"] if cfl_mult is None else cfl_mult])
        _chk(lib().orc_solver_time(self.h, self.kernel, a, 1))

    def add_sink(self, mass, pos, velocity, accretion_radius):
        rec = np.zeros(17)
        rec[0:3], rec[3:6], rec[12], rec[16] = pos, velocity, mass, accretion_radius
        _chk(lib().orc_solver_add_sink(self.h, self.kernel, rec))

    def get_sinks(self):
        n = C.c_uint32(0)
        _chk(lib().orc_solver_get_sinks(self.h, self.kernel, C.byref(n), None))
        out = np.zeros(17 * max(n.value, 1))
        _chk(lib().orc_solver_get_sinks(self.h, self.kernel, C.byref(n), out.ctypes.data))
        return [sink_dict(out[17 * i:17 * i + 17]) for i in range(n.value)]

    def evolve_once(self):
        st = StepStats()
        _chk(lib().orc_solver_evolve_once(self.h, self.kernel, C.byref(st)))
        return {k: getattr(st, k) for k, _ in StepStats._fields_}

    def density_force(self):
        st = StepStats()
        _chk(lib().orc_solver_density_force(self.h, self.kernel, C.byref(st)))
        return {k: getattr(st, k) for k, _ in StepStats._fields_}

    def sizes(self):
        s = np.zeros(6, np.uint64)
        _chk(lib().orc_solver_sizes(self.h, self.kernel, s))
        return dict(Nm=int(s[0]), G=int(s[1]), L=int(s[2]), I=int(s[3]), P=int(s[4]),
                    PL=int(s[5]))

    def get_f64(self, what, n):
        out = np.zeros(n)
        _chk(lib().orc_solver_get_f64(self.h, self.kernel, what, out))
        return out

    def get_u32(self, what, n):
        out = np.zeros(max(n, 1), np.uint32)
        _chk(lib().orc_solver_get_u32(self.h, self.kernel, what, out))
        return out[:n]

    def transient(self):
        """Intermediate products of the last step, keyed like the GPU library's getters."""
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
        d["ghost_src"] = self.get_u32(0, G)
        d["sorted_codes"] = self.get_u32(1, Nm)
        d["sort_index"] = self.get_u32(2, Nm)
        d["reduc_map"] = self.get_u32(3, L + 2)
        d["owner"] = self.get_u32(4, N)
        d["part_cnt"] = self.get_u32(5, N)
        d["part_offset"] = self.get_u32(6, N)
        d["part_ids"] = self.get_u32(7, P)
        d["leaf_cnt"] = self.get_u32(8, L)
        d["leaf_offset"] = self.get_u32(9, L)
        d["leaf_ids"] = self.get_u32(10, PL)
        d["lchild"] = self.get_u32(11, I)
        d["rchild"] = self.get_u32(12, I)
        d["endrange"] = self.get_u32(13, I)
        d["codes"] = self.get_u32(14, Nm)
        lf = np.zeros(max(I, 1), np.uint8)
        rf = np.zeros(max(I, 1), np.uint8)
        _chk(lib().orc_solver_get_u8(self.h, self.kernel, 0, lf))
        _chk(lib().orc_solver_get_u8(self.h, self.kernel, 1, rf))
        d["lflag"], d["rflag"] = lf[:I], rf[:I]
        amin = np.zeros(3 * (I + L))
        amax = np.zeros(3 * (I + L))
        _chk(lib().orc_solver_get_aabb(self.h, self.kernel, amin, amax))
        d["aabb_min"], d["aabb_max"] = amin.reshape(-1, 3), amax.reshape(-1, 3)
        return d

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solver_free(self.h)
            self.h = None
