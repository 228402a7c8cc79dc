"""ctypes binding of libshamx_sph.so (the C ABI declared in include/shamx_sph.h).

The library is loaded from shamrock_amd/_lib/ (built in-tree by shamrock_amd/build.py).  There
is no fallback: if the library is missing or cannot be loaded, importing the binding raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libshamx_sph.so")

KERNEL_IDS = {"M4": 0, "M6": 1, "M8": 2, "C2": 3, "C4": 4, "C6": 5}
EOS_IDS = {"adiabatic": 0, "isothermal": 1, "locally_isothermal_lp07": 2}
FIELD_IDS = {"xyz_x": 0, "xyz_y": 1, "xyz_z": 2, "vxyz_x": 3, "vxyz_y": 4, "vxyz_z": 5,
             "axyz_x": 6, "axyz_y": 7, "axyz_z": 8, "hpart": 9, "uint": 10, "duint": 11}

ERRORS = {1: ValueError, 2: RuntimeError, 3: RuntimeError, 4: MemoryError, 5: RuntimeError}


class ShamxConfig(C.Structure):
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
        ("mixed_precision", C.c_int32), ("av_kind", C.c_int32),
        ("alpha_min", C.c_double), ("alpha_max", C.c_double), ("sigma_decay", C.c_double),
        ("constant_G", C.c_double), ("sink_softening", C.c_double),
        ("keep_storage_order", C.c_int32), ("adjoint_fp32", C.c_int32),
    ]


class ShamxStepLog(C.Structure):
    _fields_ = [
        ("rate", C.c_double), ("t_total", C.c_double), ("t_ghost", C.c_double),
        ("t_tree", C.c_double), ("t_sort", C.c_double), ("t_neigh", C.c_double),
        ("t_hiter", C.c_double),
        ("t_omega", C.c_double), ("t_eos", C.c_double), ("t_force", C.c_double),
        ("t_corr", C.c_double),
        ("n_local", C.c_uint32), ("n_merged", C.c_uint32), ("n_leaf", C.c_uint32),
        ("subcycles", C.c_uint32), ("h_iterations", C.c_uint32),
        ("corrector_iterations", C.c_uint32), ("n_sorts", C.c_uint32),
        ("n_force_evals", C.c_uint32),
        ("n_pairs", C.c_uint64), ("n_interacting", C.c_uint64),
        ("energy", C.c_double), ("eps_v", C.c_double), ("next_dt", C.c_double),
        ("time", C.c_double), ("dt", C.c_double), ("t_reorder", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every symbol include/shamx_sph.h declares, with its ctypes signature
_P = C.c_void_p
_u32 = C.c_uint32
_d = C.c_double
_i = C.c_int
_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "shamx_last_error": (C.c_char_p, []),
    "shamx_get_device_count": (_i, [C.POINTER(C.c_int)]),
    "shamx_ctx_create": (_i, [C.POINTER(_P)]),
    "shamx_ctx_destroy": (_i, [_P]),
    "shamx_bbox": (_i, [_P, _P, _u32, _P, _P, _P, _P]),
    "shamx_morton": (_i, [_P, _P, _u32, _P, _P, _P, _P, _P]),
    "shamx_sort": (_i, [_P, _P, _u32, _u32, _P, _P]),
    "shamx_sort_noalloc": (_i, [_P, _u32, _u32, _P, _P]),
    "shamx_tree_build": (_i, [_P, _P, _u32, _P, _P, _P, _P, _u32, _d]),
    "shamx_tree_reduce_karras": (_i, [_P, _P, _u32, _P, _P, _P, _u32]),
    "shamx_tree_aabb_hmax": (_i, [_P, _P, _P, _P, _P, _P, _d]),
    "shamx_tree_sizes": (_i, [_P, C.POINTER(C.c_uint32)]),
    "shamx_tree_get": (_i, [_P, _P] + [_P] * 11),
    "shamx_neigh_build": (_i, [_P, _P, _u32, _P, _P, _P, _P, _d, _d]),
    "shamx_neigh_sizes": (_i, [_P, C.POINTER(C.c_uint32)]),
    "shamx_neigh_get": (_i, [_P, _P] + [_P] * 7),
    "shamx_h_iterate": (_i, [_P, _i, _u32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _d, _d, _d, _i,
                             _dp]),
    "shamx_omega": (_i, [_P, _i, _u32, _P, _P, _P, _P, _P, _P, _P, _P, _d]),
    "shamx_eos": (_i, [_P, _i, _u32, _i, _d, _d, _d, _d, _d, _P, _P, _P, _P, _P, _P, _P, _d]),
    "shamx_force": (_i, [_P, _i, _u32] + [_P] * 24 + [_d] * 6 + [_i, _dp]),
    "shamx_cfl": (_i, [_P, _u32] + [_P] * 5 + [_d, _d, _P, _dp]),
    "shamx_predictor": (_i, [_P, _u32, _d] + [_P] * 11 + [_i, _dp]),
    "shamx_corrector": (_i, [_P, _u32, _d] + [_P] * 12 + [_d, _dp]),
    "shamx_solver_adjoint_step": (_i, [_P, C.POINTER(_P)]),
    "shamx_solver_adjoint_density": (_i, [_P, _P, _P, _P, _P]),
    "shamx_solver_create": (_i, [C.POINTER(ShamxConfig), _u32, C.POINTER(_P)]),
    "shamx_solver_destroy": (_i, [_P]),
    "shamx_solver_set_field": (_i, [_P, _i, _P, _i]),
    "shamx_solver_get_field": (_i, [_P, _i, _P, _i]),
    "shamx_solver_get_storage_order": (_i, [_P, _P]),
    "shamx_solver_products_save": (_i, [_P, C.c_uint32]),
    "shamx_solver_products_load": (_i, [_P, C.c_uint32]),
    "shamx_solver_products_free": (_i, [_P]),
    "shamx_solver_set_storage_order": (_i, [_P, _P]),
    "shamx_solver_get_time": (_i, [_P, _dp]),
    "shamx_solver_set_time": (_i, [_P, _dp]),
    "shamx_solver_add_sink": (_i, [_P, C.c_double, _dp, _dp, C.c_double]),
    "shamx_solver_get_sinks": (_i, [_P, C.POINTER(C.c_uint32), _P]),
    "shamx_solver_set_sinks": (_i, [_P, C.c_uint32, _P]),
    "shamx_solver_evolve_once": (_i, [_P, C.POINTER(ShamxStepLog)]),
    "shamx_solver_density_force": (_i, [_P, C.POINTER(ShamxStepLog)]),
    "shamx_solver_stream": (_i, [_P, C.POINTER(_P)]),
    "shamx_solver_sizes": (_i, [_P, C.POINTER(C.c_uint64)]),
    "shamx_solver_get_f64": (_i, [_P, _i, _P]),
    "shamx_solver_get_u32": (_i, [_P, _i, _P]),
    "shamx_solver_get_u8": (_i, [_P, _i, _P]),
}

_lib = None


def _init_torch_runtime_first():
    """PyTorch ships its own HIP runtime (libamdhip64.so) next to the system one this library
    links (libamdhip64.so.7).  Both can live in one process and share the device address
    space, but torch's must initialise first; initialise it before loading this library when
    torch is importable (a process without torch just uses the system runtime)."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def lib():
    """Load libshamx_sph.so; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        _init_torch_runtime_first()
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with shamrock_amd/build.py "
                              "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        msg = lib().shamx_last_error().decode()
        raise ERRORS.get(rc, RuntimeError)(f"shamx error {rc}: {msg}")


def ptr(a) -> int:
    """Device/host address of a torch tensor or numpy array (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def dptr(v: np.ndarray):
    return v.ctypes.data_as(_dp)
