"""shamrock_amd -- MI355X-native SPH timestep behind the reference's Model/Solver interface.

    import shamrock_amd as shamrock
    ctx = shamrock.Context()
    ctx.pdata_layout_new()
    model = shamrock.get_Model_SPH(context=ctx, vector_type="f64_3", sph_kernel="M6")

The compute path is libshamx_sph.so (hand-written HIP for gfx950, C ABI in
include/shamx_sph.h); see shamrock_amd/model.py for the Python surface.  `shamrock_amd.phys`
holds the analytic solutions (SodTube) and `shamrock_amd.math` the lattice helpers
(get_ideal_hcp_box) that the reference's scripts use.
"""
from . import math, phys
from .model import (Constants, Context, EvolveUntilResults, SPHModel, SolverConfig,
                    TimestepLog, UnitSystem, get_Model_SPH)

__all__ = ["Context", "get_Model_SPH", "SPHModel", "SolverConfig", "TimestepLog",
           "EvolveUntilResults", "UnitSystem", "Constants", "phys", "math"]
