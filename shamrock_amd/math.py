"""Lattice helpers of the reference's `shamrock.math` module used by the SPH setups."""
from __future__ import annotations

import math as _m


def _hcp_index_bounds(dr, box_min, box_max):
    """LatticeHCP::get_box_index_bounds (crystalLattice.hpp:139-162): integer lattice
    coordinates covering the box, widened by one on each side."""
    scale = (2.0, _m.sqrt(3.0), 2 * _m.sqrt(6.0) / 3)
    lo = [int((box_min[k] / scale[k]) / dr) - 1 for k in range(3)]
    hi = [int((box_max[k] / scale[k]) / dr) + 1 for k in range(3)]
    return lo, hi


def get_ideal_hcp_box(dr, box_min, box_max):
    """shamrock.math.get_ideal_hcp_box (pyshammath.cpp:905-907, crystalLattice.hpp:387-401):
    the periodic box of the HCP lattice of spacing dr nearest to (box_min, box_max).  Index
    bounds around the box, then the x extent raised to at least 2 cells and the y and z bounds
    made of even sum (a periodic HCP stacking needs an even number of rows and layers)."""
    lo, hi = _hcp_index_bounds(dr, box_min, box_max)
    if hi[0] - lo[0] < 2:
        hi[0] += 1
    if (hi[1] + lo[1]) % 2 != 0:
        hi[1] += 1
    if (hi[2] + lo[2]) % 2 != 0:
        hi[2] += 1
    s3, s6 = _m.sqrt(3.0), _m.sqrt(6.0)
    bmin = (2.0 * lo[0] * dr, s3 * lo[1] * dr, 2 * s6 * lo[2] / 3 * dr)
    bmax = (2.0 * hi[0] * dr, s3 * hi[1] * dr, 2 * s6 * hi[2] / 3 * dr)
    return bmin, bmax
