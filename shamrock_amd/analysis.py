"""Analyses of the SPH model's state (the objects `SPHModel.make_analysis_*` returns).

`AnalysisSodTube` compares the particles with the exact shock-tube solution
(shammodels/sph/src/modules/AnalysisSodTube.cpp:27-120, bound at pySPHModel.cpp:409-414,
1466-1485): over the particles whose coordinate along `direction` lies strictly between x_min
and x_max, the mean of the squared differences of rho = m (hfact / h)^3, of the velocity vector
against v_x(sol) * direction, and of the adiabatic pressure (gamma - 1) rho u.
"""
from __future__ import annotations

import numpy as np

from .phys import SodTube


class AnalysisSodTube:
    def __init__(self, model, sod: SodTube, direction, time_val, x_ref, x_min, x_max):
        self.model = model
        self.sod = sod
        self.direction = np.asarray(direction, float)
        self.time_val = float(time_val)
        self.x_ref, self.x_min, self.x_max = float(x_ref), float(x_min), float(x_max)

    def compute_L2_dist(self):
        """(rho, (vx, vy, vz), P) mean squared deviations, as the reference returns them."""
        cfg = self.model.get_current_config().d
        if cfg.get("eos_kind", 0) != 0:
            raise ValueError("The sod analysis is only available for adiabatic EOS")
        data = self.model.collect_data()
        return sod_l2_dist(self.sod, data["xyz"], data["vxyz"], data["hpart"], data["uint"],
                           cfg["gpart_mass"], self.model.get_hfact(), cfg["gamma"],
                           self.time_val, self.direction, self.x_ref, self.x_min, self.x_max)


def sod_l2_dist(sod: SodTube, xyz, vxyz, hpart, uint, pmass, hfact, gamma, t,
                direction=(1.0, 0.0, 0.0), x_ref=0.0, x_min=-0.5, x_max=0.5):
    """The L2 deviations of AnalysisSodTube on plain arrays: xyz, vxyz (N, 3), hpart, uint (N)."""
    direction = np.asarray(direction, float)
    xyz, vxyz = np.asarray(xyz, float), np.asarray(vxyz, float)
    hpart, uint = np.asarray(hpart, float), np.asarray(uint, float)
    q = hfact / hpart
    rho = pmass * q * q * q
    P = (gamma - 1) * rho * uint
    coord = xyz @ direction
    sel = (coord > x_min) & (coord < x_max)
    count = int(np.count_nonzero(sel))
    if count == 0:
        raise RuntimeError("no particle in wanted region")
    rho_s, v_s, P_s = sod.values(t, coord[sel] - x_ref)
    d_rho = rho[sel] - rho_s
    d_P = P[sel] - P_s
    d_v = vxyz[sel] - v_s[:, None] * direction[None, :]
    return (float((d_rho * d_rho).sum() / count),
            tuple(float(c) for c in (d_v * d_v).sum(0) / count),
            float((d_P * d_P).sum() / count))
