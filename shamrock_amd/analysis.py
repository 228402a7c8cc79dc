"""Analyses of the reference's Python surface used by its CI runs.

`SodTube` restates shamphys::SodTube (shamphys/src/SodTube.cpp:22-133), including its Newton
solve for the post-shock pressure, which returns a float (shammath/solve.hpp:28-46: the
function's return type is float), and `sod_l2` restates AnalysisSodTube::compute_L2_dist
(AnalysisSodTube.cpp:27-120).
"""
from __future__ import annotations

import math

import numpy as np


def _f32(v: float) -> float:
    return float(np.float32(v))


class SodTube:
    def __init__(self, gamma, rho_1, P_1, rho_5, P_5):
        self.gamma, self.rho_1, self.P_1, self.rho_5, self.P_5 = gamma, rho_1, P_1, rho_5, P_5
        self.vx_1 = self.vx_5 = 0.0
        self.c_1 = self._cs(P_1, rho_1)
        self.c_5 = self._cs(P_5, rho_5)
        if P_5 > P_1:
            raise ValueError("not correct")
        self._P4 = self._solve_P_4()

    def _cs(self, P, rho):
        return math.sqrt(self.gamma * P / rho)

    def _solve_P_4(self):
        g, P_1, P_5, c_1, c_5 = self.gamma, self.P_1, self.P_5, self.c_1, self.c_5

        def f(P_4):
            z = (P_4 / P_5 - 1.)
            gm1, gp1, g2 = g - 1.0, g + 1.0, 2.0 * g
            fact1 = gm1 / g2 * (c_5 / c_1) * z / math.sqrt(1. + gp1 / g2 * z)
            fact = math.pow(1.0 - fact1, g2 / gm1)
            return P_1 * fact - P_4

        def df(x):  # derivative_upwind (shammath/derivatives.hpp:38-40)
            return (f(x + 1e-6) - f(x)) / 1e-6

        xk, eps = float(P_1), 100000.0
        while eps > 1e-6:
            xkp1 = xk - (f(xk) / df(xk))
            eps = abs(xk - xkp1)
            xk = xkp1
        return _f32(xk)

    def get_value(self, t, x):
        g = self.gamma
        P_4 = self._P4
        z = (P_4 / self.P_5 - 1.)
        gm1, gp1 = g - 1.0, g + 1.0
        gmfact1, gmfact2 = 0.5 * gm1 / g, 0.5 * gp1 / g
        fact = math.sqrt(1. + gmfact2 * z)
        vx_4 = self.c_5 * z / (g * fact)
        rho_4 = self.rho_5 * (1. + gmfact2 * z) / (1. + gmfact1 * z)
        w = self.c_5 * fact
        P_3, vx_3 = P_4, vx_4
        rho_3 = self.rho_1 * math.pow(P_3 / self.P_1, 1. / g)
        c3 = self._cs(P_3, rho_3)
        xi = 0.
        xsh = xi + w * t
        xcd = xi + vx_3 * t
        xft = xi + (vx_3 - c3) * t
        xhd = xi - self.c_1 * t
        if x < xhd:
            return self.rho_1, self.vx_1, self.P_1
        if x < xft:
            vx = 2. / gp1 * (self.c_1 + (x - xi) / t)
            locfact = 1. - 0.5 * gm1 * vx / self.c_1
            return (self.rho_1 * math.pow(locfact, 2. / gm1), vx,
                    self.P_1 * math.pow(locfact, 2. * g / gm1))
        if x < xcd:
            return rho_3, vx_3, P_3
        if x < xsh:
            return rho_4, vx_4, P_4
        return self.rho_5, self.vx_5, self.P_5


def sod_l2(sod: SodTube, fields: dict, gpart_mass: float, hfact: float, gamma: float, t: float,
           x_ref: float = 0.0, x_min: float = -0.5, x_max: float = 0.5):
    """Mean squared deviation from the analytic Sod solution of rho, v and P over the
    particles with x_min < x < x_max, direction (1, 0, 0): returns (rho, (vx, vy, vz), P)."""
    s_rho = s_P = 0.0
    s_v = [0.0, 0.0, 0.0]
    n = 0.0
    x_all, h_all, u_all = fields["x"], fields["h"], fields["u"]
    vx_all, vy_all, vz_all = fields["vx"], fields["vy"], fields["vz"]
    for i in range(x_all.size):
        h = float(h_all[i])
        rho = gpart_mass * (hfact / h) * (hfact / h) * (hfact / h)
        P = (gamma - 1) * rho * float(u_all[i])
        x = float(x_all[i]) - x_ref
        if (x + x_ref) > x_min and (x + x_ref) < x_max:
            r_rho, r_vx, r_P = sod.get_value(t, x)
            d_rho, d_P = rho - r_rho, P - r_P
            d_v = (float(vx_all[i]) - r_vx, float(vy_all[i]), float(vz_all[i]))
            s_rho += d_rho * d_rho
            s_P += d_P * d_P
            s_v = [s_v[k] + d_v[k] * d_v[k] for k in range(3)]
            n += 1
    if n == 0:
        raise RuntimeError("no particle in wanted region")
    return s_rho / n, tuple(v / n for v in s_v), s_P / n
