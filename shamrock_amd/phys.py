"""Analytic solutions used by the reference's CI runs (the `shamrock.phys` module).

`SodTube` is the exact Riemann solution of a shock tube at rest (pyShamphys.cpp:352-367,
shamphys/src/SodTube.cpp:22-133): left state 1, right state 5, rarefaction head / foot, contact
and shock.  The post-shock pressure is found by Newton's method on the pressure-matching
function with an upwind difference quotient of step 1e-6 and stops when the step falls below
1e-6; the reference's Newton helper returns a single-precision float
(shammath/solve.hpp:28-46), which moves P_4 by ~1e-8 relatively.  The CI values the reference
checks to 1e-11 depend on that rounding, so it is reproduced here.
"""
from __future__ import annotations

import math

import numpy as np


class SodTube:
    """Exact solution of Sod's shock tube; `get_value(t, x)` -> (rho, v_x, P)."""

    def __init__(self, *, gamma, rho_1, P_1, rho_5, P_5):
        if P_5 > P_1:
            raise ValueError("not correct")  # SodTube.cpp:27-29 rejects P_5 > P_1
        self.gamma = float(gamma)
        self.rho_1, self.P_1 = float(rho_1), float(P_1)
        self.rho_5, self.P_5 = float(rho_5), float(P_5)
        self.c_1 = self._sound(self.P_1, self.rho_1)
        self.c_5 = self._sound(self.P_5, self.rho_5)
        self.p_star = self._post_shock_pressure()
        self._waves()

    def _sound(self, P, rho):
        return math.sqrt(self.gamma * P / rho)

    def _pressure_mismatch(self, p4):
        """Pressure behind the rarefaction minus p4, for a shock of strength p4 / P_5."""
        g = self.gamma
        s = p4 / self.P_5 - 1.0
        a = (g - 1.0) / (2.0 * g) * (self.c_5 / self.c_1) * s
        a = a / math.sqrt(1.0 + (g + 1.0) / (2.0 * g) * s)
        return self.P_1 * math.pow(1.0 - a, 2.0 * g / (g - 1.0)) - p4

    def _post_shock_pressure(self):
        step, p = 1e-6, self.P_1
        while True:
            slope = (self._pressure_mismatch(p + step) - self._pressure_mismatch(p)) / step
            p_next = p - self._pressure_mismatch(p) / slope
            done = abs(p - p_next) <= 1e-6
            p = p_next
            if done:
                return float(np.float32(p))

    def _waves(self):
        g = self.gamma
        s = self.p_star / self.P_5 - 1.0
        root = math.sqrt(1.0 + 0.5 * (g + 1.0) / g * s)
        self.v_star = self.c_5 * s / (g * root)
        self.rho_shocked = (self.rho_5 * (1.0 + 0.5 * (g + 1.0) / g * s)
                            / (1.0 + 0.5 * (g - 1.0) / g * s))
        self.shock_speed = self.c_5 * root
        self.rho_expanded = self.rho_1 * math.pow(self.p_star / self.P_1, 1.0 / g)
        self.c_expanded = self._sound(self.p_star, self.rho_expanded)

    def get_value(self, t, x):
        """(rho, v_x, P) at time t > 0 and position x (the discontinuity starts at x = 0)."""
        rho, v, P = self.values(t, np.asarray([x], float))
        return float(rho[0]), float(v[0]), float(P[0])

    def values(self, t, x):
        """Vectorised get_value over an array of positions."""
        g = self.gamma
        x = np.asarray(x, float)
        head = -self.c_1 * t
        foot = (self.v_star - self.c_expanded) * t
        contact = self.v_star * t
        shock = self.shock_speed * t
        rho = np.full(x.shape, self.rho_5)
        v = np.zeros(x.shape)
        P = np.full(x.shape, self.P_5)
        zone4 = x < shock
        rho[zone4], v[zone4], P[zone4] = self.rho_shocked, self.v_star, self.p_star
        zone3 = x < contact
        rho[zone3], v[zone3], P[zone3] = self.rho_expanded, self.v_star, self.p_star
        fan = x < foot
        if np.any(fan):
            vf = 2.0 / (g + 1.0) * (self.c_1 + x[fan] / t)
            base = 1.0 - 0.5 * (g - 1.0) * vf / self.c_1
            rho[fan] = self.rho_1 * np.power(base, 2.0 / (g - 1.0))
            P[fan] = self.P_1 * np.power(base, 2.0 * g / (g - 1.0))
            v[fan] = vf
        left = x < head
        rho[left], v[left], P[left] = self.rho_1, 0.0, self.P_1
        return rho, v, P
