"""The oracle's Cullen & Dehnen (2010) viscosity pieces against what the mathematics fixes.

* d(div v)/dt uses matrix-corrected gradients (DiffOperatorDtDivv.cpp:89-175): for linear
  fields v = A x, a = B x the corrected gradients are exact on any particle set, so
  d(div v)/dt = tr(B) - tr(A A) to rounding.  A transposed product or a wrong 3x3 inverse fails.
* div v and curl v are the plain SPH estimates (DiffOperator.cpp:24-268): on the interior of a
  cubic lattice they approach tr(A) and the curl of the field to the lattice's normalisation
  error (about 2 %).  The reference's curl carries the opposite sign of the mathematical curl
  (-(1/(rho Omega)) sum m v_ab x grad W); only |curl v|^2 enters the update, so the oracle keeps
  the reference's sign and the test pins it.
* alpha update (UpdateViscosity.cpp:120-219): limits of the implicit Euler relaxation
  d alpha/dt = -(alpha - alpha_loc) / tau with tau = h / (sigma c_s).
"""
import numpy as np
import pytest

import oracle_lib as orc


def _lattice(n1):
    dx = 1.0 / n1
    g = np.arange(n1) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1), dx


A = np.array([[0.3, -0.7, 0.2], [0.5, 0.1, -0.4], [0.25, 0.6, 0.45]])
B = np.array([[0.1, 0.2, -0.3], [0.4, -0.5, 0.6], [-0.7, 0.8, 0.9]])


@pytest.mark.parametrize("kernel,hfact", [("M4", 1.2), ("M6", 1.0)])
def test_dtdivv_exact_for_linear_fields(kernel, hfact):
    rng = np.random.default_rng(7)
    pos, dx = _lattice(10)
    pos = pos + rng.uniform(-0.3, 0.3, pos.shape) * dx
    h = hfact * dx * rng.uniform(0.9, 1.1, len(pos))
    R = {"M4": 2.0, "M6": 3.0}[kernel]
    c = orc.neigh_cache(pos, h, len(pos), htol=1.0, Rkern=R)
    off = np.array([0.2, -0.1, 0.3])
    vel = pos @ A.T + off
    acc = pos @ B.T - off
    o = orc.cd10_operators(kernel, c, pos, vel, acc, h, np.ones(len(pos)), dx ** 3)
    expect = np.trace(B) - np.trace(A @ A)
    np.testing.assert_allclose(o["dtdivv"], expect, rtol=1e-11, atol=1e-12)


def test_divv_curl_on_lattice_interior():
    pos, dx = _lattice(14)
    h = np.full(len(pos), 1.2 * dx)
    c = orc.neigh_cache(pos, h, len(pos), htol=1.0, Rkern=2.0)
    inner = np.all((pos > 3 * dx) & (pos < 1 - 3 * dx), 1)
    om = np.ones(len(pos))
    zero = np.zeros_like(pos)
    o = orc.cd10_operators("M4", c, pos, pos @ A.T, zero, h, om, dx ** 3)
    np.testing.assert_allclose(o["divv"][inner], np.trace(A), rtol=0.03)
    W = np.array([0.3, -0.2, 0.5])
    o = orc.cd10_operators("M4", c, pos, np.cross(W, pos), zero, h, om, dx ** 3)
    curl = np.stack([o["curlx"], o["curly"], o["curlz"]], 1)[inner]
    np.testing.assert_allclose(curl, np.broadcast_to(-2 * W, curl.shape), rtol=0.03)
    assert np.abs(o["divv"][inner]).max() < 1e-12
    o = orc.cd10_operators("M4", c, pos, 0.7 * pos, zero, h, om, dx ** 3)
    np.testing.assert_allclose(o["divv"][inner], 2.1, rtol=0.03)
    assert np.abs(np.stack([o["curlx"], o["curly"], o["curlz"]])[:, inner]).max() < 1e-12
    # Omega divides both estimates
    o2 = orc.cd10_operators("M4", c, pos, 0.7 * pos, zero, h, 2 * om, dx ** 3)
    np.testing.assert_allclose(o2["divv"], o["divv"] / 2, rtol=1e-14)


def _alpha(dt, cs, h, alpha, divv, curl, dtdivv, amin=0.0, amax=1.0, sigma=0.1):
    n = 1
    f = lambda v: np.full(n, float(v))  # noqa: E731
    return orc.cd10_alpha(dt, f(cs), f(h), f(alpha), f(divv), [f(curl), f(0), f(0)], f(dtdivv),
                          amin, amax, sigma)[0]


def test_alpha_update_limits():
    # no sound speed yet (first step): alpha_loc = alpha_min and no relaxation
    assert _alpha(0.1, 0.0, 0.05, 0.3, -1.0, 0.0, -50.0) == 0.3
    # strong converging flow with no shear: the local target saturates at alpha_max and alpha
    # jumps to it (the update takes alpha_loc whenever it exceeds alpha)
    assert _alpha(1e-3, 1.0, 0.05, 0.1, -1.0, 0.0, -1e6) == 1.0
    # pure shear (div v = 0): Balsara factor 0, target 0, implicit Euler decay
    dt, cs, h, a0, sigma = 2e-3, 1.5, 0.05, 0.8, 0.1
    fact = dt * cs * sigma / h
    np.testing.assert_allclose(_alpha(dt, cs, h, a0, 0.0, 3.0, -100.0), a0 / (1 + fact),
                               rtol=1e-15)
    # dt = 0 keeps alpha when the target is below it
    assert _alpha(0.0, cs, h, a0, 0.0, 3.0, -100.0) == a0
    # a long step relaxes to the target: with d(div v)/dt >= 0 the source is off, target 0
    assert _alpha(1e6, cs, h, a0, 1.0, 0.0, 1.0) < 1e-6
    # with no shear and no convergence the Balsara factor defaults to 1 (fac + |curl|^2 = 0)
    np.testing.assert_allclose(_alpha(1e9, cs, h, 0.0, 1.0, 0.0, -1.0), 10 * h * h / cs ** 2,
                               rtol=1e-6)
    # the target grows as h^2 |d(div v)/dt| / c_s^2 (Cullen & Dehnen eq. 13), below alpha_max
    t1 = _alpha(1e9, 1.0, 0.05, 0.0, -1.0, 0.0, -1.0)
    t2 = _alpha(1e9, 1.0, 0.05, 0.0, -1.0, 0.0, -2.0)
    t3 = _alpha(1e9, 2.0, 0.05, 0.0, -1.0, 0.0, -2.0)
    np.testing.assert_allclose([t2 / t1, t2 / t3], [2.0, 4.0], rtol=1e-12)
    np.testing.assert_allclose(t1, 10 * 0.05 ** 2 * 1.0, rtol=1e-12)
