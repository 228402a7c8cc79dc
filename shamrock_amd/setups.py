"""Particle setups of the benchmark workloads (host side, numpy, seeded).

Positions follow the reference's cubic lattice generator (GeneratorLatticeCubic.hpp:41-109,
crystalLattice.hpp:426-460: r = dr * (i, j, k), kept when inside the half-open box, h = dr);
the energy injection follows Model::add_kernel_value (Model.hpp:739-768).  The reference
iterates lattice points in a remapped order (IteratorDiscontinuous); here points are emitted
with x fastest, which only permutes particle ids.
"""
from __future__ import annotations

import math

import numpy as np

KERNEL_RKERN = {"M4": 2.0, "M6": 3.0, "M8": 4.0, "C2": 2.0, "C4": 2.0, "C6": 2.0}
KERNEL_HFACT = {"M4": 1.2, "M6": 1.0, "M8": 1.0, "C2": 1.4, "C4": 1.6, "C6": 2.2}


def lattice_cubic(dr: float, box_min, box_max) -> np.ndarray:
    """Cubic lattice dr*(i,j,k) inside [box_min, box_max) (crystalLattice.hpp:410-460)."""
    bmin = np.asarray(box_min, float)
    bmax = np.asarray(box_max, float)
    lo = [int(v / dr) - 1 for v in bmin]
    hi = [int(v / dr) + 1 for v in bmax]
    axes = [np.arange(lo[k], hi[k], dtype=np.int64) for k in range(3)]
    k, j, i = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    p = dr * np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    keep = np.all((p >= bmin) & (p < bmax), axis=1)
    return np.ascontiguousarray(p[keep])


def kernel_W(name: str, r: np.ndarray, h: float) -> np.ndarray:
    """W_3d of the named kernel (sphkernels.hpp), numpy, for setup-time smoothing."""
    q = r / h
    pi = 3.141592653589793116
    if name == "M4":
        f = np.where(q < 1, 0.25 * (2 - q) ** 3 - (1 - q) ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0))
        norm = 1 / pi
    elif name == "M6":
        t1, t2, t3 = (3 - q) ** 5, -6 * (2 - q) ** 5, 15 * (1 - q) ** 5
        f = np.where(q < 1, t1 + t2 + t3, np.where(q < 2, t1 + t2, np.where(q < 3, t1, 0)))
        norm = 1 / (120 * pi)
    else:
        raise ValueError(f"setup smoothing not available for kernel {name}")
    return norm * f / h**3


def base_config(kernel: str, box_min, box_max, periodic=True, **kw) -> dict:
    cfg = dict(kernel=kernel, periodic=int(periodic), eos_kind=0, tree_reduction_level=3,
               h_iter_per_subcycles=50, h_max_subcycles=100, box_min=list(map(float, box_min)),
               box_max=list(map(float, box_max)), gpart_mass=0.0, gamma=5.0 / 3.0, eos_cs=1.0,
               eos_cs0=1.0, eos_q=0.5, eos_r0=1.0, alpha_u=1.0, alpha_AV=1.0, beta_AV=2.0,
               htol_up_coarse=1.1, htol_up_fine=1.1, epsilon_h=1e-6, cfl_cour=0.3,
               cfl_force=0.3, cfl_multiplier_stiffness=2.0, mixed_precision=0)
    cfg.update(kw)
    return cfg


def fields_from(pos, h, u, v=None, m=None) -> dict:
    n = pos.shape[0]
    v = np.zeros((n, 3)) if v is None else v
    f = dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), vx=v[:, 0].copy(),
             vy=v[:, 1].copy(), vz=v[:, 2].copy(), ax=np.zeros(n), ay=np.zeros(n), az=np.zeros(n),
             h=np.full(n, h) if np.isscalar(h) else np.asarray(h, float).copy(),
             u=np.full(n, u) if np.isscalar(u) else np.asarray(u, float).copy(), du=np.zeros(n))
    return f


def sound_wave(n_side: int = 32, kernel: str = "M4", amp: float = 1e-4) -> tuple[dict, dict]:
    """Plain FP64 sound wave: n^3 cubic lattice in the periodic unit box, rho = 1, u = 1,
    gamma = 5/3, v_x = amp sin(2 pi x), constant AV (1, 1, 2), CFL 0.3."""
    dr = 1.0 / n_side
    pos = lattice_cubic(dr, (0, 0, 0), (1, 1, 1))
    n = pos.shape[0]
    v = np.zeros((n, 3))
    v[:, 0] = amp * np.sin(2 * math.pi * pos[:, 0])
    cfg = base_config(kernel, (0, 0, 0), (1, 1, 1), gpart_mass=1.0 / n)
    return cfg, fields_from(pos, dr, 1.0, v)


def sedov(n_side: int = 128, kernel: str = "M6", E: float = 1.0, u_bg: float = 1e-5,
          mixed: bool = False) -> tuple[dict, dict]:
    """Sedov-Taylor blast: n^3 lattice in the periodic box [-0.5, 0.5)^3, rho = 1, background
    u = 1e-5, energy E deposited as u_i += E W(r_i, h_inj) / (m sum W) with the kernel support
    R h_inj = 2 dr around the origin (Model.hpp:739-768 idiom), gamma = 5/3."""
    dr = 1.0 / n_side
    pos = lattice_cubic(dr, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    n = pos.shape[0]
    m = 1.0 / n
    h_inj = 2 * dr / KERNEL_RKERN[kernel]
    r = np.sqrt((pos**2).sum(axis=1))
    W = kernel_W(kernel, r, h_inj)
    u = np.full(n, u_bg)
    u += E * W / (m * W.sum())
    cfg = base_config(kernel, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), gpart_mass=m,
                      mixed_precision=int(mixed))
    return cfg, fields_from(pos, dr, u)


def random_box(n: int, seed: int = 0, kernel: str = "M4", periodic=True, jitter=0.3):
    """Jittered lattice in the periodic unit box with random velocities: a generic parity case."""
    side = round(n ** (1 / 3))
    dr = 1.0 / side
    pos = lattice_cubic(dr, (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter, jitter, pos.shape) * dr
    pos = np.mod(pos, 1.0)
    v = rng.normal(0, 0.05, pos.shape)
    u = 1.0 + 0.2 * rng.random(pos.shape[0])
    cfg = base_config(kernel, (0, 0, 0), (1, 1, 1), periodic=periodic,
                      gpart_mass=1.0 / pos.shape[0])
    return cfg, fields_from(pos, dr, u, v)


WORKLOADS = {
    "sound_wave_32": lambda: sound_wave(32),
    "sedov_128": lambda: sedov(128),
    "sedov_128_mixed": lambda: sedov(128, mixed=True),
}
