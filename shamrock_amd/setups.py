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
               cfl_force=0.3, cfl_multiplier_stiffness=2.0, mixed_precision=0, av_kind=0,
               alpha_min=0.0, alpha_max=1.0, sigma_decay=0.1, constant_G=6.6743015e-11,
               sink_softening=0.0)
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


def plummer_positions(n: int, seed: int = 0, a: float = 1.0, M: float = 1.0,
                      r_max: float = 30.0, kernel: str = "M4"):
    """Positions (n, 3) and smoothing lengths of the Plummer sphere of `plummer`."""
    rng = np.random.default_rng(seed)
    f_max = r_max**3 / (r_max**2 + a * a) ** 1.5
    X = rng.uniform(0, f_max, n)
    r = a / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = np.ascontiguousarray(r[:, None] * d)
    rho = 3 * M / (4 * math.pi * a**3) * (1 + r * r / (a * a)) ** -2.5
    h = KERNEL_HFACT[kernel] * ((M / n) / rho) ** (1.0 / 3.0)
    return pos, h


def plummer(n: int, seed: int = 0, a: float = 1.0, M: float = 1.0, r_max: float = 30.0,
            u: float = 0.05, kernel: str = "M4", pos_h=None) -> tuple[dict, dict]:
    """Plummer sphere (scale radius a, mass M) truncated at r_max, free boundaries, zero
    velocities, isothermal gas at u = 0.05 (EOS isothermal, c_s^2 = (gamma - 1) u with
    gamma = 5/3).  Radii by the inverse of the enclosed-mass fraction r^3 / (r^2 + a^2)^(3/2),
    drawn inside the truncation; isotropic directions; h from the analytic density,
    h = hfact (m / rho)^(1/3).  pos_h: positions and h of plummer_positions already drawn."""
    pos, h = plummer_positions(n, seed, a, M, r_max, kernel) if pos_h is None else pos_h
    m = M / n
    lo, hi = pos.min(0) - 1, pos.max(0) + 1
    cfg = base_config(kernel, lo, hi, periodic=False, gpart_mass=m, eos_kind=1,
                      eos_cs=math.sqrt((5.0 / 3.0 - 1.0) * u))
    return cfg, fields_from(pos, h, u)


def knn_h(pos, kernel: str) -> np.ndarray:
    """Half the distance to the k-th nearest particle, k = (4 pi / 3) (hfact R)^3 (the
    kernel's neighbour count), scaled by R / 2: the support radius R h then holds k particles."""
    from scipy.spatial import cKDTree
    R = KERNEL_RKERN[kernel]
    k = int(round(4.0 / 3.0 * math.pi * (KERNEL_HFACT[kernel] * R) ** 3))
    k = max(1, min(k, pos.shape[0] - 1))
    d, _ = cKDTree(pos).query(pos, k=k + 1, workers=-1)
    return d[:, -1] / R


def disc(n: int, seed: int = 0, r_in: float = 1.0, r_out: float = 10.0, m_disc: float = 0.01,
         h_over_r: float = 0.05, q: float = 0.25, alpha_AV: float = 0.1, kernel: str = "M4",
         softening: float = 0.1, mixed: bool = False) -> tuple[dict, dict]:
    """Gas disc around a central sink (G = M = 1): surface density proportional to 1/r on
    [r_in, r_out] (radii uniform in r), locally isothermal LP07 sound speed
    c_s = h_over_r (r / r_in)^(-q) (spherical r, ComputeEos.cpp:726-798), Gaussian vertical
    profile of scale height H = c_s / Omega_K, Keplerian azimuthal velocity about the sink,
    constant viscosity alpha_AV (beta 2, alpha_u 1), free boundaries.  h starts at 0.6 times the
    k-nearest-neighbour estimate: in a disc thinner than the kernel the density sum falls more
    slowly than rho_h(h) beyond the root, so the reference's Newton iteration
    (density.hpp:22-41) only converges from below.  The sink (mass 1 at rest at the origin, accretion
    radius 0.1) is in cfg["sinks"]; its force on the gas is softened by `softening` (a choice of
    this library: the reference's sink force has no softening)."""
    rng = np.random.default_rng(seed)
    G = M = 1.0
    r = rng.uniform(r_in, r_out, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    cs = h_over_r * (r / r_in) ** (-q)
    omega_k = np.sqrt(G * M / r**3)
    H = cs / omega_k
    zz = rng.normal(0, 1, n) * H
    pos = np.ascontiguousarray(np.stack([r * np.cos(phi), r * np.sin(phi), zz], 1))
    m = m_disc / n
    h = 0.6 * knn_h(pos, kernel)
    vk = np.sqrt(G * M / r)
    v = np.stack([-vk * np.sin(phi), vk * np.cos(phi), np.zeros(n)], 1)
    gamma = 5.0 / 3.0
    u = cs**2 / (gamma * (gamma - 1))
    lo, hi = pos.min(0) - 1, pos.max(0) + 1
    cfg = base_config(kernel, lo, hi, periodic=False, gpart_mass=m, gamma=gamma, eos_kind=2,
                      eos_cs0=h_over_r, eos_q=q, eos_r0=r_in, alpha_AV=alpha_AV, beta_AV=2.0,
                      alpha_u=1.0, constant_G=G, sink_softening=softening,
                      mixed_precision=int(mixed))
    cfg["sinks"] = [dict(mass=M, pos=(0.0, 0.0, 0.0), velocity=(0.0, 0.0, 0.0),
                         accretion_radius=0.1)]
    return cfg, fields_from(pos, h, u, v)


def kelvin_helmholtz(n_side: int = 64, kernel: str = "M4", phase: float = 0.0,
                     seed: int = 0) -> tuple[dict, dict]:
    """Kelvin-Helmholtz shear layers in the periodic box [-0.5, 0.5)^3: a dense band |y| < 0.25
    (density 2, lattice spacing dr) between layers of density 1 (spacing dr 2^(1/3)), equal
    particle masses, pressure 2.5 everywhere, v_x = -0.5 in the band and +0.5 outside, and the
    seeded perturbation v_y = 0.01 sin(4 pi (x + phase)) (+ 1e-4 white noise)."""
    rng = np.random.default_rng(seed)
    n_d = int(round(n_side * (4.0 / 3.0) ** (1.0 / 3.0)))   # band: twice the number density
    dr_d = 1.0 / n_d
    n_o = int(round(n_d / 2 ** (1.0 / 3.0)))
    dr_o = 1.0 / n_o
    band = lattice_cubic(dr_d, (-0.5, -0.25, -0.5), (0.5, 0.25, 0.5))
    outer = lattice_cubic(dr_o, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    outer = outer[np.abs(outer[:, 1]) >= 0.25]
    pos = np.ascontiguousarray(np.concatenate([band, outer]))
    n = pos.shape[0]
    m = (2.0 * 0.5 + 1.0 * 0.5) / n
    dense = np.abs(pos[:, 1]) < 0.25
    rho = np.where(dense, 2.0, 1.0)
    gamma = 5.0 / 3.0
    u = 2.5 / ((gamma - 1) * rho)
    v = np.zeros((n, 3))
    v[:, 0] = np.where(dense, -0.5, 0.5)
    v[:, 1] = 0.01 * np.sin(4 * math.pi * (pos[:, 0] + phase)) + 1e-4 * rng.normal(size=n)
    h = np.where(dense, dr_d, dr_o) * KERNEL_HFACT[kernel]
    cfg = base_config(kernel, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), gpart_mass=m)
    return cfg, fields_from(pos, h, u, v)


def hcp_box_dim(dr: float, xcnt: int, ycnt: int, zcnt: int) -> tuple[float, float, float]:
    """Model.get_box_dim_fcc_3d (generators.hpp:31-51): extent of xcnt x ycnt x zcnt HCP cells."""
    i, j, k = xcnt, ycnt, zcnt
    r = (float(2 * i + ((j + k) % 2)), math.sqrt(3.0) * (j + (1.0 / 3.0) * (k % 2)),
         2 * math.sqrt(6.0) * k / 3)
    return r[0] * dr, r[1] * dr, r[2] * dr


def fcc_cube(dr: float, box_min, box_max) -> np.ndarray:
    """Model::add_cube_fcc_3d (Model.cpp:1093-1190, generators.hpp:82-130): points
    box_min + dr * (2 i + (j + k) % 2, sqrt(3) (j + (k % 2) / 3), 2 sqrt(6) k / 3) for
    0 <= i, j, k < ceil(extent / (dr * (2, sqrt(3), 2 sqrt(6) / 3))), kept inside the
    half-open box."""
    bmin = np.asarray(box_min, float)
    bmax = np.asarray(box_max, float)
    s3, s6 = math.sqrt(3.0), math.sqrt(6.0)
    cells = ((bmax - bmin) / np.array([2.0, s3, 2 * s6 / 3])) / dr
    ni, nj, nk = (int(math.ceil(c)) for c in cells)
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    x = (2 * i + (j + k) % 2).astype(np.float64)
    y = s3 * (j + (1.0 / 3.0) * (k % 2))
    z = 2 * s6 * k.astype(np.float64) / 3
    p = np.stack([x * dr, y * dr, z * dr], axis=1) + bmin
    keep = np.all((p >= bmin) & (p < bmax), axis=1)
    return np.ascontiguousarray(p[keep])


def hcp_lattice(dr: float, box_min, box_max) -> np.ndarray:
    """HCP lattice points inside the half-open box (crystalLattice.hpp LatticeHCP: generator,
    get_box_index_bounds, index ranges [min, max); GeneratorLatticeHCP.hpp:83-106 keeps the
    points with Patch::is_in_patch_converted).  Emitted with i fastest (the reference's
    discontinuous iterator only permutes the same points)."""
    bmin = np.asarray(box_min, float)
    bmax = np.asarray(box_max, float)
    s3, s6 = math.sqrt(3.0), math.sqrt(6.0)
    cmin = [bmin[0] / 2.0, bmin[1] / s3, bmin[2] / (2 * s6 / 3)]
    cmax = [bmax[0] / 2.0, bmax[1] / s3, bmax[2] / (2 * s6 / 3)]
    lo = [int(c / dr) - 1 for c in cmin]
    hi = [int(c / dr) + 1 for c in cmax]
    k, j, i = np.meshgrid(np.arange(lo[2], hi[2]), np.arange(lo[1], hi[1]),
                          np.arange(lo[0], hi[0]), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    x = (2 * i + (np.abs(j + k) % 2)).astype(np.float64)
    y = s3 * (j + (1.0 / 3.0) * (np.abs(k) % 2))
    z = 2 * s6 * k.astype(np.float64) / 3
    p = np.stack([dr * x, dr * y, dr * z], axis=1)
    keep = np.all((p >= bmin) & (p < bmax), axis=1)
    return np.ascontiguousarray(p[keep])


def sod_tube(resol: int = 128, kernel: str = "M6") -> tuple[dict, dict]:
    """The reference's Sod-tube CI run (examples/tests_ci/sod_tube_sph.py): HCP lattices of
    spacing dr (left, rho = 1, P = 1) and 2 dr (right, rho = 0.125, P = 0.1) in the periodic box
    [-xs, xs] x [-ys/2, ys/2] x [-zs/2, zs/2], gamma = 1.4, Cullen & Dehnen viscosity
    (alpha in [0, 1], sigma 0.1, alpha_u 1, beta 2), CFL 0.1, particle mass = total mass / N."""
    gamma = 1.4
    rho_g, rho_d = 1, 0.125
    fact = (rho_g / rho_d) ** (1.0 / 3.0)
    P_g, P_d = 1, 0.1
    u_g = P_g / ((gamma - 1) * rho_g)
    u_d = P_d / ((gamma - 1) * rho_d)
    xs, ys, zs = hcp_box_dim(1, resol, 24, 24)
    dr = 1 / xs
    xs, ys, zs = hcp_box_dim(dr, resol, 24, 24)
    bmin, bmax = (-xs, -ys / 2, -zs / 2), (xs, ys / 2, zs / 2)
    p1 = hcp_lattice(dr, (-xs, -ys / 2, -zs / 2), (0, ys / 2, zs / 2))
    p2 = hcp_lattice(dr * fact, (0, -ys / 2, -zs / 2), (xs, ys / 2, zs / 2))
    pos = np.ascontiguousarray(np.concatenate([p1, p2]))
    h = np.concatenate([np.full(len(p1), dr), np.full(len(p2), dr * fact)])
    u = np.zeros(pos.shape[0])
    left = np.all((pos >= np.array([-xs, -ys / 2, -zs / 2])) &
                  (pos < np.array([0, ys / 2, zs / 2])), axis=1)
    right = np.all((pos >= np.array([0, -ys / 2, -zs / 2])) &
                   (pos < np.array([xs, ys / 2, zs / 2])), axis=1)
    u[left] = u_g
    u[right] = u_d
    vol_b = xs * ys * zs
    totmass = (rho_d * vol_b) + (rho_g * vol_b)
    pmass = totmass / pos.shape[0]
    cfg = base_config(kernel, bmin, bmax, gpart_mass=pmass, gamma=gamma, cfl_cour=0.1,
                      cfl_force=0.1, av_kind=1, alpha_min=0.0, alpha_max=1.0, sigma_decay=0.1,
                      alpha_u=1.0, beta_AV=2.0)
    return cfg, fields_from(pos, h, u)


WORKLOADS = {
    "sound_wave_32": lambda: sound_wave(32),
    "sedov_128": lambda: sedov(128),
    "sedov_128_mixed": lambda: sedov(128, mixed=True),
}
