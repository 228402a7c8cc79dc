"""GPU tests of the hand-written adjoint of the SPH step (shamrock_amd/csrc/adjoint.hip).

The gradient of L = sum_a c_a rho_a(final) w.r.t. the initial velocities and internal energies,
through several full steps, is compared with central finite differences of the FP64 forward
(same time steps, h converged to 1e-12 so that the forward map is smooth to that level).

The conductivity signal speed sqrt(|P_a - P_b| / rho_avg) has an unbounded derivative where
P_a = P_b, which a uniform-pressure start (the Kelvin-Helmholtz setup) puts on every pair of a
lattice; the artificial viscosity switch max(-v_ab.r_ab, 0) and |v_ab.r_ab| have kinks at
v_ab.r_ab = 0, where every pair of a band of the Kelvin-Helmholtz shear flow starts.  The
finite-difference cases therefore start from internal energies perturbed by 5 % and velocities
perturbed by 1e-3, where the step is differentiable with bounded derivatives and the velocity
steps of the differences (1e-6) do not cross a kink.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from shamrock_amd import setups  # noqa: E402
from shamrock_amd.adjoint import rollout_density, rollout_dts, rollout_passes  # noqa: E402

DEV = "cuda"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _case(name):
    if name == "random_box":
        cfg, f = setups.random_box(1000, seed=4, jitter=0.3)
    else:
        cfg, f = setups.kelvin_helmholtz(10)
    cfg = dict(cfg, epsilon_h=1e-12, cfl_cour=0.1, cfl_force=0.1)
    rng = np.random.default_rng(11)
    v = np.stack([f["vx"], f["vy"], f["vz"]], 1)
    v0 = torch.tensor(v + 1e-3 * rng.standard_normal(v.shape), device=DEV)
    u0 = torch.tensor(f["u"] * (1 + 0.05 * rng.standard_normal(f["u"].size)), device=DEV)
    return cfg, f, v0, u0


def _loss(cfg, f, v0, u0, c, steps, dts, k=2):
    rho = rollout_density(cfg, f, v0, u0, steps, checkpoint_every=k, dts=dts)
    return (rho * c).sum()


@pytest.mark.parametrize("name", ["random_box", "kelvin_helmholtz"])
def test_adjoint_matches_central_differences(name):
    say(f"[adjoint fd] {name}")
    cfg, f, v0, u0 = _case(name)
    steps = 4
    dts = rollout_dts(cfg, f, v0, u0, steps)
    assert dts[0] == 0 and all(d > 0 for d in dts[1:])
    rng = np.random.default_rng(0)
    c = torch.tensor(rng.normal(size=v0.shape[0]), device=DEV)
    v = v0.clone().requires_grad_(True)
    u = u0.clone().requires_grad_(True)
    L = _loss(cfg, f, v, u, c, steps, dts)
    L.backward()
    gv, gu = v.grad.cpu().numpy(), u.grad.cpu().numpy()
    n = v0.shape[0]
    comps = [(int(i), int(d)) for i, d in zip(rng.integers(0, n, 48), rng.integers(0, 3, 48))]
    comps_u = [int(i) for i in rng.integers(0, n, 16)]
    fd_v, fd_u = [], []
    with torch.no_grad():
        for i, d in comps:
            e = 1e-6
            vp, vm = v0.clone(), v0.clone()
            vp[i, d] += e
            vm[i, d] -= e
            fd_v.append((_loss(cfg, f, vp, u0, c, steps, dts) -
                         _loss(cfg, f, vm, u0, c, steps, dts)).item() / (2 * e))
        for i in comps_u:
            e = 1e-5
            up, um = u0.clone(), u0.clone()
            up[i] += e
            um[i] -= e
            fd_u.append((_loss(cfg, f, v0, up, c, steps, dts) -
                         _loss(cfg, f, v0, um, c, steps, dts)).item() / (2 * e))
    ad_v = np.array([gv[i, d] for i, d in comps])
    ad_u = np.array([gu[i] for i in comps_u])
    fd_v, fd_u = np.array(fd_v), np.array(fd_u)
    scale_v, scale_u = np.abs(fd_v).max(), np.abs(fd_u).max()
    err_v = np.abs(ad_v - fd_v).max() / scale_v
    err_u = np.abs(ad_u - fd_u).max() / scale_u
    say(f"[adjoint fd] {name}: max |grad v| {scale_v:.3e} rel err {err_v:.2e}; "
        f"max |grad u| {scale_u:.3e} rel err {err_u:.2e}")
    assert scale_v > 0 and scale_u > 0
    # the forward map is smooth only to the h tolerance (1e-12 relative): a loss of O(30) then
    # carries O(3e-11) noise, i.e. O(1e-5) absolute in a difference quotient with step 1e-6
    assert err_v < 1e-4 and err_u < 1e-4


def test_adjoint_kh64_sixty_four_components():
    """The differentiable Kelvin-Helmholtz workload at full size: 64^3 (262,444 particles), 10
    steps, a checkpoint every 5, loss = mean squared difference of the final density from a
    target made by the plain method from a phase-shifted perturbation; the gradient w.r.t. the
    initial velocities vs central differences on 64 random components.  Same perturbed start as
    the small cases (see the module docstring).  Measured: max error 2.3e-4 of the largest
    difference quotient, median 4e-7; the tolerance leaves room for one pair crossing an
    artificial-viscosity kink inside a difference step.  The FP32 adjoint (the workload's
    precision) of the same rollout is held to rel-L2 1e-4 of the FP64 gradient and to the same
    central differences."""
    say("[adjoint kh64]")
    cfg, f = setups.kelvin_helmholtz(64)
    cfg = dict(cfg, epsilon_h=1e-12, cfl_cour=0.1, cfl_force=0.1)
    rng = np.random.default_rng(11)
    v = np.stack([f["vx"], f["vy"], f["vz"]], 1)
    v0 = torch.tensor(v + 1e-3 * rng.standard_normal(v.shape), device=DEV)
    u0 = torch.tensor(f["u"] * (1 + 0.05 * rng.standard_normal(f["u"].size)), device=DEV)
    n, steps, k = v0.shape[0], 10, 5
    assert n == 262_444
    dts = rollout_dts(cfg, f, v0, u0, steps)
    _, ft = setups.kelvin_helmholtz(64, phase=0.1)
    vt = torch.tensor(np.stack([ft["vx"], ft["vy"], ft["vz"]], 1) +
                      1e-3 * rng.standard_normal(v.shape), device=DEV)
    with torch.no_grad():
        target = rollout_density(cfg, f, vt, u0, steps, checkpoint_every=k, dts=dts)

    def loss(vv):
        rho = rollout_density(cfg, f, vv, u0, steps, checkpoint_every=k, dts=dts)
        return ((rho - target) ** 2).mean()

    vv = v0.clone().requires_grad_(True)
    loss(vv).backward()
    g = vv.grad.cpu().numpy()
    # the FP32 adjoint (pair terms in FP32, Kahan sums, FP64 r_ab) of the same FP64 rollout
    cfg32 = dict(cfg, adjoint_fp32=1)
    v32 = v0.clone().requires_grad_(True)
    rho32 = rollout_density(cfg32, f, v32, u0, steps, checkpoint_every=k, dts=dts)
    ((rho32 - target) ** 2).mean().backward()
    g32 = v32.grad.cpu().numpy()
    rel32 = float(np.linalg.norm(g32 - g) / np.linalg.norm(g))
    say(f"[adjoint kh64] FP32 adjoint vs FP64 adjoint: rel-L2 {rel32:.2e}")
    assert rel32 < 1e-4
    comps = [(int(i), int(d)) for i, d in zip(rng.integers(0, n, 64), rng.integers(0, 3, 64))]
    fd = []
    with torch.no_grad():
        for i, d in comps:
            e = 1e-6
            vp, vm = v0.clone(), v0.clone()
            vp[i, d] += e
            vm[i, d] -= e
            fd.append((loss(vp) - loss(vm)).item() / (2 * e))
    fd = np.array(fd)
    ad = np.array([g[i, d] for i, d in comps])
    scale = np.abs(fd).max()
    err = np.abs(ad - fd) / scale
    err32 = np.abs(np.array([g32[i, d] for i, d in comps]) - fd) / scale
    say(f"[adjoint kh64] max |fd| {scale:.3e} max rel err {err.max():.2e} "
        f"median {np.median(err):.2e}; FP32 adjoint max {err32.max():.2e} "
        f"median {np.median(err32):.2e}")
    assert scale > 0
    assert err.max() < 1e-3 and np.median(err) < 1e-5
    assert err32.max() < 1e-3 and np.median(err32) < 1e-4


def test_adjoint_lp07_matches_central_differences():
    """Locally isothermal LP07 gas (the disc workload's EOS, c_s depending on the distance to
    the origin), free boundaries, no sink: the gradient of sum_a c_a rho_a after 4 steps w.r.t.
    the initial velocities vs central differences on 48 random components.  The velocities move
    the particles, the positions set c_s and P, so the position derivative of the sound speed
    is exercised."""
    say("[adjoint lp07]")
    cfg, f = setups.disc(20000, seed=2)
    cfg = dict(cfg, sinks=[], epsilon_h=1e-12, cfl_cour=0.1, cfl_force=0.1)
    rng = np.random.default_rng(5)
    v = np.stack([f["vx"], f["vy"], f["vz"]], 1)
    v0 = torch.tensor(v + 1e-3 * rng.standard_normal(v.shape), device=DEV)
    u0 = torch.tensor(f["u"], device=DEV)
    steps = 4
    dts = rollout_dts(cfg, f, v0, u0, steps)
    c = torch.tensor(rng.normal(size=v0.shape[0]), device=DEV)
    vv = v0.clone().requires_grad_(True)
    _loss(cfg, f, vv, u0, c, steps, dts).backward()
    g = vv.grad.cpu().numpy()
    n = v0.shape[0]
    comps = [(int(i), int(d)) for i, d in zip(rng.integers(0, n, 48), rng.integers(0, 3, 48))]
    fd = []
    with torch.no_grad():
        for i, d in comps:
            e = 1e-6
            vp, vm = v0.clone(), v0.clone()
            vp[i, d] += e
            vm[i, d] -= e
            fd.append((_loss(cfg, f, vp, u0, c, steps, dts) -
                       _loss(cfg, f, vm, u0, c, steps, dts)).item() / (2 * e))
    fd = np.array(fd)
    ad = np.array([g[i, d] for i, d in comps])
    scale = np.abs(fd).max()
    err = np.abs(ad - fd) / scale
    say(f"[adjoint lp07] max |fd| {scale:.3e} max rel err {err.max():.2e} "
        f"median {np.median(err):.2e}")
    assert scale > 0
    assert err.max() < 1e-4


@pytest.mark.parametrize("loss_kind,tol", [("entropy", 1e-4), ("weighted_u", 1e-6)])
def test_adjoint_disc_with_sink(loss_kind, tol):
    """The disc workload's derivative at test size: 20,000 particles, LP07, central sink
    (softened), 4 steps; the gradient w.r.t. the initial velocities (through the gas and the
    sink, whose position, velocity and acceleration from the gas are part of the differentiated
    state) vs central differences on 48 random components.  Losses: the final entropy sum
    sum_a u_a / rho_a^(gamma - 1) of the workload, and sum_a c_a u_a.  Measured: 1.2e-5 of the
    largest quotient for the entropy sum, the same with the sink removed (the noise of that
    loss's difference quotients), 8e-8 for the weighted u."""
    from shamrock_amd.adjoint import rollout_state

    say(f"[adjoint disc sink] {loss_kind}")
    cfg, f = setups.disc(20000, seed=2)
    cfg = dict(cfg, epsilon_h=1e-12, cfl_cour=0.1, cfl_force=0.1)
    gm1 = cfg["gamma"] - 1
    rng = np.random.default_rng(9)
    v = np.stack([f["vx"], f["vy"], f["vz"]], 1)
    v0 = torch.tensor(v + 1e-3 * rng.standard_normal(v.shape), device=DEV)
    u0 = torch.tensor(f["u"], device=DEV)
    steps = 4
    dts = rollout_dts(cfg, f, v0, u0, steps)

    c = torch.tensor(rng.normal(size=v0.shape[0]), device=DEV)

    def loss(vv):
        rho, u = rollout_state(cfg, f, vv, u0, steps, checkpoint_every=2, dts=dts)
        return (u / rho ** gm1).sum() if loss_kind == "entropy" else (c * u).sum()

    vv = v0.clone().requires_grad_(True)
    loss(vv).backward()
    g = vv.grad.cpu().numpy()
    n = v0.shape[0]
    comps = [(int(i), int(d)) for i, d in zip(rng.integers(0, n, 48), rng.integers(0, 3, 48))]
    fd = []
    with torch.no_grad():
        for i, d in comps:
            e = 1e-6
            vp, vm = v0.clone(), v0.clone()
            vp[i, d] += e
            vm[i, d] -= e
            fd.append((loss(vp) - loss(vm)).item() / (2 * e))
    fd = np.array(fd)
    ad = np.array([g[i, d] for i, d in comps])
    scale = np.abs(fd).max()
    err = np.abs(ad - fd) / scale
    say(f"[adjoint disc sink] {loss_kind}: max |fd| {scale:.3e} max rel err {err.max():.2e} "
        f"median {np.median(err):.2e}")
    assert scale > 0
    assert err.max() < tol


def test_adjoint_rejects_lp07_with_periodic_images():
    from shamrock_amd.solver import GpuSolver

    cfg, f = setups.random_box(300, seed=4, jitter=0.3)
    cfg = dict(cfg, eos_kind=2, eos_cs0=1.0, eos_q=0.5, eos_r0=1.0)
    g = GpuSolver(cfg, f)
    g.evolve_once()
    seeds = [torch.zeros(g.n, dtype=torch.float64, device=DEV) for _ in range(11)]
    with pytest.raises(Exception, match="LP07 with free boundaries"):
        g.adjoint_step(seeds)
    g.close()


def test_adjoint_independent_of_checkpoint_interval():
    """The gradient does not depend on the checkpoint interval, nor on whether the backward
    pass takes the steps' products from the product store or from replays: bit for bit."""
    say("[adjoint ckpt]")
    cfg, f, v0, u0 = _case("random_box")
    steps = 5
    dts = rollout_dts(cfg, f, v0, u0, steps)
    c = torch.tensor(np.random.default_rng(1).normal(size=v0.shape[0]), device=DEV)
    grads = []
    for k, store in ((1, True), (2, True), (5, True), (2, False), (5, False)):
        v = v0.clone().requires_grad_(True)
        u = u0.clone().requires_grad_(True)
        rho = rollout_density(cfg, f, v, u, steps, checkpoint_every=k, dts=dts,
                              store_products=store)
        (rho * c).sum().backward()
        grads.append((v.grad.cpu().numpy(), u.grad.cpu().numpy()))
    for g in grads[1:]:
        np.testing.assert_array_equal(g[0], grads[0][0])
        np.testing.assert_array_equal(g[1], grads[0][1])


def test_fit_initial_velocity_to_target_density():
    """Gradient descent on v0 toward the final density of a phase-shifted run lowers the loss."""
    say("[adjoint fit]")
    cfg, f, v0, u0 = _case("kelvin_helmholtz")
    cfg_t, f_t = setups.kelvin_helmholtz(10, phase=0.1)
    steps = 4
    dts = rollout_dts(cfg, f, v0, u0, steps)
    vt = torch.tensor(np.stack([f_t["vx"], f_t["vy"], f_t["vz"]], 1), device=DEV)
    with torch.no_grad():
        target = rollout_density(cfg, f, vt, u0, steps, dts=dts)
    v = v0.clone().requires_grad_(True)
    losses = []
    lr = None
    for it in range(4):
        rho = rollout_density(cfg, f, v, u0, steps, dts=dts)
        loss = ((rho - target) ** 2).mean()
        losses.append(loss.item())
        (g,) = torch.autograd.grad(loss, v)
        if lr is None:  # a step that moves v by ~1e-3 of its scale
            lr = 1e-3 * v0.abs().max().item() / g.abs().max().item()
        with torch.no_grad():
            v -= lr * g
    say(f"[adjoint fit] losses {losses}")
    assert losses[-1] < losses[0]


def _rerun_case():
    """A rollout whose correctors rerun: the jittered random box (u ~ 1, v ~ 0.05) at multiples
    of the solver's time steps, else a Sedov blast at 12^3."""
    for name, factors in (("random_box", (2, 4, 8)), ("sedov", (1, 2))):
        if name == "random_box":
            cfg, f = setups.random_box(1000, seed=4, jitter=0.3)
        else:
            cfg, f = setups.sedov(12)
        cfg = dict(cfg, epsilon_h=1e-12)
        rng = np.random.default_rng(5)
        v = np.stack([f["vx"], f["vy"], f["vz"]], 1)
        v0 = torch.tensor(v + 1e-3 * rng.standard_normal(v.shape), device=DEV)
        u0 = torch.tensor(f["u"] * (1 + 0.05 * rng.standard_normal(f["u"].size)), device=DEV)
        steps = 4
        base = rollout_dts(cfg, f, v0, u0, steps)
        for factor in factors:
            dts = [factor * d for d in base]
            passes = rollout_passes(cfg, f, v0, u0, steps, dts)
            say(f"[adjoint reruns] {name} dt x {factor}: corrector passes per step {passes}")
            if max(passes) >= 2:
                return name, cfg, f, v0, u0, steps, dts, passes
    return None


def test_adjoint_through_corrector_reruns():
    """Steps whose corrector reruns (time steps a multiple of those the solver picks, until a
    pass's velocity change is above the 1e-2 tolerance) are differentiated through every pass:
    pass i evaluates the force from the velocities and u of pass i - 1.  The gradient of
    sum c_a rho_a w.r.t. v0 and u0 is compared with central differences (relative steps for u;
    a component whose perturbed runs take other numbers of passes is a jump of the map, not a
    derivative, and is left out), for the FP64 and the FP32 adjoint."""
    say("[adjoint reruns]")
    case = _rerun_case()
    assert case is not None, "no rollout with corrector reruns"
    name, cfg, f, v0, u0, steps, dts, passes = case
    rng = np.random.default_rng(3)
    c = torch.tensor(rng.normal(size=v0.shape[0]), device=DEV)
    grads = {}
    for fp32 in (0, 1):
        cf = dict(cfg, adjoint_fp32=fp32)
        v = v0.clone().requires_grad_(True)
        u = u0.clone().requires_grad_(True)
        _loss(cf, f, v, u, c, steps, dts).backward()
        grads[fp32] = (v.grad.cpu().numpy(), u.grad.cpu().numpy())
    n = v0.shape[0]
    uabs = np.abs(u0.cpu().numpy())
    comps = [(int(i), int(d)) for i, d in zip(rng.integers(0, n, 32), rng.integers(0, 3, 32))]
    comps_u = [int(i) for i in rng.choice(np.argsort(uabs)[-n // 4:], 12, replace=False)]
    fd_v, fd_u, keep_v, keep_u = [], [], [], []

    def same_passes(vv, uu):
        return rollout_passes(cfg, f, vv, uu, steps, dts) == passes

    with torch.no_grad():
        for i, d in comps:
            e = 1e-6
            vp, vm = v0.clone(), v0.clone()
            vp[i, d] += e
            vm[i, d] -= e
            keep_v.append(same_passes(vp, u0) and same_passes(vm, u0))
            fd_v.append((_loss(cfg, f, vp, u0, c, steps, dts) -
                         _loss(cfg, f, vm, u0, c, steps, dts)).item() / (2 * e))
        for i in comps_u:
            e = 1e-5 * uabs[i]
            up, um = u0.clone(), u0.clone()
            up[i] += e
            um[i] -= e
            keep_u.append(same_passes(v0, up) and same_passes(v0, um))
            # the loss change per relative change of u_i
            fd_u.append((_loss(cfg, f, v0, up, c, steps, dts) -
                         _loss(cfg, f, v0, um, c, steps, dts)).item() / (2 * e) * uabs[i])
    keep_v, keep_u = np.array(keep_v), np.array(keep_u)
    say(f"[adjoint reruns] {name}: components kept v {keep_v.sum()}/{keep_v.size}, "
        f"u {keep_u.sum()}/{keep_u.size}")
    assert keep_v.sum() >= 16 and keep_u.sum() >= 6
    fd_v, fd_u = np.array(fd_v)[keep_v], np.array(fd_u)[keep_u]
    cv = [cd for cd, k in zip(comps, keep_v) if k]
    cu = [i for i, k in zip(comps_u, keep_u) if k]
    scale_v, scale_u = np.abs(fd_v).max(), np.abs(fd_u).max()
    assert scale_v > 0 and scale_u > 0
    for fp32, (gv, gu) in grads.items():
        ad_v = np.array([gv[i, d] for i, d in cv])
        ad_u = np.array([gu[i] * uabs[i] for i in cu])
        err_v = np.abs(ad_v - fd_v).max() / scale_v
        err_u = np.abs(ad_u - fd_u).max() / scale_u
        say(f"[adjoint reruns] fp32={fp32}: rel err v {err_v:.2e}, u {err_u:.2e}")
        assert err_v < 1e-4 and err_u < 1e-4
    g64, g32 = grads[0][0], grads[1][0]
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-4
