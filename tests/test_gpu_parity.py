"""GPU parity tests: the HIP library (through its C ABI) against the CPU oracle.

Integer / index products (Morton codes, sort, reduction, Karras tree, boxes, neighbour lists,
ghost selection) must be bit-exact; FP64 fields (h, Omega, P, c_s, a, du/dt) must agree within
1e-10 relative (the reference's own regression tolerance, regression_sph_kh.py:281-300).
"""
import ctypes as C
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import oracle_lib as orc  # noqa: E402
from shamrock_amd import _lib, setups  # noqa: E402
from shamrock_amd._lib import check, lib  # noqa: E402
from shamrock_amd.solver import GpuSolver  # noqa: E402

DEV = "cuda:0"
RTOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    lib()
    n = C.c_int(0)
    check(lib().shamx_get_device_count(C.byref(n)))
    assert n.value >= 1


@pytest.fixture(scope="module")
def ctx():
    h = C.c_void_p()
    check(lib().shamx_ctx_create(C.byref(h)))
    yield h
    lib().shamx_ctx_destroy(h)


def dev(a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(DEV)


def stream():
    return torch.cuda.current_stream().cuda_stream


def relerr(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    if a.size == 0 and b.size == 0:
        return 0.0
    scale = np.maximum(np.abs(b), np.max(np.abs(b)) * 1e-3 + 1e-300)
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


# ------------------------------------------------------------------ single phases
@pytest.mark.parametrize("n,seed", [(1, 0), (1000, 1), (200_000, 2)])
def test_bbox_morton_bitexact(ctx, n, seed):
    say(f"[bbox_morton] n={n} seed={seed}")
    rng = np.random.default_rng(seed)
    pos = rng.normal(0, 1, (n, 3)) * np.array([1, 3, 0.1]) + np.array([0.5, -2, 7])
    x, y, z = (dev(pos[:, k], np.float64) for k in range(3))
    bb = torch.zeros(6, dtype=torch.float64, device=DEV)
    check(lib().shamx_bbox(ctx, stream(), n, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                           bb.data_ptr()))
    codes = torch.zeros(n, dtype=torch.int32, device=DEV)
    check(lib().shamx_morton(ctx, stream(), n, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                             bb.data_ptr(), codes.data_ptr()))
    torch.cuda.synchronize()
    bb_o = orc.bbox(pos)
    np.testing.assert_array_equal(bb.cpu().numpy(), bb_o)
    codes_o = orc.morton(pos, bb_o, n, bits=32)
    np.testing.assert_array_equal(codes.cpu().numpy().view(np.uint32), codes_o)


def test_morton_closed_forms(ctx):
    """(0,0,0) -> 0, the far corner -> 0x3FFFFFFF (mortonTests.cpp:15-53), axis codes."""
    pos = np.array([[0, 0, 0], [1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    x, y, z = (dev(pos[:, k], np.float64) for k in range(3))
    bb = dev(np.array([0, 0, 0, 1, 1, 1], float))
    codes = torch.zeros(5, dtype=torch.int32, device=DEV)
    check(lib().shamx_morton(ctx, stream(), 5, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                             bb.data_ptr(), codes.data_ptr()))
    c = codes.cpu().numpy().view(np.uint32)
    assert c[0] == 0 and c[1] == 0x3FFFFFFF
    assert c[2] == 0x24924924 and c[3] == 0x12492492 and c[4] == 0x09249249


@pytest.mark.parametrize("n,bits,seed", [(0, 30, 0), (1, 30, 0), (17, 30, 1), (4097, 30, 2),
                                         (1_000_003, 30, 3), (300_000, 8, 4),
                                         (2_500_000, 30, 5)])
def test_radix_sort_stable(ctx, n, bits, seed):
    say(f"[radix_sort] n={n} bits={bits}")
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    if n > 100:
        keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # many duplicates
    k = dev(keys.view(np.int32))
    perm = torch.zeros(n, dtype=torch.int32, device=DEV)
    check(lib().shamx_sort(ctx, stream(), n, bits, k.data_ptr(), perm.data_ptr()))
    torch.cuda.synchronize()
    ref = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(perm.cpu().numpy().view(np.uint32), ref)
    np.testing.assert_array_equal(k.cpu().numpy().view(np.uint32), keys[ref])


@pytest.mark.parametrize("items", ["8", "16"])
@pytest.mark.parametrize("n,bits,seed", [(1, 30, 0), (4097, 30, 2), (300_001, 32, 6),
                                         (1_000_003, 30, 3)])
def test_radix_sort_stable_tile_sizes(ctx, n, bits, seed, items, monkeypatch):
    """Both tile sizes of the sort (16 keys per thread, the default, and 8) give the stable
    order."""
    monkeypatch.setenv("SHAMX_SORT_ITEMS", items)
    say(f"[radix_sort] n={n} bits={bits} items={items}")
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    if n > 100:
        keys[: n // 3] = keys[n // 3: 2 * (n // 3)]
    k = dev(keys.view(np.int32))
    perm = torch.zeros(n, dtype=torch.int32, device=DEV)
    check(lib().shamx_sort(ctx, stream(), n, bits, k.data_ptr(), perm.data_ptr()))
    torch.cuda.synchronize()
    ref = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(perm.cpu().numpy().view(np.uint32), ref)
    np.testing.assert_array_equal(k.cpu().numpy().view(np.uint32), keys[ref])


def _tree_gpu(ctx, pos, h, level, htol, split=False):
    """Tree through shamx_tree_build, or (split) through its two halves
    shamx_tree_reduce_karras + shamx_tree_aabb_hmax."""
    n = pos.shape[0]
    x, y, z = (dev(pos[:, k], np.float64) for k in range(3))
    hh = dev(h, np.float64)
    if split:
        check(lib().shamx_tree_reduce_karras(ctx, stream(), n, x.data_ptr(), y.data_ptr(),
                                             z.data_ptr(), level))
        check(lib().shamx_tree_aabb_hmax(ctx, stream(), x.data_ptr(), y.data_ptr(),
                                         z.data_ptr(), hh.data_ptr(), htol))
    else:
        check(lib().shamx_tree_build(ctx, stream(), n, x.data_ptr(), y.data_ptr(),
                                     z.data_ptr(), hh.data_ptr(), level, htol))
    s = np.zeros(4, np.uint32)
    check(lib().shamx_tree_sizes(ctx, s.ctypes.data_as(C.POINTER(C.c_uint32))))
    n_obj, L, I = (int(v) for v in s[:3])
    out = dict(sorted=torch.zeros(n, dtype=torch.int32, device=DEV),
               sort_index=torch.zeros(n, dtype=torch.int32, device=DEV),
               reduc_map=torch.zeros(L + 2, dtype=torch.int32, device=DEV),
               lchild=torch.zeros(max(I, 1), dtype=torch.int32, device=DEV),
               rchild=torch.zeros(max(I, 1), dtype=torch.int32, device=DEV),
               lflag=torch.zeros(max(I, 1), dtype=torch.uint8, device=DEV),
               rflag=torch.zeros(max(I, 1), dtype=torch.uint8, device=DEV),
               endrange=torch.zeros(max(I, 1), dtype=torch.int32, device=DEV),
               aabb_min=torch.zeros(3 * (I + L), dtype=torch.float64, device=DEV),
               aabb_max=torch.zeros(3 * (I + L), dtype=torch.float64, device=DEV),
               rint=torch.zeros(I + L, dtype=torch.float64, device=DEV))
    check(lib().shamx_tree_get(ctx, stream(), *[out[k].data_ptr() for k in
                                                ("sorted", "sort_index", "reduc_map", "lchild",
                                                 "rchild", "lflag", "rflag", "endrange",
                                                 "aabb_min", "aabb_max", "rint")]))
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    for k in ("sorted", "sort_index", "reduc_map", "lchild", "rchild", "endrange"):
        res[k] = res[k].view(np.uint32)
    for k in ("lchild", "rchild", "lflag", "rflag", "endrange"):
        res[k] = res[k][:I]
    res["aabb_min"] = res["aabb_min"].reshape(-1, 3)
    res["aabb_max"] = res["aabb_max"].reshape(-1, 3)
    res.update(L=L, I=I, x=x, y=y, z=z, h=hh)
    return res


def test_tree_build_rejects_empty(ctx):
    """An empty set is refused on the host, before any launch, as the reference refuses it
    ("obj_cnt is 0, cannot build a CompressedLeafBVH")."""
    say("[tree_build] empty")
    x = dev(np.zeros(1))
    with pytest.raises(ValueError, match="obj_cnt is 0"):
        check(lib().shamx_tree_build(ctx, stream(), 0, x.data_ptr(), x.data_ptr(), x.data_ptr(),
                                     x.data_ptr(), 3, 1.1))


def _same_modulo_ties(codes_sorted, perm_a, perm_b):
    """permutations equal as multisets inside every run of equal keys"""
    starts = np.flatnonzero(np.r_[True, codes_sorted[1:] != codes_sorted[:-1]])
    ends = np.r_[starts[1:], codes_sorted.size]
    for s, e in zip(starts, ends):
        if e - s == 1:
            if perm_a[s] != perm_b[s]:
                return False
        elif not np.array_equal(np.sort(perm_a[s:e]), np.sort(perm_b[s:e])):
            return False
    return True


@pytest.mark.parametrize("case,split", [("lattice", False), ("clustered", False),
                                        ("clustered", True), ("single", False),
                                        ("coincident", False), ("pair", True)])
def test_tree_build_bitexact(ctx, case, split):
    """Whole-tree parity with the oracle.  The small cases are the one-leaf trees of the
    reference's one-cell test (KarrasRadixTreeAABBTests.cpp:241-380: every object in a single
    reduced cell, no internal cell): one particle, seven at one point (a zero-width bounding
    box), and two close particles merged by the reduction."""
    say(f"[tree_build] case={case} split={split}")
    rng = np.random.default_rng(7)
    if case == "lattice":
        pos = setups.lattice_cubic(1 / 24, (0, 0, 0), (1, 1, 1))
        pos = pos + rng.uniform(-0.1, 0.1, pos.shape) / 24
    elif case == "single":
        pos = np.array([[0.3, -0.2, 0.7]])
    elif case == "coincident":
        pos = np.tile([[0.3, -0.2, 0.7]], (7, 1))
    elif case == "pair":
        pos = np.array([[0.0, 0.0, 0.0], [1e-3, 2e-3, -1e-3]])
    else:  # heavy Morton-cell collisions: a Plummer-like core
        r = rng.pareto(2.0, 60_000)[:, None] * 1e-3
        d = rng.normal(size=(60_000, 3))
        pos = r * d / np.linalg.norm(d, axis=1, keepdims=True)
    h = rng.uniform(0.01, 0.05, pos.shape[0])
    g = _tree_gpu(ctx, pos, h, 3, 1.1, split=split)
    bb = orc.bbox(pos)
    t = orc.Tree(pos, bb, 3, bits=32)
    n = pos.shape[0]
    np.testing.assert_array_equal(g["sorted"], t.sorted_codes[:n])
    assert _same_modulo_ties(t.sorted_codes[:n], g["sort_index"], t.sort_index[:n])
    if case == "lattice":
        np.testing.assert_array_equal(g["sort_index"], t.sort_index[:n])
    np.testing.assert_array_equal(g["reduc_map"], t.reduc_map)
    assert g["L"] == t.L and g["I"] == t.I
    for k in ("lchild", "rchild", "lflag", "rflag", "endrange"):
        np.testing.assert_array_equal(g[k], getattr(t, k), err_msg=k)
    np.testing.assert_array_equal(g["aabb_min"], t.aabb_min)
    np.testing.assert_array_equal(g["aabb_max"], t.aabb_max)
    np.testing.assert_array_equal(g["rint"], t.field_max(h) * 1.1)
    if case in ("single", "coincident", "pair"):
        assert g["L"] == 1 and g["I"] == 0


def test_cfl_closed_form():
    """shamx_cfl: dt_i = min(C_cour h / vsig, C_force sqrt(h / |a|)) and the minimum, bit for
    bit against the same IEEE operations in numpy (ComputeCFLCourant / ComputeCFLForce)."""
    say("[cfl]")
    rng = np.random.default_rng(5)
    n = 100_003
    h = rng.uniform(1e-3, 1e-1, n)
    vsig = rng.uniform(0.1, 10, n)
    a = rng.normal(size=(3, n))
    a[:, 17] = 0  # |a| = 0: the force limit is +inf, the Courant one applies
    out = dev(np.zeros(n))
    dmin = np.zeros(1)
    # the device inputs stay referenced until the call has returned
    ins = [dev(h), dev(vsig), dev(a[0]), dev(a[1]), dev(a[2])]
    check(lib().shamx_cfl(stream(), n, *[t.data_ptr() for t in ins], 0.3, 0.25, out.data_ptr(),
                          _lib.dptr(dmin)))
    absa = np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    with np.errstate(divide="ignore"):
        want = np.minimum(0.3 * h / vsig, 0.25 * np.sqrt(h / absa))
    np.testing.assert_array_equal(out.cpu().numpy(), want)
    assert dmin[0] == want.min()


def _eos_gpu(eos_kind, pos, h, u, mass, gamma=5 / 3, cs=1.0, cs0=1.0, q=0.5, r0=1.0):
    n = pos.shape[0]
    ins = [dev(h), dev(u), dev(pos[:, 0]), dev(pos[:, 1]), dev(pos[:, 2])]
    P, c = dev(np.zeros(n)), dev(np.zeros(n))
    check(lib().shamx_eos(stream(), 0, n, eos_kind, gamma, cs, cs0, q, r0,
                          *[t.data_ptr() for t in ins], P.data_ptr(), c.data_ptr(), mass))
    torch.cuda.synchronize()
    return P.cpu().numpy(), c.cpu().numpy()


def test_eos_closed_forms_on_gpu():
    """shamx_eos against the closed forms of the three equations of state (eos.hpp:31-98,
    ComputeEos.cpp:152-247, 656-665, 726-798): adiabatic P = (gamma - 1) rho u,
    c_s = sqrt(gamma P / rho); isothermal P = c_s^2 rho with the configured c_s; LP07
    c_s^2 = c_s0^2 (|r|^2 / r0^2)^(-q), P = c_s^2 rho, with c_s = c_s0 at |r| = r0 and
    c_s = c_s0 / 2 at |r| = 16 r0 for q = 1/4 (rho = m (hfact / h)^3, M4 hfact 1.2)."""
    say("[eos]")
    rng = np.random.default_rng(3)
    n = 10_001
    pos = rng.normal(0, 5, (n, 3))
    pos[0], pos[1], pos[2] = (1, 0, 0), (0, -1, 0), (16, 0, 0)
    h = rng.uniform(0.05, 0.5, n)
    u = rng.uniform(0.1, 2.0, n)
    m = 1e-3
    rho = m * (1.2 / h) ** 3
    P, c = _eos_gpu(0, pos, h, u, m, gamma=1.4)
    np.testing.assert_allclose(P, 0.4 * rho * u, rtol=1e-15)
    np.testing.assert_allclose(c, np.sqrt(1.4 * P / rho), rtol=1e-15)
    P, c = _eos_gpu(1, pos, h, u, m, cs=0.37)
    assert np.all(c == 0.37)
    np.testing.assert_allclose(P, 0.37 * 0.37 * rho, rtol=1e-15)
    cs0, q = 0.8, 0.25
    P, c = _eos_gpu(2, pos, h, u, m, cs0=cs0, q=q, r0=1.0)
    assert c[0] == pytest.approx(cs0, rel=1e-15) and c[1] == pytest.approx(cs0, rel=1e-15)
    assert c[2] == pytest.approx(cs0 / 2, rel=1e-15)
    cs2 = cs0**2 * ((pos**2).sum(1)) ** (-q)
    np.testing.assert_allclose(c, np.sqrt(cs2), rtol=1e-14)
    np.testing.assert_allclose(P, cs2 * rho, rtol=1e-14)
    _, c2 = _eos_gpu(2, pos, h, u, m, cs0=cs0, q=q, r0=2.0)
    assert c2[2] ** 2 == pytest.approx((cs0 / 2) ** 2 * 4**q, rel=1e-14)


def _neigh_gpu(ctx, pos, h, n_local, htol, Rkern):
    g = _tree_gpu(ctx, pos, h, 3, htol)
    check(lib().shamx_neigh_build(ctx, stream(), n_local, g["x"].data_ptr(), g["y"].data_ptr(),
                                  g["z"].data_ptr(), g["h"].data_ptr(), htol, Rkern))
    s = np.zeros(4, np.uint32)
    check(lib().shamx_neigh_sizes(ctx, s.ctypes.data_as(C.POINTER(C.c_uint32))))
    P = int(s[1]) + (int(s[2]) << 32)
    PL = int(s[3])
    L = g["L"]
    t = {k: torch.zeros(max(m, 1), dtype=torch.int32, device=DEV)
         for k, m in (("cnt", n_local), ("offset", n_local), ("ids", P), ("owner", n_local),
                      ("leaf_cnt", L), ("leaf_offset", L), ("leaf_ids", PL))}
    check(lib().shamx_neigh_get(ctx, stream(), *[t[k].data_ptr() for k in
                                                 ("cnt", "offset", "ids", "owner", "leaf_cnt",
                                                  "leaf_offset", "leaf_ids")]))
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy().view(np.uint32) for k, v in t.items()}
    out["ids"], out["leaf_ids"] = out["ids"][:P], out["leaf_ids"][:PL]
    return out, g


@pytest.mark.parametrize("case", ["jittered", "clustered", "two_scales", "few", "coincident"])
def test_neighbour_cache_sets_match_oracle(ctx, case):
    """Per-particle neighbour sets and counts, owner leaves and leaf lists vs the oracle.  The
    last two are tiny sets: nine particles that all see each other, and nine at one point (a
    one-leaf tree with no internal cell)."""
    say(f"[neigh] case={case}")
    rng = np.random.default_rng(21)
    if case == "few":
        pos = rng.uniform(0, 1e-2, (9, 3))
        h = np.full(9, 0.05)
    elif case == "coincident":
        pos = np.tile([[0.25, 0.5, -0.75]], (9, 1))
        h = np.full(9, 0.01)
    elif case == "jittered":
        pos = setups.lattice_cubic(1 / 20, (0, 0, 0), (1, 1, 1))
        pos = pos + rng.uniform(-0.45, 0.45, pos.shape) / 20
        h = np.full(pos.shape[0], 1.2 / 20)
    elif case == "clustered":  # Plummer-like core: leaves of hundreds of particles
        from scipy.spatial import cKDTree

        n = 40_000
        r = rng.pareto(1.5, n)[:, None] * 2e-3
        d = rng.normal(size=(n, 3))
        pos = r * d / np.linalg.norm(d, axis=1, keepdims=True)
        h = cKDTree(pos).query(pos, 40)[0][:, -1] / 2  # ~60 neighbours per particle
    else:  # two smoothing-length scales side by side
        pos = rng.uniform(0, 1, (20_000, 3))
        h = np.where(pos[:, 0] < 0.5, 0.02, 0.06)
    n_local = int(pos.shape[0] * 0.8)
    go, g = _neigh_gpu(ctx, pos, h, n_local, 1.1, 2.0)
    o = orc.neigh_cache(pos, h, n_local)
    assert g["L"] == o["L"]
    np.testing.assert_array_equal(go["owner"], o["owner"])
    np.testing.assert_array_equal(go["leaf_cnt"], o["leaf_cnt"])
    np.testing.assert_array_equal(go["leaf_ids"], o["leaf_ids"])
    np.testing.assert_array_equal(go["cnt"], o["cnt"])
    np.testing.assert_array_equal(go["offset"], o["offset"])
    for a in range(n_local):
        ga = go["ids"][go["offset"][a]: go["offset"][a] + go["cnt"][a]]
        oa = o["ids"][o["offset"][a]: o["offset"][a] + o["cnt"][a]]
        if case == "jittered":  # unique Morton codes: same order, not only same set
            np.testing.assert_array_equal(ga, oa)
        else:
            np.testing.assert_array_equal(np.sort(ga), np.sort(oa))


def test_h_iterate_golden_on_gpu():
    """IterateSmoothingLengthDensityTests.cpp:250-366 sequences, through shamx_h_iterate."""
    import json
    import os
    say("[h_iterate golden]")
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "iterate_smoothing_length_density.json")))
    pos = np.array([[i, j, k] for i in range(4) for j in range(4) for k in range(4)], float)
    n = pos.shape[0]
    cnt = dev(np.full(n, n, np.int32))
    off = dev((np.arange(n) * n).astype(np.int32))
    ids = dev(np.tile(np.arange(n, dtype=np.int32), n))
    x, y, z = (dev(pos[:, k], np.float64) for k in range(3))
    new_h = dev(np.full(n, fx["h0"]))
    eps = dev(np.full(n, fx["eps0"]))
    mm = np.zeros(2)
    seq = dict(emin=[], emax=[], hmin=[], hmax=[])
    for outer in range(50):
        old_h = new_h.clone()
        eps.fill_(fx["eps0"])
        max_eps = fx["eps0"]
        for inner in range(10):
            check(lib().shamx_h_iterate(stream(), 0, n, cnt.data_ptr(), off.data_ptr(),
                                        ids.data_ptr(), x.data_ptr(), y.data_ptr(),
                                        z.data_ptr(), old_h.data_ptr(), new_h.data_ptr(),
                                        eps.data_ptr(), fx["gpart_mass"], fx["h_evol_max"],
                                        fx["h_evol_iter_max"], 0, _lib.dptr(mm)))
            hn = new_h.cpu().numpy()
            seq["emax"].append(mm[0])
            seq["emin"].append(mm[1])
            seq["hmin"].append(hn.min())
            seq["hmax"].append(hn.max())
            max_eps = mm[0]
            if max_eps < 1e-6:
                break
        if mm[1] == -1:
            continue
        if max_eps < 1e-6:
            break
    tol = fx["tol"]
    np.testing.assert_allclose(new_h.cpu().numpy(), fx["expected_h_end"], rtol=0, atol=tol)
    np.testing.assert_allclose(eps.cpu().numpy(), fx["expected_eps_end"], rtol=0, atol=tol)
    for k, ref in (("emin", "seq_eps_min"), ("emax", "seq_eps_max"), ("hmin", "seq_h_min"),
                   ("hmax", "seq_h_max")):
        assert len(seq[k]) == len(fx[ref]), k
        np.testing.assert_allclose(seq[k], fx[ref], rtol=0, atol=tol, err_msg=k)


# ------------------------------------------------------------------ full step
def _compare_transient(g, o, exact_sort=True):
    assert g["sizes"] == o["sizes"], (g["sizes"], o["sizes"])
    for k in ("ghost_src", "sorted_codes", "codes", "reduc_map", "lchild", "rchild", "lflag",
              "rflag", "endrange", "owner", "leaf_cnt", "leaf_offset", "leaf_ids"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    np.testing.assert_array_equal(g["tree_bbox"], o["tree_bbox"])
    np.testing.assert_array_equal(g["aabb_min"], o["aabb_min"])
    np.testing.assert_array_equal(g["aabb_max"], o["aabb_max"])
    if exact_sort:
        np.testing.assert_array_equal(g["rint_tree"], o["rint_tree"])
    else:
        # h from sums whose terms came in another order within runs of equal codes
        assert relerr(g["rint_tree"], o["rint_tree"]) <= RTOL
    if exact_sort:
        np.testing.assert_array_equal(g["sort_index"], o["sort_index"])
        for k in ("part_cnt", "part_offset", "part_ids"):
            np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    else:
        # the oracle's bitonic sort (the reference's) orders equal Morton codes differently
        # from the stable radix sort: same permutation inside every run of equal codes
        assert _same_modulo_ties(g["sorted_codes"], g["sort_index"], o["sort_index"])
        np.testing.assert_array_equal(g["part_cnt"], o["part_cnt"])
    for k in ("mx", "my", "mz"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    for k in ("mh", "omega", "P", "cs", "momega", "vsig"):
        assert relerr(g[k], o[k]) <= RTOL, (k, relerr(g[k], o[k]))
    # the GPU keeps v and u of the local particles in place (the corrector has updated them
    # since the ghost refresh), the oracle keeps a pre-corrector copy: compare the images only
    n = g["omega"].size
    for k in ("mu", "mvx", "mvy", "mvz"):
        assert relerr(g[k][n:], o[k][n:]) <= RTOL, (k, relerr(g[k][n:], o[k][n:]))


@pytest.mark.parametrize("workload,steps", [("sound_wave_16", 4), ("sedov_20", 3),
                                            ("random_box_4096", 3), ("plummer_2000", 3),
                                            ("kelvin_helmholtz_14", 3)])
def test_full_step_parity(workload, steps):
    """Whole steps vs the oracle.  Plummer: free boundaries, h over a decade (a realization the
    reference's subcycled Newton iteration converges on); Kelvin-Helmholtz: two lattices of
    different spacing, shear flow."""
    name, n = workload.rsplit("_", 1)
    if name == "sound_wave":
        cfg, f = setups.sound_wave(int(n))
    elif name == "sedov":
        cfg, f = setups.sedov(int(n))
    elif name == "plummer":
        cfg, f = setups.plummer(int(n), seed=1, r_max=3.0)
    elif name == "kelvin_helmholtz":
        cfg, f = setups.kelvin_helmholtz(int(n))
    else:
        cfg, f = setups.random_box(int(n), seed=3)
    g = GpuSolver(cfg, f)
    o = orc.Solver(cfg, f)
    for step in range(steps):
        say(f"[full_step] {workload} step {step}")
        lg = g.evolve_once()
        lo = o.evolve_once()
        assert lg["subcycles"] == lo["subcycles"] and lg["h_iterations"] == lo["h_iterations"]
        assert lg["n_pairs"] == lo["n_pairs"]
        _compare_transient(g.transient(), o.transient(), exact_sort=(name != "random"))
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "h", "u", "du"):
            assert relerr(g.get(k), o.get(k)) <= RTOL, (k, relerr(g.get(k), o.get(k)))
        tg, to = g.time(), o.time()
        assert abs(tg["dt"] - to["dt"]) <= RTOL * abs(to["dt"])
        assert tg["cfl_mult"] == to["cfl_mult"]
    g.close()


@pytest.mark.parametrize("workload,steps", [("sod_16", 5), ("kelvin_helmholtz_14", 4)])
def test_full_step_parity_varying_av(workload, steps):
    """Cullen & Dehnen varying viscosity (div v, curl v, d(div v)/dt, alpha update, per-particle
    alpha in the force) over whole steps vs the oracle, alpha and the saved sound speed
    included.  The first step runs with the sound speed 0 (alpha stays at its initial value);
    from the second on the source and decay terms are live."""
    name, n = workload.rsplit("_", 1)
    if name == "sod":
        cfg, f = setups.sod_tube(int(n))
    else:
        cfg, f = setups.kelvin_helmholtz(int(n))
        cfg = dict(cfg, av_kind=1, alpha_min=0.0, alpha_max=1.0, sigma_decay=0.1)
    g = GpuSolver(cfg, f)
    o = orc.Solver(cfg, f)
    alive = 0
    for step in range(steps):
        say(f"[varying_av] {workload} step {step}")
        lg = g.evolve_once()
        lo = o.evolve_once()
        assert lg["subcycles"] == lo["subcycles"] and lg["h_iterations"] == lo["h_iterations"]
        assert lg["n_pairs"] == lo["n_pairs"]
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "h", "u", "du", "alpha_AV",
                  "soundspeed"):
            assert relerr(g.get(k), o.get(k)) <= RTOL, (k, relerr(g.get(k), o.get(k)))
        alive += int(np.count_nonzero(o.get("alpha_AV")) > 0)
        tg, to = g.time(), o.time()
        assert abs(tg["dt"] - to["dt"]) <= RTOL * abs(to["dt"])
        assert tg["cfl_mult"] == to["cfl_mult"]
    assert alive >= 1, "alpha never left its initial value: the update was not exercised"
    g.close()


@pytest.mark.parametrize("periodic", [False, True])
def test_full_step_parity_sinks(periodic):
    """A gas disc around a central sink (LP07 EOS, constant AV 0.1, softened sink force) over
    whole steps vs the oracle: particle fields, the sink's position, velocity and both of its
    accelerations.  The sink's acceleration from the gas is a sum over all particles with heavy
    cancellation, so it is compared to 1e-12 of the sum of the magnitudes of its terms.  The
    periodic case puts the disc in a box it crosses, so the sink force must see the drifted
    positions before the wrap, as the reference does."""
    cfg, f = setups.disc(20000, seed=2)
    if periodic:
        cfg = dict(cfg, periodic=1, box_min=[-4.0, -4.0, -3.0], box_max=[4.0, 4.0, 3.0])
        keep = (np.abs(f["x"]) < 4.0) & (np.abs(f["y"]) < 4.0)
        f = {k: v[keep] for k, v in f.items()}
    g = GpuSolver(cfg, f)
    o = orc.Solver(cfg, f)
    for step in range(4):
        say(f"[sinks] periodic={periodic} step {step}")
        lg = g.evolve_once()
        lo = o.evolve_once()
        assert lg["subcycles"] == lo["subcycles"] and lg["h_iterations"] == lo["h_iterations"]
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "h", "u", "du"):
            assert relerr(g.get(k), o.get(k)) <= RTOL, (k, relerr(g.get(k), o.get(k)))
        sg, so = g.get_sinks()[0], o.get_sinks()[0]
        for k in ("pos", "velocity", "ext_acceleration"):
            np.testing.assert_allclose(sg[k], so[k], rtol=RTOL, atol=1e-300, err_msg=k)
        d = np.stack([o.get("x"), o.get("y"), o.get("z")], 1) - np.asarray(so["pos"])
        terms = cfg["constant_G"] * cfg["gpart_mass"] / (d * d).sum(1)
        np.testing.assert_allclose(sg["sph_acceleration"], so["sph_acceleration"], rtol=0,
                                   atol=1e-12 * terms.sum(), err_msg="sph_acceleration")
        assert g.time()["dt"] == pytest.approx(o.time()["dt"], rel=RTOL)
    g.close()


@pytest.mark.parametrize("workload", ["plummer_6000", "disc_8000"])
def test_density_force_pass_parity(workload):
    """The density/force pass (tree, cache, one h subcycle, Omega, EOS, forces, CFL) vs the
    oracle's.  Plummer truncated at r = 30: h over three decades and a subcycle that does not
    converge everywhere (the clamped particles must match too); disc: the sink's external
    force on the gas.  Run twice, the second pass on the first one's h."""
    name, n = workload.rsplit("_", 1)
    if name == "plummer":
        cfg, f = setups.plummer(int(n), seed=1, r_max=30.0)
    else:
        cfg, f = setups.disc(int(n), seed=2)
    g = GpuSolver(cfg, f)
    o = orc.Solver(cfg, f)
    for rep in range(2):
        say(f"[density_force] {workload} pass {rep}")
        lg = g.density_force()
        lo = o.density_force()
        assert lg["subcycles"] == lo["subcycles"] == 1
        assert lg["h_iterations"] == lo["h_iterations"]
        assert lg["n_pairs"] == lo["n_pairs"]
        # Plummer at r = 30: the 10-bit cells of the 60-wide box hold several particles
        _compare_transient(g.transient(), o.transient(), exact_sort=(name != "plummer"))
        for k in ("h", "ax", "ay", "az", "du"):
            assert relerr(g.get(k), o.get(k)) <= RTOL, (k, relerr(g.get(k), o.get(k)))
        for k in ("x", "vx", "u"):
            np.testing.assert_array_equal(g.get(k), f[k])
        assert lg["next_dt"] == pytest.approx(lo["next_dt"], rel=RTOL)
        assert g.time() == o.time()
    g.close()


@pytest.mark.parametrize("mode", ["packed_rows", "overflow", "batch_cull", "batch_cull_packed"])
@pytest.mark.parametrize("workload", ["disc_20000", "plummer_6000", "sedov_16"])
def test_neighbour_layouts_parity(workload, mode, monkeypatch):
    """The neighbour-cache layouts that the workloads of these tests do not reach by
    themselves.  packed_rows (SHAMX_NEIGH_ROWS=packed): fixed-capacity leaf lists and packed
    particle rows (count, scan, fill), the path of rows that u32 entry indices cannot address
    at fixed capacity.  overflow (SHAMX_NEIGH_CAP=16): capacities of 16, so that nearly every
    leaf list and row runs past its capacity and is filled again into the overflow regions
    (the path of the few outlier rows of a disc whose tail particles have 20x the median h).
    batch_cull (SHAMX_BATCH_CULL=1, alone and with packed rows): stage 2 drops candidates
    batch by batch (64 sorted positions of a leaf, on their boxes), the path of clustered
    sets.  Whole steps (or passes, Plummer) vs the oracle, lists included.  (The disc needs 20,000
    particles for its first step's h iteration to converge, in the oracle too.)"""
    if mode in ("packed_rows", "batch_cull_packed"):
        monkeypatch.setenv("SHAMX_NEIGH_ROWS", "packed")
    if mode.startswith("batch_cull"):
        monkeypatch.setenv("SHAMX_BATCH_CULL", "1")
    if mode == "overflow":
        monkeypatch.setenv("SHAMX_NEIGH_CAP", "16")
    name, n = workload.rsplit("_", 1)
    if name == "disc":
        cfg, f = setups.disc(int(n), seed=2)
    elif name == "plummer":
        cfg, f = setups.plummer(int(n), seed=1, r_max=30.0)
    else:
        cfg, f = setups.sedov(int(n))
    g = GpuSolver(cfg, f)
    o = orc.Solver(cfg, f)
    for step in range(2):
        say(f"[layouts] {workload} {mode} step {step}")
        if name == "plummer":
            lg, lo = g.density_force(), o.density_force()
        else:
            lg, lo = g.evolve_once(), o.evolve_once()
        assert lg["n_pairs"] == lo["n_pairs"]
        _compare_transient(g.transient(), o.transient(), exact_sort=(name == "sedov"))
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "h", "u", "du"):
            assert relerr(g.get(k), o.get(k)) <= RTOL, (k, relerr(g.get(k), o.get(k)))
    g.close()


def test_mixed_precision_error_bound():
    """Mixed step: rel-L2 of a and du/dt vs the FP64 step on the same state <= 1e-5."""
    cfg, f = setups.sedov(24)
    g64 = GpuSolver(cfg, f)
    cfg_m = dict(cfg, mixed_precision=1)
    g32 = GpuSolver(cfg_m, f)
    for step in range(2):
        say(f"[mixed] step {step}")
        g64.evolve_once()
        g32.evolve_once()
    # same state, one more force evaluation each: compare accelerations
    st = {k: g64.get(k) for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "h", "u",
                                  "du")}
    a = GpuSolver(cfg, st)
    b = GpuSolver(cfg_m, st)
    for s in (a, b):
        s.set_time(*[g64.time()[k] for k in ("t", "dt", "cfl_mult")])
    la = a.evolve_once()
    lb = b.evolve_once()
    # the same number of force evaluations on both sides (this step reruns its corrector: the
    # oracle takes 3 passes on the third Sedov 24^3 step), so the accelerations compared come
    # from the same sequence of corrector passes
    say(f"[mixed] corrector passes fp64={la['corrector_iterations']} "
        f"mixed={lb['corrector_iterations']}")
    assert la["corrector_iterations"] == lb["corrector_iterations"] >= 2
    A = np.stack([a.get(k) for k in ("ax", "ay", "az")])
    B = np.stack([b.get(k) for k in ("ax", "ay", "az")])
    err_a = np.linalg.norm(A - B) / np.linalg.norm(A)
    err_u = np.linalg.norm(a.get("du") - b.get("du")) / np.linalg.norm(a.get("du"))
    say(f"[mixed] rel-L2 a={err_a:.3e} du={err_u:.3e}")
    assert err_a <= 1e-5 and err_u <= 1e-5


@pytest.mark.parametrize("mode", ["fp64", "mixed", "cd10"])
def test_record_refresh_equals_full_pack(mode, monkeypatch):
    """Corrector reruns refresh v, u, P and c_s of the force records in place instead of
    rebuilding them.  SHAMX_FULL_PACK=1 forces the full rebuild: on Sedov 20^3, whose second
    and third steps rerun their corrector (2 and 3 passes in the oracle), the records of the
    last force evaluation and every field after each step must be the same bytes either way
    (FP64, mixed precision, and the varying viscosity, whose alpha the records carry too)."""
    cfg, f = setups.sedov(20)
    if mode == "mixed":
        cfg = dict(cfg, mixed_precision=1)
    if mode == "cd10":
        cfg = dict(cfg, av_kind=1, alpha_min=0.0, alpha_max=1.0, sigma_decay=0.1)
    runs = {}
    for full in (False, True):
        if full:
            monkeypatch.setenv("SHAMX_FULL_PACK", "1")
        else:
            monkeypatch.delenv("SHAMX_FULL_PACK", raising=False)
        g = GpuSolver(cfg, f)
        out = []
        for step in range(3):
            say(f"[refresh] {mode} full_pack={full} step {step}")
            lg = g.evolve_once()
            nm = g.sizes()["Nm"]
            rb = (64 if mode == "mixed" else 144) * nm
            out.append((lg["corrector_iterations"], g.get_u8(2, rb),
                        {k: g.get(k) for k in ("vx", "vy", "vz", "ax", "ay", "az", "u", "du",
                                               "h", "alpha_AV")}))
        g.close()
        runs[full] = out
    assert max(c for c, _, _ in runs[False]) >= 2, "no step reran its corrector"
    for (ca, ra, fa), (cb, rb_, fb) in zip(runs[False], runs[True]):
        assert ca == cb
        np.testing.assert_array_equal(ra, rb_)
        for k in fa:
            np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)


@pytest.mark.parametrize("lanes", ["1", "2", "16"])
def test_pair_lanes_same_bits(lanes, monkeypatch):
    """SHAMX_PAIR_LANES picks how many lanes share a particle's row in the h iteration, the
    Omega pass and the force loop; the sums take their terms in list order whatever the
    choice.  Three Sedov 20^3 steps (corrector reruns included) at the default four lanes and
    at `lanes` must leave every field the same bits."""
    cfg, f = setups.sedov(20)
    runs = {}
    for g_lanes in ("4", lanes):
        monkeypatch.setenv("SHAMX_PAIR_LANES", g_lanes)
        g = GpuSolver(cfg, f)
        out = []
        for step in range(3):
            say(f"[pair lanes] {g_lanes} step {step}")
            lg = g.evolve_once()
            out.append((lg["corrector_iterations"], lg["h_iterations"],
                        {k: g.get(k) for k in ("x", "vx", "vy", "vz", "ax", "ay", "az", "u",
                                               "du", "h")}))
        g.close()
        runs[g_lanes] = out
    monkeypatch.delenv("SHAMX_PAIR_LANES", raising=False)
    for (ca, ha, fa), (cb, hb, fb) in zip(runs["4"], runs[lanes]):
        assert (ca, ha) == (cb, hb)
        for k in fa:
            np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)
