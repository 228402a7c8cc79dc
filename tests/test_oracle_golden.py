"""Oracle pinned against the golden vectors of the reference's own unit tests.

Each fixture under tests/golden/ is transcribed from the reference test named in its "source"
field; the oracle (oracle/sph_oracle.hpp) must reproduce it exactly (tolerance 1e-6 for the
smoothing-length sequences, as the reference test itself uses).
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    with open(os.path.join(GOLD, name + ".json")) as f:
        return json.load(f)


def bb6(fx):
    return np.array(fx["bbox"][0] + fx["bbox"][1], dtype=np.float64)


def u64(a):
    return np.array([int(v) for v in a], dtype=np.uint64)


def test_morton_code_set_u64():
    fx = load("morton_code_set")
    codes = orc.morton(fx["positions"], bb6(fx), fx["morton_count"], bits=64)
    np.testing.assert_array_equal(codes, u64(fx["codes"]))


def test_morton_sorted_set_bitonic_tie_order():
    fx = load("morton_code_set")
    codes = orc.morton(fx["positions"], bb6(fx), fx["morton_count"], bits=64)
    keys, idx = orc.bitonic(codes)
    np.testing.assert_array_equal(keys, u64(fx["sorted"]))
    np.testing.assert_array_equal(idx, np.array(fx["sort_index"], np.uint32))


@pytest.mark.parametrize("name", ["morton_reduced_set", "morton_reduced_set_single_cell"])
def test_morton_reduced_set(name):
    fx = load(name)
    n = len(fx["positions"])
    codes = orc.morton(fx["positions"], bb6(fx), fx["morton_count"], bits=64)
    if "codes" in fx:
        np.testing.assert_array_equal(codes, u64(fx["codes"]))
    keys, idx = orc.bitonic(codes)
    if "sorted" in fx:
        np.testing.assert_array_equal(keys, u64(fx["sorted"]))
    np.testing.assert_array_equal(idx, np.array(fx["sort_index"], np.uint32))
    rmap = orc.reduce(keys, n, fx["level"])
    np.testing.assert_array_equal(rmap, np.array(fx["reduc_index_map"], np.uint32))
    L = rmap.size - 2
    np.testing.assert_array_equal(keys[rmap[:L]], u64(fx["reduced_codes"]))


def test_karras_u64():
    fx = load("karras_u64")
    k = orc.karras(u64(fx["codes"]))
    for key in ("lchild", "rchild", "lflag", "rflag", "endrange"):
        np.testing.assert_array_equal(k[key], np.array(fx[key]), err_msg=key)


def test_karras_one_cell():
    k = orc.karras(u64([0b100100]))
    assert all(v.size == 0 for v in k.values())


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
def test_karras_legacy(dtype):
    fx = load("karras_legacy_u32_u64")
    k = orc.karras(np.array(fx["codes"], dtype=dtype))
    for key in ("lchild", "rchild", "lflag", "rflag", "endrange"):
        np.testing.assert_array_equal(k[key], np.array(fx[key]), err_msg=key)


@pytest.mark.parametrize("name", ["karras_tree_aabb", "karras_tree_aabb_one_cell"])
def test_tree_aabb(name):
    fx = load(name)
    t = orc.Tree(fx["positions"], bb6(fx), fx["level"], bits=64, morton_count=fx["morton_count"])
    np.testing.assert_array_equal(t.codes, u64(fx["codes"]))
    np.testing.assert_array_equal(t.sort_index, np.array(fx["sort_index"], np.uint32))
    np.testing.assert_array_equal(t.reduc_map, np.array(fx["reduc_index_map"], np.uint32))
    np.testing.assert_array_equal(t.leaf_codes, u64(fx["reduced_codes"]))
    for key in ("lchild", "rchild", "lflag", "rflag", "endrange"):
        if key in fx:
            np.testing.assert_array_equal(getattr(t, key), np.array(fx[key]), err_msg=key)
    np.testing.assert_array_equal(t.aabb_min, np.array(fx["aabb_min"]))
    np.testing.assert_array_equal(t.aabb_max, np.array(fx["aabb_max"]))


def test_tree_field_max():
    fx = load("karras_tree_field")
    t = orc.Tree(fx["positions"], bb6(fx), fx["level"], bits=64)
    np.testing.assert_array_equal(t.sort_index, np.array(fx["sort_index"], np.uint32))
    np.testing.assert_array_equal(t.field_max(np.array(fx["field"])), np.array(fx["expected"],
                                                                                float))


@pytest.mark.parametrize("mode", ["all", "box"])
def test_clbvh_object_iterator(mode):
    fx = load("clbvh_object_iterator")
    t = orc.Tree(fx["positions"], bb6(fx), fx["level"], bits=64)
    assert t.I == fx["internal_cell_count"]
    counts, neigh = t.box_query(fx[mode]["s"])
    np.testing.assert_array_equal(counts, np.array(fx[mode]["counts"], np.uint32))
    np.testing.assert_array_equal(neigh, np.array(fx[mode]["neigh"], np.uint32))


def test_iterate_smoothing_length_density_sequences():
    """IterateSmoothingLengthDensityTests.cpp:25-248 harness replayed on the oracle."""
    fx = load("iterate_smoothing_length_density")
    pos = np.array([[i, j, k] for i in range(4) for j in range(4) for k in range(4)], float)
    n = pos.shape[0]
    cnt = np.full(n, n, np.uint32)
    off = (np.arange(n) * n).astype(np.uint32)
    ids = np.tile(np.arange(n, dtype=np.uint32), n)
    new_h = np.full(n, fx["h0"])
    eps = np.full(n, fx["eps0"])
    seq = dict(emin=[], emax=[], hmin=[], hmax=[])
    for outer in range(50):
        old_h = new_h.copy()
        eps[:] = fx["eps0"]
        max_eps = fx["eps0"]
        for inner in range(10):
            orc.iterate_h(fx["kernel"], cnt, off, ids, pos, old_h, new_h, eps, fx["gpart_mass"],
                          fx["h_evol_max"], fx["h_evol_iter_max"])
            seq["emin"].append(eps.min())
            seq["emax"].append(eps.max())
            seq["hmin"].append(new_h.min())
            seq["hmax"].append(new_h.max())
            assert np.all((eps >= 0) | (eps == -1))
            assert np.all((new_h > 0) & (new_h <= 10))
            max_eps = eps.max()
            if max_eps < 1e-6:
                break
        if eps.min() == -1:
            continue
        if max_eps < 1e-6:
            break
        assert outer != 49, "h iterator did not converge"
    tol = fx["tol"]
    np.testing.assert_allclose(new_h, fx["expected_h_end"], rtol=0, atol=tol)
    np.testing.assert_allclose(eps, fx["expected_eps_end"], rtol=0, atol=tol)
    for k, ref in (("emin", "seq_eps_min"), ("emax", "seq_eps_max"), ("hmin", "seq_h_min"),
                   ("hmax", "seq_h_max")):
        assert len(seq[k]) == len(fx[ref]), k
        np.testing.assert_allclose(seq[k], fx[ref], rtol=0, atol=tol, err_msg=k)
