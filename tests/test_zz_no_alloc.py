"""shamx_sort_noalloc (no reference counterpart: the sort for callers that cannot allocate)
against shamx_sort through the C ABI: two calls in a row, the second on fewer keys, must give
the same sorted keys and permutation; sizes beyond the static buffer are rejected."""
import ctypes as C
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from shamrock_amd._lib import check, lib  # noqa: E402

DEV = "cuda:0"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    lib()


def _keys(n, bits, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    k[: n // 4] = k[n // 4: 2 * (n // 4)]  # equal keys: the order must be the stable one
    return k


def _sort_ref(ctx, keys, bits):
    k = torch.from_numpy(keys.view(np.int32).copy()).to(DEV)
    p = torch.zeros(keys.size, dtype=torch.int32, device=DEV)
    check(lib().shamx_sort(ctx, torch.cuda.current_stream().cuda_stream, keys.size, bits,
                           k.data_ptr(), p.data_ptr()))
    torch.cuda.synchronize()
    return k.cpu().numpy().view(np.uint32), p.cpu().numpy().view(np.uint32)


def _sort_noalloc(keys, bits):
    k = torch.from_numpy(keys.view(np.int32).copy()).to(DEV)
    p = torch.full((keys.size,), -1, dtype=torch.int32, device=DEV)
    check(lib().shamx_sort_noalloc(torch.cuda.current_stream().cuda_stream, keys.size, bits,
                                   k.data_ptr(), p.data_ptr()))
    torch.cuda.synchronize()
    return k.cpu().numpy().view(np.uint32), p.cpu().numpy().view(np.uint32)


def test_sort_noalloc_matches_sort_twice():
    ctx = C.c_void_p()
    check(lib().shamx_ctx_create(C.byref(ctx)))
    try:
        for n, bits, seed in ((3_000_001, 30, 1), (70_001, 30, 2), (4_194_304, 32, 3)):
            say(f"[noalloc] n={n} bits={bits}")
            keys = _keys(n, bits, seed)
            ka, pa = _sort_ref(ctx, keys, bits)
            kb, pb = _sort_noalloc(keys, bits)
            np.testing.assert_array_equal(kb, ka)
            np.testing.assert_array_equal(pb, pa)
            np.testing.assert_array_equal(pa, np.argsort(keys, kind="stable"))
    finally:
        lib().shamx_ctx_destroy(ctx)


def test_sort_noalloc_rejects_oversize():
    k = torch.zeros(8, dtype=torch.int32, device=DEV)
    rc = lib().shamx_sort_noalloc(torch.cuda.current_stream().cuda_stream, 4_194_305, 30,
                                  k.data_ptr(), k.data_ptr())
    assert rc == 1 and b"4194304" in lib().shamx_last_error()
