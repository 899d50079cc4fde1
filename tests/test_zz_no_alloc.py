"""gsr_sort_pairs_noalloc (fixed per-GPU scratch, no device allocation)
against gsr_sort_pairs, through the C ABI: two calls in a row on the same
GPU, the second with fewer keys (the fixed scratch keeps the first call's
contents, so whatever must start at zero has to be cleared by the call), for
32- and 64-bit keys; sizes beyond the scratch are refused before any launch."""
import sys

import numpy as np
import pytest

from conftest import empty_dev, to_dev, to_host
from gsr import lib

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _keys(rng, n, bits):
    hi = (1 << bits) - 1 if bits < 64 else np.iinfo(np.uint64).max
    keys = rng.integers(0, hi, size=n, dtype=np.uint64, endpoint=True)
    if n > 10:
        keys[: n // 3] = keys[rng.integers(0, 16, n // 3)]    # ties
    return keys


def _sort(torch, ctx, keys, vals, key_bytes, bits, noalloc):
    n = keys.size
    kdt = np.uint32 if key_bytes == 4 else np.uint64
    k0 = to_dev(torch, keys.astype(kdt))
    v0 = to_dev(torch, vals)
    k1 = empty_dev(torch, n, kdt)
    v1 = empty_dev(torch, n, np.uint32)
    torch.cuda.synchronize()
    if noalloc:
        idx = ctx.sort_pairs_noalloc(key_bytes, k0, v0, k1, v1, n, 0, bits)
    else:
        idx = ctx.sort_pairs(key_bytes, k0, v0, k1, v1, n, 0, bits)
    return (to_host(k1 if idx else k0, kdt)[:n].astype(np.uint64), to_host(v1 if idx else v0, np.uint32)[:n])


@pytest.mark.parametrize("key_bytes,bits", [(4, 29), (4, 32), (8, 45), (8, 64)])
def test_noalloc_sort_equals_sort_pairs(torch_cuda, ctx, key_bytes, bits):
    rng = np.random.default_rng(bits + key_bytes)
    for n in (3_000_000, 700_001):          # second call smaller than the first
        _say(f"noalloc sort key_bytes={key_bytes} bits={bits} n={n}")
        keys = _keys(rng, n, bits)
        vals = rng.permutation(n).astype(np.uint32)
        k_ref, v_ref = _sort(torch_cuda, ctx, keys, vals, key_bytes, bits, noalloc=False)
        k, v = _sort(torch_cuda, ctx, keys, vals, key_bytes, bits, noalloc=True)
        assert np.array_equal(k, k_ref) and np.array_equal(v, v_ref), n
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(v, vals[order])


def test_noalloc_sort_small_and_empty(torch_cuda, ctx):
    rng = np.random.default_rng(3)
    for n in (70_000, 1, 0):
        keys = _keys(rng, n, 32) if n else np.zeros(1, np.uint64)
        vals = np.arange(max(n, 1), dtype=np.uint32)
        kdt = np.uint32
        k0, v0 = to_dev(torch_cuda, keys.astype(kdt)), to_dev(torch_cuda, vals)
        k1, v1 = empty_dev(torch_cuda, max(n, 1), kdt), empty_dev(torch_cuda, max(n, 1), np.uint32)
        torch_cuda.cuda.synchronize()
        idx = ctx.sort_pairs_noalloc(4, k0, v0, k1, v1, n, 0, 32)
        if n:
            v = to_host(v1 if idx else v0, np.uint32)[:n]
            assert np.array_equal(v, vals[np.argsort(keys[:n], kind="stable")])


def test_noalloc_sort_refuses_sizes_beyond_its_scratch(torch_cuda, ctx):
    d = empty_dev(torch_cuda, 16, np.uint64)
    for key_bytes, limit in ((4, 133_955_584), (8, 92_094_464)):
        with pytest.raises(lib.GsrError) as e:
            ctx.sort_pairs_noalloc(key_bytes, d, d, d, d, limit + 1, 0, 8 * key_bytes)
        assert e.value.code == lib.GSR_ERR_ARGUMENT
        assert f"at most {limit} keys" in str(e.value)
    with pytest.raises(lib.GsrError) as e:
        ctx.sort_pairs_noalloc(4, d, d, d, d, 10, 0, 33)
    assert e.value.code == lib.GSR_ERR_ARGUMENT
