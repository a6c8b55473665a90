"""The allocation-free sort (ysim_hip_sort_pairs_noalloc): the same permutation
as ysim_hip_sort_pairs, called through the C ABI on two inputs in a row, the
second smaller (the static scratch keeps the first call's contents), and a
refusal, before any launch, of an input larger than its scratch."""
import ctypes as C

import numpy as np
import pytest

from ysim_amd._lib import ERR_CAPACITY, YsimError, call, load

pytestmark = pytest.mark.gpu


def sort_with(ctx, fn, keys, vals, key_bits):
    dk = ctx.to_device(keys)
    dv = ctx.to_device(vals)
    call(fn, ctx.h, dk.ptr, dv.ptr, len(keys), key_bits)
    return dk.get(), dv.get()


@pytest.mark.parametrize("key_bits", [63, 30])
def test_noalloc_sort_matches_sort_pairs(ctx, key_bits):
    rng = np.random.default_rng(17 + key_bits)
    kd = np.uint64 if key_bits > 32 else np.uint32
    for n in (522_242, 77_777):  # the 512 x 512 fold's triangle count, then a smaller input
        keys = rng.integers(0, 1 << min(key_bits, 20), n, dtype=np.uint64)  # many equal keys
        keys = (keys << np.uint64(key_bits - min(key_bits, 20))).astype(kd)
        vals = rng.permutation(n).astype(np.uint32)
        k1, v1 = sort_with(ctx, "ysim_hip_sort_pairs", keys, vals, key_bits)
        k2, v2 = sort_with(ctx, "ysim_hip_sort_pairs_noalloc", keys, vals, key_bits)
        assert np.array_equal(k1, k2) and np.array_equal(v1, v2), n
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k2, keys[order]) and np.array_equal(v2, vals[order]), n


def test_noalloc_sort_refuses_oversize(ctx):
    load()
    n = 1 << 22  # 4 M 64-bit keys need ~80 MB of scratch
    with pytest.raises(YsimError) as e:
        call("ysim_hip_sort_pairs_noalloc", ctx.h, C.c_void_p(16), C.c_void_p(16), n, 64)
    assert e.value.code == ERR_CAPACITY
    assert "static scratch" in str(e.value)
