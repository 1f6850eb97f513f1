"""The allocation-free prefix sum (dio_prefix_sum_encode_static, no counterpart in
the reference) against the allocating PrefixSum through the C ABI: the same scan and
total, bit for bit, on calls in a row that share the static workspace (a large one
past the self-clearing size, then smaller ones), with a device count, and the size
limit refused on the host."""
import ctypes

import numpy as np
import pytest

import diorama_amd
from diorama_amd import scene
from tests.gpu_util import context, say

pytestmark = pytest.mark.gpu


def allocating_scan(ctx, values, capacity, count=None):
    data = ctx.upload(values)
    scan = diorama_amd.PrefixSum(ctx, data, capacity, count)
    scan.encode()
    out = data.download(np.uint32, len(values))
    total = int(scan.total.download(np.uint32, 1)[0])
    scan.destroy()
    data.destroy()
    return out, total


def test_static_scan_matches_the_allocating_one_on_consecutive_calls():
    ctx = context()
    lib = diorama_amd.lib()
    total = ctx.buffer(16)
    # 5,000,003 values: 1,221 tiles, past the self-clearing 1,024 (a clearing launch
    # follows the kernel); then 70,001 and 4,096 (one exact tile), self-cleared.
    for k, n in enumerate([5_000_003, 70_001, 4_096]):
        values = scene.tile_touch_counts(n, 30 + k)
        data = ctx.upload(values)
        say("static scan", n)
        rc = lib.dio_prefix_sum_encode_static(ctx.handle, data.handle, n, None, total.handle)
        assert rc == diorama_amd.DIO_OK, lib.dio_last_error()
        got = data.download(np.uint32, n)
        got_total = int(total.download(np.uint32, 1)[0])
        want, want_total = allocating_scan(ctx, values, n)
        assert np.array_equal(got, want)
        assert got_total == want_total == int(values.astype(np.uint64).sum()) & 0xFFFFFFFF
        exclusive = np.concatenate([[0], np.cumsum(values[:-1], dtype=np.uint64)]).astype(np.uint32)
        assert np.array_equal(got, exclusive)
        data.destroy()
    total.destroy()


def test_static_scan_with_a_device_count_and_the_size_limit():
    ctx = context()
    lib = diorama_amd.lib()
    n = 300_000
    values = scene.tile_touch_counts(n, 41)
    count = ctx.upload(np.array([123_457], dtype=np.uint32))
    data = ctx.upload(values)
    total = ctx.buffer(16)
    say("static scan with count")
    diorama_amd.PrefixSum.encode_static(ctx, data, n, total, count)
    got = data.download(np.uint32, n)
    want, want_total = allocating_scan(ctx, values, n, count)
    assert np.array_equal(got, want)
    assert int(total.download(np.uint32, 1)[0]) == want_total
    assert np.array_equal(got[123_457:], values[123_457:])  # past the count: untouched
    # beyond the static workspace: refused before any launch
    big = (1 << 26) + 1
    rc = lib.dio_prefix_sum_encode_static(ctx.handle, data.handle, big, None, total.handle)
    assert rc == diorama_amd.DIO_ERR_INVALID
    assert b"at most 67108864 values" in lib.dio_last_error()
    for b in (count, data, total):
        b.destroy()
