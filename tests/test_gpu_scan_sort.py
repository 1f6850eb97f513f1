"""PrefixSum and RadixSort on the GPU against closed forms: bit-exact.

Cases follow src/checks/depthSort.ts:46-104 (300,000 values in [0, 9] from
mulberry32(3); sort sizes 1, 1000, 1024, 100,003 with half full-range and half
small keys) plus the microbenchmark inputs: geometric tile-touch counts (mean 6,
capped at 256) for the scan and 64-bit (Zipf tile, depth bits) keys for the sort.
"""
import numpy as np
import pytest

import diorama_amd
from diorama_amd import scene
from tests.gpu_util import context, say

pytestmark = pytest.mark.gpu


def exclusive(values):
    out = np.zeros_like(values, dtype=np.uint64)
    if values.size:
        out[1:] = np.cumsum(values[:-1].astype(np.uint64))
    return (out & 0xFFFFFFFF).astype(np.uint32)


def run_scan(values, capacity=None, count=None):
    ctx = context()
    capacity = values.size if capacity is None else capacity
    data = ctx.buffer(max(capacity, 1) * 4)
    if values.size:
        data.upload(values)
    count_buf = None
    if count is not None:
        count_buf = ctx.upload(np.array([count, 0, 0, 0], dtype=np.uint32))
    scan = diorama_amd.PrefixSum(ctx, data, capacity, count_buf)
    scan.encode()
    out = data.download(np.uint32, max(capacity, 1))[:capacity]
    total = int(scan.total.download(np.uint32, 1)[0])
    scan.destroy()
    return out, total


def geometric_counts(n, seed):
    return scene.tile_touch_counts(n, seed)


@pytest.mark.parametrize("n", [0, 1, 511, 512, 4095, 4096, 4097, 100_003])
def test_scan_sizes(n):
    say("scan", n)
    values = np.random.default_rng(n).integers(0, 1000, size=n).astype(np.uint32)
    out, total = run_scan(values)
    assert np.array_equal(out, exclusive(values))
    assert total == int(values.astype(np.uint64).sum()) & 0xFFFFFFFF


def test_scan_depth_sort_check_case():
    random = scene.mulberry32(3)
    values = np.array([int(random() * 10) for _ in range(300_000)], dtype=np.uint32)
    say("scan", "mulberry32(3) 300000")
    out, total = run_scan(values)
    assert np.array_equal(out, exclusive(values))
    assert total == int(values.sum())


@pytest.mark.parametrize("n", [1 << 20, 16 << 20])
def test_scan_geometric_microbench_inputs(n):
    values = geometric_counts(n, 5)
    say("scan geometric", n)
    out, total = run_scan(values)
    assert np.array_equal(out, exclusive(values))
    assert total == int(values.astype(np.uint64).sum())


def test_scan_count_from_device():
    """Only min(count, capacity) values are scanned; the rest stay untouched."""
    values = geometric_counts(50_000, 6)
    say("scan device count")
    out, total = run_scan(values, capacity=50_000, count=30_001)
    assert np.array_equal(out[:30_001], exclusive(values[:30_001]))
    assert np.array_equal(out[30_001:], values[30_001:])
    assert total == int(values[:30_001].sum())


def run_sort(keys, values, high_bits=0, capacity=None, count=None):
    ctx = context()
    n = values.size
    capacity = n if capacity is None else capacity
    count_buf = None if count is None else ctx.upload(np.array([count, 0, 0, 0], dtype=np.uint32))
    sort = diorama_amd.RadixSort(ctx, capacity, high_bits, count_buf)
    sort.keys.upload(keys)
    sort.values.upload(values)
    sort.encode()
    width = 2 if high_bits else 1
    k = sort.keys.download(np.uint32, capacity * width)
    v = sort.values.download(np.uint32, capacity)
    sort.destroy()
    return k, v


@pytest.mark.parametrize("n", [1, 1000, 1024, 100_003, 1_100_000])
def test_sort_32bit_depth_sort_check_case(n):
    random = scene.mulberry32(n)
    keys = np.array([int(random() * 2 ** 32) if i % 2 == 0 else int(random() * 16) for i in range(n)],
                    dtype=np.uint32)
    values = np.arange(n, dtype=np.uint32)
    say("sort32", n)
    k, v = run_sort(keys, values)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(v, order.astype(np.uint32))
    assert np.array_equal(k, keys[order])


def microbench_keys(n, seed):
    return scene.tile_depth_keys(n, seed)


@pytest.mark.parametrize("n", [100_003, 8 << 20, 30_000_000])
def test_sort_64bit_microbench_keys(n):
    keys, lo, hi = microbench_keys(n, 7)
    values = np.random.default_rng(8).permutation(n).astype(np.uint32)  # arbitrary payloads
    say("sort64", n)
    k, v = run_sort(keys, values, high_bits=16)
    order = np.lexsort((lo, hi & 0xFFFF))  # stable: ties keep input order
    assert np.array_equal(v, values[order])
    assert np.array_equal(k[0::2], lo[order]) and np.array_equal(k[1::2], hi[order])


def test_sort_64bit_ignores_high_word_bits_past_high_bits():
    """Only the low 16 bits of the high word are looked at (src/radixSort.ts:103-104)."""
    n = 20_000
    rng = np.random.default_rng(3)
    lo = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    hi = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    keys = np.empty(2 * n, dtype=np.uint32)
    keys[0::2], keys[1::2] = lo, hi
    values = np.arange(n, dtype=np.uint32)
    say("sort64 high bits")
    k, v = run_sort(keys, values, high_bits=16)
    assert np.array_equal(v, np.lexsort((lo, hi & 0xFFFF)).astype(np.uint32))


def test_sort_count_from_device():
    n, count = 30_000, 12_345
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    values = np.arange(n, dtype=np.uint32)
    say("sort device count")
    k, v = run_sort(keys, values, count=count)
    order = np.argsort(keys[:count], kind="stable")
    assert np.array_equal(v[:count], order.astype(np.uint32))
    assert np.array_equal(k[:count], keys[:count][order])


def test_buffer_copy_and_timestamps():
    """copyBufferToBuffer with offsets, and timestamp queries around it."""
    ctx = context()
    src = ctx.upload(np.arange(1000, dtype=np.uint32))
    dst = ctx.upload(np.zeros(1000, dtype=np.uint32))
    say("buffer copy")
    ctx.timestamp(0)
    dst.copy_from(src, nbytes=400, dst_offset=40, src_offset=800)
    ctx.timestamp(1)
    ms = ctx.elapsed_ms(0, 1)
    out = dst.download(np.uint32)
    assert np.array_equal(out[10:110], np.arange(200, 300))
    assert not out[:10].any() and not out[110:].any()
    assert 0 <= ms < 1000
    with pytest.raises(diorama_amd.DioramaError):
        dst.copy_from(src, nbytes=4000, dst_offset=4)


def test_back_to_back_calls_reuse_their_workspace():
    """A scan or sort leaves its workspace cleared for the next call (no clearing
    launch precedes a call): 12 calls in a row on one PrefixSum and one RadixSort,
    each with its own device count, every result checked."""
    ctx = context()
    rng = np.random.default_rng(11)
    capacity = 300_000
    data = ctx.buffer(capacity * 4)
    count_buf = ctx.buffer(16)
    scan = diorama_amd.PrefixSum(ctx, data, capacity, count_buf)
    sort = diorama_amd.RadixSort(ctx, capacity, 16, count_buf)
    counts = [capacity, 1, 4097, 123_457, 0, 299_999, 4096, 70_001, capacity, 2, 150_000, 9]
    say("back-to-back scans and sorts")
    for n in counts:
        count_buf.upload(np.array([n, 0, 0, 0], dtype=np.uint32))
        values = rng.integers(0, 300, size=capacity).astype(np.uint32)
        data.upload(values)
        scan.encode()
        out = data.download(np.uint32, capacity)
        assert np.array_equal(out[:n], exclusive(values[:n])), n
        assert int(scan.total.download(np.uint32, 1)[0]) == int(values[:n].sum()), n
        low = rng.integers(0, 1 << 32, size=capacity, dtype=np.uint64).astype(np.uint32)
        high = rng.integers(0, 1 << 12, size=capacity).astype(np.uint32)
        keys = np.stack([low, high], axis=1).reshape(-1)
        idx = np.arange(capacity, dtype=np.uint32)
        sort.keys.upload(keys)
        sort.values.upload(idx)
        sort.encode()
        k = sort.keys.download(np.uint32, 2 * capacity).reshape(-1, 2)
        v = sort.values.download(np.uint32, capacity)
        full = (high[:n].astype(np.uint64) << 32) | low[:n]
        want = np.argsort(full, kind="stable").astype(np.uint32)
        assert np.array_equal(v[:n], want), n
        assert np.array_equal(k[:n, 0], low[want]) and np.array_equal(k[:n, 1], high[want]), n
    scan.destroy()
    sort.destroy()


def test_back_to_back_large_scans_reuse_their_workspace():
    """Past 1024 tiles the scan's workspace is cleared by a launch after the kernel
    instead of by its last workgroup; the next call still starts from a clean one."""
    ctx = context()
    rng = np.random.default_rng(12)
    capacity = 5_000_000
    data = ctx.buffer(capacity * 4)
    count_buf = ctx.buffer(16)
    scan = diorama_amd.PrefixSum(ctx, data, capacity, count_buf)
    say("back-to-back large scans")
    for n in [capacity, 1000, 4_200_001, 0, capacity]:
        count_buf.upload(np.array([n, 0, 0, 0], dtype=np.uint32))
        values = rng.integers(0, 100, size=capacity).astype(np.uint32)
        data.upload(values)
        scan.encode()
        out = data.download(np.uint32, capacity)
        assert np.array_equal(out[:n], exclusive(values[:n])), n
        assert int(scan.total.download(np.uint32, 1)[0]) == int(values[:n].astype(np.uint64).sum()), n
    scan.destroy()


def test_buffers_and_objects_outlive_their_context():
    """Buffers and objects destroyed after their context was closed (as a garbage
    collector may do at exit) free cleanly; the context goes with the last of them."""
    ctx = diorama_amd.Context(0)
    b = ctx.buffer(1024)
    b.upload(np.arange(256, dtype=np.uint32))
    r = diorama_amd.GaussianRenderer(ctx)
    s = diorama_amd.PrefixSum(ctx, b, 256, None)
    s.encode()
    ctx.close()
    s.destroy()
    r.destroy()
    b.destroy()
