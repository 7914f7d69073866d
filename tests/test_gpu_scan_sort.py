"""The step's device-count exclusive scan and stable LSD radix sort
(frt_scan_sort.hip) against numpy: np.cumsum for the scan, a stable argsort
on the masked keys for the sort (equal keys keep their input order, which the
per-cell sums of the deposit passes and the ray-list order rely on).  Sizes
cover empty, single, sub-block, block-boundary, multi-round and multi-block
inputs; bit widths cover one digit pass, several and a short last pass."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 63, 64, 255, 256, 257, 2047, 2048, 2049, 100_003, 3_000_001]
BITS = [1, 5, 8, 12, 21, 25, 32]


@pytest.fixture(scope="module")
def rt():
    from flexrt_imaging_reion_amd import inputs
    from flexrt_imaging_reion_amd.rt import FlexRT

    wl = inputs.make_workload("uniform32", n=16)
    g = FlexRT(wl["cfg"])
    yield g
    g.close()


@pytest.mark.parametrize("n", SIZES)
def test_scan_matches_cumsum(rt, n):
    print("CASE scan n=%d" % n, flush=True)
    rng = np.random.default_rng(n)
    c = rng.integers(0, 7, size=n).astype(np.int32)
    if n > 10:
        c[rng.integers(0, n, size=n // 10)] = 0  # runs of zeros, as in segment counts
    _, _, s, tot = rt.selftest_scan_sort(counts=c)
    ref = np.concatenate([[0], np.cumsum(c, dtype=np.int64)[:-1]]) if n else np.zeros(0, np.int64)
    assert np.array_equal(s.astype(np.int64), ref)
    assert tot == int(c.astype(np.int64).sum())


@pytest.mark.parametrize("bits", BITS)
@pytest.mark.parametrize("n", SIZES)
def test_sort_is_stable_and_sorted(rt, n, bits):
    print("CASE sort n=%d bits=%d" % (n, bits), flush=True)
    rng = np.random.default_rng(1000 * n + bits)
    # few distinct keys (long runs of equal keys) and full-range keys
    hi = min(1 << bits, 50) if n % 2 else (1 << bits)
    k = rng.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)
    if bits < 32:
        k |= rng.integers(0, 1 << (32 - bits), size=n, dtype=np.uint64).astype(np.uint32) << np.uint32(bits)
    ko, perm, _, _ = rt.selftest_scan_sort(keys=k, bits=bits)
    mask = np.uint32(0xFFFFFFFF) if bits == 32 else np.uint32((1 << bits) - 1)
    ref = np.argsort(k & mask, kind="stable")
    assert np.array_equal(perm, ref.astype(np.int32))
    assert np.array_equal(ko, k[ref])


@pytest.mark.parametrize("bits", [12, 21, 25, 32])
def test_sort_full_range_keys_many_blocks(rt, bits):
    """Every digit pass with keys spread over the whole digit range and a
    grid of many blocks: each pass's per-block counts must come from that
    pass's input order."""
    print("CASE sort full-range n=2500000 bits=%d" % bits, flush=True)
    rng = np.random.default_rng(bits)
    n = 2_500_000
    k = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    ko, perm, _, _ = rt.selftest_scan_sort(keys=k, bits=bits)
    mask = np.uint32(0xFFFFFFFF) if bits == 32 else np.uint32((1 << bits) - 1)
    ref = np.argsort(k & mask, kind="stable")
    assert np.array_equal(perm, ref.astype(np.int32))


def test_sort_of_tile_keys_in_list_order(rt):
    """Keys as the ray sort sees them: mostly ascending with local disorder."""
    rng = np.random.default_rng(7)
    n = 1_000_000
    k = (np.arange(n) // 300 + rng.integers(-2, 3, size=n)).clip(0, 4095).astype(np.uint32)
    ko, perm, _, _ = rt.selftest_scan_sort(keys=k, bits=12)
    ref = np.argsort(k, kind="stable")
    assert np.array_equal(perm, ref.astype(np.int32))


def _cells(kind, count, n3, rng):
    if kind == "uniform":
        return rng.integers(0, n3, size=count).astype(np.uint32)
    if kind == "few":  # long runs per cell: many LDS rounds per group of 64 cells
        return rng.choice(rng.integers(0, n3, size=7), size=count).astype(np.uint32)
    # as the segments of a tile-ordered ray list: ascending with local disorder
    c = np.arange(count) * n3 // max(count, 1) + rng.integers(-70, 71, size=count)
    return (c % n3).astype(np.uint32)


@pytest.mark.parametrize("n,count,kind", [
    (16, 0, "uniform"), (16, 1, "uniform"), (16, 1000, "uniform"), (16, 300_000, "uniform"),
    (16, 200_000, "few"), (16, 500_000, "listorder"), (12, 100_000, "uniform"), (5, 5_000, "uniform"),
    (5, 3_000, "few"), (4, 3_000, "uniform"), (3, 700, "listorder"), (2, 50, "uniform"),
    (36, 2_000_000, "listorder"),
])
def test_cell_sort_matches_stable_argsort(n, count, kind):
    """The step's cell sort (radix passes over the groups of 64 cells, then
    one block per group): every item's slot is its position in the stable
    sort by cell, and each cell's range is [first, last + 1) of its items
    (0, 0 when it has none).  Grids cover no radix pass (n^3 <= 64), a
    partial last group (n = 5, 12: n^3 not a multiple of 64), one and two
    passes."""
    from flexrt_imaging_reion_amd import inputs
    from flexrt_imaging_reion_amd.rt import FlexRT

    print("CASE cell sort n=%d count=%d %s" % (n, count, kind), flush=True)
    n3 = n ** 3
    rng = np.random.default_rng(n * 7919 + count)
    cells = _cells(kind, count, n3, rng)
    wl = inputs.make_workload("uniform32", n=n)
    with FlexRT(wl["cfg"]) as g:
        slot, cs, ce = g.selftest_sort_cells(cells)
    order = np.argsort(cells, kind="stable")
    ref = np.empty(count, dtype=np.int64)
    ref[order] = np.arange(count)
    assert np.array_equal(slot.astype(np.int64), ref)
    cnt = np.bincount(cells, minlength=n3).astype(np.int64)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    ref_s = np.where(cnt > 0, start, 0)
    ref_e = np.where(cnt > 0, start + cnt, 0)
    assert np.array_equal(cs.astype(np.int64), ref_s)
    assert np.array_equal(ce.astype(np.int64), ref_e)
