"""Host-side tile packing (dips_pack_tiles): no GPU needed.

Properties checked on random record layouts (rows of 0..120 entries, batches
of uneven size): every batch's rows are covered exactly once by consecutive
tiles; a tile holds at most 64 slots (one per entry, one for a row without
entries) and 64 rows unless it is a single row with more than 64 entries;
packing is greedy (the next row would not have fitted); and the batch table
indexes the boundary runs.  A brute-force
re-packing in Python gives the same boundaries.
"""
import numpy as np
import pytest

CAP = 64


def _greedy(ptr, begin, rows):
    out = [begin]
    r, end = begin, begin + rows
    while r < end:
        ln = ptr[r + 1] - ptr[r]
        if ln > CAP:
            r += 1
            out.append(r)
            continue
        n = k = 0
        while r < end and k < CAP:
            ln = ptr[r + 1] - ptr[r]
            if ln > CAP or n + max(ln, 1) > CAP:
                break
            n += max(ln, 1)
            r += 1
            k += 1
        out.append(r)
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pack_tiles_properties(seed):
    from jax_dips.records import batch_bounds, pack_tiles

    rng = np.random.default_rng(seed)
    n = 5000
    lens = rng.choice([0, 1, 7, 7, 7, 13, 40, 64, 65, 120], size=n)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    B, begins, rows = batch_bounds(n, 1234)
    tile_row, btb = pack_tiles(ptr, n, begins, rows)
    assert btb[-1] == len(tile_row) and len(btb) == len(begins) + 1
    for b, (b0, nr) in enumerate(zip(begins, rows)):
        t = tile_row[btb[b]:btb[b + 1]].astype(np.int64)
        assert t[0] == b0 and t[-1] == b0 + nr and np.all(np.diff(t) > 0)
        assert list(t) == _greedy(ptr.astype(np.int64), int(b0), int(nr))
        for r0, r1 in zip(t[:-1], t[1:]):
            ne = ptr[r1] - ptr[r0]
            if r1 - r0 == 1 and ne > CAP:
                continue
            slots = np.maximum(np.diff(ptr[r0:r1 + 1]), 1).sum()
            assert slots <= CAP and r1 - r0 <= CAP


def test_pack_tiles_all_empty_rows_and_errors():
    from jax_dips import _native as N
    from jax_dips.records import pack_tiles

    ptr = np.zeros(201, np.int32)  # 200 rows without entries: at most 64 rows per tile
    tile_row, btb = pack_tiles(ptr, 200, [0], [200])
    assert list(tile_row) == [0, 64, 128, 192, 200]
    bad = np.array([0, 5, 3, 9], np.int32)  # decreasing offsets
    with pytest.raises(N.DipsError):
        pack_tiles(bad, 3, [0], [3])
    with pytest.raises(N.DipsError):
        pack_tiles(np.arange(11, dtype=np.int32), 10, [5], [6])  # batch beyond the records
