"""The LDS-staged tiles' layout (sov_hip_tile_layout, the host builder the
device path uploads), checked on the CPU against the CSR it is built from:
every entry stored exactly once with its row, column and source position
decodable from (group, unit, key); units that cover the slices in order
(a staged unit is one slice with at least 800 entries per group, a gather
unit a run of sparser slices whose columns fit the key's 19 column bits);
staged units alternating between the two LDS buffers, slices of 64 KiB
(FP32) or 48 KiB (FP64) unless the row sums need the room; each (group, unit) tile split
into 12 row-aligned sub-ranges (one per entry wave), each a multiple of the
entries per lane (4) with padding only at its end; rows dealt to groups in 64-row
chunks, two workgroups per group; the pass list the
kernels read (every sub-range cut into passes of at most 64 lanes' entries, padded to an even count for FP32,
listed per group and wave, unit by unit); and the product the layout
encodes equal to the CSR product.  Cases cover empty rows, a row
denser than one wave pass, several slices (ragged last slice), and more
row groups than half the compute units when a group would exceed 8128
rows."""
import ctypes as C

import numpy as np
import pytest

from sovereign_amd.device import lib

SENT = 0xFFF80000          # row 8191, column 0
COLBITS = 19


def layout(ptr, idx, src, rows, cols, cus, tsize):
    L = lib()
    shape = np.zeros(8, np.int32)
    E = np.zeros(1, np.int64)
    nnz = len(idx)
    args = (rows, cols, nnz, ptr.ctypes.data, idx.ctypes.data if nnz else None, src.ctypes.data if nnz else None,
            cus, tsize, shape.ctypes.data, E.ctypes.data)
    assert L.sov_hip_tile_layout(*args, None, None, None, None) == 0
    nb, S, W, rmax, U, waves, Q, NBUF = (int(v) for v in shape)
    assert Q == 2 and NBUF == 2
    keys = np.zeros(max(int(E[0]), 1), np.uint32)
    so = np.zeros(max(int(E[0]), 1), np.int32)
    ws = np.zeros(nb * U * waves + 1, np.int32)
    units = np.zeros(4 * U + 1, np.int32)
    assert L.sov_hip_tile_layout(*args, keys.ctypes.data, so.ctypes.data, ws.ctypes.data, units.ctypes.data) == 0
    return nb, S, W, rmax, U, waves, keys[: E[0]], so[: E[0]], ws, units


def passes(ptr, idx, src, rows, cols, cus, tsize, nb, U, waves):
    L = lib()
    P = np.zeros(1, np.int64)
    nnz = len(idx)
    args = (rows, cols, nnz, ptr.ctypes.data, idx.ctypes.data if nnz else None, src.ctypes.data if nnz else None,
            cus, tsize, P.ctypes.data)
    assert L.sov_hip_tile_passes(*args, None, None) == 0
    pidx = np.zeros(nb * waves * (U + 1), np.int32)
    pdesc = np.zeros(2 * max(int(P[0]), 1), np.int32)
    assert L.sov_hip_tile_passes(*args, pidx.ctypes.data, pdesc.ctypes.data) == 0
    return pidx, pdesc[: 2 * int(P[0])].reshape(-1, 2)


def random_csr(rows, cols, lens, rng):
    lens = np.asarray(lens, np.int64)
    ptr = np.zeros(rows + 1, np.int32)
    ptr[1:] = np.cumsum(lens)
    idx = [np.sort(rng.choice(cols, size=int(k), replace=False)) for k in lens]
    idx = np.concatenate(idx).astype(np.int32) if ptr[-1] else np.zeros(0, np.int32)
    return ptr, idx


CASES = {
    "tiny": (1, 5, [3], 256),
    "staged_and_gather": (256, 81920, "mixed", 1),
    "empty_rows": (70, 40, "0-5", 4),
    "two_slices": (300, 20000, "0-40", 2),
    "dense_row": (130, 33000, [33000] + [0] * 128 + [5], 3),
    "long_rows": (200, 40000, [0] * 50 + [1500] + [7] * 149, 1),
    "many_groups": (8192 * 3 + 5, 100, "1-4", 2),
    "no_entries": (64, 10, [0] * 64, 8),
}


@pytest.mark.parametrize("tsize", [4, 8])
@pytest.mark.parametrize("name", list(CASES))
def test_layout_stores_the_csr(name, tsize):
    rows, cols, lens, cus = CASES[name]
    rng = np.random.default_rng(len(name))
    if lens == "mixed":
        # dense in the first two slices (staged), sparse in the last three (one gather unit)
        ptr = np.arange(rows + 1, dtype=np.int32) * 21
        idx = np.concatenate([np.sort(np.concatenate([rng.choice(32768, 20, replace=False),
                                                      32768 + rng.choice(49152, 1)])) for _ in range(rows)])
        idx = idx.astype(np.int32)
    else:
        if isinstance(lens, str):
            lo, hi = (int(v) for v in lens.split("-"))
            lens = rng.integers(lo, hi, rows)
        ptr, idx = random_csr(rows, cols, lens, rng)
    nnz = int(ptr[-1])
    src = rng.permutation(nnz).astype(np.int32)      # arbitrary source positions
    nb, S, W, rmax, U, WAVES, keys, so, ws, units = layout(ptr, idx, src, rows, cols, cus, tsize)
    assert WAVES == 12
    nchunks = -(-rows // 64)
    assert rmax <= 8128 and nb * rmax >= rows and rmax % 64 == 0
    # slices of 64 KiB (FP32) or 48 KiB (FP64), or the largest smaller one whose two buffers fit beside the
    # row sums in 160 KiB of LDS
    sb = next(b for b in ((65536, 49152, 32768) if tsize == 4 else (49152, 32768))
              if 2 * b + (rmax + 64) * tsize <= 163840)
    assert W == sb // tsize and S == -(-cols // W)
    assert nb == max(min(max(cus // 2, 1), nchunks), -(-nchunks // 127))
    # units: contiguous slice ranges; staged = one slice with >= 800 entries per group
    first = units[0:4 * U:4]
    staged = units[1:4 * U:4]
    last = np.append(first[1:], S) if U else first
    assert units[4 * U] == S and (U == 0 or first[0] == 0) and np.all(np.diff(first) > 0)
    per_slice = np.bincount(idx // W, minlength=S) if S else np.zeros(0, np.int64)
    for u in range(U):
        width = last[u] - first[u]
        dense = per_slice[first[u]:last[u]] >= 800 * nb
        if staged[u]:
            assert width == 1 and dense.all()
        else:
            assert 1 <= width <= min(63, 2 ** COLBITS // W - 1) and not dense.any()
    sidx = np.flatnonzero(staged)
    assert np.all(units[2:4 * U:4][sidx] == np.arange(len(sidx)) % 2)          # the two LDS buffers in turn
    for u in range(U):
        later = sidx[sidx > u]
        assert units[4 * u + 3] == (later[0] if len(later) else U)
    if name == "staged_and_gather" and tsize == 4:
        assert list(staged) == [1, 1, 0] and list(first) == [0, 1, 2]
    assert len(ws) == nb * U * WAVES + 1 and ws[-1] == len(keys)
    lane = 4                                                             # entries per lane per pass
    assert np.all(np.diff(ws) >= 0) and np.all(ws % lane == 0)
    row_of_src = np.repeat(np.arange(rows), np.diff(ptr))
    pos_of_src = np.empty(nnz, np.int64)
    pos_of_src[src] = np.arange(nnz)                 # src value -> CSR position
    seen = np.zeros(nnz, bool)
    for t in range(nb * U):
        b, u = divmod(t, U)
        tile_rows = []
        for w in range(WAVES):
            e0, e1 = ws[t * WAVES + w], ws[t * WAVES + w + 1]
            k = keys[e0:e1]
            real = k != SENT
            nreal = int(real.sum())
            assert np.all(real[:nreal]) and e1 - e0 - nreal < lane   # padding only at the end
            k, sv = k[:nreal].astype(np.int64), so[e0:e0 + nreal]
            l, q = k >> COLBITS, k & (2 ** COLBITS - 1)
            assert np.all(np.diff(l) >= 0) and np.all(l < rmax)
            p = ((l >> 6) * nb + b) * 64 + (l & 63)
            assert np.all(p < rows)
            col = first[u] * W + q
            assert np.all(col < last[u] * W) and np.all(col < cols)
            assert np.all(so[e0 + nreal:e1] == -1)
            at = pos_of_src[sv]
            assert np.all(row_of_src[at] == p) and np.all(idx[at] == col)
            assert not np.any(seen[at])
            seen[at] = True
            tile_rows.append(set(l.tolist()))
        for w in range(WAVES):                        # row-aligned: no row shared by two waves
            for v in range(w + 1, WAVES):
                assert not (tile_rows[w] & tile_rows[v])
    assert seen.all()
    # the pass list: per (group, wave), unit by unit, each sub-range cut into
    # passes of 64 lanes' entries (the last one short), plus, for the FP32 tiles,
    # an empty pass when that count is odd (their kernels add passes in
    # pairs, and a pass with a successor must fill all 64 lanes)
    pidx, pdesc = passes(ptr, idx, src, rows, cols, cus, tsize, nb, U, WAVES)
    want = []
    for b in range(nb):
        for w in range(WAVES):
            base = (b * WAVES + w) * (U + 1)
            for u in range(U):
                assert pidx[base + u] == len(want)
                t = b * U + u
                e0, e1 = ws[t * WAVES + w], ws[t * WAVES + w + 1]
                cut = [(e, min(64 * lane, e1 - e)) for e in range(e0, e1, 64 * lane)]
                want += cut + ([(e1, 0)] if len(cut) % 2 and tsize == 4 else [])
            assert pidx[base + U] == len(want)
            if tsize == 4:
                assert np.all((pidx[base + 1:base + U + 1] - pidx[base:base + U]) % 2 == 0)
    assert pdesc.tolist() == [list(p) for p in want]
    # the product the layout encodes
    rng2 = np.random.default_rng(1)
    val = rng2.standard_normal(max(nnz, 1))
    x = rng2.standard_normal(cols)
    y = np.zeros(rows)
    real = keys != SENT
    t_of = np.searchsorted(ws, np.arange(len(keys)), side="right") - 1
    g, u = np.divmod(t_of // WAVES, max(U, 1))
    k = keys.astype(np.int64)
    l = k >> COLBITS
    p = ((l >> 6) * nb + g) * 64 + (l & 63)
    np.add.at(y, p[real], val[pos_of_src[so[real]]] * x[first[u[real]] * W + (k[real] & (2 ** COLBITS - 1))])
    want = np.array([val[ptr[r]:ptr[r + 1]] @ x[idx[ptr[r]:ptr[r + 1]]] for r in range(rows)])
    np.testing.assert_allclose(y, want, rtol=1e-12, atol=1e-12)


def test_layout_rejects_bad_input():
    L = lib()
    shape = np.zeros(8, np.int32)
    E = np.zeros(1, np.int64)
    ptr = np.array([0, 2], np.int32)
    idx = np.array([0, 7], np.int32)      # column 7 out of range for cols = 5
    src = np.array([0, 1], np.int32)
    assert L.sov_hip_tile_layout(1, 5, 2, ptr.ctypes.data, idx.ctypes.data, src.ctypes.data, 4, 4, shape.ctypes.data,
                                 E.ctypes.data, None, None, None, None) != 0
    bad_ptr = np.array([0, 3], np.int32)  # row pointers disagree with nnz
    assert L.sov_hip_tile_layout(1, 5, 2, bad_ptr.ctypes.data, idx.ctypes.data, src.ctypes.data, 4, 4,
                                 shape.ctypes.data, E.ctypes.data, None, None, None, None) != 0
    assert L.sov_hip_tile_layout(0, 5, 0, ptr.ctypes.data, None, None, 4, 4, shape.ctypes.data, E.ctypes.data, None,
                                 None, None, None) != 0
    good = np.array([0, 2], np.int32)
    idx_ok = np.array([0, 3], np.int32)
    assert L.sov_hip_tile_layout(1, 5, 2, good.ctypes.data, idx_ok.ctypes.data, src.ctypes.data, 4, 6,
                                 shape.ctypes.data, E.ctypes.data, None, None, None, None) != 0   # element size


def ragged_csr():
    """A CSR whose gathered lengths are not whole 16-byte vectors in either
    orientation: n = 40003 columns (3 FP32 / 1 FP64 elements past the last
    whole vector), m = 301 rows (1 / 1).  Row 7 has an entry in every column,
    so the last columns are used."""
    rng = np.random.default_rng(11)
    m, n = 301, 40003
    lens = rng.integers(600, 900, m)
    lens[7] = n
    ptr, idx = random_csr(m, n, lens, rng)
    return m, n, ptr, idx


def transpose_csr(rows, cols, ptr, idx):
    order = np.argsort(idx, kind="stable")
    tptr = np.zeros(cols + 1, np.int32)
    tptr[1:] = np.cumsum(np.bincount(idx, minlength=cols))
    tidx = np.repeat(np.arange(rows, dtype=np.int32), np.diff(ptr))[order]
    return tptr, tidx.astype(np.int32), order.astype(np.int32)


@pytest.mark.parametrize("tsize", [4, 8])
def test_ragged_lengths_stage_the_last_slice(tsize):
    # The staging waves load whole 16-byte vectors and write the partial last
    # one's elements separately; this case (the one the GPU tiles test runs)
    # must put the partial vector in a staged unit in both orientations on a
    # 256-CU device, or that path goes untested.
    m, n, ptr, idx = ragged_csr()
    src = np.arange(len(idx), dtype=np.int32)
    tptr, tidx, tsrc = transpose_csr(m, n, ptr, idx)
    for rows, cols, p, i, s in ((m, n, ptr, idx, src), (n, m, tptr, tidx, tsrc)):
        assert cols % (16 // tsize) != 0
        nb, S, W, rmax, U, waves, keys, so, ws, units = layout(p, i, s, rows, cols, 256, tsize)
        last = [u for u in range(U) if units[4 * u] == S - 1]
        assert len(last) == 1 and units[4 * last[0] + 1] == 1, (rows, cols, S, units)
