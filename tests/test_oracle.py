"""Pins the oracle (CPU restatement) before it is trusted as the GPU checker.

Each check is chosen so that a plausible mistake fails it: a wrong SplitMix64
constant or byte order (published outputs), a '<' for '<=' in the commit rule
(closed forms, pairwise brute force, snapshot-read invariant), a dropped or
reordered op in the replay (RMW counting closed form, last-writer closed
form), and a race or ordering slip in the threaded pipeline (equality with
the serial replay).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from batches import OP_READ, OP_RMW_ADD, OP_UPDATE, op_txn, random_batch
from ltpmg_amd.engine import Batch

GOLDEN = 0x9E3779B97F4A7C15
M = (1 << 64) - 1


# ---- SplitMix64: published outputs of splitmix64.c started at state 0 ----
def test_splitmix64_known_answers():
    expect = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    assert [orc.stream(0, j) for j in range(4)] == expect


def test_payload_bytes_are_little_endian_stream():
    p = orc.payload(0, 20)
    assert p[:8].tobytes() == (0xE220A8397B1DCDAF).to_bytes(8, "little")
    assert p[8:16].tobytes() == (0x6E789E6AA1B965F4).to_bytes(8, "little")
    assert p[16:20].tobytes() == (0x06C45D188009454F).to_bytes(8, "little")[:4]


def _py_mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_fill_matches_definition():
    seed, table, col, width = 0x4C54504D47000001, 3, 5, 12
    out = np.zeros((4, width), np.uint8)
    orc.lib().orc_fill_column(seed, table, col, width, 7, 4, out.ctypes.data)
    for i in range(4):
        row = 7 + i
        cs = _py_mix64((seed + GOLDEN * (1 + ((row << 10) | (table << 6) | col))) & M)
        words = [_py_mix64((cs + (j + 1) * GOLDEN) & M) for j in range(2)]
        expect = b"".join(w.to_bytes(8, "little") for w in words)[:width]
        assert out[i].tobytes() == expect


# ---- Mode A closed forms (Appendix-A style known answers) ----
def _image(n_rows=64, widths=(4, 4)):
    return orc.Image({0: (n_rows, list(widths))}).fill(11)


def test_mode_a_all_distinct_keys_all_commit():
    txns = [[(OP_UPDATE, 0, 0, 2 * t, t + 1), (OP_READ, 0, 1, 2 * t + 1, 0)] for t in range(20)]
    b = Batch.from_txns(txns)
    im = _image()
    snap = im.copy()
    commit, res = orc.replay(im, b, orc.MODE_A, 4)
    assert commit.all()
    for t in range(20):
        assert res[2 * t + 1].tobytes() == snap.cols[0][1][2 * t + 1].tobytes()
        assert im.cols[0][0][2 * t].tobytes() == orc.payload(t + 1, 4).tobytes()


def test_mode_a_every_txn_writes_one_key_only_first_commits():
    txns = [[(OP_READ, 0, 0, 10 + t, 0), (OP_RMW_ADD, 0, 1, 5, 1)] for t in range(16)]
    commit, _ = orc.replay(_image(), Batch.from_txns(txns), orc.MODE_A, 4)
    assert commit.tolist() == [1] + [0] * 15


@pytest.mark.parametrize("w", [0, 3, 9])
def test_mode_a_readers_and_one_writer(w):
    txns = [[(OP_READ, 0, 0, 7, 0)] for _ in range(12)]
    txns[w] = [(OP_UPDATE, 0, 0, 7, 99)]
    commit, _ = orc.replay(_image(), Batch.from_txns(txns), orc.MODE_A, 4)
    assert commit.tolist() == [1 if t <= w else 0 for t in range(12)]


@pytest.mark.parametrize("seed", range(12))
def test_mode_a_commit_set_equals_pairwise_definition(seed):
    rng = np.random.default_rng(seed)
    b = random_batch(rng, n_txns=8, n_rows=6, n_cols=2, max_len=4)
    commit, _ = orc.replay(_image(8), b, orc.MODE_A, 4)
    assert (commit == orc.commit_pairwise(b)).all()


@pytest.mark.parametrize("seed", range(6))
def test_mode_a_committed_reads_see_snapshot_and_only_committed_writes_land(seed):
    rng = np.random.default_rng(100 + seed)
    b = random_batch(rng, n_txns=200, n_rows=64, n_cols=3, max_len=6)
    im = orc.Image({0: (64, [4, 8, 16])}).fill(5)
    snap = im.copy()
    commit, res = orc.replay(im, b, orc.MODE_A, 16)
    ot = op_txn(b)
    written = set()
    for i in range(b.n_ops):
        kind, col, row = int(b.meta[i] & 3), int((b.meta[i] >> 2) & 63), int(b.rows[i])
        if not commit[ot[i]]:
            continue
        w = im.schema[0][1][col]
        if kind == OP_READ:
            assert res[i, :w].tobytes() == snap.cols[0][col][row].tobytes()
            assert not res[i, w:].any()
        else:
            written.add((row, col))
    for c in range(3):
        for r in range(64):
            if (r, c) not in written:
                assert im.cols[0][c][r].tobytes() == snap.cols[0][c][r].tobytes()
    assert commit[0] == 1


# ---- Mode O closed forms ----
@pytest.mark.parametrize("seed", range(4))
def test_mode_o_rmw_counts(seed):
    rng = np.random.default_rng(200 + seed)
    b = random_batch(rng, n_txns=150, n_rows=16, n_cols=2, max_len=5, kinds=(OP_READ, OP_RMW_ADD))
    b.args[:] = np.where((b.meta & 3) == OP_RMW_ADD, 1, 0)
    im = orc.Image({0: (16, [4, 4])}).fill(9)
    snap = im.copy()
    commit, res = orc.replay(im, b, orc.MODE_O, 4)
    assert commit.all()
    count = {}
    for i in range(b.n_ops):
        kind, col, row = int(b.meta[i] & 3), int((b.meta[i] >> 2) & 63), int(b.rows[i])
        base = int(snap.cols[0][col][row].view("<u4")[0])
        if kind == OP_READ:
            assert int(res[i].view("<u4")[0]) == (base + count.get((row, col), 0)) & 0xFFFFFFFF
        else:
            count[(row, col)] = count.get((row, col), 0) + 1
    for c in range(2):
        for r in range(16):
            base = int(snap.cols[0][c][r].view("<u4")[0])
            assert int(im.cols[0][c][r].view("<u4")[0]) == (base + count.get((r, c), 0)) & 0xFFFFFFFF


@pytest.mark.parametrize("seed", range(4))
def test_mode_o_last_writer(seed):
    rng = np.random.default_rng(300 + seed)
    b = random_batch(rng, n_txns=150, n_rows=16, n_cols=2, max_len=5, kinds=(OP_READ, OP_UPDATE))
    im = orc.Image({0: (16, [100, 12])}).fill(9)
    snap = im.copy()
    _, res = orc.replay(im, b, orc.MODE_O, 128)
    last = {}
    for i in range(b.n_ops):
        kind, col, row = int(b.meta[i] & 3), int((b.meta[i] >> 2) & 63), int(b.rows[i])
        w = snap.schema[0][1][col]
        if kind == OP_READ:
            expect = orc.payload(last[(row, col)], w) if (row, col) in last else snap.cols[0][col][row]
            assert res[i, :w].tobytes() == expect.tobytes()
            assert not res[i, w:].any()
        else:
            last[(row, col)] = int(b.args[i])
    for (row, col), arg in last.items():
        assert im.cols[0][col][row].tobytes() == orc.payload(arg, snap.schema[0][1][col]).tobytes()


def _add_words(cell: np.ndarray, delta: int) -> np.ndarray:
    """Every little-endian u32 word of the cell plus delta (mod 2^32)."""
    w = cell.copy().view("<u4")
    return (w + np.uint32(delta & 0xFFFFFFFF)).astype("<u4").view(np.uint8)


@pytest.mark.parametrize("seed", range(6))
def test_mode_o_mixed_update_and_rmw(seed):
    # UPDATE and RMW_ADD interleaved on the same cells: a READ returns the
    # payload of the last UPDATE before it (or the snapshot when there is
    # none) plus the sum of the RMW deltas since that UPDATE; the final cell
    # is the same fold over all of the cell's ops.  Deltas are signed and
    # widths are not multiples of 8, so a dropped delta, a delta carried
    # across an UPDATE, a sign slip or a lost payload tail all fail.
    rng = np.random.default_rng(500 + seed)
    b = random_batch(rng, n_txns=300, n_rows=12, n_cols=3, max_len=5,
                     kinds=(OP_READ, OP_UPDATE, OP_RMW_ADD))
    widths = [4, 12, 100]
    im = orc.Image({0: (12, widths)}).fill(13 + seed)
    snap = im.copy()
    commit, res = orc.replay(im, b, orc.MODE_O, 128)
    assert commit.all()
    state = {}  # (row, col) -> (base bytes, delta since the base)
    n_mixed = 0
    for i in range(b.n_ops):
        kind, col, row = int(b.meta[i] & 3), int((b.meta[i] >> 2) & 63), int(b.rows[i])
        w = widths[col]
        base, d = state.get((row, col), (snap.cols[0][col][row], 0))
        if kind == OP_READ:
            assert res[i, :w].tobytes() == _add_words(base, d).tobytes(), (i, row, col)
            assert not res[i, w:].any()
            n_mixed += int(d != 0 and (row, col) in state and base is not snap.cols[0][col][row])
        elif kind == OP_UPDATE:
            state[(row, col)] = (orc.payload(int(b.args[i]), w), 0)
        else:
            state[(row, col)] = (base, d + int(np.int64(b.args[i].astype(np.int64))))
    assert n_mixed > 0  # some READs see an UPDATE followed by RMWs
    for (row, col), (base, d) in state.items():
        assert im.cols[0][col][row].tobytes() == _add_words(base, d).tobytes(), (row, col)


def test_replay_rejects_malformed_batches():
    im = _image()
    bad_row = Batch.from_txns([[(OP_READ, 0, 0, 64, 0)]])
    with pytest.raises(ValueError):
        orc.replay(im, bad_row, orc.MODE_A, 4)
    bad_col = Batch.from_txns([[(OP_READ, 0, 5, 1, 0)]])
    with pytest.raises(ValueError):
        orc.replay(im, bad_col, orc.MODE_O, 4)


# ---- threaded pipeline restatement == serial replay ----
@pytest.mark.parametrize("mode", [orc.MODE_A, orc.MODE_O])
@pytest.mark.parametrize("seed", range(3))
def test_cpu_pipeline_equals_replay(mode, seed):
    rng = np.random.default_rng(400 + seed)
    b = random_batch(rng, n_txns=3000, n_rows=500, n_cols=3, max_len=8, zipf=0.9 if seed else None)
    schema = {0: (500, [4, 100, 8])}
    im1 = orc.Image(schema).fill(1)
    im2 = im1.copy()
    c1, r1 = orc.replay(im1, b, mode, 128)
    c2, r2, secs = orc.cpu_pipeline(im2, b, mode, 128, n_threads=4)
    assert secs >= 0
    assert (c1 == c2).all()
    m = orc.read_mask(b, c1)
    assert (r1[m] == r2[m]).all()
    for c in range(3):
        assert (im1.cols[0][c] == im2.cols[0][c]).all()


def test_cpu_pipeline_multi_table():
    rng = np.random.default_rng(7)
    b1 = random_batch(rng, n_txns=400, n_rows=50, n_cols=2, max_len=4, table=0)
    b2 = random_batch(rng, n_txns=400, n_rows=30, n_cols=3, max_len=4, table=2)
    from batches import concat
    b = concat(b1, b2)
    schema = {0: (50, [4, 4]), 1: (1, [4]), 2: (30, [8, 32, 4])}
    for mode in (orc.MODE_A, orc.MODE_O):
        im1 = orc.Image(schema).fill(3)
        im2 = im1.copy()
        c1, r1 = orc.replay(im1, b, mode, 32)
        c2, r2, _ = orc.cpu_pipeline(im2, b, mode, 32, n_threads=3)
        assert (c1 == c2).all()
        m = orc.read_mask(b, c1)
        assert (r1[m] == r2[m]).all()
        assert im1.digest() == im2.digest()
