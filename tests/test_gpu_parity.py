"""GPU parity: the HIP engine against the oracle, bit-exact.

Every test opens a fresh engine on cuda:0, fills tables from a seed on the
device, runs a batch through the C ABI and compares commit flags, every
defined read result and the whole final table image with the oracle's serial
replay of the same batch.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from batches import OP_READ, OP_RMW_ADD, OP_UPDATE, concat, random_batch
from ltpmg_amd import MODE_A, MODE_O, Batch, Engine, LtpmgError, gen_kv

pytestmark = pytest.mark.gpu

BOTH = [MODE_A, MODE_O]


def _setup(e: Engine, schema: dict, seed: int) -> orc.Image:
    for t, (n, widths) in sorted(schema.items()):
        e.create_table(t, n, widths)
        e.fill_table(t, seed)
    return orc.Image(schema).fill(seed)


def _compare(e: Engine, im: orc.Image, batch: Batch, mode: int, check_image=True, pipeline=False):
    rs = e.result_stride
    commit, res = e.run(batch, mode, pipeline=pipeline)
    oc, ores = orc.replay(im, batch, mode, rs)
    assert commit.shape == oc.shape
    bad = np.nonzero(commit != oc)[0]
    assert bad.size == 0, f"commit flags differ at txns {bad[:10]}"
    m = orc.read_mask(batch, oc)
    if not (res[m] == ores[m]).all():
        idx = np.nonzero(m)[0]
        wrong = idx[np.any(res[m] != ores[m], axis=1)]
        pytest.fail(f"{wrong.size} read results differ, first ops {wrong[:8]}")
    if check_image:
        for t, (n, widths) in im.schema.items():
            for c in range(len(widths)):
                got = e.read_column(t, c)
                if not (got == im.cols[t][c]).all():
                    rows = np.nonzero(np.any(got != im.cols[t][c], axis=1))[0]
                    pytest.fail(f"table {t} col {c}: {rows.size} rows differ, first {rows[:8]}")
            assert e.digest(t) == orc.digest_columns(t, im.cols[t])
    st = e.stats()
    assert st["committed"] == int(oc.sum()) and st["aborted"] == batch.n_txns - int(oc.sum())
    return commit, res, st


# ---------------- tables ----------------
def test_fill_and_digest_match_oracle():
    schema = {0: (1000, [4, 8, 12, 16, 32, 100, 256]), 3: (77, [100])}
    with Engine() as e:
        im = _setup(e, schema, 0x4C54504D47000001)
        for t, (n, widths) in schema.items():
            for c in range(len(widths)):
                assert (e.read_column(t, c) == im.cols[t][c]).all(), (t, c)
            assert e.digest(t) == orc.digest_columns(t, im.cols[t])
        assert e.result_stride == 256


def test_load_and_read_column_roundtrip():
    with Engine() as e:
        e.create_table(0, 500, [4, 100])
        data = np.random.default_rng(1).integers(0, 256, (500, 100), dtype=np.uint8)
        e.load_column(0, 1, data)
        assert (e.read_column(0, 1) == data).all()
        part = e.read_column(0, 1, row0=123, n_rows=10)
        assert (part == data[123:133]).all()
        with pytest.raises(LtpmgError):
            e.read_column(0, 1, row0=495, n_rows=10)


# ---------------- building blocks (reference self-test KATs) ----------------
def test_sort_reference_kat_descending_million():
    # gpulauncher.cu:1237-1264: 1,000,000 values in descending order sort to i+1
    n = 1_000_000
    keys = np.arange(n, 0, -1, dtype=np.uint32)
    with Engine() as e:
        k, v = e.selftest_sort(keys, np.arange(n, dtype=np.uint32), 32)
    assert (k == np.arange(1, n + 1, dtype=np.uint32)).all()
    assert (v == np.arange(n - 1, -1, -1, dtype=np.uint32)).all()


@pytest.mark.parametrize("n,bits", [(1, 8), (100, 7), (4097, 8), (70000, 13), (300001, 24),
                                    (1 << 20, 28), (123457, 32), (100000, 9), (65537, 17),
                                    (250001, 18), (1 << 21, 27), ((1 << 24) + 4097, 28)])
def test_sort_is_stable_and_matches_numpy(n, bits):
    # (2^24 + 4097, 28): 7-bit digits over 8195 tiles of 2048 keys, more than
    # 64 tile rows per offset-scan chunk, so the scans take their general
    # (one thread per digit) path
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    if bits < 32:
        keys &= np.uint32((1 << bits) - 1)
    if n > 1000:  # heavy duplicates, as zipf keys have
        keys[: n // 3] = keys[0]
    vals = np.arange(n, dtype=np.uint32)
    with Engine() as e:
        k, v = e.selftest_sort(keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    assert (k == keys[order]).all()
    assert (v == vals[order]).all()


def test_segscan_reference_kat():
    # gpulauncher.cu:1267-1321: 128 values i % 64 in two segments of 64
    vals = (np.arange(128) % 64).astype(np.uint32)
    heads = np.zeros(128, np.uint8)
    heads[64] = 1
    with Engine() as e:
        out = e.selftest_segscan(vals, heads)
    seg = np.concatenate([np.cumsum(vals[:64]) - vals[:64], np.cumsum(vals[64:]) - vals[64:]])
    assert (out == seg.astype(np.uint32)).all()


@pytest.mark.parametrize("n,p", [(5000, 0.01), (200003, 0.0001), (100000, 0.5), (9000, 0.0)])
def test_segscan_random(n, p):
    rng = np.random.default_rng(n)
    vals = rng.integers(0, 1000, n).astype(np.uint32)
    heads = (rng.random(n) < p).astype(np.uint8)
    with Engine() as e:
        out = e.selftest_segscan(vals, heads)
    expect = np.empty(n, np.uint64)
    run = 0
    for i in range(n):
        if i == 0 or heads[i]:
            run = 0
        expect[i] = run
        run += int(vals[i])
    assert (out == (expect & 0xFFFFFFFF).astype(np.uint32)).all()


@pytest.mark.parametrize("n,p", [(9_000_000, 1e-4), (36_000_000, 1e-4), (9_000_000, 2e-7),
                                 (3_000_000, 0.0)])
def test_segscan_many_tiles(n, p):
    # thousands of 2048-element tiles.  A tile's carry folds the tile
    # aggregates back to the newest one holding a segment head, 64 per step:
    # at p = 2e-7 a segment spans thousands of tiles, and with no head but
    # the first the last tile reads back over all 1464 tiles before it
    rng = np.random.default_rng(9)
    vals = rng.integers(0, 1 << 20, n).astype(np.uint32)
    heads = (rng.random(n) < p).astype(np.uint8)
    for b in (4_194_304, 8192 * 2048, 2 * 8192 * 2048):
        if b < n:
            heads[b] = 0  # a segment across a tile boundary / a 64-tile look-back step
    with Engine() as e:
        out = e.selftest_segscan(vals, heads)
    excl = np.cumsum(vals, dtype=np.uint64) - vals
    start = np.maximum.accumulate(np.where((heads != 0) | (np.arange(n) == 0), np.arange(n), 0))
    expect = (excl - excl[start]) & np.uint64(0xFFFFFFFF)
    assert (out == expect.astype(np.uint32)).all()


# ---------------- closed forms on the GPU ----------------
def test_gpu_mode_a_closed_forms():
    schema = {0: (64, [4, 4])}
    with Engine() as e:
        im = _setup(e, schema, 21)
        distinct = Batch.from_txns([[(OP_UPDATE, 0, 0, 2 * t, t + 1), (OP_READ, 0, 1, 2 * t + 1, 0)]
                                    for t in range(20)])
        c, _, _ = _compare(e, im, distinct, MODE_A)
        assert c.all()
        one_key = Batch.from_txns([[(OP_READ, 0, 0, 40 + t % 20, 0), (OP_RMW_ADD, 0, 1, 5, 1)]
                                   for t in range(16)])
        c, _, _ = _compare(e, im, one_key, MODE_A)
        assert c.tolist() == [1] + [0] * 15
        readers = [[(OP_READ, 0, 0, 7, 0)] for _ in range(12)]
        readers[4] = [(OP_UPDATE, 0, 0, 7, 99)]
        c, _, _ = _compare(e, im, Batch.from_txns(readers), MODE_A)
        assert c.tolist() == [1] * 5 + [0] * 7


# ---------------- random batches vs the oracle ----------------
@pytest.mark.parametrize("mode", BOTH)
@pytest.mark.parametrize("seed", range(4))
def test_random_batches(mode, seed):
    rng = np.random.default_rng(1000 + seed)
    schema = {0: (3000, [4, 8, 100, 32])}
    b = random_batch(rng, n_txns=5000, n_rows=3000, n_cols=4, max_len=12,
                     zipf=0.9 if seed % 2 else None)
    with Engine() as e:
        im = _setup(e, schema, seed)
        _compare(e, im, b, mode)


@pytest.mark.parametrize("mode", BOTH)
def test_multi_table(mode):
    rng = np.random.default_rng(5)
    schema = {0: (2000, [4, 4, 4]), 2: (1, [8]), 5: (700, [256, 16, 100])}
    b = concat(random_batch(rng, 3000, 2000, 3, max_len=6, table=0),
               random_batch(rng, 200, 1, 1, max_len=1, table=2),
               random_batch(rng, 3000, 700, 3, max_len=6, table=5))
    perm = rng.permutation(b.n_txns)  # interleave the tables' txns
    txns = [[(int(b.meta[i] & 3), int(b.meta[i] >> 8), int((b.meta[i] >> 2) & 63), int(b.rows[i]),
              int(b.args[i])) for i in range(b.txn_off[t], b.txn_off[t + 1])] for t in perm]
    with Engine() as e:
        im = _setup(e, schema, 77)
        _compare(e, im, Batch.from_txns(txns), mode)


@pytest.mark.parametrize("mode", BOTH)
@pytest.mark.parametrize("widths", [[8, 4], [12, 4], [40, 8], [8, 96], [148, 4, 20]],
                         ids=["stride8", "stride16", "stride64", "stride96", "stride160"])
def test_result_strides(mode, widths):
    # the result stride is the widest slot: 8 B (word moves), 16 and 64 B
    # (one 16-byte chunk per lane), 96 and 160 B (lanes loop over chunks)
    rng = np.random.default_rng(len(widths) * 1000 + widths[0])
    schema = {0: (1500, widths)}
    b = random_batch(rng, n_txns=4000, n_rows=1500, n_cols=len(widths), max_len=10, zipf=0.9)
    with Engine() as e:
        im = _setup(e, schema, 0x5EED + widths[0])
        assert e.result_stride == {8: 8, 12: 16, 40: 64, 96: 96, 148: 160}[max(widths)]
        _compare(e, im, b, mode)


@pytest.mark.parametrize("n_ops", [1, 2047, 2048, 2049, 4096 + 255])
@pytest.mark.parametrize("mode,pipeline", [(MODE_A, False), (MODE_O, True), (MODE_O, False)],
                         ids=["lowest_tid_wins", "ordered_pipeline", "ordered_one_launch"])
def test_tile_boundaries(n_ops, mode, pipeline):
    # one-op txns; when there is room, a 255-op txn starts at op 1900 and
    # straddles the first 2048-op tile boundary (of the pipeline's tiles;
    # batches this small otherwise take the one-launch ordered path)
    rng = np.random.default_rng(n_ops)

    def single():
        return [(int(rng.integers(0, 3)), 0, 0, int(rng.integers(0, 5000)), 3)]

    if n_ops > 1900 + 255:
        rows = rng.choice(5000, 255, replace=False)
        big = [(int(rng.integers(0, 3)), 0, 0, int(r), int(rng.integers(0, 2**40))) for r in rows]
        txns = [single() for _ in range(1900)] + [big] + [single() for _ in range(n_ops - 2155)]
    else:
        txns = [single() for _ in range(n_ops)]
    with Engine() as e:
        im = _setup(e, {0: (5000, [16])}, 3)
        _compare(e, im, Batch.from_txns(txns), mode, pipeline=pipeline)


def test_max_length_txns():
    rng = np.random.default_rng(9)
    txns = []
    for _ in range(40):
        rows = rng.choice(20000, 256, replace=False)
        txns.append([(int(rng.integers(0, 3)), 0, 1, int(r), int(rng.integers(0, 2**60))) for r in rows])
    b = Batch.from_txns(txns)
    for mode in BOTH:
        with Engine() as e:
            im = _setup(e, {0: (20000, [4, 100])}, 4)
            _compare(e, im, b, mode)


# ---------------- BASELINE workloads ----------------
@pytest.mark.parametrize("mode", BOTH)
def test_c1_kv_int32_rmw(mode):
    # 100k rows x 10 int32; 8192 txns x 10 ops on distinct uniform keys,
    # 50% READ (whole row) / 50% RMW(+1) on one column
    b = gen_kv(8192, 100_000, 10, ops=10, read_pct=50, write_kind=OP_RMW_ADD, read_all_cols=True,
               rmw_delta=1, seed=0x4C54504D47000002)
    with Engine() as e:
        im = _setup(e, {0: (100_000, [4] * 10)}, 0x4C54504D47000001)
        c, _, st = _compare(e, im, b, mode)
    if mode == MODE_A:
        assert 0 < c.sum() < len(c)


@pytest.mark.parametrize("mode", BOTH)
def test_c2_shape_reduced(mode):
    # C2's shape at 1/10 of the rows and 1/16 of the txns: 100-byte fields,
    # 16 ops per txn, zipf 0.99, 50/50 READ/UPDATE
    b = gen_kv(1 << 16, 1_000_000, 10, ops=16, zipf_theta=0.99, read_pct=50, write_kind=OP_UPDATE,
               seed=0x4C54504D47000003)
    with Engine() as e:
        im = _setup(e, {0: (1_000_000, [100] * 10)}, 0x4C54504D47000004)
        _compare(e, im, b, mode)


def test_ordered_mode_over_many_tiles():
    # 4.8M ops: more than 2048 scan tiles, so segment carries cross rounds
    b = gen_kv(300_000, 1_000_000, 4, ops=16, zipf_theta=0.99, read_pct=50, write_kind=OP_UPDATE,
               seed=0x4C54504D47000005)
    with Engine() as e:
        im = _setup(e, {0: (1_000_000, [100] * 4)}, 0x4C54504D47000006)
        _compare(e, im, b, MODE_O)


@pytest.mark.parametrize("first_kind", [OP_UPDATE, OP_RMW_ADD])
def test_ordered_mode_tile_of_reads_ends_a_written_segment(first_kind):
    # Sorted list: cell (0,0) = one write then 2100 READs, cell (0,1) = 2500
    # READs.  The second 2048-op tile is READs only, yet the segment that
    # ends in it carries a write from the first tile: its final write must
    # still be applied, and it must not take a READ's place in that tile.
    txns = [[(first_kind, 0, 0, 0, 0x1234_5678_9ABC)]]
    txns += [[(OP_READ, 0, 0, 0, 0)] for _ in range(2100)]
    txns += [[(OP_READ, 0, 1, 0, 0)] for _ in range(2500)]
    # the same shape one table over, so a tile can hold the end of one and
    # the start of the other
    txns += [[(first_kind, 1, 0, 3, 77)]] + [[(OP_READ, 1, 0, 3, 0)] for _ in range(4095)]
    b = Batch.from_txns(txns)
    with Engine() as e:
        im = _setup(e, {0: (100, [100, 100]), 1: (10, [8])}, 0x4C54504D4700000A)
        _compare(e, im, b, MODE_O)


def test_ordered_mode_past_the_column_list():
    # 5 tables x 64 columns = 320 columns: the first 256 (in table, column
    # order) come from the kernel's LDS copy of the column list, the rest
    # from the descriptor table in global memory
    widths = [[4, 16, 100, 8][c % 4] for c in range(64)]
    schema = {t: (40 + t, widths[t:] + widths[:t]) for t in range(5)}
    rng = np.random.default_rng(17)
    txns = []
    for _ in range(6000):
        ops, seen = [], set()
        for _ in range(int(rng.integers(1, 9))):
            t = int(rng.integers(0, 5))
            r, c = int(rng.integers(0, 40 + t)), int(rng.integers(0, 64))
            if (t, r, c) in seen:
                continue
            seen.add((t, r, c))
            kind = [OP_READ, OP_UPDATE, OP_RMW_ADD][int(rng.integers(0, 3))]
            ops.append((kind, t, c, r, int(rng.integers(0, 2**62))))
        txns.append(ops)
    b = Batch.from_txns(txns)
    with Engine() as e:
        im = _setup(e, schema, 0x4C54504D4700000B)
        _compare(e, im, b, MODE_O)


def _sparse_image(schema: dict, seed: int, batches) -> orc.Image:
    """An oracle image that fills only the cells the batches touch."""
    im = orc.Image(schema)
    for t, (n, widths) in schema.items():
        for c in range(len(widths)):
            rows = np.unique(np.concatenate([b.rows[(((b.meta >> 8) & 15) == t) & (((b.meta >> 2) & 63) == c)]
                                             for b in batches]))
            if rows.size:
                im.fill_cells(seed, t, c, rows)
    return im


@pytest.mark.parametrize("rows,cols", [(34_000_000, 64), (70_000_000, 33)])
def test_ordered_mode_past_2e31_cells(rows, cols):
    # 34M rows x 64 columns = 2.18e9 cells: keys above 2^31 reach the sort and
    # decode.  70M x 33 = 2.31e9 cells, but 33 columns padded to 64 would need
    # 4.48e9 keys, so the keys keep a stride of 33 and decode with a division
    rng = np.random.default_rng(12)
    n = 60_000
    hi = rng.integers(rows - 2_000_000, rows, n // 2)
    lo = rng.integers(0, rows, n // 2)
    keys = np.concatenate([hi, lo])
    txns = []
    for t in range(n // 4):
        ops = []
        for k in range(4):
            r = int(keys[4 * t + k])
            ops.append((OP_UPDATE if k % 2 else OP_READ, 0, int(rng.integers(0, cols)), r,
                        int(rng.integers(0, 2**62))))
        # distinct (row, col) within the txn
        seen, uniq = set(), []
        for op in ops:
            if (op[3], op[2]) not in seen:
                seen.add((op[3], op[2]))
                uniq.append(op)
        txns.append(uniq)
    b1 = Batch.from_txns(txns)
    # the second batch reads every cell the first one touched, then some
    cells = sorted({(int(r), int(c)) for r, c in zip(b1.rows, (b1.meta >> 2) & 63)})
    b2 = Batch.from_txns([[(OP_READ, 0, c, r, 0)] for r, c in cells])
    schema = {0: (rows, [4] * cols)}
    seed = 0x4C54504D47000009
    with Engine() as e:
        e.create_table(0, rows, [4] * cols)
        e.fill_table(0, seed)
        im = _sparse_image(schema, seed, [b1, b2])
        rs = e.result_stride
        for b in (b1, b2):
            commit, res = e.run(b, MODE_O)
            oc, ores = orc.replay(im, b, MODE_O, rs)
            assert (commit == oc).all()
            m = orc.read_mask(b, oc)
            assert (res[m] == ores[m]).all()


def test_ordered_mode_refuses_more_than_2e32_cells():
    rows, cols = 67_200_000, 64  # 4.3e9 cells
    b = Batch.from_txns([[(OP_UPDATE, 0, 63, rows - 1, 5)], [(OP_READ, 0, 63, rows - 1, 0)]])
    with Engine() as e:
        e.create_table(0, rows, [4] * cols)
        e.submit(b)
        with pytest.raises(LtpmgError) as ei:
            e.execute(MODE_O)
        assert ei.value.code == -1 and "cells" in str(ei.value)
        commit, res = e.run(b, MODE_A)
        assert commit.tolist() == [1, 0]


def test_batch_sequence_and_determinism():
    schema = {0: (4000, [4, 100])}
    rng = np.random.default_rng(33)
    batches = [random_batch(rng, 3000, 4000, 2, max_len=10, zipf=0.8) for _ in range(5)]
    modes = [MODE_A, MODE_O, MODE_A, MODE_A, MODE_O]
    outs = []
    for _ in range(2):
        with Engine() as e:
            im = _setup(e, schema, 8)
            run = []
            for b, m in zip(batches, modes):
                c, r, _ = _compare(e, im, b, m, check_image=False)
                run.append((c, r[orc.read_mask(b, c)]))
            for c in range(2):
                assert (e.read_column(0, c) == im.cols[0][c]).all()
            outs.append((run, e.digest(0)))
    for (c1, r1), (c2, r2) in zip(outs[0][0], outs[1][0]):
        assert (c1 == c2).all() and (r1 == r2).all()
    assert outs[0][1] == outs[1][1]


@pytest.mark.parametrize("mode", BOTH)
def test_two_slots_apply_in_execute_order(mode):
    """Batches alternate between two slots without host waits in between; the
    table must still see them in execute order (the serving form's overlap)."""
    schema = {0: (3000, [4, 100])}
    rng = np.random.default_rng(44)
    batches = [random_batch(rng, 4000, 3000, 2, max_len=8, zipf=0.8) for _ in range(6)]
    got = []
    with Engine(n_slots=2) as e:
        im = _setup(e, schema, 12)
        for k, b in enumerate(batches):
            slot = k % 2
            if k >= 2:
                got.append(e.fetch(slot))
            e.submit(b, slot)
            e.execute(mode, slot)
        got.append(e.fetch(0))
        got.append(e.fetch(1))
        for k, b in enumerate(batches):
            oc, ores = orc.replay(im, b, mode, e.result_stride)
            commit, res = got[k]
            assert (commit == oc).all(), k
            m = orc.read_mask(b, oc)
            assert (res[m] == ores[m]).all(), k
        for c in range(2):
            assert (e.read_column(0, c) == im.cols[0][c]).all()


@pytest.mark.parametrize("trusted", [False, True])
def test_staged_submit(trusted):
    rng = np.random.default_rng(55)
    b = random_batch(rng, 3000, 2000, 3, max_len=10, zipf=0.7)
    with Engine() as e:
        im = _setup(e, {0: (2000, [4, 100, 8])}, 6)
        for mode in BOTH:
            e.submit_staged(b, trusted=trusted)
            e.execute(mode)
            commit, res = e.fetch()
            oc, ores = orc.replay(im, b, mode, e.result_stride)
            assert (commit == oc).all()
            m = orc.read_mask(b, oc)
            assert (res[m] == ores[m]).all()
        bad = Batch.from_txns([[(OP_READ, 0, 0, 2000, 0)]])
        with pytest.raises(LtpmgError):
            e.submit_staged(bad, trusted=False)


# ---------------- error paths ----------------
def test_invalid_batches_are_rejected():
    with Engine() as e:
        im = _setup(e, {0: (100, [4])}, 1)
        bad = [
            Batch.from_txns([[(OP_READ, 0, 0, 100, 0)]]),        # row out of range
            Batch.from_txns([[(OP_READ, 0, 1, 5, 0)]]),          # column out of range
            Batch.from_txns([[(OP_READ, 4, 0, 5, 0)]]),          # no such table
            Batch(np.array([0, 1, 1], np.uint32), np.array([1], np.uint32),
                  np.array([0], np.uint16), np.array([0], np.uint64)),  # empty txn
            Batch(np.array([0, 1], np.uint32), np.array([1], np.uint32),
                  np.array([3], np.uint16), np.array([0], np.uint64)),  # kind 3
        ]
        for b in bad:
            with pytest.raises(LtpmgError) as ei:
                e.submit(b)
            assert ei.value.code == -1
        with pytest.raises(LtpmgError):
            e.execute(5)
        # the engine is still usable
        _compare(e, im, Batch.from_txns([[(OP_UPDATE, 0, 0, 3, 5)], [(OP_READ, 0, 0, 3, 0)]]), MODE_O)


def test_written_cell_accessed_again_in_a_txn_is_rejected():
    # a txn may read a cell and then write it, but nothing may touch a cell
    # after the txn wrote it (the engine's reads see the batch-start snapshot
    # and a txn's writes of one cell would race)
    with Engine() as e:
        im = _setup(e, {0: (100, [4, 8])}, 2)
        bad = [
            [(OP_UPDATE, 0, 0, 5, 1), (OP_READ, 0, 0, 5, 0)],
            [(OP_RMW_ADD, 0, 1, 7, 1), (OP_RMW_ADD, 0, 1, 7, 2)],
            [(OP_READ, 0, 0, 9, 0), (OP_UPDATE, 0, 0, 9, 3), (OP_RMW_ADD, 0, 0, 9, 1)],
            [(OP_RMW_ADD, 0, 1, 3, 1), (OP_READ, 0, 0, 4, 0), (OP_UPDATE, 0, 1, 3, 1)],
        ]
        for ops in bad:
            b = Batch.from_txns([[(OP_READ, 0, 0, 1, 0)], ops])
            with pytest.raises(LtpmgError) as ei:
                e.submit(b)
            assert ei.value.code == -1 and "txn 1" in str(ei.value), str(ei.value)
            with pytest.raises(LtpmgError):
                e.submit_staged(b, trusted=False)
        ok = Batch.from_txns([
            [(OP_READ, 0, 0, 5, 0), (OP_UPDATE, 0, 0, 5, 1)],           # read, then write
            [(OP_READ, 0, 1, 5, 0), (OP_READ, 0, 1, 5, 0)],             # two reads
            [(OP_UPDATE, 0, 0, 6, 2), (OP_UPDATE, 0, 1, 6, 3)],         # two cells of one row
            [(OP_RMW_ADD, 0, 0, 7, 4), (OP_RMW_ADD, 0, 0, 8, 5)],
        ])
        for mode in BOTH:
            _compare(e, im, ok, mode)


def test_trusted_submit_still_checks_the_txn_offsets():
    # a trusted submit skips the per-op checks, but the kernels derive every
    # op range from txn_off, so its structure is always checked
    with Engine() as e:
        _setup(e, {0: (100, [4])}, 3)
        bad = [
            Batch(np.array([0, 1, 1], np.uint32), np.array([1], np.uint32),
                  np.array([0], np.uint16), np.array([0], np.uint64)),   # empty txn
            Batch(np.array([0, 2, 1], np.uint32), np.array([1, 2], np.uint32),
                  np.array([0, 0], np.uint16), np.array([0, 0], np.uint64)),  # offsets go back
            Batch(np.array([1, 2], np.uint32), np.array([1, 2], np.uint32),
                  np.array([0, 0], np.uint16), np.array([0, 0], np.uint64)),  # txn_off[0] != 0
        ]
        for b in bad:
            with pytest.raises(LtpmgError) as ei:
                e.submit_staged(b, trusted=True)
            assert ei.value.code == -1


def test_hbm_budget_is_enforced():
    with Engine(hbm_budget=64 << 20) as e:
        with pytest.raises(LtpmgError) as ei:
            e.create_table(0, 1_000_000, [100])
        assert ei.value.code == -3
        e.create_table(1, 1000, [4])
