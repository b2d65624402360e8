"""GPU parity of the row ops (INSERT, DELETE) on append tables against the
oracle's serial replay, bit-exact: commit flags, per-op status, read and
INSERT result slots, presence bytes, the append cursor and every cell.
The reference's insert / delete executors (cuda/src/gpulauncher.cu:205-356)
are what these ops replace."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from batches import OP_READ, OP_ROW, ROW_DELETE, ROW_INSERT, random_batch, random_row_batch
from ltpmg_amd import MODE_A, MODE_O, Batch, Engine, LtpmgError, Stream, StreamPlan

pytestmark = pytest.mark.gpu

BOTH = [MODE_A, MODE_O]
SEED = 0x5EED0005


def _setup(e: Engine, schema: dict, append: dict, seed: int = SEED) -> orc.Image:
    for t, (n, widths) in sorted(schema.items()):
        if t in append:
            e.create_append_table(t, n, widths, append[t])
        else:
            e.create_table(t, n, widths)
        e.fill_table(t, seed)
    return orc.Image(schema, append=append).fill(seed)


def _check_tables(e: Engine, im: orc.Image, what: str = ""):
    for t, (n, widths) in im.schema.items():
        for c in range(len(widths)):
            got = e.read_column(t, c)
            if not (got == im.cols[t][c]).all():
                rows = np.nonzero(np.any(got != im.cols[t][c], axis=1))[0]
                pytest.fail(f"{what}table {t} col {c}: {rows.size} rows differ, first {rows[:8]}")
        assert e.digest(t) == orc.digest_columns(t, im.cols[t]), f"{what}table {t} digest"
        if t in im.present:
            assert e.table_rows(t) == im.cursor[t], f"{what}table {t} cursor"
            pres = e.read_presence(t)
            bad = np.nonzero(pres != im.present[t])[0]
            assert bad.size == 0, f"{what}table {t}: presence differs at rows {bad[:8]}"


def _run_compare(e: Engine, im: orc.Image, batch: Batch, mode: int, what: str = "", pipeline: bool = False):
    rs = e.result_stride
    commit, res = e.run(batch, mode, pipeline=pipeline)
    status = e.fetch_status()
    oc, ores, ofound = orc.replay(im, batch, mode, rs, with_found=True)
    bad = np.nonzero(commit != oc)[0]
    assert bad.size == 0, f"{what}commit flags differ at txns {bad[:10]}"
    bad = np.nonzero(status != ofound)[0]
    assert bad.size == 0, f"{what}status differs at ops {bad[:10]}"
    m = orc.read_mask(batch, oc)
    if not (res[m] == ores[m]).all():
        idx = np.nonzero(m)[0]
        wrong = idx[np.any(res[m] != ores[m], axis=1)]
        pytest.fail(f"{what}{wrong.size} result slots differ, first ops {wrong[:8]}")
    st = e.stats()
    is_ins = (batch.meta & 0xFF) == OP_ROW
    assert st["inserted"] == int(ofound[is_ins].sum()), what
    assert st["committed"] == int(oc.sum())
    return oc, ofound


# batches of up to LTPMG_SMALL_MAX_OPS ops take the one-launch ordered
# kernel unless the multi-kernel pipeline is asked for
PATHS = [(MODE_A, False), (MODE_O, True), (MODE_O, False)]
PATH_IDS = ["lowest_tid_wins", "ordered_pipeline", "ordered_one_launch"]


@pytest.mark.parametrize("mode,pipeline", PATHS, ids=PATH_IDS)
@pytest.mark.parametrize("seed", range(3))
def test_row_ops_random_batches(mode, pipeline, seed):
    rng = np.random.default_rng(seed)
    schema = {0: (12_000, [4, 8, 16, 100, 32])}
    with Engine() as e:
        im = _setup(e, schema, {0: 3000})
        for k in range(3):
            b = random_row_batch(rng, 2000, 4000, 5, max_len=6, insert_pct=15, delete_pct=10)
            assert b.n_ops <= 8192
            oc, found = _run_compare(e, im, b, mode, f"batch {k}: ", pipeline)
            assert found.sum() > 0 and (mode == MODE_O or 0 < oc.sum() < len(oc))
        _check_tables(e, im)


@pytest.mark.parametrize("mode", BOTH)
def test_row_ops_over_many_tiles(mode):
    # about 130k ops: many 2048-op insert tiles and CC tiles, INSERTs spread
    # over all of them, DELETEs of rows that earlier ops read and write
    rng = np.random.default_rng(7)
    schema = {0: (200_000, [8, 16])}
    b = random_row_batch(rng, 40_000, 120_000, 2, max_len=6, insert_pct=10, delete_pct=5)
    with Engine() as e:
        im = _setup(e, schema, {0: 100_000})
        _run_compare(e, im, b, mode)
        _check_tables(e, im)


@pytest.mark.parametrize("mode", BOTH)
def test_inserts_into_six_append_tables_in_one_batch(mode):
    # the insert writer ranks four tables per pass: six tables take two
    # passes per 2048-op tile, over about 40k ops
    rng = np.random.default_rng(13)
    schema = {t: (6000 + 500 * t, [4, 8 + 4 * t, 32][: 1 + t % 3]) for t in range(6)}
    parts = [random_row_batch(rng, 1500, 2000 + 200 * t, len(schema[t][1]), max_len=6, table=t, insert_pct=40,
                              delete_pct=5) for t in range(6)]
    txns = []
    for p in parts:
        for t in range(p.n_txns):
            a, z = int(p.txn_off[t]), int(p.txn_off[t + 1])
            txns.append([(int(p.meta[i] & 3), int((p.meta[i] >> 8) & 15), int((p.meta[i] >> 2) & 63),
                          int(p.rows[i]), int(p.args[i])) for i in range(a, z)])
    order = rng.permutation(len(txns))
    b = Batch.from_txns([txns[i] for i in order])
    assert b.n_ops > 4 * 2048
    with Engine() as e:
        im = _setup(e, schema, {t: 2000 + 200 * t for t in range(6)})
        _, found = _run_compare(e, im, b, mode, pipeline=True)
        is_ins = (b.meta & 0xFF) == OP_ROW
        tabs = (b.meta >> 8) & 15
        assert all(found[is_ins & (tabs == t)].sum() > 0 for t in range(6))
        _check_tables(e, im)


@pytest.mark.parametrize("mode,pipeline", PATHS, ids=PATH_IDS)
def test_row_ops_with_ordinary_and_two_append_tables(mode, pipeline):
    rng = np.random.default_rng(11)
    schema = {0: (5000, [4, 12]), 3: (4000, [8, 100]), 5: (3000, [16])}
    parts = [random_batch(rng, 600, 5000, 2, table=0),
             random_row_batch(rng, 600, 2500, 2, table=3, insert_pct=25),
             random_row_batch(rng, 600, 2000, 1, table=5, delete_pct=30)]
    # interleave the txns of the three tables
    txns = []
    for p in parts:
        for t in range(p.n_txns):
            a, z = int(p.txn_off[t]), int(p.txn_off[t + 1])
            txns.append([(int(p.meta[i] & 3), int((p.meta[i] >> 8) & 15), int((p.meta[i] >> 2) & 63),
                          int(p.rows[i]), int(p.args[i])) for i in range(a, z)])
    order = rng.permutation(len(txns))
    b = Batch.from_txns([txns[i] for i in order])
    with Engine() as e:
        im = _setup(e, schema, {3: 2000, 5: 1500})
        _run_compare(e, im, b, mode, pipeline=pipeline)
        _check_tables(e, im)


def test_inserts_fill_the_table_then_efull():
    schema = {0: (100, [8, 4])}
    ins = lambda n, a0: Batch.from_txns([[(OP_ROW, 0, ROW_INSERT, 0, a0 + i)] for i in range(n)])
    with Engine() as e:
        im = _setup(e, schema, {0: 90})
        _run_compare(e, im, ins(10, 1), MODE_O)
        assert e.table_rows(0) == 100
        with pytest.raises(LtpmgError) as ei:
            e.run(ins(1, 50), MODE_O)
        assert ei.value.code == -7
        with pytest.raises(LtpmgError):
            e.run(ins(1, 51), MODE_A)
        # the refused batches changed nothing; DELETEs and reads still run
        b = Batch.from_txns([[(OP_ROW, 0, ROW_DELETE, 95, 0)], [(OP_READ, 0, 1, 99, 0)]])
        _run_compare(e, im, b, MODE_A)
        _check_tables(e, im)


def test_status_of_an_ordinary_batch_is_its_commit_flags():
    rng = np.random.default_rng(3)
    with Engine() as e:
        e.create_table(0, 1000, [8])
        e.fill_table(0, 1)
        b = random_batch(rng, 3000, 200, 1, max_len=4)
        commit, _ = e.run(b, MODE_A)
        status = e.fetch_status()
        op_txn = np.repeat(np.arange(b.n_txns), np.diff(b.txn_off.astype(np.int64)))
        assert (status == commit[op_txn]).all()
        assert 0 < commit.sum() < b.n_txns


def test_invalid_row_ops_are_rejected():
    with Engine() as e:
        e.create_table(0, 100, [8])
        e.create_append_table(1, 100, [8], 10)
        bad = [
            [(OP_ROW, 0, ROW_INSERT, 0, 1)],          # not an append table
            [(OP_ROW, 1, ROW_INSERT, 3, 1)],          # INSERT names row 0
            [(OP_ROW, 1, ROW_DELETE, 100, 0)],        # row past the capacity
            [(OP_ROW, 1, 2, 1, 0)],                   # no such row op
            [(OP_ROW, 1, ROW_DELETE, 5, 0), (OP_READ, 1, 0, 5, 0)],  # txn reads the row it deletes
        ]
        for txn in bad:
            with pytest.raises(LtpmgError):
                e.submit(Batch.from_txns([txn]))
        ok = Batch.from_txns([[(OP_ROW, 1, ROW_DELETE, 5, 0), (OP_ROW, 1, ROW_INSERT, 0, 9),
                               (OP_READ, 1, 0, 6, 0)]])
        e.run(ok, MODE_O)
        assert e.fetch_status().tolist() == [1, 1, 1]
        assert e.table_rows(1) == 11


@pytest.mark.parametrize("plan", ["default", "plain"])
def test_stream_of_row_op_batches(plan):
    rng = np.random.default_rng(21)
    schema = {0: (20_000, [8, 16]), 1: (8000, [4, 32])}
    batches, modes = [], []
    for i in range(12):
        if i % 3 == 2:
            batches.append(random_batch(rng, int(rng.integers(100, 1500)), 20_000, 2, table=0))
        else:
            batches.append(random_row_batch(rng, int(rng.integers(100, 1500)), 5000, 2, table=1,
                                            insert_pct=12, delete_pct=8))
        modes.append(MODE_O if i % 4 < 2 else MODE_A)
    p = StreamPlan.default() if plan == "default" else StreamPlan.plain()
    with Engine() as e:
        im = _setup(e, schema, {1: 4000})
        st = Stream(e, batches, modes, keep_results=True)
        st.run(p)
        rs = e.result_stride
        for k, (b, m) in enumerate(zip(batches, modes)):
            oc, ores = orc.replay(im, b, m, rs)
            c, r = st.fetch(k)
            assert (c == oc).all(), f"batch {k}: commit flags differ"
            mask = orc.read_mask(b, oc)
            assert (r[mask] == ores[mask]).all(), f"batch {k}: result slots differ"
        _check_tables(e, im)
        st.close()


def test_static_entry_refuses_engines_with_append_tables():
    import torch
    from ltpmg_amd.engine import _Batch
    import ctypes as C
    with Engine() as e:
        e.create_append_table(0, 100, [8], 10)
        dev = torch.device("cuda:0")
        off = torch.tensor([0, 1], dtype=torch.int32, device=dev)
        rows = torch.zeros(1, dtype=torch.int32, device=dev)
        meta = torch.zeros(1, dtype=torch.int16, device=dev)
        args = torch.zeros(1, dtype=torch.int64, device=dev)
        commit = torch.zeros(1, dtype=torch.uint8, device=dev)
        res = torch.zeros(e.result_stride, dtype=torch.uint8, device=dev)
        b = _Batch(1, 1, C.cast(off.data_ptr(), C.POINTER(C.c_uint32)), C.cast(rows.data_ptr(), C.POINTER(C.c_uint32)),
                   C.cast(meta.data_ptr(), C.POINTER(C.c_uint16)), C.cast(args.data_ptr(), C.POINTER(C.c_uint64)))
        rc = e._lib.ltpmg_execute_static(e._h, MODE_O, C.byref(b), C.c_void_p(commit.data_ptr()),
                                         C.c_void_p(res.data_ptr()), None)
        assert rc == -5  # LTPMG_ESTATE
        torch.cuda.synchronize()


def test_ordered_groups_merge_past_inserts_only_below_the_start_cursor():
    # batches that INSERT and then read / write / delete only rows below the
    # cursor at the start of the run merge into one group; a batch that
    # reads a row the group appended starts a new one; every batch equals
    # the oracle's replay in stream order
    rng = np.random.default_rng(31)
    n_present, cap = 3000, 20_000
    low = [random_row_batch(rng, 400, n_present, 2, table=0, insert_pct=20, delete_pct=10)
           for _ in range(5)]
    reader = Batch.from_txns([[(OP_READ, 0, 1, n_present + 3, 0)], [(OP_READ, 0, 0, 17, 0)]])
    batches = low[:3] + [reader] + low[3:]
    plan = StreamPlan(merge_max_ops=1 << 24, fused_max_ops=0, use_graph=False)
    with Engine() as e:
        im = _setup(e, {0: (cap, [8, 16])}, {0: n_present})
        with Stream(e, batches, MODE_O, keep_results=True) as st:
            assert st.groups(plan) == [0, 3]
            st.run(plan)
            rs = e.result_stride
            for k, b in enumerate(batches):
                oc, ores = orc.replay(im, b, MODE_O, rs)
                c, r = st.fetch(k)
                assert (c == oc).all()
                mask = orc.read_mask(b, oc)
                assert (r[mask] == ores[mask]).all(), f"batch {k}: result slots differ"
            _check_tables(e, im)
        assert im.present[0][n_present + 3] == 1  # the reader saw a row an earlier batch appended
