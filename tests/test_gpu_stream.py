"""Batching form on the GPU: a stream of uneven batches resident in HBM, run
under different plans (merged ordered-mode launch groups, the persistent
fused lowest-TID-wins kernel, graph capture and replay, the plain per-batch
form), each checked bit for bit against the oracle's serial replay of the
same batches in stream order."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from ltpmg_amd import MODE_A, MODE_O, OP_RMW_ADD, OP_UPDATE, Engine, LtpmgError, Stream, StreamPlan, gen_kv
from ltpmg_amd.engine import OeStream, create_oe_tables

pytestmark = pytest.mark.gpu

KV = 9  # the key-value table's id beside the nine order-entry tables


def _kv_stream(seed: int, n: int, rows: int, cols: int, *, lo=200, hi=20000, table=0,
               theta=0.9, write_kind=OP_UPDATE):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        t = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        out.append(gen_kv(t, rows, cols, ops=(1, 64), zipf_theta=theta if i % 2 else 0.5,
                          read_pct=50, write_kind=write_kind, table_id=table, seed=seed * 1000 + i))
    return out


def _setup(e: Engine, schema: dict, seed: int) -> orc.Image:
    for t, (n, widths) in sorted(schema.items()):
        e.create_table(t, n, widths)
        e.fill_table(t, seed)
    return orc.Image(schema).fill(seed)


def _check(e: Engine, im: orc.Image, st: Stream, batches, modes, results=True):
    rs = e.result_stride
    expect = []
    for b, (batch, m) in enumerate(zip(batches, modes)):
        oc, ores = orc.replay(im, batch, m, rs)
        expect.append(int(oc.sum()))
        c, r = st.fetch(b, results=results)
        bad = np.nonzero(c != oc)[0]
        assert bad.size == 0, f"batch {b}: commit flags differ at txns {bad[:8]}"
        if results:
            mask = orc.read_mask(batch, oc)
            assert (r[mask] == ores[mask]).all(), f"batch {b}: read results differ"
    assert st.committed().tolist() == expect
    for t in im.schema:
        assert e.digest(t) == orc.digest_columns(t, im.cols[t]), f"table {t} differs"


def test_ordered_stream_merged_into_one_group():
    schema = {0: (30_000, [100] * 6 + [4, 8])}
    batches = _kv_stream(1, 10, 30_000, 8, lo=100, hi=6000, write_kind=OP_RMW_ADD)
    modes = [MODE_O] * len(batches)
    with Engine() as e:
        im = _setup(e, schema, 21)
        with Stream(e, batches, MODE_O, keep_results=True) as st:
            plan = StreamPlan(merge_max_ops=1 << 28)
            assert st.groups(plan) == [0]
            assert st.groups(StreamPlan(merge_max_ops=0)) == list(range(len(batches)))
            st.run(plan)
            _check(e, im, st, batches, modes)


@pytest.mark.parametrize("tile", [1024, 2048, 4096])
def test_lowest_tid_wins_stream_fused_kernel(tile):
    schema = {0: (50_000, [100] * 4 + [8])}
    # one large batch spreads several tiles over each workgroup
    batches = _kv_stream(2, 6, 50_000, 5, lo=50, hi=4000, theta=0.6)
    batches.append(gen_kv(14_000, 50_000, 5, ops=(1, 64), zipf_theta=0.3, seed=77))
    modes = [MODE_A] * len(batches)
    with Engine() as e:
        im = _setup(e, schema, 22)
        with Stream(e, batches, MODE_A, keep_results=True) as st:
            plan = StreamPlan(fused_max_ops=1 << 30, fused_tile_ops=tile, fused_wgs_per_cu=1)
            st.run(plan)
            _check(e, im, st, batches, modes)


def test_mixed_modes_plain_and_planned_agree():
    schema = {0: (20_000, [100] * 3 + [4]), 1: (5000, [16, 32])}
    b0 = _kv_stream(3, 6, 20_000, 4, lo=100, hi=5000)
    b1 = _kv_stream(4, 4, 5000, 2, lo=100, hi=3000, table=1, write_kind=OP_RMW_ADD)
    batches = [b0[0], b1[0], b0[1], b0[2], b1[1], b1[2], b0[3], b0[4], b1[3], b0[5]]
    modes = [MODE_O, MODE_O, MODE_A, MODE_O, MODE_O, MODE_A, MODE_A, MODE_O, MODE_O, MODE_A]
    for plan in (StreamPlan.plain(), StreamPlan.default(),
                 StreamPlan(merge_max_ops=1 << 16, fused_max_ops=1 << 16, use_graph=False)):
        with Engine() as e:
            im = _setup(e, schema, 23)
            with Stream(e, batches, modes, keep_results=True) as st:
                st.run(plan)
                _check(e, im, st, batches, modes)


def test_graph_replays_continue_from_the_tables():
    schema = {0: (8000, [100, 100, 8])}
    batches = _kv_stream(5, 8, 8000, 3, lo=100, hi=3000, write_kind=OP_RMW_ADD)
    modes = [MODE_A, MODE_O] * 4
    with Engine() as e:
        im = _setup(e, schema, 24)
        with Stream(e, batches, modes, keep_results=True) as st:
            plan = StreamPlan(merge_max_ops=1 << 20, fused_max_ops=1 << 20, use_graph=True)
            for _ in range(3):
                st.run(plan)
                _check(e, im, st, batches, modes)


def test_order_entry_and_kv_in_one_stream():
    # order-entry batches with INSERTs and Delivery DELETEs between
    # key-value batches; a rewinding stream run twice gives the same rows
    W, cap, pre = 2, 60_000, 50
    oe = OeStream(W, neworder_pct=40, payment_pct=40, delivery_pct=10, preload=pre, seed=9)
    batches, modes = [], []
    kv = _kv_stream(6, 4, 40_000, 5, lo=300, hi=6000, table=KV)
    for i in range(4):
        batches += [oe.next(1500 + 700 * i), kv[i]]
        modes += [MODE_A if i % 2 else MODE_O, MODE_O]
    with Engine() as e:
        schema, append = create_oe_tables(e, W, cap, 25, preload=pre)
        schema = dict(schema)
        schema[KV] = (40_000, [100] * 5)
        e.create_table(KV, 40_000, [100] * 5)
        e.fill_table(KV, 25)
        im = orc.Image(schema, append=append).fill(25)
        with Stream(e, batches, modes, keep_results=True, rewind=True) as st:
            st.run(StreamPlan(merge_max_ops=1 << 22, fused_max_ops=1 << 22))
            _check(e, im, st, batches, modes)
            for t in append:
                assert e.table_rows(t) == im.cursor[t] and (e.read_presence(t) == im.present[t]).all()
            cursors = {t: e.table_rows(t) for t in append}
            st.run(StreamPlan.plain())  # rewinds to the rows at creation, appends the same count
            assert {t: e.table_rows(t) for t in append} == cursors


def test_tuner_picks_a_plan_that_runs_correctly():
    schema = {0: (20_000, [100] * 4)}
    batches = _kv_stream(7, 8, 20_000, 4, lo=100, hi=8000)
    modes = [MODE_A, MODE_O, MODE_O, MODE_A, MODE_O, MODE_A, MODE_O, MODE_O]
    with Engine() as e:
        im = _setup(e, schema, 26)
        with Stream(e, batches, modes, keep_results=True) as st:
            plan, trials = st.tune()
            assert trials >= 2 and plan.use_graph
            assert plan.fused_tile_ops in (1024, 2048, 4096)
            e.fill_table(0, 26)  # the trials applied the stream's writes
            st.run()  # the tuned plan
            _check(e, im, st, batches, modes)


def test_results_kept_only_on_request_and_bad_input_rejected():
    schema = {0: (5000, [100, 4])}
    batches = _kv_stream(8, 3, 5000, 2, lo=100, hi=1000)
    with Engine() as e:
        im = _setup(e, schema, 27)
        with Stream(e, batches, MODE_O) as st:
            with pytest.raises(LtpmgError):
                st.fetch(0)  # nothing ran yet
            st.run()
            with pytest.raises(LtpmgError):
                st.fetch(0, results=True)
            _check(e, im, st, batches, [MODE_O] * 3, results=False)
            with pytest.raises(LtpmgError):
                st.run(StreamPlan(fused_tile_ops=3000))
        with pytest.raises(LtpmgError):
            Stream(e, batches, [MODE_O, 7, MODE_O])
