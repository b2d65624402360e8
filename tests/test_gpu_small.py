"""The one-launch ordered path: a batch of at most LTPMG_SMALL_MAX_OPS ops
runs as one workgroup (LDS radix sort, segmented scan, cell moves).  Each
case is checked bit for bit against the oracle's serial replay (commit
flags, every READ result, the whole final table image) and against the
multi-kernel pipeline on the same input, and the per-kernel record shows
which path ran."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from batches import OP_READ, OP_RMW_ADD, OP_UPDATE, random_batch, random_row_batch
from ltpmg_amd import MODE_A, MODE_O, Batch, Engine, Stream, StreamPlan, gen_kv
from ltpmg_amd.engine import SMALL_MAX_OPS

pytestmark = pytest.mark.gpu


def log(msg: str) -> None:
    print(f"[small] {msg}", flush=True)


def _setup(e: Engine, schema: dict, seed: int) -> orc.Image:
    for t, (n, widths) in sorted(schema.items()):
        e.create_table(t, n, widths)
        e.fill_table(t, seed)
    return orc.Image(schema).fill(seed)


def _run(schema: dict, seed: int, b: Batch, pipeline: bool):
    """Runs b on a fresh engine; returns commit flags, results, every column
    and the kernels that ran, after checking them against the oracle."""
    with Engine() as e:
        im = _setup(e, schema, seed)
        e.submit(b)
        e.execute(MODE_O, profile=True, pipeline=pipeline)
        commit, res = e.fetch()
        kernels = set(e.stats()["kernels"])
        oc, ores = orc.replay(im, b, MODE_O, e.result_stride)
        bad = np.nonzero(commit != oc)[0]
        assert bad.size == 0, f"commit flags differ at txns {bad[:10]}"
        m = orc.read_mask(b, oc)
        if not (res[m] == ores[m]).all():
            idx = np.nonzero(m)[0]
            wrong = idx[np.any(res[m] != ores[m], axis=1)]
            pytest.fail(f"{wrong.size} READ results differ, first ops {wrong[:8]}")
        cols = {}
        for t, (n, widths) in im.schema.items():
            for c in range(len(widths)):
                got = e.read_column(t, c)
                if not (got == im.cols[t][c]).all():
                    rows = np.nonzero(np.any(got != im.cols[t][c], axis=1))[0]
                    pytest.fail(f"table {t} col {c}: {rows.size} rows differ, first {rows[:8]}")
                cols[(t, c)] = got
            assert e.digest(t) == orc.digest_columns(t, im.cols[t])
        return commit, np.where(m[:, None], res, 0), cols, kernels


def _both_paths(schema: dict, seed: int, b: Batch) -> None:
    log(f"{b.n_txns} txns, {b.n_ops} ops")
    one = _run(schema, seed, b, pipeline=False)
    pipe = _run(schema, seed, b, pipeline=True)
    assert one[3] == {"mo_small"}, f"one-launch path not taken: {one[3]}"
    assert "mo_small" not in pipe[3] and "mo_apply_reads" in pipe[3]
    assert (one[0] == pipe[0]).all() and (one[1] == pipe[1]).all()
    for k in one[2]:
        assert (one[2][k] == pipe[2][k]).all(), f"column {k} differs between the paths"


def _single_op_txns(rng, n_ops: int, n_rows: int, n_cols: int, hot: float = 0.0):
    txns = []
    for _ in range(n_ops):
        row = 0 if rng.random() < hot else int(rng.integers(0, n_rows))
        kind = int(rng.integers(0, 3))
        arg = int(rng.integers(0, 2**63)) if kind == OP_UPDATE else int(rng.integers(-3, 4)) & (2**64 - 1)
        txns.append([(kind, 0, int(rng.integers(0, n_cols)), row, arg)])
    return Batch.from_txns(txns)


@pytest.mark.parametrize("n_ops", [1, 2, 63, 64, 65, 1000, 1024, 1025, 4096, 8191, SMALL_MAX_OPS])
def test_sizes_match_oracle_and_pipeline(n_ops):
    # one-op txns, a third of them on one hot row, so segments cross chunk,
    # wave and thread boundaries of the one-launch kernel
    rng = np.random.default_rng(n_ops)
    schema = {0: (3000, [100, 4, 16])}
    _both_paths(schema, n_ops, _single_op_txns(rng, n_ops, 3000, 3, hot=0.3))


@pytest.mark.parametrize("theta", [None, 0.99])
@pytest.mark.parametrize("kinds", ["read_update", "read_rmw", "all"])
def test_mixed_txns(theta, kinds):
    rng = np.random.default_rng(7 + len(kinds) + (0 if theta is None else 1))
    ks = {"read_update": (OP_READ, OP_UPDATE), "read_rmw": (OP_READ, OP_RMW_ADD),
          "all": (OP_READ, OP_UPDATE, OP_RMW_ADD)}[kinds]
    schema = {0: (5000, [4, 8, 100, 32])}
    b = random_batch(rng, n_txns=1200, n_rows=5000, n_cols=4, max_len=12, kinds=ks, zipf=theta)
    b = b if b.n_ops <= SMALL_MAX_OPS else Batch(b.txn_off[:501], b.rows[:b.txn_off[500]],
                                                 b.meta[:b.txn_off[500]], b.args[:b.txn_off[500]])
    _both_paths(schema, 31, b)


def test_one_hot_cell():
    # every op on one cell: one segment of 8192 elements, key bits 0 (one
    # table of one row and column: no sort pass at all)
    rng = np.random.default_rng(3)
    txns = []
    for t in range(SMALL_MAX_OPS):
        kind = [OP_READ, OP_UPDATE, OP_RMW_ADD][int(rng.integers(0, 3))]
        txns.append([(kind, 0, 0, 0, int(rng.integers(0, 2**40)))])
    _both_paths({0: (1, [100])}, 5, Batch.from_txns(txns))


def test_kv_zipf_batch():
    # the C2 workload's shape at 512 txns x 16 ops
    b = gen_kv(512, 1_000_000, 10, ops=16, zipf_theta=0.99, read_pct=50, write_kind=OP_UPDATE, seed=99)
    assert b.n_ops == SMALL_MAX_OPS
    _both_paths({0: (1_000_000, [100] * 10)}, 0x4C54504D47000001, b)


@pytest.mark.parametrize("widths", [[8, 4], [12, 4], [40, 8], [8, 96], [148, 4, 20]],
                         ids=["stride8", "stride16", "stride64", "stride96", "stride160"])
def test_result_strides(widths):
    rng = np.random.default_rng(len(widths) * 100 + widths[0])
    schema = {0: (700, widths)}
    b = random_batch(rng, n_txns=900, n_rows=700, n_cols=len(widths), max_len=8, zipf=0.9)
    assert b.n_ops <= SMALL_MAX_OPS
    _both_paths(schema, 0x5EED + widths[0], b)


def test_tables_past_the_column_list():
    # 5 tables x 64 columns: the last 64 columns' descriptors come from
    # global memory, not the kernel's LDS column list
    widths = [[4, 16, 100, 8][c % 4] for c in range(64)]
    schema = {t: (40 + t, widths[t:] + widths[:t]) for t in range(5)}
    rng = np.random.default_rng(17)
    txns = []
    for _ in range(1500):
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
    assert b.n_ops <= SMALL_MAX_OPS
    _both_paths(schema, 0x4C54504D4700000B, b)


def test_larger_batches_take_the_pipeline():
    rng = np.random.default_rng(11)
    big = _single_op_txns(rng, SMALL_MAX_OPS + 1, 3000, 3)
    _, _, _, kernels = _run({0: (3000, [100, 4, 16])}, 1, big, pipeline=False)
    assert "mo_small" not in kernels and "mo_apply_reads" in kernels


def _run_rows(schema: dict, append: dict, batches, pipeline: bool):
    """Runs the batches in order on fresh append tables; checks each against
    the oracle (commit flags, per-op status, result slots, inserted count)
    and returns what every batch left: statuses, results, cursors, presence
    and columns."""
    with Engine() as e:
        for t, (n, widths) in sorted(schema.items()):
            if t in append:
                e.create_append_table(t, n, widths, append[t])
            else:
                e.create_table(t, n, widths)
            e.fill_table(t, 0x5EED0006)
        im = orc.Image(schema, append=append).fill(0x5EED0006)
        out = []
        for b in batches:
            e.submit(b)
            e.execute(MODE_O, profile=True, pipeline=pipeline)
            commit, res = e.fetch()
            status = e.fetch_status()
            st = e.stats()
            assert ("mo_small" in st["kernels"]) == (not pipeline)
            oc, ores, ofound = orc.replay(im, b, MODE_O, e.result_stride, with_found=True)
            assert (commit == oc).all()
            bad = np.nonzero(status != ofound)[0]
            assert bad.size == 0, f"status differs at ops {bad[:10]}"
            m = orc.read_mask(b, oc)
            if not (res[m] == ores[m]).all():
                idx = np.nonzero(m)[0]
                wrong = idx[np.any(res[m] != ores[m], axis=1)]
                pytest.fail(f"{wrong.size} result slots differ, first ops {wrong[:8]}")
            is_row = (b.meta & 3) == 3
            assert st["inserted"] == int((ofound & is_row & (((b.meta >> 2) & 63) == 0)).sum())
            out.append((status, np.where(m[:, None], res, 0)))
        tables = {}
        for t, (n, widths) in im.schema.items():
            for c in range(len(widths)):
                got = e.read_column(t, c)
                assert (got == im.cols[t][c]).all(), f"table {t} col {c} differs from the oracle"
                tables[(t, c)] = got
            if t in im.present:
                assert e.table_rows(t) == im.cursor[t]
                assert (e.read_presence(t) == im.present[t]).all()
                tables[(t, "present")] = e.read_presence(t)
        return out, tables


@pytest.mark.parametrize("seed", range(3))
def test_row_op_batches_match_oracle_and_pipeline(seed):
    # INSERTs, DELETEs (several of one row in a batch), cell ops on an append
    # table and an ordinary one, batch after batch, both paths
    rng = np.random.default_rng(100 + seed)
    schema = {0: (3000, [4, 12]), 2: (6000, [8, 100, 16]), 4: (2500, [16])}
    append = {2: 2500, 4: 1200}
    batches = []
    for k in range(4):
        parts = [random_batch(rng, 150, 3000, 2, table=0),
                 random_row_batch(rng, 500, 2500 + 300 * k, 3, table=2, insert_pct=20, delete_pct=15),
                 random_row_batch(rng, 300, 1200, 1, table=4, delete_pct=35)]
        txns = []
        for p in parts:
            for t in range(p.n_txns):
                a, z = int(p.txn_off[t]), int(p.txn_off[t + 1])
                txns.append([(int(p.meta[i] & 3), int((p.meta[i] >> 8) & 15), int((p.meta[i] >> 2) & 63),
                              int(p.rows[i]), int(p.args[i])) for i in range(a, z)])
        b = Batch.from_txns([txns[i] for i in rng.permutation(len(txns))])
        assert b.n_ops <= SMALL_MAX_OPS
        batches.append(b)
    log(f"{len(batches)} row-op batches, {[b.n_ops for b in batches]} ops")
    one = _run_rows(schema, append, batches, pipeline=False)
    pipe = _run_rows(schema, append, batches, pipeline=True)
    for (s1, r1), (s2, r2) in zip(one[0], pipe[0]):
        assert (s1 == s2).all() and (r1 == r2).all()
    for k in one[1]:
        assert (one[1][k] == pipe[1][k]).all(), f"{k} differs between the paths"


def test_stream_of_small_batches():
    # small ordered batches under the tuned plan, under one-launch groups
    # without merging, and merged up to the one-launch size, each against the
    # oracle's replay of the batches in order
    rng = np.random.default_rng(23)
    schema = {0: (20_000, [100] * 4 + [4, 8])}
    batches = []
    for i in range(24):
        n = int(rng.integers(20, 500))
        batches.append(gen_kv(n, 20_000, 6, ops=(1, 16), zipf_theta=0.9 if i % 2 else 0.5, read_pct=50,
                              write_kind=OP_RMW_ADD if i % 3 == 0 else OP_UPDATE, seed=500 + i))
    modes = [MODE_O] * len(batches)
    plans = [StreamPlan(merge_max_ops=0, small_max_ops=SMALL_MAX_OPS, use_graph=False),
             StreamPlan(merge_max_ops=SMALL_MAX_OPS, small_max_ops=SMALL_MAX_OPS, use_graph=True),
             None]
    for plan in plans:
        with Engine() as e:
            im = _setup(e, schema, 41)
            with Stream(e, batches, MODE_O, keep_results=True) as st:
                if plan is None:
                    plan, trials = st.tune()
                    log(f"tuned plan {plan} after {trials} runs")
                    # the tuning runs applied the stream's writes; start over
                    for t, (n, widths) in schema.items():
                        e.fill_table(t, 41)
                groups = st.groups(plan)
                log(f"plan {plan}: {len(groups)} groups")
                st.run(plan)
                rs = e.result_stride
                for b, batch in enumerate(batches):
                    oc, ores = orc.replay(im, batch, MODE_O, rs)
                    c, r = st.fetch(b, results=True)
                    assert (c == oc).all(), f"batch {b}: commit flags differ"
                    mask = orc.read_mask(batch, oc)
                    assert (r[mask] == ores[mask]).all(), f"batch {b}: read results differ"
                for t in schema:
                    assert e.digest(t) == orc.digest_columns(t, im.cols[t]), f"table {t} differs"


@pytest.mark.parametrize("mode,pipeline", [(MODE_O, False), (MODE_O, True), (MODE_A, False)],
                         ids=["ordered_one_launch", "ordered_pipeline", "lowest_tid_wins"])
def test_result_slots_are_zero_past_the_column(mode, pipeline):
    # The ordered kernels write a READ's result slot only up to its column's
    # slot and the fetch zeroes the rest on the host: a 4-byte READ after a
    # 256-byte READ in the same result slot must come back with zeros past
    # its 4 bytes, in both paths (and in lowest-TID-wins, which writes whole
    # slots).
    schema = {0: (50, [256, 4, 24])}
    wide = Batch.from_txns([[(OP_READ, 0, 0, r, 0)] for r in range(20)])
    narrow = Batch.from_txns([[(OP_READ, 0, 1 + r % 2, r, 0)] for r in range(20)])
    with Engine() as e:
        im = _setup(e, schema, 9)
        stride = e.result_stride
        assert stride == 256
        for b in (wide, narrow, wide, narrow):
            commit, res = e.run(b, mode, pipeline=pipeline)
            oc, ores = orc.replay(im, b, mode, stride)
            assert commit.all()
            assert (res == ores).all(), "result slots differ from the oracle's (padding included)"
        for r in range(20):
            w = 4 if r % 2 == 0 else 24
            assert not res[r, w:].any(), f"op {r}: bytes past its {w}-byte column are not zero"
