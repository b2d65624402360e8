"""Bit-exact parity at the sizes the benchmark reports.

  C2  the exact key-value batch bench.py times: 10,000,000 rows x 10 fields
      of 100 B, 2^20 txns x 16 ops, Zipf 0.99, 50% READ / 50% UPDATE (both
      modes), plus the same shape with uniform keys in the lowest-TID-wins
      mode, where ~10^5 txns commit;
  C3  one order-entry batch over W = 64 partitions, its NewOrders and
      Payments appending rows (both modes);
  C4  a stream of 20 uneven batches, 1,000 to 1,000,000 txns, alternately
      the 10M-row key-value table (1-64 ops, Zipf 0.5 / 0.99) and W = 64
      order entry, run under the plan the tuner picks (both modes);
  C5  a served run over W = 1024 order entry beside a 100M-row key-value
      table (about 176 GiB of tables), every recorded batch replayed.

Every check is against the oracle's serial replay on a sparse image that
holds only the cells the batches touch: commit flags, every defined READ and
INSERT result, every touched cell (C2), each whole table through the
engine's additive digest, digest(after) - digest(before) = the sum over the
touched cells of term(after) - term(before), which covers the untouched
cells too, and the append tables' cursors and presence bytes.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

import oracle_ffi as orc
from ltpmg_amd import MODE_A, MODE_O, OP_UPDATE, Engine, Stream, gen_kv
from ltpmg_amd.engine import OeStream, oe_append, oe_schema, take_txns

pytestmark = pytest.mark.gpu

TABLE_SEED = 0x4C54504D47000001
BATCH_SEED = 0x4C54504D47000002
KV_ROWS, KV_COLS, KV_WIDTH = 10_000_000, 10, 100
C5_PRELOAD = 100  # orders per district at W = 1024: the ordered mode's 2^32-cell key space
M64 = (1 << 64) - 1


def log(msg: str) -> None:
    print(f"[fullsize] {msg}", flush=True)


def _create(e: Engine, schema: dict, append: dict | None = None) -> dict:
    for t, (n, widths) in sorted(schema.items()):
        if append and t in append:
            e.create_append_table(t, n, widths, append[t])
        else:
            e.create_table(t, n, widths)
        e.fill_table(t, TABLE_SEED)
    return {t: e.digest(t) for t in schema}


def _check_append(e: Engine, im: orc.Image, what: str) -> None:
    for t in im.present:
        assert e.table_rows(t) == im.cursor[t], f"{what}: table {t} cursor"
        bad = np.nonzero(e.read_presence(t) != im.present[t])[0]
        assert bad.size == 0, f"{what}: table {t} presence differs at rows {bad[:8]}"


def _compare(batch, mode, commit, res, oc, ores, what: str) -> None:
    bad = np.nonzero(commit != oc)[0]
    assert bad.size == 0, f"{what}: commit flags differ at txns {bad[:10]}"
    m = orc.read_mask(batch, oc)
    idx = np.nonzero(m)[0]
    for lo in range(0, idx.size, 1 << 20):  # bounded temporaries on 2 GB result arrays
        part = idx[lo:lo + (1 << 20)]
        diff = np.any(res[part] != ores[part], axis=1)
        if diff.any():
            pytest.fail(f"{what}: {int(diff.sum())} READ results differ, first ops {part[diff][:8]}")


def _snapshot(im: orc.Image, cells: dict) -> dict:
    return {k: im.cols[k[0]][k[1]][rows].copy() for k, rows in cells.items()}


def _digest_delta(im: orc.Image, cells: dict, before: dict) -> dict:
    out = {}
    for (t, c), rows in cells.items():
        d = orc.digest_rows(t, c, rows, im.cols[t][c][rows]) - orc.digest_rows(t, c, rows, before[(t, c)])
        out[t] = (out.get(t, 0) + d) & M64
    return out


def _check_digests(e: Engine, d0: dict, delta: dict) -> None:
    for t, before in d0.items():
        got = (e.digest(t) - before) & M64
        assert got == delta.get(t, 0), f"table {t}: digest change differs from the oracle's"


def _run_kv_c2(batch, mode: int, check_cells: bool = True):
    schema = {0: (KV_ROWS, [KV_WIDTH] * KV_COLS)}
    with Engine() as e:
        t0 = time.time()
        d0 = _create(e, schema)
        e.submit(batch)
        e.execute(mode)
        commit, res = e.fetch()
        st = e.stats()
        log(f"mode {mode}: {st['committed']}/{batch.n_txns} committed, {st['exec_ms']:.3f} ms device, "
            f"engine side {time.time() - t0:.1f} s")
        t0 = time.time()
        im, cells = orc.sparse_image(schema, TABLE_SEED, [batch])
        before = _snapshot(im, cells)
        oc, ores = orc.replay(im, batch, mode, e.result_stride)
        log(f"oracle: {sum(r.size for r in cells.values())} touched cells, {time.time() - t0:.1f} s")
        _compare(batch, mode, commit, res, oc, ores, f"C2 mode {mode}")
        del res, ores
        if check_cells:
            for c in range(KV_COLS):
                rows = cells.get((0, c))
                if rows is None:
                    continue
                got = e.read_column(0, c)
                wrong = np.nonzero(np.any(got[rows] != im.cols[0][c][rows], axis=1))[0]
                assert wrong.size == 0, f"column {c}: {wrong.size} touched cells differ, rows {rows[wrong[:8]]}"
                del got
        _check_digests(e, d0, _digest_delta(im, cells, before))
        assert st["committed"] == int(oc.sum())
    return oc


@pytest.mark.parametrize("mode", [MODE_O, MODE_A], ids=["ordered", "lowest_tid_wins"])
def test_c2_full_size(mode):
    batch = gen_kv(1 << 20, KV_ROWS, KV_COLS, ops=16, zipf_theta=0.99, read_pct=50,
                   write_kind=OP_UPDATE, seed=BATCH_SEED)
    assert batch.n_ops == 16 << 20
    oc = _run_kv_c2(batch, mode)
    if mode == MODE_O:
        assert oc.all()


def test_c2_full_size_uniform_keys_lowest_tid_wins():
    # uniform keys: every row is written by ~0.84 txns of the batch, so a txn
    # at position t commits with probability ~exp(-128 t / 10^7) and about
    # 10^7 / 128 txns commit -- a commit set with real content at full size
    batch = gen_kv(1 << 20, KV_ROWS, KV_COLS, ops=16, zipf_theta=0.0, read_pct=50,
                   write_kind=OP_UPDATE, seed=BATCH_SEED + 1)
    oc = _run_kv_c2(batch, MODE_A)
    assert 20_000 < int(oc.sum()) < 200_000


@pytest.mark.parametrize("mode", [MODE_O, MODE_A], ids=["ordered", "lowest_tid_wins"])
def test_c3_order_entry_w64(mode):
    W, cap = 64, 1 << 18
    schema, append = oe_schema(W, cap), oe_append(W, cap)
    batch = OeStream(W, seed=BATCH_SEED + 3, byname_pct=15, table_seed=TABLE_SEED).next(1 << 18)
    with Engine() as e:
        d0 = _create(e, schema, append)
        e.submit(batch)
        e.execute(mode)
        commit, res = e.fetch()
        status = e.fetch_status()
        st = e.stats()
        log(f"C3 mode {mode}: {st['committed']}/{batch.n_txns} committed, {batch.n_ops} ops, "
            f"{st['inserted']} rows appended, {st['exec_ms']:.3f} ms device")
        im, cells = orc.sparse_image(schema, TABLE_SEED, [batch], append)
        before = _snapshot(im, cells)
        oc, ores, found = orc.replay(im, batch, mode, e.result_stride, with_found=True)
        _compare(batch, mode, commit, res, oc, ores, f"C3 mode {mode}")
        assert (status == found).all()
        _check_digests(e, d0, _digest_delta(im, cells, before))
        _check_append(e, im, f"C3 mode {mode}")
        ins = (batch.meta & 0xFF) == 3
        assert st["inserted"] == int(found[ins].sum())
    if mode == MODE_O:  # an order, its new-order row and 5-15 lines per NewOrder
        assert st["inserted"] > batch.n_txns
    else:
        assert 0 < oc.sum() < len(oc) and st["inserted"] > 0


def _c4_batches():
    """20 batches: 1,000 and 1,000,000 txns at positions 0 and 7, the others
    floor(exp(U(ln 1e3, ln 3e5))) txns (about 2M txns, 70M ops in all, so the
    serial oracle finishes in well under a minute); even positions key-value
    (Zipf 0.5 and 0.99 in turn), odd positions order entry."""
    rng = np.random.default_rng(BATCH_SEED + 4)
    n = 20
    sizes = np.floor(np.exp(rng.uniform(np.log(1e3), np.log(3e5), n))).astype(np.int64)
    sizes[0], sizes[7] = 1000, 1_000_000
    W, cap = 64, int(sizes.sum())
    oe = OeStream(W, seed=BATCH_SEED + 5)
    batches = []
    for i, s in enumerate(sizes):
        if i % 2 == 0:
            theta = 0.5 if i % 4 == 0 else 0.99
            batches.append(gen_kv(int(s), KV_ROWS, KV_COLS, ops=(1, 64), zipf_theta=theta, read_pct=50,
                                  write_kind=OP_UPDATE, table_id=9, seed=BATCH_SEED + 100 + i))
        else:
            batches.append(oe.next(int(s)))
    schema = dict(oe_schema(W, cap))
    schema[9] = (KV_ROWS, [KV_WIDTH] * KV_COLS)
    return batches, schema, oe_append(W, cap)


@pytest.mark.parametrize("mode", [MODE_O, MODE_A], ids=["ordered", "lowest_tid_wins"])
def test_c4_stream_under_the_tuned_plan(mode):
    batches, schema, append = _c4_batches()
    log(f"C4: {len(batches)} batches, {sum(b.n_txns for b in batches)} txns, "
        f"{sum(b.n_ops for b in batches)} ops")
    with Engine() as e:
        _create(e, schema, append)
        with Stream(e, batches, mode, keep_results=True, rewind=True) as st:
            plan, trials = st.tune()
            log(f"C4 mode {mode}: tuned plan {plan} after {trials} runs, {len(st.groups())} groups")
            for t in sorted(schema):  # the tuning runs applied the stream's writes
                e.fill_table(t, TABLE_SEED)
            d0 = {t: e.digest(t) for t in schema}
            ms = st.run()
            log(f"C4 mode {mode}: one tuned pass {ms:.1f} ms device")
            im, cells = orc.sparse_image(schema, TABLE_SEED, batches, append)
            before = _snapshot(im, cells)
            rs = e.result_stride
            for k, b in enumerate(batches):
                commit, res = st.fetch(k)
                oc, ores = orc.replay(im, b, mode, rs)
                _compare(b, mode, commit, res, oc, ores, f"C4 batch {k} ({b.n_txns} txns)")
                del res, ores
            _check_digests(e, d0, _digest_delta(im, cells, before))
            _check_append(e, im, f"C4 mode {mode}")


def test_c5_served_run_replayed():
    # the served form of bench.py's large-schema leg, recorded; every batch it
    # cut is replayed in cut order through the oracle
    W, cap, pre, kv_rows = 1024, 1 << 20, C5_PRELOAD, 100_000_000
    schema = dict(oe_schema(W, cap, pre))
    schema[9] = (kv_rows, [KV_WIDTH] * KV_COLS)
    append = oe_append(W, cap, pre)
    pool = OeStream(W, seed=BATCH_SEED + 9, byname_pct=15, table_seed=TABLE_SEED, preload=pre).next(1 << 20)
    with Engine(n_slots=2, hbm_budget=240 << 30) as e:
        t0 = time.time()
        d0 = _create(e, schema, append)
        log(f"C5: {e.stats_hbm() / 2**30:.1f} GiB allocated in {time.time() - t0:.1f} s")
        # recording copies every batch's result slots back, so the run is
        # host-bound; batches of at most 2^14 keep it to 64 or more of them
        out = e.serve(pool, arrival_rate=2e7, n_arrivals=pool.n_txns, max_batch=1 << 14, max_wait_us=500.0,
                      n_inflight=2, mode=MODE_O, seed=11, record=True, record_results=True)
        assert not out["record_overflow"] and out["committed"] == pool.n_txns
        off, ids, flags = out["batch_off"], out["batch_txns"], out["batch_commit"]
        log(f"C5: {out['batches']} batches, p50 {out['p50_ms']:.2f} ms, p99 {out['p99_ms']:.2f} ms")
        assert out["batches"] >= 50
        batches = [take_txns(pool, ids[off[k]:off[k + 1]] % pool.n_txns) for k in range(len(off) - 1)]
        oe_only = {t: schema[t] for t in schema if t != 9}
        im, cells = orc.sparse_image(oe_only, TABLE_SEED, batches, append)
        before = _snapshot(im, cells)
        rs = e.result_stride
        results = out["results"]
        for k, b in enumerate(batches):
            oc, ores = orc.replay(im, b, MODE_O, rs)
            assert (flags[off[k]:off[k + 1]] == oc).all(), f"batch {k}: commit flags differ"
            sel = ids[off[k]:off[k + 1]] % pool.n_txns
            lens = np.diff(pool.txn_off.astype(np.int64))[sel]
            pool_op = np.repeat(pool.txn_off[sel].astype(np.int64) - np.concatenate([[0], np.cumsum(lens)[:-1]]),
                                lens) + np.arange(int(lens.sum()))
            m = orc.read_mask(b, oc)
            assert (results[pool_op[m]] == ores[m]).all(), f"batch {k}: served READ results differ"
        _check_digests(e, {t: d0[t] for t in oe_only}, _digest_delta(im, cells, before))
        _check_append(e, im, "C5")
        assert e.digest(9) == d0[9]  # the order-entry mix never names the key-value table
