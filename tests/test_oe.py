"""Order-entry adapter (include/ltpmg_oe.h): generator and decomposition
properties on the CPU, and GPU parity against the oracle over the nine-table
schema with its append tables (INSERTs of NewOrder and Payment, DELETEs of
Delivery)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from ltpmg_amd import MODE_A, MODE_O, Engine
from ltpmg_amd.engine import (OP_READ, OP_ROW, OP_UPDATE, ROW_DELETE, ROW_INSERT, OeStream, create_oe_tables,
                              oe_append, oe_last_name_index, oe_schema)

WAREHOUSE, DISTRICT, CUSTOMER, NEWORDER, HISTORY, ORDER, ORDERLINE, STOCK, ITEM = range(9)
C_BALANCE, C_LAST, C_DELIVERY_CNT = 7, 11, 4
O_CARRIER_ID, OL_DELIVERY_D, OL_AMOUNT = 4, 6, 8
PRE_LINES = 5
N_NAMES = 15


def _name_codes(seed: int, rows: np.ndarray) -> np.ndarray:
    """Last-name code of customer rows from the oracle's fill of the c_last
    column: first little-endian u64 of the cell mod 15."""
    im = orc.Image({CUSTOMER: (int(rows.max()) + 1, [4] * 11 + [32])})
    im.fill_cells(seed, CUSTOMER, C_LAST, rows)
    cells = im.cols[CUSTOMER][C_LAST][rows]
    return (cells[:, :8].copy().view("<u8")[:, 0] % N_NAMES).astype(np.int64)


def test_last_name_index_matches_the_filled_cells():
    seed, W = 0x4C54504D47000007, 2
    rows = np.arange(W * 10 * 3000, dtype=np.uint32)
    codes = _name_codes(seed, rows).reshape(W * 10, 3000)
    expect = np.zeros((W * 10, N_NAMES), np.uint32)
    for dist in range(W * 10):
        for k in range(N_NAMES):
            hit = np.nonzero(codes[dist] == k)[0]
            expect[dist, k] = hit[0] if hit.size else 0  # first match, C_ID 0 when none
    assert (oe_last_name_index(W, seed) == expect).all()


def test_payments_by_last_name_address_the_first_match():
    seed, W = 0x4C54504D47000008, 3
    first = oe_last_name_index(W, seed)
    by_name = OeStream(W, neworder_pct=0, payment_pct=100, seed=5, byname_pct=100,
                       table_seed=seed).next(600)
    by_id = OeStream(W, neworder_pct=0, payment_pct=100, seed=5).next(600)

    def customers(b):
        sel = (((b.meta >> 8) & 15) == CUSTOMER) & (((b.meta >> 2) & 63) == C_BALANCE)
        return b.rows[sel].astype(np.int64)

    for b, want in ((by_name, True), (by_id, False)):
        r = customers(b)
        assert r.size == b.n_txns
        dist, cid = r // 3000, r % 3000
        hit = first[dist, _name_codes(seed, r.astype(np.uint32))] == cid
        assert hit.all() if want else not hit.all()


def test_schema_shapes():
    s = oe_schema(64, 100_000)
    pre = 64 * 10 * 3000
    assert s[WAREHOUSE][0] == 64 and s[DISTRICT][0] == 640
    assert s[CUSTOMER][0] == 64 * 30_000 and s[STOCK][0] == 64 * 100_000 and s[ITEM][0] == 100_000
    assert sum(s[CUSTOMER][1]) == 644 and sum(s[STOCK][1]) == 296
    assert s[ORDER][0] == s[NEWORDER][0] == pre + 100_000 and s[HISTORY][0] == 100_000
    assert s[ORDERLINE][0] == PRE_LINES * pre + 15 * 100_000
    assert oe_append(64, 100_000) == {NEWORDER: pre, HISTORY: 0, ORDER: pre, ORDERLINE: PRE_LINES * pre}
    assert oe_append(2, 10, preload=7) == {NEWORDER: 140, HISTORY: 0, ORDER: 140, ORDERLINE: 700}


def _kinds(b):
    """Per txn: 'no' (inserts an order), 'pay' (inserts a history row),
    'del' (deletes neworder rows) or 'ro' (writes nothing)."""
    tab = (b.meta >> 8) & 15
    op_t = np.repeat(np.arange(b.n_txns), np.diff(b.txn_off.astype(np.int64)))
    ins = (b.meta & 0xFF) == OP_ROW
    dele = (b.meta & 0xFF) == (OP_ROW | (ROW_DELETE << 2))
    out = np.full(b.n_txns, "ro", dtype=object)
    out[op_t[dele]] = "del"
    out[op_t[ins & (tab == HISTORY)]] = "pay"
    out[op_t[ins & (tab == ORDER)]] = "no"
    writes = np.zeros(b.n_txns, bool)
    writes[op_t[(b.meta & 3) != OP_READ]] = True
    assert not writes[out == "ro"].any()
    return out


def test_generator_mix_and_determinism():
    a = OeStream(8, seed=3).next(20_000, n_threads=1)
    b = OeStream(8, seed=3).next(20_000, n_threads=6)
    for f in ("txn_off", "rows", "meta", "args"):
        assert (getattr(a, f) == getattr(b, f)).all()
    k = _kinds(a)
    assert abs((k == "no").mean() - 0.45) < 0.02 and abs((k == "pay").mean() - 0.43) < 0.02
    assert abs((k == "ro").mean() - 0.12) < 0.02 and not (k == "del").any()
    assert np.diff(a.txn_off.astype(np.int64)).max() <= 256
    # the reference's full mix (src/param.cpp:53-57): 44 / 44 / 4 / 4 / 4
    c = _kinds(OeStream(8, neworder_pct=44, payment_pct=44, delivery_pct=4, seed=3).next(20_000))
    assert abs((c == "del").mean() - 0.04) < 0.01 and abs((c == "ro").mean() - 0.08) < 0.01


def test_each_txn_keeps_the_engine_rules():
    # a written cell is not touched again; a deleted row is not touched by
    # another op of its txn; INSERTs name row 0
    b = OeStream(2, seed=4, remote_bp=5000, neworder_pct=40, payment_pct=30, delivery_pct=20,
                 preload=40).next(3000)
    assert (_kinds(b) == "del").sum() > 300
    for t in range(b.n_txns):
        lo, hi = b.txn_off[t], b.txn_off[t + 1]
        m = b.meta[lo:hi]
        tabs, cols, rows, kinds = ((m >> 8) & 15).tolist(), ((m >> 2) & 63).tolist(), b.rows[lo:hi].tolist(), (m & 3).tolist()
        written, deleted, touched = set(), set(), set()
        for tb, c, r, k in zip(tabs, cols, rows, kinds):
            if k == OP_ROW:
                if c == ROW_INSERT:
                    assert r == 0
                    continue
                assert (tb, r) not in touched and (tb, r) not in deleted, (t, tb, r)
                deleted.add((tb, r))
                continue
            assert (tb, c, r) not in written and (tb, r) not in deleted, (t, tb, c, r)
            touched.add((tb, r))
            if k != OP_READ:
                written.add((tb, c, r))


def test_neworder_inserts_and_remote_lines():
    b = OeStream(16, seed=5, remote_bp=10_000, neworder_pct=100, payment_pct=0).next(500)
    tab = (b.meta >> 8) & 15
    col = (b.meta >> 2) & 63
    ins = (b.meta & 0xFF) == OP_ROW
    assert ((tab == STOCK) & (col == 5)).sum() > 0  # s_remote_cnt
    op_t = np.repeat(np.arange(b.n_txns), np.diff(b.txn_off.astype(np.int64)))
    for t in range(b.n_txns):
        mine = op_t == t
        lines = int((mine & (tab == ITEM) & (col == 2)).sum())  # one i_price READ per line
        assert int((mine & ins & (tab == ORDERLINE)).sum()) == lines
        assert int((mine & ins & (tab == ORDER)).sum()) == 1 and int((mine & ins & (tab == NEWORDER)).sum()) == 1
    assert len(np.unique(b.args[ins])) == int(ins.sum())  # every appended row its own payload


def test_delivery_and_read_only_rows_follow_the_preload_layout():
    pre = 40
    b = OeStream(3, neworder_pct=0, payment_pct=0, delivery_pct=50, preload=pre, seed=8).next(400)
    tab = (b.meta >> 8) & 15
    col = (b.meta >> 2) & 63
    op_t = np.repeat(np.arange(b.n_txns), np.diff(b.txn_off.astype(np.int64)))
    k = _kinds(b)
    for t in np.flatnonzero(k == "del")[:30]:
        mine = op_t == t
        dels = b.rows[mine & (tab == NEWORDER)]
        assert len(dels) == 10 and len(set(dels.tolist())) == 10
        assert sorted(b.rows[mine & (tab == ORDER) & (col == O_CARRIER_ID)].tolist()) == sorted(dels.tolist())
        lines = sorted(b.rows[mine & (tab == ORDERLINE) & (col == OL_DELIVERY_D)].tolist())
        assert lines == sorted(int(r) * PRE_LINES + l for r in dels for l in range(PRE_LINES))
        assert int((mine & (tab == ORDERLINE) & (col == OL_AMOUNT)).sum()) == 10 * PRE_LINES
        assert int((mine & (tab == CUSTOMER) & (col == C_DELIVERY_CNT)).sum()) <= 10
        assert (dels < 3 * 10 * pre).all()
    for t in np.flatnonzero(k == "ro")[:30]:
        mine = op_t == t
        orows = b.rows[mine & (tab == ORDER)]
        lrows = b.rows[mine & (tab == ORDERLINE)]
        if orows.size:  # order status: the order and its own lines
            assert len(set(orows.tolist())) == 1
            assert set((lrows // PRE_LINES).tolist()) == {int(orows[0])}
            cust = b.rows[mine & (tab == CUSTOMER)]
            assert set((cust // 3000).tolist()) == {int(orows[0]) // pre}  # same district
        else:  # stock level: lines of the district's orders
            dist = int(b.rows[mine & (tab == DISTRICT)][0])
            assert set((lrows // (pre * PRE_LINES)).tolist()) == {dist}


def test_oracle_mode_a_matches_pairwise_on_small_batches():
    W, cap, pre = 1, 2000, 20
    b = OeStream(W, seed=6, neworder_pct=40, payment_pct=40, delivery_pct=10, preload=pre).next(60)
    im = orc.Image(oe_schema(W, cap, pre), append=oe_append(W, cap, pre)).fill(1)
    c, _ = orc.replay(im, b, orc.MODE_A, 256)
    assert (c == orc.commit_pairwise(b)).all()
    assert c[0] == 1 and c.sum() < len(c)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [MODE_A, MODE_O])
def test_gpu_order_entry_parity(mode):
    # NewOrder / Payment INSERTs, Delivery DELETEs of preloaded neworder
    # rows (some deleted twice across batches), read-only txns on the
    # preloaded orders; commit flags, result and INSERT slots, per-op status,
    # presence, cursors and every cell against the oracle
    W, cap, pre = 4, 30_000, 60
    st = OeStream(W, neworder_pct=40, payment_pct=40, delivery_pct=10, preload=pre, seed=7,
                  byname_pct=60, table_seed=11)
    with Engine() as e:
        sch, app = create_oe_tables(e, W, cap, 11, preload=pre)
        im = orc.Image(sch, append=app).fill(11)
        rs = e.result_stride
        took, missed = 0, 0
        for _ in range(3):
            b = st.next(8000)
            commit, res = e.run(b, mode)
            status = e.fetch_status()
            oc, ores, found = orc.replay(im, b, mode, rs, with_found=True)
            assert (commit == oc).all()
            assert (status == found).all()
            m = orc.read_mask(b, oc)
            assert (res[m] == ores[m]).all()
            dels = (b.meta & 0xFF) == (OP_ROW | (ROW_DELETE << 2))
            took += int(found[dels].sum())
            missed += int(dels.sum() - found[dels].sum())
        assert took > 0 and missed > 0  # rows deleted, and rows already gone (or txns aborted)
        for t in sch:
            assert e.digest(t) == orc.digest_columns(t, im.cols[t]), t
        for t in app:
            assert e.table_rows(t) == im.cursor[t], t
            assert (e.read_presence(t) == im.present[t]).all(), t
