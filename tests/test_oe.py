"""Order-entry adapter (include/ltpmg_oe.h): generator properties on the CPU,
and GPU parity against the oracle over the nine-table schema."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from ltpmg_amd import MODE_A, MODE_O, Engine
from ltpmg_amd.engine import OP_READ, OP_RMW_ADD, OP_UPDATE, OeStream, oe_schema

WAREHOUSE, DISTRICT, CUSTOMER, NEWORDER, HISTORY, ORDER, ORDERLINE, STOCK, ITEM = range(9)


def test_schema_shapes():
    s = oe_schema(64, 100_000)
    assert s[WAREHOUSE][0] == 64 and s[DISTRICT][0] == 640
    assert s[CUSTOMER][0] == 64 * 30_000 and s[STOCK][0] == 64 * 100_000 and s[ITEM][0] == 100_000
    assert sum(s[CUSTOMER][1]) == 644 and sum(s[STOCK][1]) == 296
    assert s[ORDERLINE][0] == 15 * 100_000


def test_generator_mix_and_determinism():
    a = OeStream(8, 50_000, seed=3).next(20_000, n_threads=1)
    b = OeStream(8, 50_000, seed=3).next(20_000, n_threads=6)
    for f in ("txn_off", "rows", "meta", "args"):
        assert (getattr(a, f) == getattr(b, f)).all()
    tab = (a.meta >> 8) & 15
    kind = a.meta & 3
    op_t = np.repeat(np.arange(a.n_txns), np.diff(a.txn_off.astype(np.int64)))
    new_order = np.zeros(a.n_txns, bool)
    new_order[op_t[(tab == NEWORDER)]] = True
    payment = np.zeros(a.n_txns, bool)
    payment[op_t[(tab == HISTORY)]] = True
    assert abs(new_order.mean() - 0.45) < 0.02 and abs(payment.mean() - 0.43) < 0.02
    # read-only txns write nothing
    writes = np.zeros(a.n_txns, bool)
    writes[op_t[kind != OP_READ]] = True
    assert abs((~writes).mean() - 0.12) < 0.02
    lens = np.diff(a.txn_off.astype(np.int64))
    assert lens.max() <= 256


def test_written_cells_are_not_touched_again_in_a_txn():
    b = OeStream(2, 20_000, seed=4, remote_bp=5000).next(3000)
    for t in range(b.n_txns):
        lo, hi = b.txn_off[t], b.txn_off[t + 1]
        cells = list(zip(((b.meta[lo:hi] >> 8) & 15).tolist(), ((b.meta[lo:hi] >> 2) & 63).tolist(),
                         b.rows[lo:hi].tolist()))
        kinds = (b.meta[lo:hi] & 3).tolist()
        written = set()
        for cell, k in zip(cells, kinds):
            assert cell not in written, (t, cell)
            if k != OP_READ:
                written.add(cell)


def test_remote_lines_and_append_rows():
    st = OeStream(16, 10_000, seed=5, remote_bp=10_000, neworder_pct=100, payment_pct=0)
    b = st.next(500)
    tab = (b.meta >> 8) & 15
    col = (b.meta >> 2) & 63
    remote = (tab == STOCK) & (col == 5)  # s_remote_cnt
    assert remote.sum() > 0
    assert list(st.spec.append_next) == [500, 500, 500, 7500]
    ol_rows = b.rows[(tab == ORDERLINE)]
    assert ol_rows.max() < 500 * 15


def test_oracle_mode_a_matches_pairwise_on_small_batches():
    sch = oe_schema(1, 2000)
    b = OeStream(1, 2000, seed=6).next(60)
    im = orc.Image(sch).fill(1)
    c, _ = orc.replay(im, b, orc.MODE_A, 256)
    assert (c == orc.commit_pairwise(b)).all()
    assert c[0] == 1 and c.sum() < len(c)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [MODE_A, MODE_O])
def test_gpu_order_entry_parity(mode):
    W, cap = 4, 30_000
    sch = oe_schema(W, cap)
    st = OeStream(W, cap, seed=7)
    with Engine() as e:
        for t, (n, widths) in sch.items():
            e.create_table(t, n, widths)
            e.fill_table(t, 11)
        im = orc.Image(sch).fill(11)
        for _ in range(3):
            b = st.next(8000)
            commit, res = e.run(b, mode)
            oc, ores = orc.replay(im, b, mode, e.result_stride)
            assert (commit == oc).all()
            m = orc.read_mask(b, oc)
            assert (res[m] == ores[m]).all()
        for t in sch:
            assert e.digest(t) == orc.digest_columns(t, im.cols[t])
