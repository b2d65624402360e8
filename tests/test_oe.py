"""Order-entry adapter (include/ltpmg_oe.h): generator properties on the CPU,
and GPU parity against the oracle over the nine-table schema."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from ltpmg_amd import MODE_A, MODE_O, Engine
from ltpmg_amd.engine import OP_READ, OP_RMW_ADD, OP_UPDATE, OeStream, oe_last_name_index, oe_schema

WAREHOUSE, DISTRICT, CUSTOMER, NEWORDER, HISTORY, ORDER, ORDERLINE, STOCK, ITEM = range(9)
C_BALANCE, C_LAST = 7, 11
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
    by_name = OeStream(W, 5000, neworder_pct=0, payment_pct=100, seed=5, byname_pct=100,
                       table_seed=seed).next(600)
    by_id = OeStream(W, 5000, neworder_pct=0, payment_pct=100, seed=5).next(600)

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
    st = OeStream(W, cap, seed=7, byname_pct=60, table_seed=11)
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
