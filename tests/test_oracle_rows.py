"""Pins the oracle's INSERT / DELETE semantics on append tables before it is
trusted as the GPU checker (the reference's insert and delete executors,
cuda/src/gpulauncher.cu:205-356).

Each check is chosen so that a plausible mistake fails it: a wrong row seed
(inserted bytes reduce to the published SplitMix64 outputs for a chosen
arg), INSERTs appended in TID order instead of op order or before the
batch's other ops (row ids and the read of an inserted row in the same
batch), an aborted txn's INSERT taking a row, a DELETE not counted as a
write by the lowest-TID-wins rule (pairwise brute force), a second DELETE
or an op on a deleted row taking effect, and an overflow that writes past
the capacity.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from batches import (OP_READ, OP_RMW_ADD, OP_ROW, OP_UPDATE, ROW_DELETE, ROW_INSERT, op_txn,
                     random_row_batch)
from ltpmg_amd.engine import Batch

K_ROW = 0xD1B54A32D192ED03  # rowseed(arg, c) = mix64(arg + (c + 1) * K_ROW)
M = (1 << 64) - 1
SPLITMIX0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


def _ins(arg, table=0):
    return (OP_ROW, table, ROW_INSERT, 0, arg & M)


def _del(row, table=0):
    return (OP_ROW, table, ROW_DELETE, row, 0)


def _image(cap=16, widths=(8, 16), n_present=6):
    return orc.Image({0: (cap, list(widths))}, append={0: n_present}).fill(5)


def _row_id(slot: np.ndarray) -> int:
    return int.from_bytes(slot[:4].tobytes(), "little")


@pytest.mark.parametrize("mode", [orc.MODE_A, orc.MODE_O])
def test_inserted_bytes_reduce_to_published_splitmix(mode):
    # arg = -(c+1) * K_ROW makes column c's row seed mix64(0) = 0, whose
    # stream is SplitMix64 from state 0
    im = _image()
    b = Batch.from_txns([[_ins(-1 * K_ROW)], [_ins(-2 * K_ROW)]])
    commit, res, found = orc.replay(im, b, mode, 16, with_found=True)
    assert commit.tolist() == [1, 1] and found.tolist() == [1, 1]
    assert [_row_id(res[0]), _row_id(res[1])] == [6, 7]
    assert not res[:, 4:].any()
    assert im.cols[0][0][6].tobytes() == SPLITMIX0[0].to_bytes(8, "little")
    assert im.cols[0][1][7].tobytes() == b"".join(w.to_bytes(8, "little") for w in SPLITMIX0[:2])
    assert im.cursor[0] == 8
    assert im.present[0].tolist() == [1] * 8 + [0] * 8


def test_committed_inserts_take_rows_in_op_order_after_the_batch():
    im = _image(cap=64, n_present=10)
    rng = np.random.default_rng(3)
    txns = []
    for t in range(12):
        ops = [(OP_READ, 0, 0, int(rng.integers(0, 10)), 0)]
        for _ in range(int(rng.integers(0, 3))):
            ops.insert(int(rng.integers(0, len(ops) + 1)), _ins(int(rng.integers(0, 2**63))))
        txns.append(ops)
    # a READ of the first row the batch appends: not present yet
    txns.append([(OP_READ, 0, 1, 10, 0)])
    b = Batch.from_txns(txns)
    before = im.copy()
    commit, res, found = orc.replay(im, b, orc.MODE_O, 16, with_found=True)
    is_ins = (b.meta & 0xFF) == OP_ROW
    k = int(is_ins.sum())
    assert k > 4
    expect = 10 + np.cumsum(is_ins) - 1
    for i in np.flatnonzero(is_ins):
        assert _row_id(res[i]) == expect[i]
        assert found[i] == 1
    assert im.cursor[0] == 10 + k
    assert im.present[0][:10 + k].all() and not im.present[0][10 + k:].any()
    last = len(b.rows) - 1
    assert found[last] == 0 and not res[last].any()
    for i in np.flatnonzero(~is_ins)[:-1]:  # reads see the tables before the appends
        assert found[i] == 1
        assert res[i][:8].tobytes() == before.cols[0][0][b.rows[i]].tobytes()


def test_aborted_txn_takes_no_row():
    im = _image(n_present=4)
    b = Batch.from_txns([[(OP_UPDATE, 0, 0, 2, 9)],
                         [(OP_UPDATE, 0, 0, 2, 7), _ins(1)],
                         [_ins(2)]])
    commit, res, found = orc.replay(im, b, orc.MODE_A, 16, with_found=True)
    assert commit.tolist() == [1, 0, 1]
    assert found.tolist() == [1, 0, 0, 1]
    assert _row_id(res[3]) == 4 and im.cursor[0] == 5


def test_insert_locks_no_row():
    # INSERT's row field is 0 by convention; it conflicts with nothing
    im = _image()
    b = Batch.from_txns([[(OP_UPDATE, 0, 0, 0, 3)], [_ins(5), (OP_READ, 0, 1, 1, 0)],
                         [_ins(6)], [(OP_READ, 0, 0, 1, 0), _ins(7)]])
    commit, _ = orc.replay(im, b, orc.MODE_A, 16)
    assert commit.tolist() == [1, 1, 1, 1]
    assert orc.commit_pairwise(b).tolist() == [1, 1, 1, 1]


def test_ordered_delete_then_ops_on_the_deleted_row():
    im = _image()
    r = 3
    b = Batch.from_txns([[(OP_READ, 0, 0, r, 0)], [_del(r)], [(OP_READ, 0, 0, r, 0)],
                         [(OP_UPDATE, 0, 0, r, 123)], [(OP_RMW_ADD, 0, 1, r, 1)], [_del(r)],
                         [(OP_READ, 0, 1, r + 1, 0)]])
    before = im.copy()
    commit, res, found = orc.replay(im, b, orc.MODE_O, 16, with_found=True)
    assert commit.all()
    assert found.tolist() == [1, 1, 0, 0, 0, 0, 1]
    assert res[0][:8].tobytes() == before.cols[0][0][r].tobytes()
    assert not res[2].any()
    assert im.cols[0][0][r].tobytes() == before.cols[0][0][r].tobytes()
    assert im.cols[0][1][r].tobytes() == before.cols[0][1][r].tobytes()
    assert im.present[0][r] == 0 and im.present[0][r + 1] == 1
    assert im.cursor[0] == 6  # a DELETE does not move the cursor


def test_lowest_tid_wins_counts_delete_as_a_write():
    im = _image()
    b = Batch.from_txns([[_del(2)], [(OP_READ, 0, 0, 2, 0)],     # reader after: aborts
                         [(OP_READ, 0, 0, 4, 0)], [_del(4)],     # reader before: both commit
                         [_del(5)], [_del(5)]])                  # second deleter aborts
    before = im.copy()
    commit, res, found = orc.replay(im, b, orc.MODE_A, 16, with_found=True)
    assert commit.tolist() == [1, 0, 1, 1, 1, 0]
    assert found.tolist() == [1, 0, 1, 1, 1, 0]
    assert res[2][:8].tobytes() == before.cols[0][0][4].tobytes()
    assert im.present[0][[2, 4, 5]].tolist() == [0, 0, 0]


@pytest.mark.parametrize("seed", range(10))
def test_lowest_tid_wins_with_row_ops_equals_pairwise_definition(seed):
    rng = np.random.default_rng(100 + seed)
    b = random_row_batch(rng, 40, 24, 2, max_len=4)
    im = _image(cap=200, n_present=16)
    commit, _ = orc.replay(im, b, orc.MODE_A, 16)
    assert commit.tolist() == orc.commit_pairwise(b).tolist()
    assert 0 < commit.sum() < len(commit)


@pytest.mark.parametrize("mode", [orc.MODE_A, orc.MODE_O])
@pytest.mark.parametrize("seed", range(6))
def test_presence_and_delete_status_closed_form(mode, seed):
    # presence after = presence before, minus the present rows that a
    # committed DELETE names, plus one row per committed INSERT; the DELETE
    # that takes effect on a row is its first committed one in op order
    rng = np.random.default_rng(200 + seed)
    cap, n_present = 400, 200
    b = random_row_batch(rng, 120, 260, 2, max_len=5, delete_pct=20)
    im = _image(cap=cap, n_present=n_present)
    im.present[0][rng.choice(n_present, 30, replace=False)] = 0  # earlier deletes
    before = im.present[0].copy()
    commit, res, found = orc.replay(im, b, mode, 16, with_found=True)
    ok = commit[op_txn(b)] != 0
    is_del = (b.meta & 0xFF) == (OP_ROW | (ROW_DELETE << 2))
    is_ins = (b.meta & 0xFF) == OP_ROW
    deleted = np.unique(b.rows[is_del & ok])
    deleted = deleted[before[deleted] == 1]
    k = int((is_ins & ok).sum())
    expect = before.copy()
    expect[deleted] = 0
    expect[n_present:n_present + k] = 1
    assert im.present[0].tolist() == expect.tolist()
    assert im.cursor[0] == n_present + k
    first = {}
    for i in np.flatnonzero(is_del & ok):
        first.setdefault(int(b.rows[i]), int(i))
    for i in np.flatnonzero(is_del):
        assert found[i] == int(ok[i] and before[b.rows[i]] == 1 and first[int(b.rows[i])] == i), i
    assert (found[is_ins] == ok[is_ins]).all()
    assert len(deleted) > 5 and k > 5


def test_inserted_rows_are_ordinary_rows_in_the_next_batch():
    im = _image(n_present=2)
    orc.replay(im, Batch.from_txns([[_ins(-1 * K_ROW)]]), orc.MODE_O, 16)
    b = Batch.from_txns([[(OP_READ, 0, 0, 2, 0)], [(OP_RMW_ADD, 0, 0, 2, 1)], [_del(2)],
                         [(OP_READ, 0, 0, 2, 0)]])
    commit, res, found = orc.replay(im, b, orc.MODE_O, 16, with_found=True)
    assert found.tolist() == [1, 1, 1, 0]
    assert res[0][:8].tobytes() == SPLITMIX0[0].to_bytes(8, "little")
    w = np.frombuffer(SPLITMIX0[0].to_bytes(8, "little"), "<u4")  # RMW adds to each u32 word
    assert im.cols[0][0][2].view("<u4").tolist() == [(int(x) + 1) & 0xFFFFFFFF for x in w]


def test_overflow_appends_what_fits_and_reports():
    im = _image(cap=8, n_present=6)
    b = Batch.from_txns([[_ins(1)], [_ins(2)], [_ins(3)]])
    with pytest.raises(orc.AppendFull):
        orc.replay(im, b, orc.MODE_O, 16)
    assert im.cursor[0] == 8 and im.present[0].all()


def test_replay_rejects_malformed_row_ops():
    ordinary = orc.Image({0: (16, [8])}).fill(1)
    with pytest.raises(ValueError):
        orc.replay(ordinary, Batch.from_txns([[_ins(1)]]), orc.MODE_O, 16)
    with pytest.raises(ValueError):
        orc.replay(_image(), Batch.from_txns([[(OP_ROW, 0, ROW_INSERT, 1, 0)]]), orc.MODE_O, 16)
    with pytest.raises(ValueError):
        orc.replay(_image(), Batch.from_txns([[_del(16)]]), orc.MODE_A, 16)
    with pytest.raises(ValueError):
        orc.replay(_image(), Batch.from_txns([[(OP_ROW, 0, 2, 1, 0)]]), orc.MODE_O, 16)
    with pytest.raises(ValueError):  # a row id needs a 4-byte slot
        orc.replay(_image(widths=(2,)), Batch.from_txns([[_ins(1)]]), orc.MODE_O, 2)


def test_cpu_pipeline_refuses_append_tables():
    im = _image()
    with pytest.raises(ValueError):
        orc.cpu_pipeline(im, Batch.from_txns([[(OP_READ, 0, 0, 1, 0)]]), orc.MODE_O, 16)
