"""Small batch builders used by the oracle and GPU tests (numpy RNG, kept
independent of the product's generator)."""
from __future__ import annotations

import numpy as np

from ltpmg_amd.engine import OP_READ, OP_RMW_ADD, OP_ROW, OP_UPDATE, ROW_DELETE, ROW_INSERT, Batch, meta


def random_batch(rng: np.random.Generator, n_txns: int, n_rows: int, n_cols: int, *,
                 max_len: int = 4, kinds=(OP_READ, OP_UPDATE, OP_RMW_ADD),
                 table: int = 0, distinct_rows: bool = True, zipf: float | None = None) -> Batch:
    """Random txns; a written cell is never touched twice within a txn."""
    txns = []
    for _ in range(n_txns):
        length = int(rng.integers(1, max_len + 1))
        ops, used_rows, used_cells = [], set(), set()
        for _ in range(length):
            for _attempt in range(50):
                if zipf:
                    row = int(min(rng.zipf(1.0 + zipf), n_rows) - 1)
                else:
                    row = int(rng.integers(0, n_rows))
                col = int(rng.integers(0, n_cols))
                if distinct_rows and row in used_rows:
                    continue
                if (row, col) in used_cells:
                    continue
                break
            else:
                continue
            kind = int(rng.choice(kinds))
            used_rows.add(row)
            used_cells.add((row, col))
            arg = int(rng.integers(0, 2**63)) if kind == OP_UPDATE else int(rng.integers(-3, 4)) & 0xFFFFFFFFFFFFFFFF
            ops.append((kind, table, col, row, arg))
        if not ops:
            ops.append((OP_READ, table, 0, 0, 0))
        txns.append(ops)
    return Batch.from_txns(txns)


def concat(*batches: Batch) -> Batch:
    off = [0]
    for b in batches:
        off.extend((b.txn_off[1:].astype(np.int64) + off[-1]).tolist())
    return Batch(np.array(off, np.uint32), np.concatenate([b.rows for b in batches]),
                 np.concatenate([b.meta for b in batches]), np.concatenate([b.args for b in batches]))


def op_txn(batch: Batch) -> np.ndarray:
    return np.repeat(np.arange(batch.n_txns), np.diff(batch.txn_off.astype(np.int64)))


def random_row_batch(rng: np.random.Generator, n_txns: int, n_rows: int, n_cols: int, *,
                     table: int = 0, max_len: int = 4, insert_pct: int = 20, delete_pct: int = 15,
                     kinds=(OP_READ, OP_UPDATE, OP_RMW_ADD)) -> Batch:
    """Random txns on an append table of n_rows capacity: INSERTs (row 0,
    a random arg), DELETEs and cell ops, each non-INSERT op of a txn on its
    own row (so no txn accesses a row it deletes or a cell it wrote)."""
    txns = []
    for _ in range(n_txns):
        length = int(rng.integers(1, max_len + 1))
        ops, used = [], set()
        for _ in range(length):
            u = int(rng.integers(0, 100))
            if u < insert_pct:
                ops.append((OP_ROW, table, ROW_INSERT, 0, int(rng.integers(0, 2**63))))
                continue
            row = int(rng.integers(0, n_rows))
            if row in used:
                continue
            used.add(row)
            if u < insert_pct + delete_pct:
                ops.append((OP_ROW, table, ROW_DELETE, row, 0))
                continue
            kind = int(rng.choice(kinds))
            arg = int(rng.integers(0, 2**63)) if kind == OP_UPDATE else int(rng.integers(-3, 4)) & 0xFFFFFFFFFFFFFFFF
            ops.append((kind, table, int(rng.integers(0, n_cols)), row, arg))
        if not ops:
            ops.append((OP_ROW, table, ROW_INSERT, 0, int(rng.integers(0, 2**63))))
        txns.append(ops)
    return Batch.from_txns(txns)


__all__ = ["random_batch", "random_row_batch", "concat", "op_txn", "meta", "OP_READ", "OP_UPDATE",
           "OP_RMW_ADD", "OP_ROW", "ROW_INSERT", "ROW_DELETE"]
