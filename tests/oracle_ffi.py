"""ctypes binding of the oracle (oracle/_build/liboracle.so).

Test infrastructure only: the oracle is the checker for the GPU engine and
the CPU baseline, never part of the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "_build" / "liboracle.so"
MAX_COLS = 64
MODE_A, MODE_O = 0, 1


class _OrcTable(C.Structure):
    _fields_ = [("n_rows", C.c_uint32), ("n_cols", C.c_uint32),
                ("width", C.c_uint32 * MAX_COLS), ("col", C.c_void_p * MAX_COLS),
                ("present", C.c_void_p), ("cursor", C.c_uint32)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    L = C.CDLL(str(ORACLE_SO))
    L.orc_mix64.restype = C.c_uint64
    L.orc_mix64.argtypes = [C.c_uint64]
    L.orc_stream.restype = C.c_uint64
    L.orc_stream.argtypes = [C.c_uint64, C.c_uint32]
    L.orc_cellseed.restype = C.c_uint64
    L.orc_cellseed.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_bytes_from_stream.restype = None
    L.orc_bytes_from_stream.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p]
    L.orc_fill_column.restype = None
    L.orc_fill_column.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32, C.c_void_p]
    L.orc_fill_rows.restype = None
    L.orc_fill_rows.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                C.c_uint32, C.c_void_p]
    batch_args = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                  C.c_void_p, C.c_void_p, C.c_uint32]
    L.orc_replay.restype = C.c_int
    L.orc_replay.argtypes = [C.POINTER(_OrcTable), C.c_uint32, *batch_args, C.c_void_p]
    L.orc_commit_pairwise.restype = C.c_int
    L.orc_commit_pairwise.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_cpu_pipeline.restype = C.c_double
    L.orc_cpu_pipeline.argtypes = [C.POINTER(_OrcTable), C.c_uint32, *batch_args, C.c_uint32]
    _lib = L
    return L


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def mix64(z: int) -> int:
    return lib().orc_mix64(z & 0xFFFFFFFFFFFFFFFF)


def stream(s: int, j: int) -> int:
    return lib().orc_stream(s & 0xFFFFFFFFFFFFFFFF, j)


def payload(arg: int, width: int) -> np.ndarray:
    out = np.empty(width, np.uint8)
    lib().orc_bytes_from_stream(arg & 0xFFFFFFFFFFFFFFFF, width, _p(out))
    return out


class Image:
    """Host image of the tables: images[t][c] is an (n_rows, width) uint8 array.
    append={t: n_present} makes table t an append table: presence bytes
    present[t] (1 for rows [0, n_present)) and cursor[t] = n_present."""

    def __init__(self, schema: dict[int, tuple[int, list[int]]], append: dict[int, int] | None = None):
        self.schema = schema
        self.cols = {t: [np.zeros((n, w), np.uint8) for w in widths]
                     for t, (n, widths) in schema.items()}
        self.present, self.cursor = {}, {}
        for t, k in (append or {}).items():
            n = schema[t][0]
            self.present[t] = np.zeros(n, np.uint8)
            self.present[t][:k] = 1
            self.cursor[t] = k

    @property
    def n_tables(self) -> int:
        return max(self.schema) + 1

    def fill(self, seed: int) -> "Image":
        for t, (n, widths) in self.schema.items():
            for c, w in enumerate(widths):
                lib().orc_fill_column(seed & 0xFFFFFFFFFFFFFFFF, t, c, w, 0, n, _p(self.cols[t][c]))
        return self

    def fill_rows(self, seed: int, table: int, rows: np.ndarray) -> "Image":
        """Fills only the given rows (the rest stay zero): for sparse samples."""
        r = np.ascontiguousarray(rows, np.uint32)
        for c, w in enumerate(self.schema[table][1]):
            lib().orc_fill_rows(seed & 0xFFFFFFFFFFFFFFFF, table, c, w, _p(r), len(r),
                                _p(self.cols[table][c]))
        return self

    def fill_cells(self, seed: int, table: int, col: int, rows: np.ndarray) -> "Image":
        """Fills one column at the given rows only."""
        r = np.ascontiguousarray(rows, np.uint32)
        w = self.schema[table][1][col]
        lib().orc_fill_rows(seed & 0xFFFFFFFFFFFFFFFF, table, col, w, _p(r), len(r),
                            _p(self.cols[table][col]))
        return self

    def copy(self) -> "Image":
        im = Image(self.schema)
        im.cols = {t: [a.copy() for a in cs] for t, cs in self.cols.items()}
        im.present = {t: a.copy() for t, a in self.present.items()}
        im.cursor = dict(self.cursor)
        return im

    def tables(self):
        arr = (_OrcTable * self.n_tables)()
        for t, (n, widths) in self.schema.items():
            arr[t].n_rows = n
            arr[t].n_cols = len(widths)
            for c, w in enumerate(widths):
                arr[t].width[c] = w
                arr[t].col[c] = self.cols[t][c].ctypes.data
            if t in self.present:
                arr[t].present = self.present[t].ctypes.data
                arr[t].cursor = self.cursor[t]
        return arr

    def sync(self, arr) -> None:
        """Reads the append cursors back from the struct array a call changed."""
        for t in self.cursor:
            self.cursor[t] = arr[t].cursor

    def digest(self) -> int:
        """Same definition as the engine's table digest, restated with numpy."""
        total = 0
        for t, (n, widths) in self.schema.items():
            total = (total + digest_columns(t, self.cols[t])) & 0xFFFFFFFFFFFFFFFF
        return total


M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _np_mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def digest_columns(table: int, cols: list[np.ndarray]) -> int:
    """Sum over cells of h, h0 = mix64(golden * (1 + (row<<10 | table<<6 | col))),
    h = mix64(h ^ le64 word) for each 8-byte word of the width (last zero padded)."""
    total = np.uint64(0)
    golden = np.uint64(0x9E3779B97F4A7C15)
    for c, a in enumerate(cols):
        n, w = a.shape
        rows = np.arange(n, dtype=np.uint64)
        ids = (rows << np.uint64(10)) | np.uint64(table << 6) | np.uint64(c)
        with np.errstate(over="ignore"):
            h = _np_mix64(golden * (ids + np.uint64(1)))
        padded = np.zeros((n, (w + 7) // 8 * 8), np.uint8)
        padded[:, :w] = a
        words = padded.view("<u8")
        for k in range(words.shape[1]):
            h = _np_mix64(h ^ words[:, k])
        with np.errstate(over="ignore"):
            total = total + h.sum(dtype=np.uint64)
    return int(total)


def digest_rows(table: int, col: int, rows: np.ndarray, cells: np.ndarray) -> int:
    """The engine's table-digest terms of the given cells only: the sum over
    rows[i] of the same per-cell hash digest_columns sums (cells[i] holds
    the cell's bytes).  With it, digest(after) - digest(before) of a whole
    table equals the sum over the touched cells of (term(after) -
    term(before)), which checks a table too large to compare cell by cell."""
    rows = np.asarray(rows, np.uint64)
    n, w = cells.shape
    golden = np.uint64(0x9E3779B97F4A7C15)
    ids = (rows << np.uint64(10)) | np.uint64(table << 6) | np.uint64(col)
    with np.errstate(over="ignore"):
        h = _np_mix64(golden * (ids + np.uint64(1)))
    padded = np.zeros((n, (w + 7) // 8 * 8), np.uint8)
    padded[:, :w] = cells
    words = padded.view("<u8")
    for k in range(words.shape[1]):
        h = _np_mix64(h ^ words[:, k])
    with np.errstate(over="ignore"):
        return int(h.sum(dtype=np.uint64))


def touched_cells(batches, schema: dict, append: dict | None = None) -> dict:
    """{(table, col): sorted unique rows} that any of the batches names, and
    for each append table in append ({table: rows present at the start})
    every column of the rows its INSERTs can take: [present, present + the
    INSERT ops of the batches)."""
    out = {}
    is_row = [(b.meta & 3) == 3 for b in batches]
    for t, (n, widths) in schema.items():
        n_ins = 0
        if append and t in append:
            n_ins = sum(int(np.count_nonzero(r & (((b.meta >> 8) & 15) == t) & (((b.meta >> 2) & 63) == 0)))
                        for b, r in zip(batches, is_row))
        extra = np.arange(append[t], min(n, append[t] + n_ins), dtype=np.uint32) if n_ins else None
        for c in range(len(widths)):
            parts = [b.rows[~r & (((b.meta >> 8) & 15) == t) & (((b.meta >> 2) & 63) == c)]
                     for b, r in zip(batches, is_row)]
            if extra is not None:
                parts.append(extra)
            rows = np.unique(np.concatenate(parts)) if parts else np.zeros(0, np.uint32)
            if rows.size:
                out[(t, c)] = rows.astype(np.uint32)
    return out


def sparse_image(schema: dict, seed: int, batches, append: dict | None = None) -> tuple[Image, dict]:
    """An image of the tables the batches name that fills only the cells they
    touch (untouched pages of the zeroed columns stay unmapped), and those
    cells as {(table, col): rows}.  Append tables (append = {table: rows
    present at the start}) keep full presence arrays; the rows their INSERTs
    can take count as touched."""
    cells = touched_cells(batches, schema, append)
    used = {t for t, _ in cells} | {t for b in batches for t in np.unique((b.meta >> 8) & 15).tolist()
                                     if t in schema}
    im = Image({t: schema[t] for t in sorted(used)},
               append={t: k for t, k in (append or {}).items() if t in used})
    for (t, c), rows in cells.items():
        im.fill_cells(seed, t, c, rows)
    return im, cells


class AppendFull(ValueError):
    """The INSERTs of a batch did not all fit their append tables."""


def replay(image: Image, batch, mode: int, res_stride: int, with_found: bool = False):
    """Serial replay in TID order; mutates image.  Returns (commit, results),
    or (commit, results, found) with with_found."""
    n_txns = len(batch.txn_off) - 1
    commit = np.zeros(n_txns, np.uint8)
    results = np.zeros((len(batch.rows), res_stride), np.uint8)
    found = np.zeros(len(batch.rows), np.uint8)
    tabs = image.tables()
    rc = lib().orc_replay(tabs, image.n_tables, n_txns, _p(batch.txn_off), _p(batch.rows),
                          _p(batch.meta), _p(batch.args), mode, _p(commit), _p(results), res_stride,
                          _p(found))
    image.sync(tabs)
    if rc == -5:
        raise AppendFull("orc_replay: an append table is full")
    if rc != 0:
        raise ValueError(f"orc_replay rejected the batch ({rc})")
    return (commit, results, found) if with_found else (commit, results)


def commit_pairwise(batch) -> np.ndarray:
    n_txns = len(batch.txn_off) - 1
    commit = np.zeros(n_txns, np.uint8)
    lib().orc_commit_pairwise(n_txns, _p(batch.txn_off), _p(batch.rows), _p(batch.meta), _p(commit))
    return commit


def cpu_pipeline(image: Image, batch, mode: int, res_stride: int, n_threads: int = 0):
    n_txns = len(batch.txn_off) - 1
    commit = np.zeros(n_txns, np.uint8)
    results = np.zeros((len(batch.rows), res_stride), np.uint8)
    tabs = image.tables()
    secs = lib().orc_cpu_pipeline(tabs, image.n_tables, n_txns, _p(batch.txn_off), _p(batch.rows),
                                  _p(batch.meta), _p(batch.args), mode, _p(commit), _p(results),
                                  res_stride, n_threads)
    if secs < 0:
        raise ValueError("orc_cpu_pipeline rejected the batch")
    return commit, results, secs


def read_mask(batch, commit: np.ndarray) -> np.ndarray:
    """Ops whose result slot is defined: READs and INSERTs of committed txns."""
    n_txns = len(batch.txn_off) - 1
    op_txn = np.repeat(np.arange(n_txns), np.diff(batch.txn_off.astype(np.int64)))
    defined = ((batch.meta & 3) == 0) | ((batch.meta & 0xFF) == 3)
    return defined & (commit[op_txn] != 0)
