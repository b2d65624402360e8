"""ctypes binding of the engine's C ABI (include/ltpmg.h, include/ltpmg_stream.h)
and of the workload generators (include/ltpmg_workload.h, include/ltpmg_oe.h).

The engine runs only on the GPU: there is no CPU fallback, and loading fails
with a RuntimeError when libltpmg_hip.so has not been built.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

MAX_TABLES = 16
MAX_COLS = 64
MAX_TXN_OPS = 256
MAX_SLOTS = 8
OP_READ, OP_UPDATE, OP_RMW_ADD, OP_ROW = 0, 1, 2, 3
ROW_INSERT, ROW_DELETE = 0, 1  # column field of an OP_ROW op
MODE_A, MODE_O = 0, 1
EXEC_PROFILE = 0x100
EXEC_PIPELINE = 0x200  # ordered batches always take the multi-kernel pipeline
SMALL_MAX_OPS = 8192   # ordered batches up to this size (no row ops) run as one launch
CFG_PROFILE = 1

STATUS = {0: "OK", -1: "EINVAL", -2: "ENOMEM", -3: "EBUDGET", -4: "EHIP", -5: "ESTATE",
          -6: "EDEVICE", -7: "EFULL"}


def meta(kind: int, table: int, col: int) -> int:
    return kind | (col << 2) | (table << 8)


class LtpmgError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} -> {STATUS.get(code, code)}: {msg}")
        self.code = code


class _Cfg(C.Structure):
    _fields_ = [("device", C.c_int32), ("n_slots", C.c_uint32), ("max_txns", C.c_uint32),
                ("max_ops", C.c_uint32), ("hbm_budget", C.c_uint64), ("flags", C.c_uint32)]


class _TableDesc(C.Structure):
    _fields_ = [("n_rows", C.c_uint32), ("n_cols", C.c_uint32),
                ("col_bytes", C.c_uint32 * MAX_COLS)]


class _Batch(C.Structure):
    _fields_ = [("n_txns", C.c_uint32), ("n_ops", C.c_uint32),
                ("txn_off", C.POINTER(C.c_uint32)), ("rows", C.POINTER(C.c_uint32)),
                ("meta", C.POINTER(C.c_uint16)), ("args", C.POINTER(C.c_uint64))]


class _Stats(C.Structure):
    _fields_ = [("n_txns", C.c_uint32), ("n_ops", C.c_uint32), ("committed", C.c_uint32),
                ("aborted", C.c_uint32), ("mode", C.c_uint32), ("exec_ms", C.c_float),
                ("upload_ms", C.c_float), ("n_kernels", C.c_uint32),
                ("kernel_name", (C.c_char * 32) * 24), ("kernel_ms", C.c_float * 24),
                ("hbm_allocated", C.c_uint64), ("inserted", C.c_uint32)]


class _Staging(C.Structure):
    _fields_ = [("txn_off", C.POINTER(C.c_uint32)), ("rows", C.POINTER(C.c_uint32)),
                ("meta", C.POINTER(C.c_uint16)), ("args", C.POINTER(C.c_uint64))]


class _KvSpec(C.Structure):
    _fields_ = [("n_txns", C.c_uint32), ("ops_min", C.c_uint32), ("ops_max", C.c_uint32),
                ("n_rows", C.c_uint32), ("zipf_theta", C.c_double), ("n_cols", C.c_uint32),
                ("read_pct", C.c_uint32), ("write_kind", C.c_uint32),
                ("read_all_cols", C.c_uint32), ("table_id", C.c_uint32),
                ("rmw_delta", C.c_int32), ("seed", C.c_uint64)]


# Every symbol include/ltpmg.h declares, with its ctypes signature.
ENGINE_API = {
    "ltpmg_open": (C.c_int, [C.POINTER(_Cfg), C.POINTER(C.c_void_p)]),
    "ltpmg_close": (C.c_int, [C.c_void_p]),
    "ltpmg_shared_engine": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "ltpmg_last_error": (C.c_char_p, [C.c_void_p]),
    "ltpmg_create_table": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(_TableDesc)]),
    "ltpmg_create_append_table": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(_TableDesc),
                                            C.c_uint32]),
    "ltpmg_table_rows": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "ltpmg_read_presence": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_void_p]),
    "ltpmg_truncate_table": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "ltpmg_fill_table": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64]),
    "ltpmg_load_column": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_void_p]),
    "ltpmg_read_column": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_void_p]),
    "ltpmg_table_digest": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64)]),
    "ltpmg_result_stride": (C.c_uint32, [C.c_void_p]),
    "ltpmg_submit_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(_Batch)]),
    "ltpmg_check_batch": (C.c_int, [C.c_void_p, C.POINTER(_Batch)]),
    "ltpmg_stage_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.POINTER(_Staging)]),
    "ltpmg_submit_staged": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "ltpmg_submit_runs": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(_Batch), C.c_void_p, C.c_uint32]),
    "ltpmg_pinned_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "ltpmg_pinned_free": (C.c_int, [C.c_void_p]),
    "ltpmg_execute": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "ltpmg_query": (C.c_int, [C.c_void_p, C.c_uint32]),
    "ltpmg_wait": (C.c_int, [C.c_void_p, C.c_uint32]),
    "ltpmg_fetch_results": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                      C.c_uint64]),
    "ltpmg_fetch_status": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "ltpmg_stats": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(_Stats)]),
    "ltpmg_selftest_sort": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.c_uint32]),
    "ltpmg_selftest_segscan": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_uint32]),
    "ltpmg_selftest_copy": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32,
                                      C.POINTER(C.c_double)]),
    "ltpmg_execute_static": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(_Batch), C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
}
STATIC_MAX_OPS = 1 << 18  # LTPMG_STATIC_MAX_OPS

class _StreamPlan(C.Structure):
    _fields_ = [("merge_max_ops", C.c_uint32), ("fused_max_ops", C.c_uint32),
                ("fused_tile_ops", C.c_uint32), ("fused_wgs_per_cu", C.c_uint32),
                ("use_graph", C.c_uint32), ("flags", C.c_uint32), ("small_max_ops", C.c_uint32)]


STREAM_KEEP_RESULTS = 0x1
STREAM_REWIND = 0x2
PLAN_SYNC_EACH = 0x1

# Every symbol include/ltpmg_stream.h declares (also in libltpmg_hip.so).
STREAM_API = {
    "ltpmg_stream_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                      C.POINTER(_Batch), C.c_uint32, C.POINTER(C.c_void_p)]),
    "ltpmg_stream_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ltpmg_stream_run": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_StreamPlan),
                                   C.POINTER(C.c_double)]),
    "ltpmg_stream_tune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_StreamPlan),
                                    C.POINTER(C.c_uint32)]),
    "ltpmg_stream_default_plan": (None, [C.POINTER(_StreamPlan)]),
    "ltpmg_stream_fetch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_uint64]),
    "ltpmg_stream_committed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "ltpmg_stream_groups": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_StreamPlan), C.c_void_p,
                                      C.c_uint32]),
}


class _ServeCfg(C.Structure):
    _fields_ = [("arrival_rate", C.c_double), ("n_arrivals", C.c_uint32), ("max_batch", C.c_uint32),
                ("min_batch", C.c_uint32), ("max_wait_us", C.c_double), ("n_inflight", C.c_uint32),
                ("mode", C.c_uint32), ("seed", C.c_uint64), ("hbm_inflight_bytes", C.c_uint64)]


class _ServeStats(C.Structure):
    _fields_ = [("offered", C.c_uint64), ("committed", C.c_uint64), ("aborted_attempts", C.c_uint64),
                ("batches", C.c_uint64), ("wall_s", C.c_double), ("committed_per_s", C.c_double),
                ("p50_ms", C.c_double), ("p90_ms", C.c_double), ("p99_ms", C.c_double),
                ("p999_ms", C.c_double), ("max_ms", C.c_double), ("mean_batch_txns", C.c_double),
                ("host_pack_s", C.c_double), ("device_ms", C.c_double),
                ("peak_inflight_bytes", C.c_uint64), ("budget_holds", C.c_uint64)]


class _ServeRecord(C.Structure):
    _fields_ = [("batch_off", C.c_void_p), ("batch_txns", C.c_void_p), ("batch_commit", C.c_void_p),
                ("cap_batches", C.c_uint32), ("cap_entries", C.c_uint32), ("n_batches", C.c_uint32),
                ("n_entries", C.c_uint32), ("overflow", C.c_uint32), ("results", C.c_void_p)]


WORKLOAD_API = {
    "ltpmg_gen_kv_max_ops": (C.c_uint64, [C.POINTER(_KvSpec)]),
    "ltpmg_gen_kv": (C.c_int, [C.POINTER(_KvSpec), C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint32]),
    "ltpmg_zeta": (C.c_double, [C.c_uint64, C.c_double]),
}

class _OeSpec(C.Structure):
    _fields_ = [("n_txns", C.c_uint32), ("warehouses", C.c_uint32), ("neworder_pct", C.c_uint32),
                ("payment_pct", C.c_uint32), ("delivery_pct", C.c_uint32), ("lines_min", C.c_uint32),
                ("lines_max", C.c_uint32), ("remote_bp", C.c_uint32), ("preload", C.c_uint32),
                ("byname_pct", C.c_uint32), ("seed", C.c_uint64), ("table_seed", C.c_uint64)]


OE_APPEND_MASK = 0x78  # LTPMG_OE_APPEND_MASK: neworder, history, order, orderline
OE_PRELOAD = 3000      # the reference's orders per district


OE_API = {
    "ltpmg_oe_schema": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "ltpmg_oe_last_name_index": (C.c_int, [C.c_uint32, C.c_uint64, C.c_void_p]),
    "ltpmg_gen_oe_max_ops": (C.c_uint64, [C.POINTER(_OeSpec)]),
    "ltpmg_gen_oe": (C.c_int, [C.POINTER(_OeSpec), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_uint64, C.POINTER(C.c_uint32), C.c_uint32]),
}

class _FixtureInfo(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("mode", C.c_uint32), ("n_tables", C.c_uint32),
                ("tables", _TableDesc * MAX_TABLES), ("append_mask", C.c_uint32),
                ("n_present", C.c_uint32 * MAX_TABLES)]


FIXTURE_NO_MODE = 0xFFFFFFFF

FIXTURE_API = {
    "ltpmg_fixture_write": (C.c_int, [C.c_char_p, C.POINTER(_FixtureInfo), C.POINTER(_Batch)]),
    "ltpmg_fixture_probe": (C.c_int, [C.c_char_p, C.POINTER(_FixtureInfo), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32)]),
    "ltpmg_fixture_read": (C.c_int, [C.c_char_p, C.POINTER(_Batch)]),
}

SERVE_API = {
    "ltpmg_serve_run": (C.c_int, [C.c_void_p, C.POINTER(_ServeCfg), C.POINTER(_Batch),
                                  C.POINTER(_ServeStats), C.POINTER(_ServeRecord)]),
}

_libs: dict[str, C.CDLL] = {}


def _load(name: str, api: dict) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = LIB_DIR / name
    if not path.exists():
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() "
                           "(python -m ltpmg_amd.build) first")
    lib = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    for fn, (res, args) in api.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    _libs[name] = lib
    return lib


def hip_lib() -> C.CDLL:
    return _load("libltpmg_hip.so", {**ENGINE_API, **STREAM_API})


def host_lib() -> C.CDLL:
    hip_lib()  # the host library links against it
    return _load("libltpmg_host.so", {**WORKLOAD_API, **SERVE_API, **OE_API, **FIXTURE_API})


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


@dataclass
class Batch:
    """SoA batch: txn t owns ops [txn_off[t], txn_off[t+1]); its TID is t."""
    txn_off: np.ndarray  # uint32 [n_txns + 1]
    rows: np.ndarray     # uint32 [n_ops]
    meta: np.ndarray     # uint16 [n_ops]
    args: np.ndarray     # uint64 [n_ops]

    @property
    def n_txns(self) -> int:
        return len(self.txn_off) - 1

    @property
    def n_ops(self) -> int:
        return len(self.rows)

    def normalized(self) -> "Batch":
        return Batch(np.ascontiguousarray(self.txn_off, np.uint32),
                     np.ascontiguousarray(self.rows, np.uint32),
                     np.ascontiguousarray(self.meta, np.uint16),
                     np.ascontiguousarray(self.args, np.uint64))

    def c_struct(self) -> _Batch:
        return _Batch(self.n_txns, self.n_ops,
                      self.txn_off.ctypes.data_as(C.POINTER(C.c_uint32)),
                      self.rows.ctypes.data_as(C.POINTER(C.c_uint32)),
                      self.meta.ctypes.data_as(C.POINTER(C.c_uint16)),
                      self.args.ctypes.data_as(C.POINTER(C.c_uint64)))

    @staticmethod
    def from_txns(txns: list[list[tuple[int, int, int, int, int]]]) -> "Batch":
        """txns: per txn a list of (kind, table, col, row, arg)."""
        off = [0]
        rows, metas, args = [], [], []
        for ops in txns:
            for kind, table, col, row, arg in ops:
                rows.append(row)
                metas.append(meta(kind, table, col))
                args.append(arg & 0xFFFFFFFFFFFFFFFF)
            off.append(len(rows))
        return Batch(np.array(off, np.uint32), np.array(rows, np.uint32),
                     np.array(metas, np.uint16), np.array(args, np.uint64))


def gen_kv(n_txns: int, n_rows: int, n_cols: int, *, ops: int | tuple[int, int] = 16,
           zipf_theta: float = 0.0, read_pct: int = 50, write_kind: int = OP_UPDATE,
           read_all_cols: bool = False, table_id: int = 0, rmw_delta: int = 1,
           seed: int = 1, n_threads: int = 0) -> Batch:
    lo, hi = (ops, ops) if isinstance(ops, int) else ops
    spec = _KvSpec(n_txns, lo, hi, n_rows, zipf_theta, n_cols, read_pct, write_kind,
                   int(read_all_cols), table_id, rmw_delta, seed & 0xFFFFFFFFFFFFFFFF)
    lib = host_lib()
    cap = int(lib.ltpmg_gen_kv_max_ops(C.byref(spec)))
    off = np.zeros(n_txns + 1, np.uint32)
    rows = np.empty(cap, np.uint32)
    metas = np.empty(cap, np.uint16)
    args = np.empty(cap, np.uint64)
    n = C.c_uint32(0)
    rc = lib.ltpmg_gen_kv(C.byref(spec), _ptr(off), _ptr(rows), _ptr(metas), _ptr(args), cap,
                          C.byref(n), n_threads)
    if rc != 0:
        raise LtpmgError("ltpmg_gen_kv", rc, "invalid workload spec")
    k = n.value
    return Batch(off, rows[:k].copy(), metas[:k].copy(), args[:k].copy())


def _oe_descs(warehouses: int, append_cap: int, preload: int):
    descs = (_TableDesc * 9)()
    present = (C.c_uint32 * 9)()
    rc = host_lib().ltpmg_oe_schema(warehouses, append_cap, preload, C.cast(descs, C.c_void_p),
                                    C.cast(present, C.c_void_p))
    if rc != 0:
        raise LtpmgError("ltpmg_oe_schema", rc, "bad order-entry sizes")
    return descs, present


def oe_schema(warehouses: int, append_cap: int, preload: int = OE_PRELOAD) -> dict[int, tuple[int, list[int]]]:
    """Order-entry tables (include/ltpmg_oe.h): {table id: (rows, column bytes)}."""
    descs, _ = _oe_descs(warehouses, append_cap, preload)
    return {t: (d.n_rows, list(d.col_bytes[:d.n_cols])) for t, d in enumerate(descs)}


def oe_append(warehouses: int, append_cap: int, preload: int = OE_PRELOAD) -> dict[int, int]:
    """The append tables of the order-entry schema: {table id: rows present at creation}."""
    _, present = _oe_descs(warehouses, append_cap, preload)
    return {t: int(present[t]) for t in range(9) if (OE_APPEND_MASK >> t) & 1}


def create_oe_tables(e: "Engine", warehouses: int, append_cap: int, seed: int,
                     preload: int = OE_PRELOAD, first_table: int = 0):
    """Creates and fills the nine order-entry tables on an engine; returns
    (schema, append) as oe_schema and oe_append give them."""
    schema = oe_schema(warehouses, append_cap, preload)
    append = oe_append(warehouses, append_cap, preload)
    for t, (n, widths) in sorted(schema.items()):
        if t in append:
            e.create_append_table(first_table + t, n, widths, append[t])
        else:
            e.create_table(first_table + t, n, widths)
        e.fill_table(first_table + t, seed)
    return schema, append


class OeStream:
    """Successive order-entry batches (each with its own seed)."""

    def __init__(self, warehouses: int, *, neworder_pct: int = 45, payment_pct: int = 43,
                 delivery_pct: int = 0, lines: tuple[int, int] = (5, 15), remote_bp: int = 100,
                 preload: int = OE_PRELOAD, seed: int = 1, byname_pct: int = 0, table_seed: int = 0):
        """byname_pct of Payments and Orderstatus select the customer by last
        name, resolved against tables filled from table_seed; preload is the
        orders per district the tables were created with."""
        self.spec = _OeSpec(0, warehouses, neworder_pct, payment_pct, delivery_pct, lines[0], lines[1],
                            remote_bp, preload, byname_pct, seed & 0xFFFFFFFFFFFFFFFF,
                            table_seed & 0xFFFFFFFFFFFFFFFF)
        self.batches = 0

    def next(self, n_txns: int, n_threads: int = 0) -> Batch:
        lib = host_lib()
        self.spec.n_txns = n_txns
        self.spec.seed = (self.spec.seed + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        cap = int(lib.ltpmg_gen_oe_max_ops(C.byref(self.spec)))
        off = np.zeros(n_txns + 1, np.uint32)
        rows = np.empty(cap, np.uint32)
        metas = np.empty(cap, np.uint16)
        args = np.empty(cap, np.uint64)
        n = C.c_uint32(0)
        rc = lib.ltpmg_gen_oe(C.byref(self.spec), _ptr(off), _ptr(rows), _ptr(metas), _ptr(args),
                              cap, C.byref(n), n_threads)
        if rc != 0:
            raise LtpmgError("ltpmg_gen_oe", rc, "invalid order-entry spec")
        self.batches += 1
        k = n.value
        return Batch(off, rows[:k].copy(), metas[:k].copy(), args[:k].copy())


def oe_last_name_index(warehouses: int, table_seed: int) -> np.ndarray:
    """[W * 10, 15] C_IDs a last-name lookup finds (include/ltpmg_oe.h)."""
    out = np.zeros(warehouses * 10 * 15, np.uint32)
    rc = host_lib().ltpmg_oe_last_name_index(warehouses, table_seed & 0xFFFFFFFFFFFFFFFF, _ptr(out))
    if rc != 0:
        raise LtpmgError("ltpmg_oe_last_name_index", rc, "bad arguments")
    return out.reshape(warehouses * 10, 15)


def write_fixture(path: str, batch: Batch, *, seed: int, schema: dict[int, tuple[int, list[int]]],
                  mode: int | None = None, append: dict[int, int] | None = None) -> None:
    """Writes a versioned binary batch fixture (include/ltpmg_fixture.h);
    append = {table id: rows present at creation} of the append tables."""
    info = _FixtureInfo()
    info.seed = seed & 0xFFFFFFFFFFFFFFFF
    info.mode = FIXTURE_NO_MODE if mode is None else mode
    info.n_tables = (max(schema) + 1) if schema else 0
    for t, (rows, widths) in schema.items():
        info.tables[t].n_rows = rows
        info.tables[t].n_cols = len(widths)
        for c, w in enumerate(widths):
            info.tables[t].col_bytes[c] = w
    for t, k in (append or {}).items():
        info.append_mask |= 1 << t
        info.n_present[t] = k
    b = batch.normalized()
    rc = host_lib().ltpmg_fixture_write(str(path).encode(), C.byref(info), C.byref(b.c_struct()))
    if rc != 0:
        raise LtpmgError("ltpmg_fixture_write", rc, f"cannot write {path}")


def read_fixture(path: str) -> tuple[Batch, dict]:
    """(batch, {"seed", "mode", "schema", "append"}) from a binary batch fixture."""
    lib = host_lib()
    info = _FixtureInfo()
    nt, no = C.c_uint32(0), C.c_uint32(0)
    rc = lib.ltpmg_fixture_probe(str(path).encode(), C.byref(info), C.byref(nt), C.byref(no))
    if rc != 0:
        raise LtpmgError("ltpmg_fixture_probe", rc, f"{path} is not a readable batch fixture")
    b = Batch(np.zeros(nt.value + 1, np.uint32), np.zeros(no.value, np.uint32),
              np.zeros(no.value, np.uint16), np.zeros(no.value, np.uint64))
    rc = lib.ltpmg_fixture_read(str(path).encode(), C.byref(b.c_struct()))
    if rc != 0:
        raise LtpmgError("ltpmg_fixture_read", rc, f"{path}: bad arrays or checksum")
    schema = {t: (info.tables[t].n_rows, list(info.tables[t].col_bytes[:info.tables[t].n_cols]))
              for t in range(info.n_tables) if info.tables[t].n_cols}
    mode = None if info.mode == FIXTURE_NO_MODE else info.mode
    append = {t: int(info.n_present[t]) for t in range(info.n_tables) if (info.append_mask >> t) & 1}
    return b, {"seed": info.seed, "mode": mode, "schema": schema, "append": append}


def zeta(n: int, theta: float) -> float:
    return float(host_lib().ltpmg_zeta(n, theta))


class Engine:
    """One MI355X engine (one GPU).  Mirrors the C ABI one call per method."""

    def __init__(self, device: int = 0, n_slots: int = 1, max_txns: int = 0, max_ops: int = 0,
                 hbm_budget: int = 0, profile: bool = False):
        self._lib = hip_lib()
        cfg = _Cfg(device, n_slots, max_txns, max_ops, hbm_budget, CFG_PROFILE if profile else 0)
        h = C.c_void_p()
        rc = self._lib.ltpmg_open(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise LtpmgError("ltpmg_open", rc, "cannot open the GPU engine")
        self._h = h
        self._batches: dict[int, Batch] = {}
        self.tables: dict[int, tuple[int, list[int]]] = {}

    def _check(self, fn: str, rc: int) -> None:
        if rc != 0:
            msg = self._lib.ltpmg_last_error(self._h)
            raise LtpmgError(fn, rc, msg.decode() if msg else "")

    def close(self) -> None:
        if self._h:
            self._lib.ltpmg_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def create_table(self, table_id: int, n_rows: int, col_bytes: list[int]) -> None:
        d = _TableDesc(n_rows, len(col_bytes))
        for i, w in enumerate(col_bytes):
            d.col_bytes[i] = w
        self._check("ltpmg_create_table", self._lib.ltpmg_create_table(self._h, table_id, C.byref(d)))
        self.tables[table_id] = (n_rows, list(col_bytes))

    def create_append_table(self, table_id: int, capacity: int, col_bytes: list[int],
                            n_present: int = 0) -> None:
        """A table of `capacity` rows whose rows [0, n_present) are present;
        INSERT ops append after them, DELETE ops clear presence."""
        d = _TableDesc(capacity, len(col_bytes))
        for i, w in enumerate(col_bytes):
            d.col_bytes[i] = w
        self._check("ltpmg_create_append_table",
                    self._lib.ltpmg_create_append_table(self._h, table_id, C.byref(d), n_present))
        self.tables[table_id] = (capacity, list(col_bytes))

    def table_rows(self, table_id: int) -> int:
        """An append table's cursor: rows allocated so far."""
        v = C.c_uint32(0)
        self._check("ltpmg_table_rows", self._lib.ltpmg_table_rows(self._h, table_id, C.byref(v)))
        return v.value

    def truncate_table(self, table_id: int, n_rows: int) -> None:
        """An append table's cursor back to n_rows; rows from there on are gone."""
        self._check("ltpmg_truncate_table", self._lib.ltpmg_truncate_table(self._h, table_id, n_rows))

    def read_presence(self, table_id: int, row0: int = 0, n_rows: int | None = None) -> np.ndarray:
        n = self.tables[table_id][0] - row0 if n_rows is None else n_rows
        out = np.empty(n, np.uint8)
        self._check("ltpmg_read_presence",
                    self._lib.ltpmg_read_presence(self._h, table_id, row0, n, _ptr(out)))
        return out

    def fill_table(self, table_id: int, seed: int) -> None:
        self._check("ltpmg_fill_table", self._lib.ltpmg_fill_table(self._h, table_id, seed))

    def load_column(self, table_id: int, col: int, data: np.ndarray, row0: int = 0) -> None:
        w = self.tables[table_id][1][col]
        a = np.ascontiguousarray(data, np.uint8).reshape(-1, w)
        self._check("ltpmg_load_column",
                    self._lib.ltpmg_load_column(self._h, table_id, col, row0, a.shape[0], _ptr(a)))

    def read_column(self, table_id: int, col: int, row0: int = 0, n_rows: int | None = None) -> np.ndarray:
        rows, widths = self.tables[table_id]
        n = rows - row0 if n_rows is None else n_rows
        out = np.empty((n, widths[col]), np.uint8)
        self._check("ltpmg_read_column",
                    self._lib.ltpmg_read_column(self._h, table_id, col, row0, n, _ptr(out)))
        return out

    def stats_hbm(self) -> int:
        """Bytes of HBM the engine holds (tables, lock words, slots)."""
        s = _Stats()
        self._lib.ltpmg_stats(self._h, 0, C.byref(s))  # fills hbm_allocated even before any batch
        return int(s.hbm_allocated)

    def digest(self, table_id: int) -> int:
        v = C.c_uint64(0)
        self._check("ltpmg_table_digest", self._lib.ltpmg_table_digest(self._h, table_id, C.byref(v)))
        return v.value

    @property
    def result_stride(self) -> int:
        return int(self._lib.ltpmg_result_stride(self._h))

    def submit(self, batch: Batch, slot: int = 0) -> None:
        b = batch.normalized()
        self._check("ltpmg_submit_batch", self._lib.ltpmg_submit_batch(self._h, slot, C.byref(b.c_struct())))
        self._batches[slot] = b

    def submit_runs(self, pool: Batch, ids: np.ndarray, slot: int = 0) -> Batch:
        """Submits pool transactions `ids` in that order through
        ltpmg_submit_runs (consecutive pool txns form one run of ops) and
        returns the batch that went up."""
        ids = np.asarray(ids, np.int64)
        b = take_txns(pool, ids)
        st = _Staging()
        self._check("ltpmg_stage_batch", self._lib.ltpmg_stage_batch(self._h, slot, b.n_txns, b.n_ops, C.byref(st)))
        C.memmove(st.txn_off, b.txn_off.ctypes.data, b.txn_off.nbytes)
        brk = np.flatnonzero(np.diff(ids) != 1) + 1
        starts = np.concatenate([[0], brk])
        ends = np.concatenate([brk, [len(ids)]])
        off = pool.txn_off.astype(np.int64)
        runs = np.zeros((len(starts), 3), np.uint32)
        runs[:, 0] = off[ids[starts]]
        runs[:, 1] = b.txn_off[starts]
        runs[:, 2] = off[ids[ends - 1] + 1] - off[ids[starts]]
        keep = runs[:, 2] > 0
        runs = np.ascontiguousarray(runs[keep])
        pc = pool.normalized()
        self._check("ltpmg_submit_runs", self._lib.ltpmg_submit_runs(self._h, slot, C.byref(pc.c_struct()),
                                                                     _ptr(runs), len(runs)))
        self._batches[slot] = b
        return b

    def submit_staged(self, batch: Batch, slot: int = 0, trusted: bool = False) -> None:
        """Writes the batch straight into the slot's pinned staging, then uploads."""
        b = batch.normalized()
        st = _Staging()
        self._check("ltpmg_stage_batch",
                    self._lib.ltpmg_stage_batch(self._h, slot, b.n_txns, b.n_ops, C.byref(st)))
        C.memmove(st.txn_off, b.txn_off.ctypes.data, b.txn_off.nbytes)
        if b.n_ops:
            C.memmove(st.rows, b.rows.ctypes.data, b.rows.nbytes)
            C.memmove(st.meta, b.meta.ctypes.data, b.meta.nbytes)
            C.memmove(st.args, b.args.ctypes.data, b.args.nbytes)
        self._check("ltpmg_submit_staged",
                    self._lib.ltpmg_submit_staged(self._h, slot, 1 if trusted else 0))
        self._batches[slot] = b

    def execute(self, mode: int, slot: int = 0, profile: bool = False, pipeline: bool = False) -> None:
        """pipeline: an ordered batch takes the multi-kernel pipeline even when
        it is small enough for the one-launch kernel."""
        m = mode | (EXEC_PROFILE if profile else 0) | (EXEC_PIPELINE if pipeline else 0)
        self._check("ltpmg_execute", self._lib.ltpmg_execute(self._h, slot, m))

    def idle(self, slot: int = 0) -> bool:
        rc = self._lib.ltpmg_query(self._h, slot)
        if rc not in (0, 1):
            self._check("ltpmg_query", rc)
        return rc == 0

    def wait(self, slot: int = 0) -> None:
        self._check("ltpmg_wait", self._lib.ltpmg_wait(self._h, slot))

    def fetch(self, slot: int = 0, results: bool = True) -> tuple[np.ndarray, np.ndarray | None]:
        b = self._batches[slot]
        commit = np.empty(b.n_txns, np.uint8)
        res = np.empty((b.n_ops, self.result_stride), np.uint8) if results else None
        self._check("ltpmg_fetch_results",
                    self._lib.ltpmg_fetch_results(self._h, slot, _ptr(commit),
                                                  _ptr(res) if res is not None else None,
                                                  res.nbytes if res is not None else 0))
        return commit, res

    def fetch_status(self, slot: int = 0) -> np.ndarray:
        """Per op: 1 when it took effect (committed, and its row present or
        its INSERT given a row), else 0."""
        out = np.empty(self._batches[slot].n_ops, np.uint8)
        self._check("ltpmg_fetch_status", self._lib.ltpmg_fetch_status(self._h, slot, _ptr(out)))
        return out

    def stats(self, slot: int = 0) -> dict:
        s = _Stats()
        self._check("ltpmg_stats", self._lib.ltpmg_stats(self._h, slot, C.byref(s)))
        kernels, launches = {}, {}
        for i in range(s.n_kernels):
            name = bytes(s.kernel_name[i]).split(b"\0", 1)[0].decode()
            kernels[name] = kernels.get(name, 0.0) + float(s.kernel_ms[i])
            launches[name] = launches.get(name, 0) + 1
        return {"n_txns": s.n_txns, "n_ops": s.n_ops, "committed": s.committed,
                "aborted": s.aborted, "mode": s.mode, "exec_ms": float(s.exec_ms),
                "upload_ms": float(s.upload_ms), "kernels": kernels,
                "kernel_launches": launches, "hbm_allocated": int(s.hbm_allocated),
                "inserted": int(s.inserted)}

    def run(self, batch: Batch, mode: int, slot: int = 0, results: bool = True, pipeline: bool = False):
        self.submit(batch, slot)
        self.execute(mode, slot, pipeline=pipeline)
        return self.fetch(slot, results)

    def selftest_sort(self, keys: np.ndarray, vals: np.ndarray, key_bits: int):
        k = np.ascontiguousarray(keys, np.uint32).copy()
        v = np.ascontiguousarray(vals, np.uint32).copy()
        self._check("ltpmg_selftest_sort",
                    self._lib.ltpmg_selftest_sort(self._h, _ptr(k), _ptr(v), len(k), key_bits))
        return k, v

    def selftest_segscan(self, vals: np.ndarray, heads: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(vals, np.uint32)
        h = np.ascontiguousarray(heads, np.uint8)
        out = np.empty_like(v)
        self._check("ltpmg_selftest_segscan",
                    self._lib.ltpmg_selftest_segscan(self._h, _ptr(v), _ptr(h), _ptr(out), len(v)))
        return out

    def copy_bandwidth(self, n_bytes: int = 1 << 30, iters: int = 20) -> float:
        """GB/s (read + write) of a plain streaming copy of n_bytes."""
        g = C.c_double(0.0)
        self._check("ltpmg_selftest_copy",
                    self._lib.ltpmg_selftest_copy(self._h, n_bytes, iters, C.byref(g)))
        return g.value

    def serve(self, pool: Batch, *, arrival_rate: float, n_arrivals: int, max_batch: int,
              min_batch: int = 0, max_wait_us: float = 200.0, n_inflight: int = 2,
              mode: int = MODE_O, seed: int = 1, hbm_inflight_bytes: int = 0, record: bool = False,
              record_results: bool = False, record_cap: int | None = None) -> dict:
        """Open-loop serving run (include/ltpmg_serve.h).  Returns the stats,
        plus the batch record when asked for (record_cap entries; retried
        txns of the lowest-TID-wins mode take one entry per attempt).
        hbm_inflight_bytes bounds the device bytes of the batches in flight."""
        p = pool.normalized()
        cfg = _ServeCfg(arrival_rate, n_arrivals, max_batch, min_batch, max_wait_us, n_inflight,
                        mode, seed & 0xFFFFFFFFFFFFFFFF, hbm_inflight_bytes)
        st = _ServeStats()
        rec = None
        keep = {}
        if record:
            cap_e = record_cap or (n_arrivals * (8 if mode == MODE_A else 1) + 16)
            keep["off"] = np.zeros(cap_e + 1, np.uint32)
            keep["txns"] = np.zeros(cap_e, np.uint32)
            keep["commit"] = np.zeros(cap_e, np.uint8)
            if record_results:
                keep["results"] = np.zeros((p.n_ops, self.result_stride), np.uint8)
            rec = _ServeRecord(keep["off"].ctypes.data, keep["txns"].ctypes.data,
                               keep["commit"].ctypes.data, cap_e, cap_e, 0, 0, 0,
                               keep["results"].ctypes.data if record_results else None)
        rc = host_lib().ltpmg_serve_run(self._h, C.byref(cfg), C.byref(p.c_struct()), C.byref(st),
                                        C.byref(rec) if rec is not None else None)
        self._check("ltpmg_serve_run", rc)
        out = {f: getattr(st, f) for f, _ in _ServeStats._fields_}
        if rec is not None:
            nb, ne = rec.n_batches, rec.n_entries
            out["record_overflow"] = bool(rec.overflow)
            out["batch_off"] = keep["off"][:nb + 1].copy()
            out["batch_txns"] = keep["txns"][:ne].copy()
            out["batch_commit"] = keep["commit"][:ne].copy()
            if record_results:
                out["results"] = keep["results"]
        return out


@dataclass
class StreamPlan:
    """Plan of a batch stream (include/ltpmg_stream.h)."""
    merge_max_ops: int = 1 << 24
    fused_max_ops: int = 1 << 18
    fused_tile_ops: int = 2048
    fused_wgs_per_cu: int = 2
    use_graph: bool = True
    sync_each: bool = False
    small_max_ops: int = SMALL_MAX_OPS

    def c_struct(self) -> _StreamPlan:
        return _StreamPlan(self.merge_max_ops, self.fused_max_ops, self.fused_tile_ops,
                           self.fused_wgs_per_cu, int(self.use_graph),
                           PLAN_SYNC_EACH if self.sync_each else 0, self.small_max_ops)

    @staticmethod
    def from_c(p: _StreamPlan) -> "StreamPlan":
        return StreamPlan(p.merge_max_ops, p.fused_max_ops, p.fused_tile_ops, p.fused_wgs_per_cu,
                          bool(p.use_graph), bool(p.flags & PLAN_SYNC_EACH), p.small_max_ops)

    @staticmethod
    def plain() -> "StreamPlan":
        """Every batch alone as ltpmg_execute runs it (three-kernel pipeline in
        lowest-TID-wins; ordered pipeline, or the one-launch kernel for small
        ordered batches), host wait after each."""
        return StreamPlan(0, 0, 2048, 1, False, True, SMALL_MAX_OPS)

    @staticmethod
    def default() -> "StreamPlan":
        p = _StreamPlan()
        hip_lib().ltpmg_stream_default_plan(C.byref(p))
        return StreamPlan.from_c(p)


def concat_batches(batches: list[Batch]) -> Batch:
    """The batches back to back, txn offsets running over all of them."""
    offs, base = [np.zeros(1, np.uint64)], 0
    for b in batches:
        offs.append(b.txn_off[1:].astype(np.uint64) + base)
        base += b.n_ops
    if base >= 1 << 32:
        raise ValueError("a stream holds fewer than 2^32 ops")
    return Batch(np.concatenate(offs).astype(np.uint32),
                 np.concatenate([b.rows for b in batches]).astype(np.uint32),
                 np.concatenate([b.meta for b in batches]).astype(np.uint16),
                 np.concatenate([b.args for b in batches]).astype(np.uint64))


class Stream:
    """A stream of batches resident in HBM (the batching form)."""

    def __init__(self, engine: "Engine", batches: list[Batch], modes: list[int] | int,
                 keep_results: bool = False, rewind: bool = False):
        """rewind: every run first truncates the append tables the stream
        inserts into back to their rows at creation (LTPMG_STREAM_REWIND)."""
        self.engine = engine
        self.batch_txns = np.array([b.n_txns for b in batches], np.uint32)
        self.batch_ops = np.array([b.n_ops for b in batches], np.int64)
        m = [modes] * len(batches) if isinstance(modes, int) else list(modes)
        self.modes = np.array(m, np.uint8)
        self.keep_results = keep_results
        allb = concat_batches(batches).normalized()
        h = C.c_void_p()
        lib = engine._lib
        engine._check("ltpmg_stream_create",
                      lib.ltpmg_stream_create(engine._h, len(batches), _ptr(self.batch_txns),
                                              _ptr(self.modes), C.byref(allb.c_struct()),
                                              (STREAM_KEEP_RESULTS if keep_results else 0) |
                                              (STREAM_REWIND if rewind else 0),
                                              C.byref(h)))
        self._h = h

    @property
    def n_batches(self) -> int:
        return len(self.batch_txns)

    def run(self, plan: StreamPlan | None = None) -> float:
        """Runs every batch once; returns the device milliseconds."""
        ms = C.c_double(0.0)
        p = plan.c_struct() if plan is not None else None
        self.engine._check("ltpmg_stream_run",
                           self.engine._lib.ltpmg_stream_run(self.engine._h, self._h,
                                                             C.byref(p) if p is not None else None,
                                                             C.byref(ms)))
        return ms.value

    def tune(self) -> tuple[StreamPlan, int]:
        p = _StreamPlan()
        n = C.c_uint32(0)
        self.engine._check("ltpmg_stream_tune",
                           self.engine._lib.ltpmg_stream_tune(self.engine._h, self._h, C.byref(p),
                                                              C.byref(n)))
        return StreamPlan.from_c(p), n.value

    def groups(self, plan: StreamPlan | None = None) -> list[int]:
        p = plan.c_struct() if plan is not None else None
        first = np.zeros(self.n_batches, np.uint32)
        n = self.engine._lib.ltpmg_stream_groups(self.engine._h, self._h,
                                                 C.byref(p) if p is not None else None,
                                                 _ptr(first), self.n_batches)
        if n < 0:
            self.engine._check("ltpmg_stream_groups", n)
        return first[:n].tolist()

    def fetch(self, batch: int, results: bool = True) -> tuple[np.ndarray, np.ndarray | None]:
        commit = np.empty(int(self.batch_txns[batch]), np.uint8)
        res = (np.empty((int(self.batch_ops[batch]), self.engine.result_stride), np.uint8)
               if results else None)
        self.engine._check("ltpmg_stream_fetch",
                           self.engine._lib.ltpmg_stream_fetch(
                               self.engine._h, self._h, batch, _ptr(commit),
                               _ptr(res) if res is not None else None,
                               res.nbytes if res is not None else 0))
        return commit, res

    def committed(self) -> np.ndarray:
        out = np.zeros(self.n_batches, np.uint32)
        self.engine._check("ltpmg_stream_committed",
                           self.engine._lib.ltpmg_stream_committed(self.engine._h, self._h,
                                                                   _ptr(out)))
        return out

    def close(self) -> None:
        if self._h and self.engine._h:
            self.engine._lib.ltpmg_stream_destroy(self.engine._h, self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def take_txns(pool: Batch, ids: np.ndarray) -> Batch:
    """The batch made of pool transactions `ids`, in that order (TID = position)."""
    ids = np.asarray(ids, np.int64)
    lens = np.diff(pool.txn_off.astype(np.int64))[ids]
    off = np.zeros(len(ids) + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    starts = pool.txn_off.astype(np.int64)[ids]
    idx = np.repeat(starts - off[:-1], lens) + np.arange(off[-1])
    return Batch(off.astype(np.uint32), pool.rows[idx], pool.meta[idx], pool.args[idx])
