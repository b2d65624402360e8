"""The allocation-free entry point, ltpmg_execute_static, against the
existing path (ltpmg_submit_batch + ltpmg_execute + ltpmg_fetch_results).

Two engines on the GPU hold identical tables.  The same two batches run in a
row on each -- the second smaller than the first, so the static scratch
still holds the first call's contents when the second runs -- through the
existing path on one engine and through ltpmg_execute_static on the other,
with the batch, the commit flags and the result slots in device buffers the
test allocates itself through the HIP runtime.  Commit flags, every defined
READ result and the final tables must be equal bit for bit, and both must
equal the oracle's serial replay.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_ffi as orc
from ltpmg_amd import MODE_A, MODE_O, OP_RMW_ADD, OP_UPDATE, Engine, LtpmgError, gen_kv
from ltpmg_amd.engine import STATIC_MAX_OPS, Batch, _Batch, concat_batches, hip_lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class DeviceBuffers:
    """Device memory for the test's inputs and outputs, straight from the
    HIP runtime the engine library is linked against (test plumbing)."""

    def __init__(self):
        hip_lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptrs = []

    def _ok(self, rc: int, what: str) -> None:
        assert rc == 0, f"{what} failed with HIP error {rc}"

    def alloc(self, nbytes: int) -> C.c_void_p:
        p = C.c_void_p()
        self._ok(self.hip.hipMalloc(C.byref(p), max(nbytes, 16)), "hipMalloc")
        self.ptrs.append(p)
        return p

    def upload(self, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self._ok(self.hip.hipMemcpy(p, C.c_void_p(a.ctypes.data), a.nbytes, H2D), "hipMemcpy H2D")
        return p

    def download(self, p: C.c_void_p, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        self._ok(self.hip.hipMemcpy(C.c_void_p(out.ctypes.data), p, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def sync(self) -> None:
        self._ok(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self) -> None:
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _run_static(e: Engine, dev: DeviceBuffers, batch: Batch, mode: int):
    b = batch.normalized()
    d_off, d_rows = dev.upload(b.txn_off), dev.upload(b.rows)
    d_meta, d_args = dev.upload(b.meta), dev.upload(b.args)
    d_commit = dev.alloc(b.n_txns)
    d_res = dev.alloc(b.n_ops * e.result_stride)
    desc = _Batch(b.n_txns, b.n_ops, C.cast(d_off, C.POINTER(C.c_uint32)), C.cast(d_rows, C.POINTER(C.c_uint32)),
                  C.cast(d_meta, C.POINTER(C.c_uint16)), C.cast(d_args, C.POINTER(C.c_uint64)))
    rc = e._lib.ltpmg_execute_static(e._h, mode, C.byref(desc), d_commit, d_res, None)
    e._check("ltpmg_execute_static", rc)
    dev.sync()
    commit = dev.download(d_commit, b.n_txns, np.uint8)
    res = dev.download(d_res, (b.n_ops, e.result_stride), np.uint8)
    return commit, res


SCHEMA = {0: (200_000, [100, 100, 8, 4]), 3: (5000, [16, 32])}
SEED = 0x4C54504D47000077


def _batches():
    big = concat_batches([
        gen_kv(12_000, 200_000, 4, ops=(1, 16), zipf_theta=0.9, read_pct=50, write_kind=OP_UPDATE, seed=71),
        gen_kv(3000, 5000, 2, ops=(1, 8), zipf_theta=0.5, read_pct=40, write_kind=OP_RMW_ADD, table_id=3,
               seed=72)])
    small = gen_kv(2500, 200_000, 4, ops=(1, 12), zipf_theta=0.99, read_pct=50, write_kind=OP_RMW_ADD, seed=73)
    assert big.n_ops <= STATIC_MAX_OPS and small.n_ops < big.n_ops // 4
    return big, small


@pytest.mark.parametrize("mode", [MODE_O, MODE_A], ids=["ordered", "lowest_tid_wins"])
def test_static_entry_point_matches_execute(mode):
    dev = DeviceBuffers()
    with Engine() as ref, Engine() as sta:
        for e in (ref, sta):
            for t, (n, widths) in sorted(SCHEMA.items()):
                e.create_table(t, n, widths)
                e.fill_table(t, SEED)
        im = orc.Image(SCHEMA).fill(SEED)
        try:
            for k, b in enumerate(_batches()):
                c_ref, r_ref = ref.run(b, mode)
                c_sta, r_sta = _run_static(sta, dev, b, mode)
                oc, ores = orc.replay(im, b, mode, ref.result_stride)
                assert (c_sta == c_ref).all(), f"call {k}: commit flags differ from ltpmg_execute"
                assert (c_ref == oc).all(), f"call {k}: commit flags differ from the oracle"
                m = orc.read_mask(b, oc)
                assert (r_sta[m] == r_ref[m]).all(), f"call {k}: READ results differ from ltpmg_execute"
                assert (r_ref[m] == ores[m]).all(), f"call {k}: READ results differ from the oracle"
                for t, (n, widths) in SCHEMA.items():
                    assert sta.digest(t) == ref.digest(t) == orc.digest_columns(t, im.cols[t]), \
                        f"call {k}: table {t} differs"
                if mode == MODE_A:
                    assert 0 < int(oc.sum()) < b.n_txns
        finally:
            dev.sync()
            dev.free()


def test_static_entry_point_rejects_oversized_batches():
    with Engine() as e:
        e.create_table(0, 100, [4])
        fake = _Batch(STATIC_MAX_OPS // 2, STATIC_MAX_OPS + 1, None, None, None, None)
        with pytest.raises(LtpmgError) as ei:
            e._check("ltpmg_execute_static",
                     e._lib.ltpmg_execute_static(e._h, MODE_O, C.byref(fake), None, None, None))
        assert ei.value.code == -1 and "static scratch" in str(ei.value)
