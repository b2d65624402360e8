"""Serving form: results in the overlapped, continuously batched stream are the
same as running each recorded batch alone (plain path, one at a time) and as
the oracle's serial replay of the same batch sequence."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ffi as orc
from ltpmg_amd import MODE_A, MODE_O, OP_RMW_ADD, OP_UPDATE, Batch, Engine, LtpmgError, gen_kv
from ltpmg_amd.engine import take_txns

pytestmark = pytest.mark.gpu

SCHEMA = {0: (50_000, [100, 100, 4, 4])}
SEED = 0x5E57


def _engine(n_slots=2):
    e = Engine(n_slots=n_slots)
    e.create_table(0, *SCHEMA[0])
    e.fill_table(0, SEED)
    return e


def _check_record(pool, out, mode):
    off, ids, flags = out["batch_off"], out["batch_txns"], out["batch_commit"]
    assert len(off) - 1 == out["batches"]
    im = orc.Image(SCHEMA).fill(SEED)
    committed = np.zeros(pool.n_txns, np.int64)
    with _engine(1) as alone:
        stride = alone.result_stride
        for k in range(len(off) - 1):
            sel = ids[off[k]:off[k + 1]] % pool.n_txns
            b = take_txns(pool, sel)
            c_alone, r_alone = alone.run(b, mode)
            c_orc, r_orc = orc.replay(im, b, mode, stride)
            assert (c_alone == flags[off[k]:off[k + 1]]).all(), f"batch {k}: commit differs alone"
            assert (c_orc == c_alone).all(), f"batch {k}: commit differs from the oracle"
            m = orc.read_mask(b, c_orc)
            assert (r_alone[m] == r_orc[m]).all()
            # results recorded in the mix, at pool op positions, for txns committed here
            if "results" in out:
                op = 0
                for j, t in enumerate(sel):
                    lo, hi = int(pool.txn_off[t]), int(pool.txn_off[t + 1])
                    if c_orc[j]:
                        mk = (pool.meta[lo:hi] & 3) == 0
                        assert (out["results"][lo:hi][mk] == r_orc[op:op + hi - lo][mk]).all(), \
                            f"batch {k} txn {t}: mixed result differs"
                        committed[t] += 1
                    op += hi - lo
        for c in range(4):
            assert (alone.read_column(0, c) == im.cols[0][c]).all()
    return im, committed


@pytest.mark.parametrize("mode", [MODE_O, MODE_A])
def test_served_equals_alone_and_oracle(mode):
    # lowest-TID-wins retries aborted txns, and under heavy skew a batch then
    # commits only a few of them; uniform keys keep that test short
    theta = 0.9 if mode == MODE_O else 0.0
    pool = gen_kv(20_000, 50_000, 4, ops=(1, 16), zipf_theta=theta, read_pct=50,
                  write_kind=OP_UPDATE, seed=31)
    with _engine(2) as e:
        out = e.serve(pool, arrival_rate=4e6, n_arrivals=pool.n_txns, max_batch=1500,
                      max_wait_us=50.0, n_inflight=2, mode=mode, seed=9, record=True,
                      record_results=True, record_cap=400_000)
        assert not out["record_overflow"]
        assert out["committed"] == pool.n_txns
        assert out["batches"] > 5
        im, committed = _check_record(pool, out, mode)
        assert (committed == 1).all()
        for c in range(4):
            assert (e.read_column(0, c) == im.cols[0][c]).all()
        assert out["p99_ms"] >= out["p50_ms"] > 0
        if mode == MODE_A:
            assert out["aborted_attempts"] > 0


def test_plain_form_and_rmw():
    pool = gen_kv(6000, 50_000, 4, ops=(2, 8), zipf_theta=0.99, read_pct=40,
                  write_kind=OP_RMW_ADD, seed=32)
    with _engine(1) as e:
        out = e.serve(pool, arrival_rate=0.0, n_arrivals=pool.n_txns, max_batch=1024, min_batch=1024,
                      n_inflight=1, mode=MODE_O, record=True, record_results=True)
        assert out["batches"] == 6
        im, committed = _check_record(pool, out, MODE_O)
        assert (committed == 1).all()
        assert e.digest(0) == im.digest()


def test_many_inflight_and_reused_pool():
    pool = gen_kv(3000, 50_000, 4, ops=(1, 4), zipf_theta=0.5, seed=33)
    with _engine(4) as e:
        out = e.serve(pool, arrival_rate=1e6, n_arrivals=12_000, max_batch=700, max_wait_us=30.0,
                      n_inflight=4, mode=MODE_O, record=True)
        assert out["committed"] == 12_000 and out["batch_off"][-1] == 12_000
        im, _ = _check_record(pool, out, MODE_O)
        assert e.digest(0) == im.digest()


@pytest.mark.parametrize("mode", [MODE_A, MODE_O])
def test_submit_runs_equals_submit(mode):
    # a batch uploaded as runs of a pinned pool (long runs by DMA straight
    # from the pool, short ones gathered in staging) runs exactly like the
    # same batch submitted whole
    import ctypes as C
    from ltpmg_amd.engine import take_txns
    rng = np.random.default_rng(5)
    pool = gen_kv(60_000, 50_000, 4, ops=(1, 16), zipf_theta=0.8, read_pct=50, write_kind=OP_UPDATE, seed=77)
    pinned = {}
    with Engine(n_slots=2) as e:
        for f, dt in (("rows", np.uint32), ("meta", np.uint16), ("args", np.uint64)):
            a = getattr(pool, f)
            p = C.c_void_p()
            assert e._lib.ltpmg_pinned_alloc(a.nbytes, C.byref(p)) == 0
            view = np.ctypeslib.as_array((C.c_uint8 * a.nbytes).from_address(p.value)).view(dt)
            view[:] = a
            pinned[f] = (p, view)
        pp = Batch(pool.txn_off, pinned["rows"][1], pinned["meta"][1], pinned["args"][1])
        # one long run, scattered single txns (retries), the pool's wrap
        ids = np.concatenate([rng.choice(60_000, 40, replace=False), np.arange(10_000, 30_000),
                              np.arange(59_000, 60_000), np.arange(0, 700)])
        e.create_table(0, 50_000, [8, 16, 4, 12])
        e.fill_table(0, 3)
        im = orc.Image({0: (50_000, [8, 16, 4, 12])}).fill(3)
        b = e.submit_runs(pp, ids, slot=1)
        assert (b.rows == take_txns(pool, ids).rows).all()
        e.execute(mode, slot=1)
        commit, res = e.fetch(slot=1)
        oc, ores = orc.replay(im, b, mode, e.result_stride)
        assert (commit == oc).all()
        m = orc.read_mask(b, oc)
        assert (res[m] == ores[m]).all()
        assert e.digest(0) == orc.digest_columns(0, im.cols[0])
        with pytest.raises(LtpmgError):  # runs that do not tile the batch
            st_runs = np.array([[0, 1, 5]], np.uint32)
            from ltpmg_amd.engine import _Staging
            st = _Staging()
            e._check("ltpmg_stage_batch", e._lib.ltpmg_stage_batch(e._h, 0, 1, 5, C.byref(st)))
            e._check("ltpmg_submit_runs", e._lib.ltpmg_submit_runs(e._h, 0, C.byref(pp.normalized().c_struct()),
                                                                    st_runs.ctypes.data, 1))
        for p, _ in pinned.values():
            assert e._lib.ltpmg_pinned_free(p) == 0
