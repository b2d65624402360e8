"""The C++ drop-in host API (Scheduler/Database/Query over the C ABI), run
through the ltpmg driver on the GPU: epoch 0's batch and outputs are dumped
and replayed by the oracle on the seeded initial image."""
from __future__ import annotations

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_ffi as orc
from ltpmg_amd.engine import LIB_DIR, read_fixture

ROOT = Path(__file__).resolve().parent.parent


def run_driver(tmp_path: Path, *flags: str) -> dict:
    prefix = tmp_path / "dump"
    cmd = [str(LIB_DIR / "ltpmg"), *flags, f"--dump_path={prefix}", "--minloglevel=1"]
    proc = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert proc.returncode == 0, proc.stderr[-2000:]
    meta = json.loads((tmp_path / "dump.json").read_text())
    batch, info = read_fixture(f"{prefix}.batch")
    meta.update(info)
    meta["batch"] = batch
    meta["commit"] = np.fromfile(f"{prefix}.commit.u8", dtype=np.uint8)
    meta["results"] = np.fromfile(f"{prefix}.results.u8", dtype=np.uint8).reshape(-1, meta["result_stride"])
    meta["status"] = np.fromfile(f"{prefix}.status.u8", dtype=np.uint8)
    meta["present"] = {t: np.fromfile(f"{prefix}.present{t}.u8", dtype=np.uint8) for t in meta["append"]}
    return meta


def check_against_oracle(meta: dict):
    im = orc.Image(meta["schema"], append=meta["append"]).fill(meta["seed"])
    b = meta["batch"]
    oc, ores, found = orc.replay(im, b, meta["mode"], meta["result_stride"], with_found=True)
    assert (meta["commit"] == oc).all()
    m = orc.read_mask(b, oc)
    assert (meta["results"][m] == ores[m]).all()
    assert (meta["status"] == found).all()
    for t in sorted(meta["schema"]):
        assert meta["digests"][t] == orc.digest_columns(t, im.cols[t])
    for t in meta["append"]:
        assert meta["rows"][str(t)] == im.cursor[t]
        assert (meta["present"][t] == im.present[t]).all()
    assert meta["committed"] == int(oc.sum())
    return oc, found


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["A", "O"])
def test_driver_kv_c1(tmp_path, mode):
    meta = run_driver(tmp_path, "--benchmark=KV_C1", "--kv_rows=100000", "--kv_fields=10",
                      "--kv_ops=10", "--batch_size=8192", "--epoch_tp=2", f"--cc_mode={mode}")
    assert meta["batch"].n_txns == 8192
    check_against_oracle(meta)


@pytest.mark.gpu
@pytest.mark.parametrize("bench", ["YCSB_A", "YCSB_B", "YCSB_C", "YCSB_E"])
def test_driver_ycsb(tmp_path, bench):
    meta = run_driver(tmp_path, f"--benchmark={bench}", "--table_size=1", "--batch_size=4096",
                      "--epoch_tp=1", "--data_distribution=zipf", "--zipf_config=0.9", "--cc_mode=A")
    b = meta["batch"]
    lens = np.diff(b.txn_off.astype(np.int64))
    expect = {"YCSB_A": 55, "YCSB_B": 91, "YCSB_C": 100, "YCSB_E": 181}[bench]
    assert (lens == expect).all()
    _, found = check_against_oracle(meta)
    if bench == "YCSB_E":  # one INSERT per committed query, appended after the preloaded rows
        ins = (b.meta & 0xFF) == 3
        assert meta["append"] == {0: meta["schema"][0][0] - 4096}
        assert meta["rows"]["0"] == meta["append"][0] + int(found[ins].sum()) > meta["append"][0]


@pytest.mark.gpu
@pytest.mark.parametrize("bench,mode", [("TPCC_PART", "A"), ("TPCC_PART", "O"), ("OE", "O"),
                                        ("TPCC_ALL", "A"), ("TPCC_ALL", "O")])
def test_driver_order_entry(tmp_path, bench, mode):
    meta = run_driver(tmp_path, f"--benchmark={bench}", "--table_size=2", "--append_cap=20000",
                      "--oe_preload=200", "--batch_size=4096", "--epoch_tp=2", f"--cc_mode={mode}")
    assert len(meta["schema"]) == 9 and set(meta["append"]) == {3, 4, 5, 6}
    oc, found = check_against_oracle(meta)
    b = meta["batch"]
    dels = (b.meta & 0xFF) == 7
    if bench == "TPCC_ALL":  # 4 % deliveries of 10 neworder rows each
        assert dels.sum() > 1000 and found[dels].sum() > 0
    else:
        assert not dels.any()


@pytest.mark.gpu
def test_driver_kv_zipf_ordered(tmp_path):
    meta = run_driver(tmp_path, "--benchmark=KV", "--kv_rows=200000", "--kv_fields=10",
                      "--kv_field_bytes=100", "--kv_ops=16", "--batch_size=16384", "--epoch_tp=3",
                      "--data_distribution=zipf", "--zipf_config=0.99", "--cc_mode=O")
    check_against_oracle(meta)


def test_driver_refuses_without_gpu_or_bad_flags(tmp_path):
    exe = LIB_DIR / "ltpmg"
    bad = subprocess.run([str(exe), "--no_such_flag=1"], capture_output=True, text=True)
    assert bad.returncode == 2 and "unknown flag" in bad.stderr
    if not Path("/dev/kfd").exists():
        proc = subprocess.run([str(exe), "--benchmark=KV_C1", "--batch_size=16", "--epoch_tp=1",
                               "--minloglevel=3"], capture_output=True, text=True)
        assert proc.returncode != 0 and "no usable MI355X" in proc.stderr
