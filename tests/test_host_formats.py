"""Host-side file formats, on the CPU: the binary batch fixture
(include/ltpmg_fixture.h) and the JSON table schemas the drop-in Database
reads (include/ltpmg_host/row_parser.h)."""
from __future__ import annotations

import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from batches import random_batch
from ltpmg_amd import MODE_A, gen_kv
from ltpmg_amd.engine import LIB_DIR, LtpmgError, read_fixture, write_fixture

SCHEMAS = Path(__file__).resolve().parent.parent / "ltpmg_amd" / "schemas"


@pytest.mark.parametrize("n_txns", [1, 7, 3001])
def test_fixture_round_trip(tmp_path, n_txns):
    rng = np.random.default_rng(n_txns)
    b = random_batch(rng, n_txns, 500, 3, max_len=9)
    schema = {0: (500, [4, 100, 8]), 2: (10, [16])}
    p = tmp_path / "b.batch"
    write_fixture(p, b, seed=0xABCDEF, schema=schema, mode=MODE_A)
    got, info = read_fixture(p)
    for f in ("txn_off", "rows", "meta", "args"):
        assert (getattr(got, f) == getattr(b, f)).all()
    assert info == {"seed": 0xABCDEF, "mode": MODE_A, "schema": schema, "append": {}}


def test_fixture_records_append_tables(tmp_path):
    b = gen_kv(300, 1000, 2, ops=(1, 5), seed=4)
    schema = {0: (1000, [8, 8]), 1: (50, [4]), 3: (7000, [12])}
    p = tmp_path / "a.batch"
    write_fixture(p, b, seed=5, schema=schema, append={1: 0, 3: 6500})
    _, info = read_fixture(p)
    assert info["append"] == {1: 0, 3: 6500} and info["schema"] == schema
    with pytest.raises(LtpmgError):  # more rows present than the table holds
        write_fixture(tmp_path / "x.batch", b, seed=5, schema=schema, append={1: 51})


def test_fixture_version_1_files_still_read(tmp_path):
    # version 1 had no append mask (the word after n_tables) and no n_present
    b = gen_kv(100, 1000, 2, ops=(1, 5), seed=6)
    p = tmp_path / "v2.batch"
    write_fixture(p, b, seed=9, schema={0: (1000, [8, 8])}, mode=MODE_A)
    raw = bytearray(p.read_bytes())
    assert int.from_bytes(raw[8:12], "little") == 2 and raw[32:36] == bytes(4)
    v1 = raw[:32] + raw[36:]
    v1[8:12] = (1).to_bytes(4, "little")
    v1[12:16] = (int.from_bytes(raw[12:16], "little") - 4).to_bytes(4, "little")
    q = tmp_path / "v1.batch"
    q.write_bytes(bytes(v1))
    got, info = read_fixture(q)
    assert (got.args == b.args).all() and info["append"] == {} and info["mode"] == MODE_A


def test_fixture_rejects_damage(tmp_path):
    b = gen_kv(2000, 10_000, 4, ops=(1, 20), seed=3)
    p = tmp_path / "b.batch"
    write_fixture(p, b, seed=1, schema={0: (10_000, [4] * 4)})
    raw = bytearray(p.read_bytes())
    assert raw[:8] == b"LTPMGFX\0"
    _, info = read_fixture(p)
    assert info["mode"] is None

    def broken(mutate) -> Path:
        q = tmp_path / "x.batch"
        data = bytearray(raw)
        mutate(data)
        q.write_bytes(bytes(data))
        return q

    for mutate in (lambda d: d.__setitem__(-5, d[-5] ^ 1),            # payload bit -> checksum
                   lambda d: d.__setitem__(0, ord("X")),              # magic
                   lambda d: d.__setitem__(8, 99),                    # a newer version
                   lambda d: d.__delitem__(slice(-8, None)),          # short file
                   lambda d: d.extend(b"\0")):                        # trailing byte
        with pytest.raises(LtpmgError):
            read_fixture(broken(mutate))


def _driver(tmp_path: Path, *flags: str) -> subprocess.CompletedProcess:
    # cwd has no ../data, so the schema comes from --schema_dir or the package
    return subprocess.run([str(LIB_DIR / "ltpmg"), *flags, "--batch_size=64", "--epoch_tp=1",
                           "--minloglevel=0"], capture_output=True, text=True, cwd=tmp_path, timeout=120)


def test_ycsb_column_counts_come_from_the_schema(tmp_path):
    out = _driver(tmp_path, "--benchmark=YCSB_A", "--table_size=1")
    assert "int 3, double 3, string 4" in out.stderr  # the packaged YCSB.json
    doc = json.loads((SCHEMAS / "YCSB.json").read_text())
    doc["tables"][0]["attributes"] = [{"name": f"k{i}", "type": t, "data_distribution": "x"}
                                      for i, t in enumerate(["int"] * 5 + ["string"] * 2)]
    d = tmp_path / "custom"
    d.mkdir()
    (d / "YCSB.json").write_text(json.dumps(doc))
    out = _driver(tmp_path, "--benchmark=YCSB_C", "--table_size=1", f"--schema_dir={d}")
    assert "int 5, double 0, string 2" in out.stderr


def test_bad_schemas_stop_the_driver(tmp_path):
    d = tmp_path / "bad"
    d.mkdir()
    (d / "YCSB.json").write_text('{"tables": [{"name": "ycsb", "size": 1, "attributes": [')
    out = _driver(tmp_path, "--benchmark=YCSB_A", f"--schema_dir={d}")
    assert out.returncode != 0 and "schema" in out.stderr and "unexpected end" in out.stderr
    tpcc = json.loads((SCHEMAS / "TPCC.json").read_text())
    tpcc["tables"][0], tpcc["tables"][1] = tpcc["tables"][1], tpcc["tables"][0]
    (d / "TPCC.json").write_text(json.dumps(tpcc))
    out = _driver(tmp_path, "--benchmark=TPCC_PART", f"--schema_dir={d}")
    assert out.returncode != 0 and "expected Warehouse" in out.stderr


def test_packaged_order_entry_schema_matches_the_adapter():
    from ltpmg_amd.engine import oe_schema
    tpcc = json.loads((SCHEMAS / "TPCC.json").read_text())["tables"]
    widths = oe_schema(1, 1000, 30)
    assert len(tpcc) == len(widths) == 9
    for t, table in enumerate(tpcc):
        assert len(table["attributes"]) == len(widths[t][1]), table["name"]
        for a, w in zip(table["attributes"], widths[t][1]):
            assert (a["type"] == "int") == (w == 4), (table["name"], a["name"], w)
