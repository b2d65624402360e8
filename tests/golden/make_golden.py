#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/ from the seeded generator and
the oracle.  Run from the repository root after a build:

    python tests/golden/make_golden.py

Each fixture holds the workload spec, a SHA-256 of the generated batch (so a
change to the generator is caught), and the oracle's outputs: the commit
flags (bit-packed, base64), a SHA-256 of every defined read result, and the
final image digest (the engine's table-digest definition).  The tiny fixture
lists its whole batch and outputs inline.
"""
from __future__ import annotations

import base64
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

import oracle_ffi as orc  # noqa: E402
from ltpmg_amd.engine import OP_READ, OP_RMW_ADD, OP_UPDATE, Batch, gen_kv  # noqa: E402

SPECS = {
    # 100k rows x 10 int32; 8192 txns x 10 ops on distinct uniform keys,
    # 50% READ of the whole row / 50% RMW(+1) of one column
    "c1_kv_int32": dict(table_rows=100_000, col_bytes=[4] * 10, table_seed=0x4C54504D47000001,
                        gen=dict(n_txns=8192, n_rows=100_000, n_cols=10, ops=10, read_pct=50,
                                 write_kind=OP_RMW_ADD, read_all_cols=True, rmw_delta=1,
                                 seed=0x4C54504D47000002)),
    # the 100-byte-field key-value shape at reduced size: 200k rows x 10 x 100 B,
    # 16384 txns x 16 ops, zipf 0.99, 50% READ / 50% UPDATE
    "c2_kv_100b_reduced": dict(table_rows=200_000, col_bytes=[100] * 10, table_seed=0x4C54504D47000004,
                               gen=dict(n_txns=16384, n_rows=200_000, n_cols=10, ops=16, zipf_theta=0.99,
                                        read_pct=50, write_kind=OP_UPDATE, seed=0x4C54504D47000003)),
    # uneven txn lengths 1..64 over a small hot table (batching shape)
    "c4_uneven_small": dict(table_rows=5000, col_bytes=[8, 32], table_seed=0x4C54504D47000005,
                            gen=dict(n_txns=3000, n_rows=5000, n_cols=2, ops=(1, 64), zipf_theta=0.5,
                                     read_pct=60, write_kind=OP_UPDATE, seed=0x4C54504D47000006)),
}


def batch_sha(b: Batch) -> str:
    h = hashlib.sha256()
    for a in (b.txn_off, b.rows, b.meta, b.args):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def outputs(commit: np.ndarray, res: np.ndarray, b: Batch, image: orc.Image) -> dict:
    m = orc.read_mask(b, commit)
    return {
        "committed": int(commit.sum()),
        "commit_bits": base64.b64encode(np.packbits(commit).tobytes()).decode(),
        "reads_sha256": hashlib.sha256(np.ascontiguousarray(res[m]).tobytes()).hexdigest(),
        "n_defined_reads": int(m.sum()),
        "image_digest": image.digest(),
    }


def make(name: str, spec: dict) -> dict:
    b = gen_kv(**spec["gen"])
    schema = {0: (spec["table_rows"], spec["col_bytes"])}
    stride = max(4 if w <= 4 else 8 if w <= 8 else 16 if w <= 16 else (w + 31) // 32 * 32
                 for w in spec["col_bytes"])
    out = {"name": name, "spec": spec, "result_stride": stride, "n_txns": b.n_txns,
           "n_ops": b.n_ops, "batch_sha256": batch_sha(b), "modes": {}}
    for mode_name, mode in (("A", orc.MODE_A), ("O", orc.MODE_O)):
        im = orc.Image(schema).fill(spec["table_seed"])
        commit, res = orc.replay(im, b, mode, stride)
        out["modes"][mode_name] = outputs(commit, res, b, im)
    return out


def tiny() -> dict:
    """12 txns on a 16-row table with every outcome written out."""
    txns = [
        [(OP_READ, 0, 0, 3, 0), (OP_UPDATE, 0, 1, 5, 111)],   # T0 writes row 5
        [(OP_READ, 0, 1, 5, 0)],                              # T1 reads row 5: aborts in A
        [(OP_RMW_ADD, 0, 0, 7, 1), (OP_READ, 0, 0, 8, 0)],    # T2 writes row 7
        [(OP_RMW_ADD, 0, 0, 7, 2)],                           # T3 writes row 7: aborts in A
        [(OP_READ, 0, 0, 9, 0), (OP_READ, 0, 1, 9, 0)],
        [(OP_UPDATE, 0, 0, 9, 222)],                          # T5 writes row 9 read by T4
        [(OP_READ, 0, 0, 9, 0)],                              # aborts in A (T5 < T6)
        [(OP_RMW_ADD, 0, 1, 12, 5), (OP_RMW_ADD, 0, 0, 12, 5)],
        [(OP_READ, 0, 1, 12, 0)],                             # aborts in A
        [(OP_UPDATE, 0, 0, 15, 333), (OP_READ, 0, 1, 14, 0)],
        [(OP_READ, 0, 0, 0, 0)],
        [(OP_UPDATE, 0, 1, 0, 444)],                          # T11 writes row 0 after T10 read it
    ]
    b = Batch.from_txns(txns)
    schema = {0: (16, [4, 8])}
    out = {"name": "tiny", "table_rows": 16, "col_bytes": [4, 8], "table_seed": 99, "result_stride": 8,
           "txns": txns, "modes": {}}
    for mode_name, mode in (("A", orc.MODE_A), ("O", orc.MODE_O)):
        im = orc.Image(schema).fill(99)
        commit, res = orc.replay(im, b, mode, 8)
        m = orc.read_mask(b, commit)
        out["modes"][mode_name] = {
            "commit": commit.tolist(),
            "reads": {str(i): res[i].tobytes().hex() for i in np.nonzero(m)[0]},
            "image": [im.cols[0][c].tobytes().hex() for c in range(2)],
        }
    return out


def main() -> None:
    (HERE / "tiny.json").write_text(json.dumps(tiny(), indent=1) + "\n")
    for name, spec in SPECS.items():
        (HERE / f"{name}.json").write_text(json.dumps(make(name, spec), indent=1) + "\n")
        print("wrote", name)


if __name__ == "__main__":
    main()
