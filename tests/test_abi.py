"""CPU-only checks of the native libraries: they load, export every symbol the
public headers declare, refuse to run without a GPU, and the seeded workload
generator behaves as documented.  No compute runs here."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ltpmg_amd import engine as eng

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path) -> set[str]:
    text = header.read_text()
    return set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s+\**\s*(ltpmg_[a-z_0-9]+)\s*\(", text, re.M))


def _exported(so: Path) -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)], check=True,
                         capture_output=True, text=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_engine_exports_every_declared_symbol():
    decl = _declared(ROOT / "include" / "ltpmg.h")
    assert len(decl) >= 16
    assert decl <= _exported(eng.LIB_DIR / "libltpmg_hip.so")
    assert decl == set(eng.ENGINE_API)


def test_stream_exports_every_declared_symbol():
    decl = _declared(ROOT / "include" / "ltpmg_stream.h")
    decl |= set(re.findall(r"^void (ltpmg_[a-z_0-9]+)\(", (ROOT / "include" / "ltpmg_stream.h").read_text(), re.M))
    assert len(decl) >= 8
    assert decl <= _exported(eng.LIB_DIR / "libltpmg_hip.so")
    assert decl == set(eng.STREAM_API)


def test_serving_and_order_entry_export_every_declared_symbol():
    exported = _exported(eng.LIB_DIR / "libltpmg_host.so")
    for header, api in (("ltpmg_serve.h", eng.SERVE_API), ("ltpmg_oe.h", eng.OE_API),
                        ("ltpmg_fixture.h", eng.FIXTURE_API)):
        decl = _declared(ROOT / "include" / header)
        assert decl and decl <= exported
        assert decl == set(api)


# The reference's GPU library is called through exactly these eight C++
# functions (include/database.h:66-70, include/query.h:63-70,
# include/scheduler.h:33-37); their signatures are the interface.
REFERENCE_EXTERNS = {
    "transfer_database_to_GPU(std::shared_ptr<Param>, Global_Table_Info*, Global_Table*, Global_Table_Index*)",
    "transfer_database_to_CPU(std::shared_ptr<Param>, Global_Table_Info*, Global_Table*, Global_Table_Index*)",
    "initial_query_on_GPU(std::shared_ptr<Param>, std::shared_ptr<std::vector<std::any, std::allocator<std::any> > >, Global_Txn_Info*)",
    "transfer_query_to_GPU(std::shared_ptr<Param>, std::shared_ptr<std::vector<std::any, std::allocator<std::any> > >, Global_Txn_Info*)",
    "transfer_query_to_CPU(std::shared_ptr<Param>, std::shared_ptr<std::vector<std::any, std::allocator<std::any> > >, Global_Txn_Info*)",
    "initial_launcher_on_GPU(std::shared_ptr<Param>)",
    "execute_on_GPU(std::shared_ptr<Param>)",
    "free_launcher_on_GPU(std::shared_ptr<Param>)",
}


def test_gpu_library_exports_exactly_the_eight_reference_externs():
    so = eng.LIB_DIR / "libltpmg_gpu.so"
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", str(so)], check=True,
                         capture_output=True, text=True).stdout
    names = {l.split(" ", 2)[2] for l in out.splitlines() if l.strip()}
    assert names == REFERENCE_EXTERNS
    for forbidden in ("Scheduler", "Database", "Query::", "Param::", "ltpmg_"):
        assert not any(forbidden in n for n in names)


def test_gpu_library_links_only_the_engine():
    out = subprocess.run(["readelf", "-d", str(eng.LIB_DIR / "libltpmg_gpu.so")], check=True,
                         capture_output=True, text=True).stdout
    needed = set(re.findall(r"NEEDED.*\[(.+)\]", out))
    ours = {n for n in needed if "ltpmg" in n}
    assert ours == {"libltpmg_hip.so"}, needed


# The reference's Param members in declaration order (include/param.h:13-96):
# the object crosses the eight externs by shared_ptr, so this order, the
# member types and Result (include/param.h:7-11) as the last member fix its
# layout.
PARAM_FIELDS = """benchmark table_size batch_size epoch_tp epoch_sync deviceIDs neworder_percent filepath
table_cnt device_cnt device_IDs data_distribution zipf_config test_1_size test_2_size warehouse_size
district_size customer_size neworder_size history_size order_size orderline_size stock_size item_size
ycsb_size bitmap_row_cnt bitmap_table_cnt test_query_batch_size test_query_2_batch_size
neworder_query_batch_size payment_query_batch_size orderstatus_query_batch_size
delivery_query_batch_size stocklevel_query_batch_size ycsb_a_query_batch_size ycsb_b_query_batch_size
ycsb_c_query_batch_size ycsb_d_query_batch_size ycsb_e_query_batch_size test_query_subtxn_cnt
test_query_2_subtxn_cnt neworder_query_subtxn_cnt payment_query_subtxn_cnt orderstatus_query_subtxn_cnt
delivery_query_subtxn_cnt stocklevel_query_subtxn_cnt ycsb_a_query_subtxn_cnt ycsb_b_query_subtxn_cnt
ycsb_c_query_subtxn_cnt ycsb_d_query_subtxn_cnt ycsb_e_query_subtxn_cnt test_sub_txn_size
tpcc_all_sub_txn_size tpcc_part_sub_txn_size ycsb_a_sub_txn_size ycsb_b_sub_txn_size ycsb_c_sub_txn_size
ycsb_d_sub_txn_size ycsb_e_sub_txn_size test_query_subtxn_kinds tpcc_all_query_subtxn_kinds
tpcc_part_query_subtxn_kinds ycsb_a_query_subtxn_kinds ycsb_b_query_subtxn_kinds ycsb_c_query_subtxn_kinds
ycsb_d_query_subtxn_kinds ycsb_e_query_subtxn_kinds select_batch_size insert_batch_size update_batch_size
scan_batch_size delete_batch_size result""".split()
PARAM_TYPES = {"benchmark": "std::string", "deviceIDs": "std::string", "filepath": "std::string",
               "data_distribution": "std::string", "zipf_config": "double", "device_IDs": "uint32_t *",
               "result": "Result"}


def test_param_has_the_reference_layout(tmp_path):
    # a compiled probe: member types, declaration order and packing as the
    # reference declares them, Result = {double cost; uint32_t epoch, batch_size}
    # every member right after its predecessor (at the next offset its
    # alignment allows), so nothing can be inserted or reordered unnoticed
    checks = ["static_assert(offsetof(Param, benchmark) == 0, \"benchmark first\");"]
    for a, b in zip(PARAM_FIELDS, PARAM_FIELDS[1:]):
        end = f"(offsetof(Param, {a}) + sizeof(Param::{a}))"
        al = f"alignof(decltype(Param::{b}))"
        checks.append(f"static_assert(offsetof(Param, {b}) == ({end} + {al} - 1) / {al} * {al}, "
                      f"\"{b} right after {a}\");")
    for f in PARAM_FIELDS:
        t = PARAM_TYPES.get(f, "uint32_t")
        checks.append(f"static_assert(std::is_same<decltype(Param::{f}), {t}>::value, \"type of {f}\");")
    src = tmp_path / "probe.cpp"
    src.write_text("""
#include <cstddef>
#include <type_traits>
#include "param.h"
static_assert(sizeof(Result) == 16 && offsetof(Result, cost) == 0 && offsetof(Result, epoch) == 8 &&
              offsetof(Result, batch_size) == 12, "Result is the reference's {double, uint32_t, uint32_t}");
static_assert(std::is_same<decltype(Result::cost), double>::value, "Result::cost");
static_assert(offsetof(Param, result) + sizeof(Result) == sizeof(Param), "result is the last member");
static_assert(!std::is_polymorphic<Param>::value, "no vtable");
""" + "\n".join(checks) + "\nint main() { return 0; }\n")
    proc = subprocess.run(["g++", "-std=c++17", "-Wno-invalid-offsetof", "-fsyntax-only",
                           f"-I{ROOT / 'include'}", f"-I{ROOT / 'include' / 'ltpmg_host'}", str(src)],
                          capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-3000:]


def test_concat_batches_keeps_each_batch():
    a = eng.Batch.from_txns([[(0, 0, 1, 5, 0)], [(1, 0, 2, 6, 9), (0, 0, 0, 7, 0)]])
    b = eng.Batch.from_txns([[(2, 1, 3, 8, 4)]])
    c = eng.concat_batches([a, b])
    assert c.txn_off.tolist() == [0, 1, 3, 4]
    assert c.rows.tolist() == [5, 6, 7, 8] and c.meta.tolist() == a.meta.tolist() + b.meta.tolist()


def test_workload_exports_every_declared_symbol():
    decl = _declared(ROOT / "include" / "ltpmg_workload.h")
    assert decl <= _exported(eng.LIB_DIR / "libltpmg_host.so")
    assert decl == set(eng.WORKLOAD_API)


def test_libraries_load_with_signatures():
    lib = eng.hip_lib()
    for name in eng.ENGINE_API:
        assert getattr(lib, name).restype is not None or name == "ltpmg_close"
    eng.host_lib()


def test_engine_contains_gfx950_code():
    # the embedded offload bundle names its target triple
    data = (eng.LIB_DIR / "libltpmg_hip.so").read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


@pytest.mark.skipif(Path("/dev/kfd").exists(), reason="a GPU is present")
def test_open_without_gpu_fails_loudly():
    with pytest.raises(eng.LtpmgError) as ei:
        eng.Engine()
    assert ei.value.code == -6  # LTPMG_EDEVICE: no CPU fallback


# ---- generator ----
def test_generator_is_thread_count_independent():
    a = eng.gen_kv(3000, 5000, 10, ops=(1, 12), zipf_theta=0.8, seed=42, n_threads=1)
    b = eng.gen_kv(3000, 5000, 10, ops=(1, 12), zipf_theta=0.8, seed=42, n_threads=7)
    for f in ("txn_off", "rows", "meta", "args"):
        assert (getattr(a, f) == getattr(b, f)).all()
    c = eng.gen_kv(3000, 5000, 10, ops=(1, 12), zipf_theta=0.8, seed=43)
    assert not (a.rows[:100] == c.rows[:100]).all()


def test_generator_distinct_rows_and_lengths():
    b = eng.gen_kv(2000, 1000, 4, ops=(1, 16), zipf_theta=0.99, seed=3)
    lens = np.diff(b.txn_off.astype(np.int64))
    assert lens.min() >= 1 and lens.max() <= 16
    assert set(np.unique(lens)) == set(range(1, 17))
    for t in range(b.n_txns):
        r = b.rows[b.txn_off[t]:b.txn_off[t + 1]]
        assert len(set(r.tolist())) == len(r)
    assert b.rows.max() < 1000


def test_generator_read_all_cols_expansion_and_kinds():
    b = eng.gen_kv(500, 100000, 10, ops=10, read_pct=50, write_kind=eng.OP_RMW_ADD,
                   read_all_cols=True, seed=9, rmw_delta=1)
    kinds = b.meta & 3
    cols = (b.meta >> 2) & 63
    i = 0
    n_read, n_write = 0, 0
    while i < b.n_ops:
        if kinds[i] == eng.OP_READ:
            assert (cols[i:i + 10] == np.arange(10)).all() and (b.rows[i:i + 10] == b.rows[i]).all()
            i += 10
            n_read += 1
        else:
            assert kinds[i] == eng.OP_RMW_ADD and b.args[i] == 1
            i += 1
            n_write += 1
    assert n_read + n_write == 5000
    assert abs(n_read / 5000 - 0.5) < 0.05


def _gray_cdf(m: np.ndarray, n: int, theta: float, zetan: float) -> np.ndarray:
    """P(key < m) under Gray et al.'s generator: keys 0 and 1 exact
    (1/zeta_n, 2^-theta/zeta_n), then u -> n * (eta*u - eta + 1)^(1/(1-theta))."""
    zeta2 = 1.0 + 0.5 ** theta
    eta = (1.0 - (2.0 / n) ** (1.0 - theta)) / (1.0 - zeta2 / zetan)
    tail = ((m / n) ** (1.0 - theta) - 1.0) / eta + 1.0
    cdf = np.maximum(zeta2 / zetan, tail)
    cdf = np.where(m <= 1, 1.0 / zetan, cdf)
    return np.clip(cdf, 0.0, 1.0)


@pytest.mark.parametrize("theta", [0.5, 0.99])
def test_zipf_follows_gray_et_al(theta):
    # single-op txns: no distinctness redraws
    n, draws = 1000, 400000
    b = eng.gen_kv(draws, n, 1, ops=1, zipf_theta=theta, seed=5)
    freq = np.bincount(b.rows, minlength=n) / draws
    z = eng.zeta(n, theta)
    # keys 0 and 1 have exactly the Zipf probabilities
    for k in (0, 1):
        p = (k + 1) ** -theta / z
        assert abs(freq[k] - p) < 4 * np.sqrt(p * (1 - p) / draws)
    # the empirical CDF follows the method's closed form
    m = np.array([2, 3, 5, 10, 30, 100, 300, 999], dtype=np.float64)
    expect = _gray_cdf(m, n, theta, z)
    got = np.array([freq[:int(k)].sum() for k in m])
    assert np.all(np.abs(got - expect) < 4 * np.sqrt(expect * (1 - expect) / draws) + 2e-3), (got, expect)


def test_zeta_matches_direct_sum():
    for n, th in [(10, 0.5), (1000, 0.99), (3 << 20, 0.5)]:
        direct = float(np.sum(np.arange(1, n + 1, dtype=np.float64) ** -th))
        assert abs(eng.zeta(n, th) - direct) < 1e-9 * direct


def test_generator_rejects_bad_specs():
    with pytest.raises(eng.LtpmgError):
        eng.gen_kv(10, 5, 1, ops=8)  # not enough rows for distinct keys
    with pytest.raises(eng.LtpmgError):
        eng.gen_kv(10, 100, 10, ops=30, read_all_cols=True)  # > 256 device ops per txn
