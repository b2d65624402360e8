// gpu_shim.cpp -- the reference's eight extern C++ entry points
// (include/database.h:66-70, include/query.h:63-70, include/scheduler.h:33-37;
// defined in cuda/src/gpuscheduler.cu:3-47) on top of the C ABI of
// libltpmg_hip.so (include/ltpmg.h).
//
// The reference's GPUScheduler singleton (gpuscheduler.cuh:54) becomes one
// engine handle here.  Failures keep the reference's behaviour of its CHECK
// macro (print and exit, cuda/include/define.cuh:57-65): LOG(FATAL).
//
// Query decomposition follows the reference's parse<T> kernels
// (cuda/src/gpuquery.cu:2929-3303): YCSB_A reads rows 0-4 and updates 5-9;
// B and D read 0-8 and update 9; C reads all ten; E scans [r, r+2) for 0-8
// and inserts 9.  A read or scan of a row reads every column (one READ per
// cell, consecutive result slots); an update writes column (TID + j) mod
// n_cols with a payload seeded by (seed, TID, j); the YCSB_E insert writes
// every column of its row (the engine's tables are fixed-size, so an insert
// is an upsert of an existing row id).  The reference's row operators are
// no-ops (gpulauncher.cu:8-165), so these value rules are this build's own.
#include "database.h"
#include "ltpmg.h"
#include "ltpmg_oe.h"
#include "oe.h"
#include "query.h"
#include "scheduler.h"

namespace {

struct ShimState {
    ltpmg_engine *e = nullptr;
    std::vector<Global_Table_Info> info;
    KV_Batch last;
    bool submitted = false, executed = false;
    uint64_t seed = 0;
} g;

void check(int rc, const char *what) {
    if (rc != LTPMG_OK) LOG(FATAL) << what << " failed (" << rc << "): " << ltpmg_last_error(g.e);
}

void ensure_open(const std::shared_ptr<Param> &param) {
    if (g.e) return;
    ltpmg_cfg cfg{};
    cfg.device = param->device_IDs && param->device_cnt ? (int32_t)param->device_IDs[0] : 0;
    cfg.n_slots = 1;
    int rc = ltpmg_open(&cfg, &g.e);
    if (rc != LTPMG_OK) LOG(FATAL) << "ltpmg_open failed (" << rc << "): no usable MI355X";
}

uint32_t n_cols(const Global_Table_Info &t) { return t.int_size + t.double_size + t.string_size; }

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Packer {
    KV_Batch &b;
    uint32_t cols;
    uint64_t seed;
    void read_row(uint32_t row) {
        for (uint32_t c = 0; c < cols; ++c) op(LTPMG_OP_READ, c, row, 0);
    }
    void op(uint32_t kind, uint32_t col, uint32_t row, uint64_t arg) {
        b.rows.push_back(row);
        b.meta.push_back(LTPMG_META(kind, 0, col));
        b.args.push_back(arg);
    }
    uint64_t payload(uint32_t tid, uint32_t j, uint32_t c) const {
        return mix64(seed ^ ((uint64_t)tid << 16) ^ ((uint64_t)j << 8) ^ c);
    }
    void end_txn() { b.txn_off.push_back((uint32_t)b.rows.size()); }
};

template <typename Q>
bool pack_ycsb(const std::any &a, Packer &p, int variant) {
    const Q *q = std::any_cast<Q>(&a);
    if (!q) return false;
    const uint32_t tid = (uint32_t)(p.b.txn_off.size() - 1);
    for (uint32_t j = 0; j < 10; ++j) {
        const uint32_t r = q->ROW_ID[j];
        bool update = (variant == 'A' && j >= 5) || ((variant == 'B' || variant == 'D') && j == 9);
        if (variant == 'E') {
            if (j < 9) {
                p.read_row(r);
                p.read_row(r + 1);
            } else {
                for (uint32_t c = 0; c < p.cols; ++c) p.op(LTPMG_OP_UPDATE, c, r, p.payload(tid, j, c));
            }
        } else if (update) {
            uint32_t c = (tid + j) % p.cols;
            p.op(LTPMG_OP_UPDATE, c, r, p.payload(tid, j, c));
        } else {
            p.read_row(r);
        }
    }
    p.end_txn();
    return true;
}

// order-entry queries: include/ltpmg_oe.h (the parse<Neworder_Query> ... of
// cuda/src/gpuquery.cu:2473-2927, with this build's keys and columns)
template <typename Q>
bool pack_oe(const std::any &a, oe::Sink &sink, KV_Batch &out) {
    const Q *q = std::any_cast<Q>(&a);
    if (!q) return false;
    oe::decompose(*q, sink);
    out.txn_off.push_back((uint32_t)out.rows.size());
    return true;
}

// vector<any> -> SoA batch
void pack(const std::vector<std::any> &batch, KV_Batch &out) {
    out = KV_Batch{};
    if (batch.size() == 1 && batch[0].type() == typeid(KV_Batch)) {
        out = std::any_cast<const KV_Batch &>(batch[0]);
        return;
    }
    if (g.info.empty()) LOG(FATAL) << "transfer_query_to_GPU before transfer_database_to_GPU";
    Packer p{out, n_cols(g.info[0]), g.seed ^ 0x5943534Full};
    oe::Sink sink{out.rows, out.meta, out.args, g.seed ^ 0x4F45ull, g.seed};
    out.txn_off.push_back(0);
    for (const std::any &a : batch) {
        if (pack_ycsb<YCSB_A_Query>(a, p, 'A') || pack_ycsb<YCSB_B_Query>(a, p, 'B') ||
            pack_ycsb<YCSB_C_Query>(a, p, 'C') || pack_ycsb<YCSB_D_Query>(a, p, 'D') ||
            pack_ycsb<YCSB_E_Query>(a, p, 'E') || pack_oe<Neworder_Query>(a, sink, out) ||
            pack_oe<Payment_Query>(a, sink, out) || pack_oe<Orderstatus_Query>(a, sink, out) ||
            pack_oe<Stocklevel_Query>(a, sink, out))
            continue;
        LOG(FATAL) << "query type " << a.type().name() << " is not handled by this engine";
    }
}

// column c of table t: offset and width inside the row-major host arrays
struct HostCol {
    const uint8_t *base;
    size_t row_stride, offset;
    uint32_t width;
};

HostCol host_col(const Global_Table_Info &t, const Global_Table &gt, uint32_t c) {
    if (c < t.int_size)
        return {reinterpret_cast<const uint8_t *>(gt.int_data), 4ull * t.int_size, 4ull * c, 4};
    c -= t.int_size;
    if (c < t.double_size)
        return {reinterpret_cast<const uint8_t *>(gt.double_data), 8ull * t.double_size, 8ull * c, 8};
    c -= t.double_size;
    const uint32_t sw = 4 * t.string_length;
    return {reinterpret_cast<const uint8_t *>(gt.string_data), (size_t)sw * t.string_size, (size_t)sw * c, sw};
}

}  // namespace

void transfer_database_to_GPU(std::shared_ptr<Param> param, Global_Table_Info *info,
                              Global_Table *tables, Global_Table_Index *) {
    ensure_open(param);
    g.seed = param->seed;
    g.info.assign(info, info + param->table_cnt);
    if (param->is_oe()) {  // include/ltpmg_oe.h: created and filled on the device
        ltpmg_table_desc d[LTPMG_OE_TABLES];
        check(ltpmg_oe_schema(param->warehouse_size, param->oe_append_cap, d), "ltpmg_oe_schema");
        for (uint32_t t = 0; t < LTPMG_OE_TABLES; ++t) {
            check(ltpmg_create_table(g.e, t, &d[t]), "ltpmg_create_table");
            check(ltpmg_fill_table(g.e, t, param->seed), "ltpmg_fill_table");
        }
        LOG(INFO) << "transfer_database_to_GPU: order-entry tables resident in HBM";
        return;
    }
    for (uint32_t t = 0; t < param->table_cnt; ++t) {
        ltpmg_table_desc d{};
        d.n_rows = info[t].table_size;
        d.n_cols = n_cols(info[t]);
        for (uint32_t c = 0; c < d.n_cols; ++c) d.col_bytes[c] = host_col(info[t], tables[t], c).width;
        check(ltpmg_create_table(g.e, t, &d), "ltpmg_create_table");
        const bool host_data = tables[t].int_data || tables[t].double_data || tables[t].string_data;
        if (!host_data) {
            check(ltpmg_fill_table(g.e, t, param->seed), "ltpmg_fill_table");
            continue;
        }
        std::vector<uint8_t> packed;
        for (uint32_t c = 0; c < d.n_cols; ++c) {
            HostCol hc = host_col(info[t], tables[t], c);
            packed.resize((size_t)hc.width * d.n_rows);
            for (size_t r = 0; r < d.n_rows; ++r)
                std::memcpy(&packed[r * hc.width], hc.base + r * hc.row_stride + hc.offset, hc.width);
            check(ltpmg_load_column(g.e, t, c, 0, d.n_rows, packed.data()), "ltpmg_load_column");
        }
    }
    LOG(INFO) << "transfer_database_to_GPU: " << param->table_cnt << " table(s) resident in HBM";
}

void transfer_database_to_CPU(std::shared_ptr<Param> param, Global_Table_Info *info,
                              Global_Table *tables, Global_Table_Index *) {
    // The reference frees the device copy without reading it back
    // (gpudatabase.cu:240-339); this build copies the final image into the
    // host arrays first, when the caller has them, then releases the engine.
    if (!g.e) return;
    for (uint32_t t = 0; t < param->table_cnt; ++t) {
        if (!(tables[t].int_data || tables[t].double_data || tables[t].string_data)) continue;
        const uint32_t cols = n_cols(info[t]), rows = info[t].table_size;
        std::vector<uint8_t> packed;
        for (uint32_t c = 0; c < cols; ++c) {
            HostCol hc = host_col(info[t], tables[t], c);
            packed.resize((size_t)hc.width * rows);
            check(ltpmg_read_column(g.e, t, c, 0, rows, packed.data()), "ltpmg_read_column");
            uint8_t *base = const_cast<uint8_t *>(hc.base);
            for (size_t r = 0; r < rows; ++r)
                std::memcpy(base + r * hc.row_stride + hc.offset, &packed[r * hc.width], hc.width);
        }
    }
    ltpmg_close(g.e);
    g = ShimState{};
}

void initial_query_on_GPU(std::shared_ptr<Param> param, std::shared_ptr<std::vector<std::any>>,
                          Global_Txn_Info *) {
    ensure_open(param);  // per-slot buffers are sized on first submit
}

void transfer_query_to_GPU(std::shared_ptr<Param> param,
                           std::shared_ptr<std::vector<std::any>> batch, Global_Txn_Info *) {
    ensure_open(param);
    pack(*batch, g.last);
    ltpmg_batch b{(uint32_t)(g.last.txn_off.size() - 1), (uint32_t)g.last.rows.size(),
                  g.last.txn_off.data(), g.last.rows.data(), g.last.meta.data(), g.last.args.data()};
    check(ltpmg_submit_batch(g.e, 0, &b), "ltpmg_submit_batch");
    g.submitted = true;
    g.executed = false;
}

void transfer_query_to_CPU(std::shared_ptr<Param>, std::shared_ptr<std::vector<std::any>>,
                           Global_Txn_Info *) {
    // results stay on the device until fetch_results_from_GPU asks for them
    g.submitted = false;
}

void initial_launcher_on_GPU(std::shared_ptr<Param> param) { ensure_open(param); }

void execute_on_GPU(std::shared_ptr<Param> param) {
    if (!g.submitted) LOG(FATAL) << "execute_on_GPU without a submitted batch";
    check(ltpmg_execute(g.e, 0, param->cc_mode), "ltpmg_execute");
    check(ltpmg_wait(g.e, 0), "ltpmg_wait");
    ltpmg_batch_stats st{};
    check(ltpmg_stats(g.e, 0, &st), "ltpmg_stats");
    param->result.cost = st.exec_ms / 1000.0;
    param->result.committed = st.committed;
    param->result.aborted = st.aborted;
    param->result.batch_size = st.n_txns;
    g.executed = true;
}

void free_launcher_on_GPU(std::shared_ptr<Param>) {}

bool fetch_results_from_GPU(std::shared_ptr<Param>, Batch_Results *out) {
    if (!g.e || !g.executed || !out) return false;
    out->batch = g.last;
    out->result_stride = ltpmg_result_stride(g.e);
    out->commit.resize(g.last.txn_off.size() - 1);
    out->results.resize((size_t)g.last.rows.size() * out->result_stride);
    check(ltpmg_fetch_results(g.e, 0, out->commit.data(), out->results.data(), out->results.size()),
          "ltpmg_fetch_results");
    return true;
}

UINT64 table_digest_on_GPU(std::shared_ptr<Param>, UINT32 table_id) {
    uint64_t d = 0;
    if (g.e) check(ltpmg_table_digest(g.e, table_id, &d), "ltpmg_table_digest");
    return d;
}
