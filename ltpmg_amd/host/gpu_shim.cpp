// gpu_shim.cpp -- the reference's eight extern C++ entry points
// (include/database.h:66-70, include/query.h:63-70, include/scheduler.h:33-37;
// defined in cuda/src/gpuscheduler.cu:3-47) on top of the C ABI of
// libltpmg_hip.so (include/ltpmg.h).  Built alone into libltpmg_gpu.so, the
// counterpart of the reference's libgpu (cuda/CMakeLists.txt:28), which
// exports these eight functions and nothing else.
//
// Only the reference's Param members are read (its layout is the
// reference's, include/ltpmg_host/param.h).  Settings the reference has no
// field for come from the environment, read once:
//   LTPMG_CC_MODE     "O" (ordered, every txn commits; the default) or "A"
//                     (lowest TID wins)
//   LTPMG_SEED        table fill and payload seed (decimal or 0x hex;
//                     default 0x4C54504D47000001)
//   LTPMG_APPEND_CAP  order-entry append tables, in transactions (2^20)
//   LTPMG_OE_PRELOAD  order-entry orders per district at creation (3000)
// The reference's GPUScheduler singleton (gpuscheduler.cuh:54) becomes the
// process-wide engine of the device (ltpmg_shared_engine), which host code
// beside these functions can also open to read what they ran.  Failures keep
// the reference's behaviour of its CHECK macro (print and exit,
// cuda/include/define.cuh:57-65).
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "database.h"
#include "ltpmg.h"
#include "oe.h"
#include "pack.h"
#include "query.h"
#include "scheduler.h"

namespace {

struct ShimState {
    ltpmg_engine *e = nullptr;
    std::vector<Global_Table_Info> info;
    KV_Batch last;
    bool submitted = false;
    bool have_opts = false;
    uint32_t mode = LTPMG_MODE_O;
    uint64_t seed = 0x4C54504D47000001ull;
    uint32_t append_cap = 1u << 20;
    uint32_t preload = 3000;
} g;

[[noreturn]] void die(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::fprintf(stderr, "F gpu_shim] ");
    std::vfprintf(stderr, fmt, ap);
    std::fprintf(stderr, "\n");
    va_end(ap);
    std::fflush(stderr);
    std::abort();
}

void check(int rc, const char *what) {
    if (rc != LTPMG_OK) die("%s failed (%d): %s", what, rc, ltpmg_last_error(g.e));
}

void read_options() {
    if (g.have_opts) return;
    g.have_opts = true;
    if (const char *m = std::getenv("LTPMG_CC_MODE")) {
        if (m[0] == 'A' || m[0] == 'a') g.mode = LTPMG_MODE_A;
        else if (m[0] == 'O' || m[0] == 'o') g.mode = LTPMG_MODE_O;
        else die("LTPMG_CC_MODE=%s: expected O or A", m);
    }
    if (const char *s = std::getenv("LTPMG_SEED")) g.seed = std::strtoull(s, nullptr, 0);
    if (const char *c = std::getenv("LTPMG_APPEND_CAP")) g.append_cap = (uint32_t)std::strtoul(c, nullptr, 0);
    if (const char *c = std::getenv("LTPMG_OE_PRELOAD")) g.preload = (uint32_t)std::strtoul(c, nullptr, 0);
}

bool is_oe(const Param &p) { return p.benchmark == "TPCC_PART" || p.benchmark == "TPCC_ALL" || p.benchmark == "OE"; }

void ensure_open(const std::shared_ptr<Param> &param) {
    read_options();
    if (g.e) return;
    const int32_t dev = param->device_IDs && param->device_cnt ? (int32_t)param->device_IDs[0] : 0;
    int rc = ltpmg_shared_engine(dev, &g.e);
    if (rc != LTPMG_OK) die("ltpmg_open failed (%d): no usable MI355X", rc);
}

uint32_t n_cols(const Global_Table_Info &t) { return t.int_size + t.double_size + t.string_size; }

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
    g.info.assign(info, info + param->table_cnt);
    if (is_oe(*param)) {  // include/ltpmg_oe.h: created and filled on the device
        ltpmg_table_desc d[LTPMG_OE_TABLES];
        uint32_t present[LTPMG_OE_TABLES];
        check(oe::schema(param->warehouse_size, g.append_cap, g.preload, d, present), "order-entry schema");
        for (uint32_t t = 0; t < LTPMG_OE_TABLES; ++t) {
            if ((LTPMG_OE_APPEND_MASK >> t) & 1u)
                check(ltpmg_create_append_table(g.e, t, &d[t], present[t]), "ltpmg_create_append_table");
            else
                check(ltpmg_create_table(g.e, t, &d[t]), "ltpmg_create_table");
            check(ltpmg_fill_table(g.e, t, g.seed), "ltpmg_fill_table");
        }
        return;
    }
    for (uint32_t t = 0; t < param->table_cnt; ++t) {
        ltpmg_table_desc d{};
        d.n_rows = info[t].table_size;
        d.n_cols = n_cols(info[t]);
        for (uint32_t c = 0; c < d.n_cols; ++c) d.col_bytes[c] = host_col(info[t], tables[t], c).width;
        if (t == 0 && param->benchmark == "YCSB_E") {
            // E's insert appends a row per query and epoch
            // (cuda/src/gpuquery.cu:3243-3264): rows [0, table_size) are
            // present, the rest of the capacity takes the inserts
            const uint32_t cap = ltpmg_pack::ycsb_e_capacity(*param, d.n_rows);
            if (cap == 0) die("YCSB_E: the table and its inserts exceed 2^32 rows");
            const uint32_t present = d.n_rows;
            d.n_rows = cap;
            check(ltpmg_create_append_table(g.e, t, &d, present), "ltpmg_create_append_table");
        } else {
            check(ltpmg_create_table(g.e, t, &d), "ltpmg_create_table");
        }
        const bool host_data = tables[t].int_data || tables[t].double_data || tables[t].string_data;
        if (!host_data) {
            check(ltpmg_fill_table(g.e, t, g.seed), "ltpmg_fill_table");
            continue;
        }
        std::vector<uint8_t> packed;
        const uint32_t host_rows = info[t].table_size;
        for (uint32_t c = 0; c < d.n_cols; ++c) {
            HostCol hc = host_col(info[t], tables[t], c);
            packed.resize((size_t)hc.width * host_rows);
            for (size_t r = 0; r < host_rows; ++r)
                std::memcpy(&packed[r * hc.width], hc.base + r * hc.row_stride + hc.offset, hc.width);
            check(ltpmg_load_column(g.e, t, c, 0, host_rows, packed.data()), "ltpmg_load_column");
        }
    }
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
    if (g.info.empty()) die("transfer_query_to_GPU before transfer_database_to_GPU");
    std::string why;
    if (!ltpmg_pack::pack_queries(*batch, g.info[0], g.seed, g.preload, g.last, &why))
        die("query type %s is not handled by this engine", why.c_str());
    ltpmg_batch b{(uint32_t)(g.last.txn_off.size() - 1), (uint32_t)g.last.rows.size(),
                  g.last.txn_off.data(), g.last.rows.data(), g.last.meta.data(), g.last.args.data()};
    check(ltpmg_submit_batch(g.e, 0, &b), "ltpmg_submit_batch");
    g.submitted = true;
}

void transfer_query_to_CPU(std::shared_ptr<Param>, std::shared_ptr<std::vector<std::any>>,
                           Global_Txn_Info *) {
    // results stay on the device; the engine's ltpmg_fetch_results reads them
    g.submitted = false;
}

void initial_launcher_on_GPU(std::shared_ptr<Param> param) { ensure_open(param); }

void execute_on_GPU(std::shared_ptr<Param> param) {
    if (!g.submitted) die("execute_on_GPU without a submitted batch");
    check(ltpmg_execute(g.e, 0, g.mode), "ltpmg_execute");
    check(ltpmg_wait(g.e, 0), "ltpmg_wait");
    ltpmg_batch_stats st{};
    check(ltpmg_stats(g.e, 0, &st), "ltpmg_stats");
    param->result.cost = st.exec_ms / 1000.0;
    param->result.batch_size = st.n_txns;
}

void free_launcher_on_GPU(std::shared_ptr<Param>) {}
