// database.cpp -- host image of the tables (the reference's src/database.cpp).
//
// Schemas: YCSB = 3 int, 3 double, 4 string columns (data/YCSB.json lists
// a..j as int,double,string,int,double,string,int,double,string,string; they
// are stored grouped by type as the reference stores them); KV_C1 = kv_fields
// int columns; KV = kv_fields "string" columns of kv_field_bytes bytes.
// Engine column order is ints, then doubles, then strings.  The KV table can
// be many GB, so it has no host copy: the engine fills it on the device from
// the same seed (ltpmg_fill_table), which gives the same bytes.
#include "database.h"

#include <dlfcn.h>

#include "ltpmg.h"
#include "ltpmg_oe.h"

namespace {
// Param's schema path ("../data/<benchmark>.json", relative as in the
// reference) if it exists, else that file name under --schema_dir or the
// schemas/ directory beside this library; "" for benchmarks without one.
std::string resolve_schema(const std::string &path) {
    if (path.empty() || std::ifstream(path).good()) return path;
    std::string dir = FLAGS_schema_dir;
    if (dir.empty()) {
        Dl_info info;
        if (dladdr(reinterpret_cast<void *>(&resolve_schema), &info) && info.dli_fname) {
            const std::string lib = info.dli_fname;
            dir = lib.substr(0, lib.find_last_of('/') + 1) + "../schemas";
        }
    }
    const std::string alt = dir + "/" + path.substr(path.find_last_of('/') + 1);
    if (std::ifstream(alt).good()) return alt;
    LOG(WARNING) << "schema " << path << " not found (nor " << alt << ")";
    return "";
}

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// the engine's seeded cell bytes (include/ltpmg.h), written to dst
void cell_bytes(uint64_t seed, uint32_t table, uint32_t row, uint32_t col, uint32_t width,
                uint8_t *dst) {
    uint64_t id = ((uint64_t)row << 10) | ((uint64_t)table << 6) | col;
    uint64_t s = mix64(seed + kGolden * (id + 1));
    for (uint32_t b = 0; b < width; b += 8) {
        uint64_t v = mix64(s + (uint64_t)(b / 8 + 1) * kGolden);
        for (uint32_t k = 0; k < 8 && b + k < width; ++k) dst[b + k] = (uint8_t)(v >> (8 * k));
    }
}
}  // namespace

Database::Database(const std::string &path) : filepath(resolve_schema(path)) {
    if (!filepath.empty()) rows = parser.parse(filepath);
}

Database::~Database() {}

std::vector<UINT32> Database::string_to_uint(const std::string &input) {
    // 4 characters per word, first character in the high byte (database.cpp:154-174)
    std::vector<UINT32> out;
    for (size_t i = 0; i < input.size(); i += 4) {
        UINT32 v = 0;
        for (size_t j = 0; j < 4 && i + j < input.size(); ++j)
            v |= (UINT32)(unsigned char)input[i + j] << (24 - 8 * j);
        out.push_back(v);
    }
    return out;
}

void Database::generate(std::shared_ptr<Param> param) { param->set_table_cnt(param->table_cnt); }

void Database::malloc_global_row(std::shared_ptr<Param> param) {
    n_tables = param->table_cnt;
    tables_for_GPU_info = new Global_Table_Info[n_tables];
    tables_for_GPU = new Global_Table[n_tables];
    index_for_GPU = new Global_Table_Index[n_tables];
    for (uint32_t i = 0; i < n_tables; ++i) {
        tables_for_GPU_info[i].table_cnt = n_tables;
        tables_for_GPU[i] = Global_Table{nullptr, nullptr, nullptr};
        index_for_GPU[i].index = nullptr;
    }
    if (ltpmg_host::is_oe(*param)) {
        // nine order-entry tables with the widths of include/ltpmg_oe.h; the
        // per-type counts of Global_Table_Info cannot express them, so only
        // the row counts are recorded and the engine fills the tables.  The
        // schema file must list the same tables in the same order.
        static const char *const kNames[LTPMG_OE_TABLES] = {"Warehouse", "District", "Customer",
                                                           "Neworder", "History", "Order",
                                                           "Orderline", "Stock", "Item"};
        if (rows.empty()) LOG(FATAL) << "order-entry benchmarks need their schema file (TPCC.json)";
        if (rows.size() != LTPMG_OE_TABLES) LOG(FATAL) << filepath << ": " << rows.size() << " tables, 9 expected";
        for (uint32_t i = 0; i < LTPMG_OE_TABLES; ++i)
            if (rows[i].get_table_name() != kNames[i])
                LOG(FATAL) << filepath << ": table " << i << " is " << rows[i].get_table_name() << ", expected "
                           << kNames[i];
        ltpmg_table_desc d[LTPMG_OE_TABLES];
        const ltpmg_host::RunOptions &opt = ltpmg_host::options();
        if (n_tables != LTPMG_OE_TABLES ||
            ltpmg_oe_schema(param->warehouse_size, opt.oe_append_cap, opt.oe_preload, d, nullptr))
            LOG(FATAL) << "order-entry schema: " << param->warehouse_size << " partitions, append capacity "
                       << opt.oe_append_cap << ", " << opt.oe_preload << " orders per district";
        for (uint32_t i = 0; i < n_tables; ++i) {
            tables_for_GPU_info[i].table_size = d[i].n_rows;
            tables_for_GPU_info[i].int_size = d[i].n_cols;
        }
        LOG(INFO) << "order-entry tables for " << param->warehouse_size << " partitions are filled on the device";
        return;
    }
    if (!ltpmg_host::is_ycsb(*param) && !ltpmg_host::is_kv(*param))
        LOG(FATAL) << "benchmark " << param->benchmark << ": schema not built in this engine";
    Global_Table_Info &t = tables_for_GPU_info[0];
    if (ltpmg_host::is_ycsb(*param)) {
        // column counts by type from the schema (src/database.cpp:86-95)
        if (rows.size() != 1) LOG(FATAL) << "YCSB needs a one-table schema file (YCSB.json)";
        for (const Attribute &a : rows[0].getAttributes()) {
            if (a.getType() == "int") ++t.int_size;
            else if (a.getType() == "double") ++t.double_size;
            else ++t.string_size;
        }
        if (t.int_size + t.double_size + t.string_size > LTPMG_MAX_COLS)
            LOG(FATAL) << filepath << ": more than " << LTPMG_MAX_COLS << " columns";
        t.string_length = 8;
        t.table_size = param->ycsb_size;
    } else if (param->benchmark == "KV_C1") {
        t.int_size = ltpmg_host::options().kv_fields;
        t.table_size = ltpmg_host::options().kv_size;
    } else {  // KV: device-filled, no host copy
        t.string_size = ltpmg_host::options().kv_fields;
        t.string_length = ltpmg_host::options().kv_field_bytes / 4;
        t.table_size = ltpmg_host::options().kv_size;
        LOG(INFO) << "KV table " << t.table_size << " x " << t.string_size << " x "
                  << ltpmg_host::options().kv_field_bytes << " B is filled on the device";
        return;
    }
    const size_t n = t.table_size;
    tables_for_GPU[0].int_data = t.int_size ? new INT32[n * t.int_size] : nullptr;
    tables_for_GPU[0].double_data = t.double_size ? new DOUBLE[n * t.double_size] : nullptr;
    tables_for_GPU[0].string_data = t.string_size ? new UINT32[n * t.string_size * t.string_length] : nullptr;
    LOG(INFO) << "table 0: " << n << " rows, int " << t.int_size << ", double " << t.double_size
              << ", string " << t.string_size;
}

void Database::copy_to_global_row(std::shared_ptr<Param>) {
    for (uint32_t tb = 0; tb < n_tables; ++tb) {
        const Global_Table_Info &t = tables_for_GPU_info[tb];
        Global_Table &g = tables_for_GPU[tb];
        if (!g.int_data && !g.double_data && !g.string_data) continue;
        const uint64_t seed = ltpmg_host::options().seed;
        const uint32_t sw = t.string_length * 4;
        auto fill = [&](size_t r0, size_t r1) {
            for (size_t r = r0; r < r1; ++r) {
                uint32_t col = 0;
                for (uint32_t c = 0; c < t.int_size; ++c, ++col)
                    cell_bytes(seed, tb, (uint32_t)r, col, 4, reinterpret_cast<uint8_t *>(&g.int_data[r * t.int_size + c]));
                for (uint32_t c = 0; c < t.double_size; ++c, ++col)
                    cell_bytes(seed, tb, (uint32_t)r, col, 8, reinterpret_cast<uint8_t *>(&g.double_data[r * t.double_size + c]));
                for (uint32_t c = 0; c < t.string_size; ++c, ++col)
                    cell_bytes(seed, tb, (uint32_t)r, col, sw,
                               reinterpret_cast<uint8_t *>(&g.string_data[(r * t.string_size + c) * t.string_length]));
            }
        };
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        std::vector<std::thread> th;
        const size_t n = t.table_size, chunk = (n + hw - 1) / hw;
        for (unsigned w = 0; w < hw; ++w) {
            size_t a = w * chunk, b = std::min(n, a + chunk);
            if (a < b) th.emplace_back(fill, a, b);
        }
        for (auto &x : th) x.join();
    }
}

void Database::free_global_row(std::shared_ptr<Param>) {
    for (uint32_t i = 0; i < n_tables; ++i) {
        delete[] tables_for_GPU[i].int_data;
        delete[] tables_for_GPU[i].double_data;
        delete[] tables_for_GPU[i].string_data;
    }
    delete[] tables_for_GPU;
    delete[] tables_for_GPU_info;
    delete[] index_for_GPU;
    tables_for_GPU = nullptr;
    tables_for_GPU_info = nullptr;
    index_for_GPU = nullptr;
    n_tables = 0;
}

void Database::print() const {
    for (uint32_t i = 0; i < n_tables; ++i)
        std::cout << "table " << i << ": " << tables_for_GPU_info[i].table_size << " rows\n";
}

void Database::print_global_row() const {
    for (uint32_t i = 0; i < n_tables; ++i) {
        const Global_Table_Info &t = tables_for_GPU_info[i];
        const Global_Table &g = tables_for_GPU[i];
        if (!g.int_data) continue;
        for (size_t j = 0; j < std::min<size_t>(t.table_size * t.int_size, 64); ++j) std::cout << g.int_data[j] << " ";
        std::cout << "\n";
    }
}

void Database::test() {}

void Database::transfer_to_GPU(std::shared_ptr<Param> param) {
    transfer_database_to_GPU(param, tables_for_GPU_info, tables_for_GPU, index_for_GPU);
}

void Database::transfer_to_CPU(std::shared_ptr<Param> param) {
    transfer_database_to_CPU(param, tables_for_GPU_info, tables_for_GPU, index_for_GPU);
}
