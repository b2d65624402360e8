// scheduler.cpp -- the batch loop of the host API (the reference's
// src/scheduler.cpp:27-73: same order of calls; the same batch is executed
// once per epoch).  The log reports committed txn/s, which equals the
// reference's batch_size / cost only when nothing aborts.
#include "scheduler.h"

#include "ltpmg.h"
#include "ltpmg_fixture.h"
#include "ltpmg_oe.h"
#include "pack.h"

namespace {

template <typename T>
void write_vec(const std::string &path, const std::vector<T> &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    if (!f) LOG(FATAL) << "cannot write " << path;
}

// The engine the reference-interface entry points run on.
ltpmg_engine *gpu_engine(const std::shared_ptr<Param> &param) {
    const int32_t dev = param->device_IDs && param->device_cnt ? (int32_t)param->device_IDs[0] : 0;
    ltpmg_engine *e = nullptr;
    if (ltpmg_shared_engine(dev, &e) != LTPMG_OK) LOG(FATAL) << "no usable MI355X";
    return e;
}

void check(ltpmg_engine *e, int rc, const char *what) {
    if (rc != LTPMG_OK) LOG(FATAL) << what << " failed (" << rc << "): " << ltpmg_last_error(e);
}

// Epoch 0's batch as a binary fixture (include/ltpmg_fixture.h: the batch,
// the fill seed, the mode and the table shapes) plus its outputs, so a
// checker can replay the batch on the seeded initial image.  The batch is
// packed again from the same queries with the same rules the entry points
// use (pack.h), and the outputs are read from the engine they ran on.
void dump(const std::shared_ptr<Param> &param, const Global_Table_Info *info, const std::vector<std::any> &queries) {
    const ltpmg_host::RunOptions &opt = ltpmg_host::options();
    ltpmg_engine *e = gpu_engine(param);
    KV_Batch batch;
    std::string why;
    if (!ltpmg_pack::pack_queries(queries, info[0], opt.seed, opt.oe_preload, batch, &why))
        LOG(FATAL) << "cannot pack " << why;
    const uint32_t n_txns = (uint32_t)(batch.txn_off.size() - 1), n_ops = (uint32_t)batch.rows.size();
    const uint32_t stride = ltpmg_result_stride(e);
    std::vector<uint8_t> commit(n_txns), results((size_t)n_ops * stride);
    check(e, ltpmg_fetch_results(e, 0, commit.data(), results.data(), results.size()), "ltpmg_fetch_results");
    ltpmg_batch_stats st{};
    check(e, ltpmg_stats(e, 0, &st), "ltpmg_stats");
    if (st.n_txns != n_txns || st.n_ops != n_ops) LOG(FATAL) << "the engine ran another batch than the one packed";
    const std::string &p = FLAGS_dump_path;
    ltpmg_fixture_info fx;
    std::memset(&fx, 0, sizeof fx);
    fx.seed = opt.seed;
    fx.mode = opt.cc_mode;
    fx.n_tables = param->table_cnt;
    if (ltpmg_host::is_oe(*param)) {
        uint32_t present[LTPMG_OE_TABLES];
        ltpmg_oe_schema(param->warehouse_size, opt.oe_append_cap, opt.oe_preload, fx.tables, present);
        fx.append_mask = LTPMG_OE_APPEND_MASK;
        for (uint32_t t = 0; t < LTPMG_OE_TABLES; ++t) fx.n_present[t] = present[t];
    } else {
        for (uint32_t t = 0; t < param->table_cnt; ++t) {
            const Global_Table_Info &ti = info[t];
            ltpmg_table_desc &d = fx.tables[t];
            d.n_rows = ti.table_size;
            for (uint32_t c = 0; c < ti.int_size; ++c) d.col_bytes[d.n_cols++] = 4;
            for (uint32_t c = 0; c < ti.double_size; ++c) d.col_bytes[d.n_cols++] = 8;
            for (uint32_t c = 0; c < ti.string_size; ++c) d.col_bytes[d.n_cols++] = 4 * ti.string_length;
        }
        if (param->benchmark == "YCSB_E") {
            fx.append_mask = 1u;
            fx.n_present[0] = info[0].table_size;
            fx.tables[0].n_rows = ltpmg_pack::ycsb_e_capacity(*param, info[0].table_size);
        }
    }
    ltpmg_batch b{n_txns, n_ops, batch.txn_off.data(), batch.rows.data(), batch.meta.data(), batch.args.data()};
    if (ltpmg_fixture_write((p + ".batch").c_str(), &fx, &b) != LTPMG_OK) LOG(FATAL) << "cannot write " << p << ".batch";
    write_vec(p + ".commit.u8", commit);
    write_vec(p + ".results.u8", results);
    std::vector<uint8_t> status(n_ops);
    check(e, ltpmg_fetch_status(e, 0, status.data()), "ltpmg_fetch_status");
    write_vec(p + ".status.u8", status);
    std::ofstream j(p + ".json");
    j << "{\"benchmark\": \"" << param->benchmark << "\", \"result_stride\": " << stride
      << ", \"committed\": " << st.committed << ", \"digests\": [";
    for (uint32_t t = 0; t < param->table_cnt; ++t) {
        uint64_t d = 0;
        check(e, ltpmg_table_digest(e, t, &d), "ltpmg_table_digest");
        j << (t ? ", " : "") << d;
    }
    // append tables: the cursor after the batch and the presence bytes
    j << "], \"rows\": {";
    bool first = true;
    for (uint32_t t = 0; t < param->table_cnt; ++t) {
        if (!((fx.append_mask >> t) & 1u)) continue;
        uint32_t rows = 0;
        check(e, ltpmg_table_rows(e, t, &rows), "ltpmg_table_rows");
        std::vector<uint8_t> pres(fx.tables[t].n_rows);
        check(e, ltpmg_read_presence(e, t, 0, fx.tables[t].n_rows, pres.data()), "ltpmg_read_presence");
        write_vec(p + ".present" + std::to_string(t) + ".u8", pres);
        j << (first ? "" : ", ") << "\"" << t << "\": " << rows;
        first = false;
    }
    j << "}}\n";
    LOG(INFO) << "dumped epoch 0 to " << p << ".*";
}

}  // namespace

Scheduler::Scheduler() {}

Scheduler::~Scheduler() {}

void Scheduler::execute(std::shared_ptr<Param> param) { execute_on_GPU(param); }

void Scheduler::init_launcher(std::shared_ptr<Param> param) { initial_launcher_on_GPU(param); }

void Scheduler::free_launcher(std::shared_ptr<Param> param) { free_launcher_on_GPU(param); }

void Scheduler::start() {
    param->print();
    param->analyse_deviceIDs();

    db->generate(param);
    db->malloc_global_row(param);
    db->copy_to_global_row(param);

    query->malloc_global_txn(param);
    query->generate_txn(param);
    db->transfer_to_GPU(param);
    query->initial_on_GPU(param);

    this->init_launcher(param);

    double cost_sum = 0.0;
    uint64_t committed_sum = 0;
    for (uint32_t i = 0; i < param->epoch_tp; ++i) {
        query->transfer_to_GPU(param);
        this->execute(param);
        if (i == 0 && !FLAGS_dump_path.empty()) dump(param, db->get_info(), query->get_batch());
        ltpmg_engine *e = gpu_engine(param);
        ltpmg_batch_stats st{};
        check(e, ltpmg_stats(e, 0, &st), "ltpmg_stats");
        cost_sum += param->result.cost;
        committed_sum += st.committed;
        LOG(INFO) << "epoch " << i << ": " << st.committed << " committed, " << st.aborted << " aborted, "
                  << param->result.cost * 1e3 << " ms, "
                  << (param->result.cost > 0 ? st.committed / param->result.cost : 0.0) << " committed txn/s";
    }
    if (param->epoch_tp)
        LOG(INFO) << "TPS (committed, device time): " << (cost_sum > 0 ? committed_sum / cost_sum : 0.0);

    this->free_launcher(param);
    query->transfer_to_CPU(param);
    db->transfer_to_CPU(param);
    query->free_global_txn();
    db->free_global_row(param);
    param->free_deviceIDs();
}

void Scheduler::end() {
    param.reset();
    db.reset();
    query.reset();
}
