// scheduler.cpp -- the batch loop of the host API (the reference's
// src/scheduler.cpp:27-73: same order of calls; the same batch is executed
// once per epoch).  The log reports committed txn/s, which equals the
// reference's batch_size / cost only when nothing aborts.
#include "scheduler.h"

#include "ltpmg.h"
#include "ltpmg_fixture.h"
#include "ltpmg_oe.h"

namespace {

template <typename T>
void write_vec(const std::string &path, const std::vector<T> &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    if (!f) LOG(FATAL) << "cannot write " << path;
}

// Epoch 0's batch as a binary fixture (include/ltpmg_fixture.h: the batch,
// the fill seed, the mode and the table shapes) plus its outputs, so a
// checker can replay the batch on the seeded initial image.
void dump(const std::shared_ptr<Param> &param, const Global_Table_Info *info) {
    Batch_Results r;
    if (!fetch_results_from_GPU(param, &r)) LOG(FATAL) << "no results to dump";
    const std::string &p = FLAGS_dump_path;
    ltpmg_fixture_info fx;
    std::memset(&fx, 0, sizeof fx);
    fx.seed = param->seed;
    fx.mode = param->cc_mode;
    fx.n_tables = param->table_cnt;
    if (param->is_oe()) {
        ltpmg_oe_schema(param->warehouse_size, param->oe_append_cap, fx.tables);
    } else {
        for (uint32_t t = 0; t < param->table_cnt; ++t) {
            const Global_Table_Info &ti = info[t];
            ltpmg_table_desc &d = fx.tables[t];
            d.n_rows = ti.table_size;
            for (uint32_t c = 0; c < ti.int_size; ++c) d.col_bytes[d.n_cols++] = 4;
            for (uint32_t c = 0; c < ti.double_size; ++c) d.col_bytes[d.n_cols++] = 8;
            for (uint32_t c = 0; c < ti.string_size; ++c) d.col_bytes[d.n_cols++] = 4 * ti.string_length;
        }
    }
    ltpmg_batch b{(uint32_t)(r.batch.txn_off.size() - 1), (uint32_t)r.batch.rows.size(), r.batch.txn_off.data(),
                  r.batch.rows.data(), r.batch.meta.data(), r.batch.args.data()};
    if (ltpmg_fixture_write((p + ".batch").c_str(), &fx, &b) != LTPMG_OK) LOG(FATAL) << "cannot write " << p << ".batch";
    write_vec(p + ".commit.u8", r.commit);
    write_vec(p + ".results.u8", r.results);
    std::ofstream j(p + ".json");
    j << "{\"benchmark\": \"" << param->benchmark << "\", \"result_stride\": " << r.result_stride
      << ", \"committed\": " << param->result.committed << ", \"digests\": [";
    for (uint32_t t = 0; t < param->table_cnt; ++t) j << (t ? ", " : "") << table_digest_on_GPU(param, t);
    j << "]}\n";
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
        if (i == 0 && !FLAGS_dump_path.empty()) dump(param, db->get_info());
        cost_sum += param->result.cost;
        committed_sum += param->result.committed;
        LOG(INFO) << "epoch " << i << ": " << param->result.committed << " committed, "
                  << param->result.aborted << " aborted, " << param->result.cost * 1e3 << " ms, "
                  << (param->result.cost > 0 ? param->result.committed / param->result.cost : 0.0)
                  << " committed txn/s";
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
