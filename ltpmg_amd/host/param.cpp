// param.cpp -- run parameters from the flags (the reference's src/param.cpp:
// the same per-benchmark sizing, src/param.cpp:3-134, plus the KV benchmarks).
#include "param.h"

#include "ltpmg.h"

namespace {
// ops per txn of each YCSB mix as the engine sees them: a read or scan of a
// whole row becomes one READ per column (10 columns in the YCSB schema)
constexpr uint32_t kYcsbCols = 10;
}  // namespace

Param::Param()
    : benchmark(FLAGS_benchmark), table_size(FLAGS_table_size), batch_size(FLAGS_batch_size),
      epoch_tp(FLAGS_epoch_tp), epoch_sync(FLAGS_epoch_sync), deviceIDs(FLAGS_deviceIDs),
      neworder_percent(FLAGS_neworder_percent), data_distribution(FLAGS_data_distribution),
      zipf_config(FLAGS_zipf_config) {
    result.cost = 0.0;
    result.epoch = epoch_tp;
    result.batch_size = batch_size;
    ltpmg_host::RunOptions &opt = ltpmg_host::options();
    if (benchmark == "TEST") {
        filepath = "../data/TEST.json";
        test_query_batch_size = batch_size / 2;
        test_query_2_batch_size = batch_size / 2;
        test_query_subtxn_cnt = 5;
        test_query_2_subtxn_cnt = 5;
        test_sub_txn_size = 5 * test_query_batch_size + 5 * test_query_2_batch_size;
        test_query_subtxn_kinds = 10;
        test_1_size = 8;
        test_2_size = 24;
        table_cnt = 2;
        bitmap_table_cnt = 2;
    } else if (ltpmg_host::is_oe(*this)) {
        filepath = "../data/TPCC.json";
        neworder_query_subtxn_cnt = 5 + 15 * 3;
        payment_query_subtxn_cnt = 7;
        orderstatus_query_subtxn_cnt = 3;
        delivery_query_subtxn_cnt = 5 * 10;
        stocklevel_query_subtxn_cnt = 1 + 2 * 10;
        if (benchmark == "OE") {
            neworder_query_batch_size = batch_size * 45 / 100;
            payment_query_batch_size = batch_size * 43 / 100;
            orderstatus_query_batch_size = stocklevel_query_batch_size =
                (batch_size - neworder_query_batch_size - payment_query_batch_size) / 2;
        } else if (benchmark == "TPCC_PART") {
            neworder_query_batch_size = batch_size * neworder_percent / 100;
            payment_query_batch_size = batch_size - neworder_query_batch_size;
            tpcc_part_sub_txn_size = neworder_query_subtxn_cnt * neworder_query_batch_size +
                                     payment_query_subtxn_cnt * payment_query_batch_size;
            tpcc_part_query_subtxn_kinds = 8 + 7;
        } else {
            neworder_query_batch_size = batch_size * 44 / 100;
            payment_query_batch_size = batch_size * 44 / 100;
            orderstatus_query_batch_size = batch_size * 4 / 100;
            delivery_query_batch_size = batch_size * 4 / 100;
            stocklevel_query_batch_size = batch_size * 4 / 100;
            tpcc_all_sub_txn_size = neworder_query_subtxn_cnt * neworder_query_batch_size +
                                    payment_query_subtxn_cnt * payment_query_batch_size +
                                    orderstatus_query_subtxn_cnt * orderstatus_query_batch_size +
                                    delivery_query_subtxn_cnt * delivery_query_batch_size +
                                    stocklevel_query_subtxn_cnt * stocklevel_query_batch_size;
            tpcc_all_query_subtxn_kinds = 8 + 7 + 3 + 5 + 3;
        }
        warehouse_size = table_size;
        district_size = warehouse_size * 10;
        customer_size = district_size * 3000;
        neworder_size = district_size * 3000;
        history_size = district_size * 3000;
        order_size = district_size * 3000;
        orderline_size = district_size * 45000;
        stock_size = table_size * 100000;
        item_size = 100000;
        table_cnt = 9;
        bitmap_table_cnt = 2;
    } else if (ltpmg_host::is_kv(*this)) {
        filepath = "";
        table_cnt = 1;
        if (benchmark == "KV_C1") opt.kv_field_bytes = 4;
        uint32_t per_read = benchmark == "KV_C1" ? opt.kv_fields : 1;
        opt.kv_sub_txn_size = batch_size * opt.kv_ops * per_read;
    } else {
        filepath = "../data/YCSB.json";
        ycsb_size = table_size * 1000000;
        table_cnt = 1;
        bitmap_row_cnt = 100;
        bitmap_table_cnt = 1;
        // logical ops per txn are 10 for every mix (src/param.cpp:105-132)
        const uint32_t ops = 10;
        if (benchmark == "YCSB_A") {
            ycsb_a_query_batch_size = batch_size;
            ycsb_a_query_subtxn_cnt = ops;
            ycsb_a_sub_txn_size = batch_size * (5 * kYcsbCols + 5);
            ycsb_a_query_subtxn_kinds = ops;
        } else if (benchmark == "YCSB_B") {
            ycsb_b_query_batch_size = batch_size;
            ycsb_b_query_subtxn_cnt = ops;
            ycsb_b_sub_txn_size = batch_size * (9 * kYcsbCols + 1);
            ycsb_b_query_subtxn_kinds = ops;
        } else if (benchmark == "YCSB_C") {
            ycsb_c_query_batch_size = batch_size;
            ycsb_c_query_subtxn_cnt = ops;
            ycsb_c_sub_txn_size = batch_size * 10 * kYcsbCols;
            ycsb_c_query_subtxn_kinds = ops;
        } else if (benchmark == "YCSB_D") {
            ycsb_d_query_batch_size = batch_size;
            ycsb_d_query_subtxn_cnt = ops;
            ycsb_d_sub_txn_size = batch_size * (9 * kYcsbCols + 1);
            ycsb_d_query_subtxn_kinds = ops;
        } else if (benchmark == "YCSB_E") {
            ycsb_e_query_batch_size = batch_size;
            ycsb_e_query_subtxn_cnt = ops;
            ycsb_e_sub_txn_size = batch_size * (9 * 2 * kYcsbCols + kYcsbCols);
            ycsb_e_query_subtxn_kinds = ops;
        }
    }
}

Param::~Param() {}

void Param::print() {
    LOG(INFO) << "benchmark: " << benchmark;
    if (benchmark == "TPCC_ALL" || benchmark == "TPCC_PART") LOG(INFO) << "warehouse_size: " << table_size;
    else if (ltpmg_host::is_kv(*this))
        LOG(INFO) << "table_size: " << ltpmg_host::options().kv_size << " x " << ltpmg_host::options().kv_fields
                  << " x " << ltpmg_host::options().kv_field_bytes << " B";
    else LOG(INFO) << "table_size: " << ycsb_size;
    LOG(INFO) << "batch_size: " << batch_size;
    LOG(INFO) << "epoch_tp: " << epoch_tp;
    LOG(INFO) << "deviceIDs: " << deviceIDs;
    LOG(INFO) << "zipf_config: " << zipf_config;
    LOG(INFO) << "data_distribution: " << data_distribution;
    LOG(INFO) << "cc_mode: " << (ltpmg_host::options().cc_mode == LTPMG_MODE_A ? "A (lowest TID wins)" : "O (ordered)");
    LOG(INFO) << "seed: " << ltpmg_host::options().seed;
}

void Param::analyse_deviceIDs() {
    std::vector<uint32_t> ids;
    std::istringstream in(deviceIDs);
    std::string tok;
    while (std::getline(in, tok, ','))
        if (!tok.empty()) ids.push_back((uint32_t)std::stoul(tok));
    device_cnt = (uint32_t)ids.size();
    device_IDs = new uint32_t[device_cnt ? device_cnt : 1];
    for (uint32_t i = 0; i < device_cnt; ++i) device_IDs[i] = ids[i];
    if (device_cnt != 1) LOG(WARNING) << "this build runs one GPU; using device " << (device_cnt ? ids[0] : 0);
}

void Param::free_deviceIDs() {
    delete[] device_IDs;
    device_IDs = nullptr;
}

uint32_t Param::get_subtxn_kinds() {
    if (benchmark == "TEST") return test_query_subtxn_kinds;
    if (benchmark == "TPCC_PART") return tpcc_part_query_subtxn_kinds;
    if (benchmark == "TPCC_ALL") return tpcc_all_query_subtxn_kinds;
    if (benchmark == "YCSB_A") return ycsb_a_query_subtxn_kinds;
    if (benchmark == "YCSB_B") return ycsb_b_query_subtxn_kinds;
    if (benchmark == "YCSB_C") return ycsb_c_query_subtxn_kinds;
    if (benchmark == "YCSB_D") return ycsb_d_query_subtxn_kinds;
    if (benchmark == "YCSB_E") return ycsb_e_query_subtxn_kinds;
    if (ltpmg_host::is_kv(*this)) return ltpmg_host::options().kv_ops;
    return 0;
}

uint32_t Param::get_sub_txn_size() {
    if (benchmark == "TEST") return test_sub_txn_size;
    if (benchmark == "TPCC_PART") return tpcc_part_sub_txn_size;
    if (benchmark == "TPCC_ALL") return tpcc_all_sub_txn_size;
    if (benchmark == "YCSB_A") return ycsb_a_sub_txn_size;
    if (benchmark == "YCSB_B") return ycsb_b_sub_txn_size;
    if (benchmark == "YCSB_C") return ycsb_c_sub_txn_size;
    if (benchmark == "YCSB_D") return ycsb_d_sub_txn_size;
    if (benchmark == "YCSB_E") return ycsb_e_sub_txn_size;
    if (ltpmg_host::is_kv(*this)) return ltpmg_host::options().kv_sub_txn_size;
    return 0;
}

uint32_t Param::get_subtxn_cnt(const std::type_info &q) {
    if (q == typeid(Test_Query)) return test_query_subtxn_cnt;
    if (q == typeid(Test_Query_2)) return test_query_2_subtxn_cnt;
    if (q == typeid(Neworder_Query)) return neworder_query_subtxn_cnt;
    if (q == typeid(Payment_Query)) return payment_query_subtxn_cnt;
    if (q == typeid(Orderstatus_Query)) return orderstatus_query_subtxn_cnt;
    if (q == typeid(Delivery_Query)) return delivery_query_subtxn_cnt;
    if (q == typeid(Stocklevel_Query)) return stocklevel_query_subtxn_cnt;
    if (q == typeid(YCSB_A_Query)) return ycsb_a_query_subtxn_cnt;
    if (q == typeid(YCSB_B_Query)) return ycsb_b_query_subtxn_cnt;
    if (q == typeid(YCSB_C_Query)) return ycsb_c_query_subtxn_cnt;
    if (q == typeid(YCSB_D_Query)) return ycsb_d_query_subtxn_cnt;
    if (q == typeid(YCSB_E_Query)) return ycsb_e_query_subtxn_cnt;
    if (q == typeid(KV_Batch)) return ltpmg_host::options().kv_ops;
    return 0;
}

uint32_t Param::get_txn_batch_size(const std::type_info &q) {
    if (q == typeid(Test_Query)) return test_query_batch_size;
    if (q == typeid(Test_Query_2)) return test_query_2_batch_size;
    if (q == typeid(Neworder_Query)) return neworder_query_batch_size;
    if (q == typeid(Payment_Query)) return payment_query_batch_size;
    if (q == typeid(Orderstatus_Query)) return orderstatus_query_batch_size;
    if (q == typeid(Delivery_Query)) return delivery_query_batch_size;
    if (q == typeid(Stocklevel_Query)) return stocklevel_query_batch_size;
    if (q == typeid(YCSB_A_Query)) return ycsb_a_query_batch_size;
    if (q == typeid(YCSB_B_Query)) return ycsb_b_query_batch_size;
    if (q == typeid(YCSB_C_Query)) return ycsb_c_query_batch_size;
    if (q == typeid(YCSB_D_Query)) return ycsb_d_query_batch_size;
    if (q == typeid(YCSB_E_Query)) return ycsb_e_query_batch_size;
    if (q == typeid(KV_Batch)) return batch_size;
    return 0;
}

namespace ltpmg_host {

RunOptions &options() {
    static RunOptions o = [] {
        RunOptions r;
        r.cc_mode = (FLAGS_cc_mode == "A" || FLAGS_cc_mode == "a") ? LTPMG_MODE_A : LTPMG_MODE_O;
        r.seed = FLAGS_seed;
        r.kv_size = FLAGS_kv_rows;
        r.kv_fields = FLAGS_kv_fields;
        r.kv_field_bytes = FLAGS_kv_field_bytes;
        r.kv_ops = FLAGS_kv_ops;
        r.read_percent = FLAGS_read_percent;
        r.oe_append_cap = FLAGS_append_cap;
        r.oe_preload = FLAGS_oe_preload;
        return r;
    }();
    return o;
}

bool is_ycsb(const Param &p) { return p.benchmark.rfind("YCSB_", 0) == 0; }
bool is_kv(const Param &p) { return p.benchmark == "KV" || p.benchmark == "KV_C1"; }
bool is_oe(const Param &p) { return p.benchmark == "TPCC_PART" || p.benchmark == "TPCC_ALL" || p.benchmark == "OE"; }

void export_gpu_options() {
    const RunOptions &o = options();
    setenv("LTPMG_CC_MODE", o.cc_mode == LTPMG_MODE_A ? "A" : "O", 1);
    setenv("LTPMG_SEED", std::to_string(o.seed).c_str(), 1);
    setenv("LTPMG_APPEND_CAP", std::to_string(o.oe_append_cap).c_str(), 1);
    setenv("LTPMG_OE_PRELOAD", std::to_string(o.oe_preload).c_str(), 1);
}

}  // namespace ltpmg_host
