// txn_generator.cpp -- seeded YCSB and KV generators of the host API.
//
// YCSB_A..E: ten distinct row ids per query, drawn uniformly or from
// Zipf(zipf_config) over the YCSB table (src/txn_generator.cpp:602-993 draws
// the same shape with rand()).  Query t uses its own Random stream, started
// from (seed, type, t), so the batch does not depend on the thread count.
// KV_Batch: the whole batch in SoA form from ltpmg_gen_kv (ltpmg_workload.h)
// or, for the OE benchmark, ltpmg_gen_oe (ltpmg_oe.h).  Neworder, Payment,
// Orderstatus, Delivery and Stocklevel queries come from the order-entry
// builders (oe.h).  The TEST generators are declared but not built.
#include "txn_generator.h"

#include <algorithm>

#include "ltpmg.h"
#include "ltpmg_oe.h"
#include "ltpmg_workload.h"
#include "oe.h"

namespace {

template <typename Q>
void gen_ycsb(std::shared_ptr<Param> param, size_t size, std::vector<Q> &out, uint64_t type_tag) {
    out.assign(size, Q{});
    const bool zipf = param->data_distribution == "zipf";
    const uint64_t rows = param->ycsb_size;
    if (rows < 10) LOG(FATAL) << "YCSB needs at least 10 rows";
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t n_threads = std::min<size_t>(hw, std::max<size_t>(1, size / 4096));
    std::vector<std::thread> th;
    for (size_t w = 0; w < n_threads; ++w)
        th.emplace_back([&, w] {
            Random r;
            if (zipf) r.init_rand_zipf(rows, param->zipf_config);
            for (size_t i = w; i < size; i += n_threads) {
                r.init_seed(ltpmg_host::options().seed ^ (type_tag << 56) ^ (uint64_t)i);
                Q &q = out[i];
                q.TID = (UINT32)i;
                for (int j = 0; j < 10; ++j) {
                    UINT32 k;
                    bool dup;
                    do {
                        k = zipf ? r.rand_zipf(0, rows - 1) : (UINT32)r.uniform_dist(0, rows - 1);
                        dup = false;
                        for (int p = 0; p < j; ++p) dup |= q.ROW_ID[p] == k;
                        // YCSB_E scans [k, k+2): keep both rows inside the table
                        if (type_tag == 5 && k + 1 >= rows) dup = true;
                    } while (dup);
                    q.ROW_ID[j] = k;
                }
            }
        });
    for (auto &t : th) t.join();
}

}  // namespace

template <typename T>
void Txn_Generator<T>::initialize(std::shared_ptr<Param>) {
    random.init_seed(ltpmg_host::options().seed);
    transactions.clear();
}

template void Txn_Generator<Test_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<Test_Query_2>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<Neworder_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<Payment_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<Orderstatus_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<Delivery_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<Stocklevel_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<YCSB_A_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<YCSB_B_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<YCSB_C_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<YCSB_D_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<YCSB_E_Query>::initialize(std::shared_ptr<Param>);
template void Txn_Generator<KV_Batch>::initialize(std::shared_ptr<Param>);

template <>
void Txn_Generator<YCSB_A_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_ycsb(param, size, transactions, 1);
    LOG(INFO) << "YCSB_A_Query transactions.size()=" << transactions.size();
}
template <>
void Txn_Generator<YCSB_B_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_ycsb(param, size, transactions, 2);
    LOG(INFO) << "YCSB_B_Query transactions.size()=" << transactions.size();
}
template <>
void Txn_Generator<YCSB_C_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_ycsb(param, size, transactions, 3);
    LOG(INFO) << "YCSB_C_Query transactions.size()=" << transactions.size();
}
template <>
void Txn_Generator<YCSB_D_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_ycsb(param, size, transactions, 4);
    LOG(INFO) << "YCSB_D_Query transactions.size()=" << transactions.size();
}
template <>
void Txn_Generator<YCSB_E_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_ycsb(param, size, transactions, 5);
    LOG(INFO) << "YCSB_E_Query transactions.size()=" << transactions.size();
}

// Order-entry queries (the reference's src/txn_generator.cpp:231-599 draws
// them with rand(); here each query has its own seeded stream).  TIDs are
// set by Query::generate_txn once the kinds are interleaved; the rows that
// INSERTs append are handed out by the engine.
namespace {
template <typename Q, typename Make>
void gen_oe_queries(std::shared_ptr<Param> param, size_t size, std::vector<Q> &out, uint64_t tag,
                    Make make) {
    // 15 % of payments by last name, as src/txn_generator.cpp:346-351 draws them
    const oe::Mix mix{param->warehouse_size, 5, 15, 100, std::max(1u, ltpmg_host::options().oe_preload), 15};
    out.resize(size);
    for (size_t i = 0; i < size; ++i) {
        ltpmg_host::Rng r(ltpmg_host::options().seed ^ (tag << 48), i);
        out[i] = make(r, mix);
        out[i].TID = (UINT32)i;
    }
}
}  // namespace

template <>
void Txn_Generator<Neworder_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_oe_queries(param, size, transactions, 11, oe::make_neworder);
    LOG(INFO) << "Neworder_Query transactions.size()=" << transactions.size();
}
template <>
void Txn_Generator<Payment_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_oe_queries(param, size, transactions, 12, oe::make_payment);
    LOG(INFO) << "Payment_Query transactions.size()=" << transactions.size();
}
template <>
void Txn_Generator<Orderstatus_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_oe_queries(param, size, transactions, 13, oe::make_orderstatus);
    LOG(INFO) << "Orderstatus_Query transactions.size()=" << transactions.size();
}
template <>
void Txn_Generator<Delivery_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_oe_queries(param, size, transactions, 15, oe::make_delivery);
    LOG(INFO) << "Delivery_Query transactions.size()=" << transactions.size();
}
template <>
void Txn_Generator<Stocklevel_Query>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    gen_oe_queries(param, size, transactions, 14, oe::make_stocklevel);
    LOG(INFO) << "Stocklevel_Query transactions.size()=" << transactions.size();
}

template <>
void Txn_Generator<KV_Batch>::generate_txn(std::shared_ptr<Param> param, size_t size) {
    if (param->benchmark == "OE") {  // the order-entry mix, straight to SoA
        ltpmg_oe_spec sp{};
        sp.n_txns = (uint32_t)size;
        sp.warehouses = param->warehouse_size;
        sp.neworder_pct = 45;
        sp.payment_pct = 43;
        sp.lines_min = 5;
        sp.lines_max = 15;
        sp.remote_bp = 100;
        sp.preload = ltpmg_host::options().oe_preload;
        sp.seed = ltpmg_host::options().seed ^ 0x4F45ull;
        sp.byname_pct = 15;  // src/txn_generator.cpp:346-351
        sp.table_seed = ltpmg_host::options().seed;  // the tables are filled from the run's seed
        KV_Batch b;
        uint64_t cap = ltpmg_gen_oe_max_ops(&sp);
        b.txn_off.resize(size + 1);
        b.rows.resize(cap);
        b.meta.resize(cap);
        b.args.resize(cap);
        uint32_t n = 0;
        if (ltpmg_gen_oe(&sp, b.txn_off.data(), b.rows.data(), b.meta.data(), b.args.data(), cap, &n, 0) != 0)
            LOG(FATAL) << "ltpmg_gen_oe rejected the order-entry spec";
        b.rows.resize(n);
        b.meta.resize(n);
        b.args.resize(n);
        transactions.clear();
        transactions.push_back(std::move(b));
        LOG(INFO) << "order-entry batch: " << size << " txns, " << n << " ops";
        return;
    }
    const bool c1 = param->benchmark == "KV_C1";
    ltpmg_kv_spec sp{};
    sp.n_txns = (uint32_t)size;
    sp.ops_min = sp.ops_max = ltpmg_host::options().kv_ops;
    sp.n_rows = ltpmg_host::options().kv_size;
    sp.zipf_theta = param->data_distribution == "zipf" ? std::min(param->zipf_config, 0.99) : 0.0;
    sp.n_cols = ltpmg_host::options().kv_fields;
    sp.read_pct = ltpmg_host::options().read_percent;
    sp.write_kind = c1 ? LTPMG_OP_RMW_ADD : LTPMG_OP_UPDATE;
    sp.read_all_cols = c1 ? 1 : 0;
    sp.table_id = 0;
    sp.rmw_delta = 1;
    sp.seed = ltpmg_host::options().seed ^ 0x4B56ull;
    KV_Batch b;
    uint64_t cap = ltpmg_gen_kv_max_ops(&sp);
    b.txn_off.resize(size + 1);
    b.rows.resize(cap);
    b.meta.resize(cap);
    b.args.resize(cap);
    uint32_t n = 0;
    int rc = ltpmg_gen_kv(&sp, b.txn_off.data(), b.rows.data(), b.meta.data(), b.args.data(), cap, &n, 0);
    if (rc != 0) LOG(FATAL) << "ltpmg_gen_kv rejected the KV spec (" << rc << ")";
    b.rows.resize(n);
    b.meta.resize(n);
    b.args.resize(n);
    transactions.clear();
    transactions.push_back(std::move(b));
    LOG(INFO) << "KV batch: " << size << " txns, " << n << " ops";
}
