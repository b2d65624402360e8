// query.cpp -- the batch of the host API (the reference's src/query.cpp).
// generate_txn runs once; TIDs are positions in the batch (query.cpp:284-292
// of the reference, which shuffles with a time-seeded engine first; the
// queries here are already i.i.d. draws, so no shuffle is needed).
#include "query.h"

#include "oe.h"
#include "rng.h"

namespace {
// The batch is handed to the extern functions without a copy; the reference
// copies the whole vector on every call (src/query.cpp:503-519).
std::shared_ptr<std::vector<std::any>> alias(std::vector<std::any> &v) {
    return std::shared_ptr<std::vector<std::any>>(&v, [](std::vector<std::any> *) {});
}

template <typename Q>
void append(std::shared_ptr<Txn_Generator<Q>> &gen, std::shared_ptr<Param> param, size_t n,
            std::vector<std::any> &batch) {
    gen = std::make_shared<Txn_Generator<Q>>();
    gen->initialize(param);
    gen->generate_txn(param, n);
    batch.reserve(batch.size() + gen->get_txn().size());
    for (auto &q : gen->get_txn()) batch.emplace_back(q);
    gen.reset();
}
}  // namespace

Query::Query() {}

Query::~Query() {}

void Query::generate_txn(std::shared_ptr<Param> param) {
    transactions_batch.clear();
    const std::string &b = param->benchmark;
    if (b == "YCSB_A") append(txn_generator_ycsb_a_query, param, param->ycsb_a_query_batch_size, transactions_batch);
    else if (b == "YCSB_B") append(txn_generator_ycsb_b_query, param, param->ycsb_b_query_batch_size, transactions_batch);
    else if (b == "YCSB_C") append(txn_generator_ycsb_c_query, param, param->ycsb_c_query_batch_size, transactions_batch);
    else if (b == "YCSB_D") append(txn_generator_ycsb_d_query, param, param->ycsb_d_query_batch_size, transactions_batch);
    else if (b == "YCSB_E") append(txn_generator_ycsb_e_query, param, param->ycsb_e_query_batch_size, transactions_batch);
    else if (ltpmg_host::is_kv(*param) || b == "OE") {
        Txn_Generator<KV_Batch> gen;
        gen.initialize(param);
        gen.generate_txn(param, param->batch_size);
        transactions_batch.emplace_back(std::move(gen.get_txn().front()));
    } else if (b == "TPCC_PART" || b == "TPCC_ALL") {
        append(txn_generator_neworder_query, param, param->neworder_query_batch_size, transactions_batch);
        append(txn_generator_payment_query, param, param->payment_query_batch_size, transactions_batch);
        if (b == "TPCC_ALL") {  // src/param.cpp:53-57 of the reference: 44/44/4/4/4
            append(txn_generator_orderstatus_query, param, param->orderstatus_query_batch_size, transactions_batch);
            append(txn_generator_delivery_query, param, param->delivery_query_batch_size, transactions_batch);
            append(txn_generator_stocklevel_query, param, param->stocklevel_query_batch_size, transactions_batch);
        }
        // interleave the kinds with a seeded Fisher-Yates shuffle (the
        // reference sorts by type-name pointer, src/query.cpp:115-116), then
        // TID = position
        ltpmg_host::Rng r(ltpmg_host::options().seed, 0x5348554646ull);
        for (size_t i = transactions_batch.size(); i > 1; --i)
            std::swap(transactions_batch[i - 1], transactions_batch[r.below(i)]);
        for (size_t t = 0; t < transactions_batch.size(); ++t) {
            std::any &a = transactions_batch[t];
            if (auto *q = std::any_cast<Neworder_Query>(&a)) q->TID = (UINT32)t;
            else if (auto *p = std::any_cast<Payment_Query>(&a)) p->TID = (UINT32)t;
            else if (auto *o = std::any_cast<Orderstatus_Query>(&a)) o->TID = (UINT32)t;
            else if (auto *d = std::any_cast<Delivery_Query>(&a)) d->TID = (UINT32)t;
            else if (auto *s = std::any_cast<Stocklevel_Query>(&a)) s->TID = (UINT32)t;
        }
    } else {
        LOG(FATAL) << "benchmark " << b << ": the TEST queries are not built in this engine";
    }
}

void Query::malloc_global_txn(std::shared_ptr<Param> param) { global_txn_info_size = param->get_subtxn_kinds(); }

void Query::free_global_txn() {}

void Query::initial_on_GPU(std::shared_ptr<Param> param) {
    initial_query_on_GPU(param, alias(transactions_batch), global_txn_info);
}

void Query::transfer_to_GPU(std::shared_ptr<Param> param) {
    transfer_query_to_GPU(param, alias(transactions_batch), global_txn_info);
}

void Query::transfer_to_CPU(std::shared_ptr<Param> param) {
    transfer_query_to_CPU(param, alias(transactions_batch), global_txn_info);
}
