// order_entry.cpp -- the order-entry adapter: schema, query generation and
// query -> op decomposition (include/ltpmg_oe.h).
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "oe.h"
#include "rng.h"

using ltpmg_host::Rng;

namespace oe {

Neworder_Query make_neworder(Rng &r, const Mix &m) {
    Neworder_Query q{};
    q.W_ID = (uint32_t)r.below(m.W);
    q.D_ID = (uint32_t)r.below(kDistricts);
    q.C_ID = (uint32_t)r.nurand(1023, 0, kCustomers - 1);
    q.O_OL_CNT = (uint32_t)r.range(m.lines_min, m.lines_max);
    for (uint32_t l = 0; l < q.O_OL_CNT; ++l) {
        uint32_t item, sw;
        bool dup;
        do {  // a stock cell is written once per txn
            item = (uint32_t)r.nurand(8191, 0, kItems - 1);
            sw = (m.W > 1 && r.below(10000) < m.remote_bp) ? (uint32_t)((q.W_ID + 1 + r.below(m.W - 1)) % m.W)
                                                         : q.W_ID;
            dup = false;
            for (uint32_t p = 0; p < l; ++p) dup |= q.INFO[p].OL_I_ID == item && q.INFO[p].OL_SUPPLY_W_ID == sw;
        } while (dup);
        q.INFO[l].OL_I_ID = item;
        q.INFO[l].OL_SUPPLY_W_ID = sw;
        q.INFO[l].OL_QUANTITY = (uint32_t)r.range(1, 10);
    }
    return q;
}

Payment_Query make_payment(Rng &r, const Mix &m) {
    Payment_Query q{};
    q.W_ID = (uint32_t)r.below(m.W);
    q.D_ID = (uint32_t)r.below(kDistricts);
    q.C_ID = (uint32_t)r.nurand(1023, 0, kCustomers - 1);
    q.C_W_ID = q.W_ID;
    q.C_D_ID = q.D_ID;
    q.H_AMOUNT = (uint32_t)r.range(100, 500000);
    if (m.byname_pct) {  // drawn only when asked, so by-id streams stay as they were
        q.isName = r.below(100) < m.byname_pct ? 1u : 0u;
        q.C_LAST = (uint32_t)r.below(kLastNames);
    }
    return q;
}

Orderstatus_Query make_orderstatus(Rng &r, const Mix &m) {
    Orderstatus_Query q{};
    q.W_ID = (uint32_t)r.below(m.W);
    q.D_ID = (uint32_t)r.below(kDistricts);
    q.C_ID = (uint32_t)r.nurand(1023, 0, kCustomers - 1);
    q.O_ID = (uint32_t)r.below(m.preload);  // a preloaded order of the district
    if (m.byname_pct) {
        q.isName = r.below(100) < m.byname_pct ? 1u : 0u;
        q.C_LAST = (uint32_t)r.below(kLastNames);
    }
    return q;
}

// Ten entries as the reference draws them (src/txn_generator.cpp:466-482):
// distinct order ids, each entry's partition, district and customer
// uniform; the order ids are drawn among the preloaded orders.
Delivery_Query make_delivery(Rng &r, const Mix &m) {
    Delivery_Query q{};
    for (uint32_t j = 0; j < kDeliveries; ++j) {
        bool dup;
        do {
            q.NO_O_ID[j] = (uint32_t)r.below(m.preload);
            dup = false;
            for (uint32_t k = 0; k < j; ++k) dup |= q.NO_O_ID[k] == q.NO_O_ID[j];
        } while (dup && j < m.preload);
    }
    for (uint32_t j = 0; j < kDeliveries; ++j) {
        q.NO_W_ID[j] = (uint32_t)r.below(m.W);
        q.NO_D_ID[j] = (uint32_t)r.below(kDistricts);
        q.NO_C_ID[j] = (uint32_t)r.below(kCustomers);
    }
    return q;
}

Stocklevel_Query make_stocklevel(Rng &r, const Mix &m) {
    Stocklevel_Query q{};
    q.W_ID = (uint32_t)r.below(m.W);
    q.D_ID = (uint32_t)r.below(kDistricts);
    for (uint32_t k = 0; k < 10; ++k) {
        q.I_ID[k] = (uint32_t)r.nurand(8191, 0, kItems - 1);
        q.O_OL_ID[k] = (uint32_t)r.below((uint64_t)m.preload * kPreLines);  // a line of the district
    }
    return q;
}

}  // namespace oe

using namespace oe;

extern "C" int ltpmg_oe_schema(uint32_t W, uint32_t cap, uint32_t preload, ltpmg_table_desc out[LTPMG_OE_TABLES],
                               uint32_t n_present[LTPMG_OE_TABLES]) {
    return oe::schema(W, cap, preload, out, n_present);
}

extern "C" int ltpmg_oe_last_name_index(uint32_t W, uint64_t table_seed, uint32_t *first) {
    if (!first || W == 0 || (uint64_t)W * kDistricts * kCustomers > 0xFFFFFFFFull) return LTPMG_EINVAL;
    for (uint32_t w = 0; w < W; ++w)
        for (uint32_t dd = 0; dd < kDistricts; ++dd)
            for (uint32_t k = 0; k < kLastNames; ++k)
                first[(w * kDistricts + dd) * kLastNames + k] = customer_by_last_name(table_seed, w, dd, k);
    return LTPMG_OK;
}

extern "C" uint64_t ltpmg_gen_oe_max_ops(const ltpmg_oe_spec *sp) {
    // a Neworder of 15 remote lines: 9 + 15 * 7 ops; a Delivery: 10 * 13
    const uint64_t per = sp && sp->delivery_pct ? kDeliveries * (3 + 2 * kPreLines) : 9 + kMaxLines * 7;
    return sp ? (uint64_t)sp->n_txns * per : 0;
}

extern "C" int ltpmg_gen_oe(ltpmg_oe_spec *sp, uint32_t *txn_off, uint32_t *rows, uint16_t *meta,
                            uint64_t *args, uint64_t max_ops, uint32_t *n_ops_out, uint32_t n_threads) {
    if (!sp || !txn_off || !n_ops_out || sp->warehouses == 0 ||
        sp->neworder_pct + sp->payment_pct + sp->delivery_pct > 100 || sp->lines_min < 1 ||
        sp->lines_max > kMaxLines || sp->lines_min > sp->lines_max || sp->remote_bp > 10000 ||
        sp->byname_pct > 100 || sp->preload > LTPMG_OE_MAX_PRELOAD ||
        (sp->preload == 0 && sp->neworder_pct + sp->payment_pct < 100))  // read-only and deliveries need orders
        return LTPMG_EINVAL;
    if (n_threads == 0) n_threads = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t n_chunks = std::min<uint32_t>(n_threads, std::max<uint32_t>(1, sp->n_txns));
    struct Chunk {
        std::vector<uint32_t> len, rows;
        std::vector<uint16_t> meta;
        std::vector<uint64_t> args;
    };
    std::vector<Chunk> chunks(n_chunks);
    const ltpmg_oe_spec spec = *sp;
    auto gen = [&](uint32_t ci) {
        Chunk &ch = chunks[ci];
        Sink sink{ch.rows, ch.meta, ch.args, spec.seed ^ 0x4F45ull, spec.table_seed, spec.preload};
        const uint64_t t0 = (uint64_t)spec.n_txns * ci / n_chunks, t1 = (uint64_t)spec.n_txns * (ci + 1) / n_chunks;
        const Mix mix{spec.warehouses, spec.lines_min, spec.lines_max, spec.remote_bp, spec.preload, spec.byname_pct};
        const uint32_t no = spec.neworder_pct, pay = no + spec.payment_pct, del = pay + spec.delivery_pct;
        for (uint64_t t = t0; t < t1; ++t) {
            Rng r(spec.seed, t);
            const size_t before = ch.rows.size();
            const uint32_t kind = (uint32_t)r.below(100);
            if (kind < no) {
                Neworder_Query q = make_neworder(r, mix);
                q.TID = (UINT32)t;
                decompose(q, sink);
            } else if (kind < pay) {
                Payment_Query q = make_payment(r, mix);
                q.TID = (UINT32)t;
                decompose(q, sink);
            } else if (kind < del) {
                Delivery_Query q = make_delivery(r, mix);
                q.TID = (UINT32)t;
                decompose(q, sink);
            } else if (r.below(2) == 0) {
                Orderstatus_Query q = make_orderstatus(r, mix);
                q.TID = (UINT32)t;
                decompose(q, sink);
            } else {
                Stocklevel_Query q = make_stocklevel(r, mix);
                q.TID = (UINT32)t;
                decompose(q, sink);
            }
            ch.len.push_back((uint32_t)(ch.rows.size() - before));
        }
    };
    std::vector<std::thread> th;
    for (uint32_t c = 0; c < n_chunks; ++c) th.emplace_back(gen, c);
    for (auto &t : th) t.join();

    uint64_t total = 0;
    for (auto &ch : chunks) total += ch.rows.size();
    if (total > max_ops || total >= (1ull << 29)) return LTPMG_EINVAL;
    uint64_t op = 0, t = 0;
    txn_off[0] = 0;
    for (auto &ch : chunks) {
        if (!ch.rows.empty()) {
            if (!rows || !meta || !args) return LTPMG_EINVAL;
            std::memcpy(rows + op, ch.rows.data(), ch.rows.size() * 4);
            std::memcpy(meta + op, ch.meta.data(), ch.meta.size() * 2);
            std::memcpy(args + op, ch.args.data(), ch.args.size() * 8);
        }
        for (uint32_t l : ch.len) {
            op += l;
            txn_off[++t] = (uint32_t)op;
        }
    }
    *n_ops_out = (uint32_t)total;
    return LTPMG_OK;
}
