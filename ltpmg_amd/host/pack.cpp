// pack.cpp -- the reference's query structs -> the engine's SoA batch
// (pack.h).
//
// Query decomposition follows the reference's parse<T> kernels
// (cuda/src/gpuquery.cu:2929-3303): YCSB_A reads rows 0-4 and updates 5-9;
// B and D read 0-8 and update 9; C reads all ten; E scans [r, r+2) for 0-8
// and inserts 9.  A read or scan of a row reads every column (one READ per
// cell, consecutive result slots); an update writes column (TID + j) mod
// n_cols with a payload seeded by (seed, TID, j); the YCSB_E insert is an
// INSERT into the YCSB table (an append table for that benchmark) whose row
// payloads are seeded by (seed, TID, 9); the engine hands out the row, so the
// query's ROW_ID[9] names none.  Order-entry queries go through the
// adapter's decomposition (include/ltpmg_oe.h, oe_decompose.cpp).  The
// reference's row operators are no-ops (gpulauncher.cu:8-165), so these
// value rules are this build's own.
#include "pack.h"

#include "ltpmg.h"
#include "oe.h"

namespace ltpmg_pack {

namespace {

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
                p.op(LTPMG_OP_ROW, LTPMG_ROW_INSERT, 0, p.payload(tid, j, 0));
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

}  // namespace

uint32_t ycsb_e_capacity(const Param &p, uint32_t table_size) {
    const uint64_t cap = (uint64_t)table_size + (uint64_t)p.ycsb_e_query_batch_size * (p.epoch_tp ? p.epoch_tp : 1);
    return cap > 0xFFFFFFFFull ? 0u : (uint32_t)cap;
}

bool pack_queries(const std::vector<std::any> &batch, const Global_Table_Info &info0, uint64_t seed,
                  uint32_t preload, KV_Batch &out, std::string *why) {
    out = KV_Batch{};
    if (batch.size() == 1 && batch[0].type() == typeid(KV_Batch)) {
        out = std::any_cast<const KV_Batch &>(batch[0]);
        return true;
    }
    Packer p{out, n_cols(info0), seed ^ 0x5943534Full};
    oe::Sink sink{out.rows, out.meta, out.args, seed ^ 0x4F45ull, seed, preload};
    out.txn_off.push_back(0);
    for (const std::any &a : batch) {
        if (pack_ycsb<YCSB_A_Query>(a, p, 'A') || pack_ycsb<YCSB_B_Query>(a, p, 'B') ||
            pack_ycsb<YCSB_C_Query>(a, p, 'C') || pack_ycsb<YCSB_D_Query>(a, p, 'D') ||
            pack_ycsb<YCSB_E_Query>(a, p, 'E') || pack_oe<Neworder_Query>(a, sink, out) ||
            pack_oe<Payment_Query>(a, sink, out) || pack_oe<Orderstatus_Query>(a, sink, out) ||
            pack_oe<Delivery_Query>(a, sink, out) || pack_oe<Stocklevel_Query>(a, sink, out))
            continue;
        if (why) *why = a.type().name();
        return false;
    }
    return true;
}

}  // namespace ltpmg_pack
