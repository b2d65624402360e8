// oe.h -- order-entry query decomposition shared by the generator
// (order_entry.cpp) and the host-API shim (gpu_shim.cpp).  The table and
// column layout is documented in include/ltpmg_oe.h.
#pragma once

#include <cstdint>
#include <vector>

#include "ltpmg.h"
#include "ltpmg_oe.h"
#include "rng.h"
#include "txn_structure.h"

namespace oe {

enum Table : uint32_t {
    WAREHOUSE = 0, DISTRICT = 1, CUSTOMER = 2, NEWORDER = 3, HISTORY = 4,
    ORDER = 5, ORDERLINE = 6, STOCK = 7, ITEM = 8
};

// column indices (ints first, then strings, as the schema lists them)
namespace w { enum : uint32_t { ID, TAX, YTD, NAME }; }
namespace d { enum : uint32_t { ID, W_ID, NEXT_O_ID, TAX, YTD, NAME }; }
namespace c { enum : uint32_t { ID, D_ID, W_ID, PAYMENT_CNT, DELIVERY_CNT, DISCOUNT, CREDIT_LIM,
                                BALANCE, YTD_PAYMENT, FIRST, MIDDLE, LAST, STREET_1, STREET_2, CITY,
                                STATE, ZIP, PHONE, SINCE, CREDIT, DATA }; }
namespace no { enum : uint32_t { O_ID, D_ID, W_ID }; }
namespace h { enum : uint32_t { C_ID, C_D_ID, C_W_ID, D_ID, W_ID, AMOUNT, DATE, DATA }; }
namespace o { enum : uint32_t { ID, D_ID, W_ID, C_ID, CARRIER_ID, OL_CNT, ALL_LOCAL, ENTRY_D }; }
namespace ol { enum : uint32_t { O_ID, D_ID, W_ID, NUMBER, I_ID, SUPPLY_W_ID, DELIVERY_D, QUANTITY,
                                 AMOUNT, DIST_INFO }; }
namespace s { enum : uint32_t { I_ID, W_ID, QUANTITY, YTD, ORDER_CNT, REMOTE_CNT, DIST_01 }; }
namespace i { enum : uint32_t { ID, IM_ID, PRICE, NAME, DATA }; }

constexpr uint32_t kDistricts = 10, kCustomers = 3000, kStock = 100000, kItems = 100000;
constexpr uint32_t kMaxLines = 15;
constexpr uint32_t kLastNames = LTPMG_OE_LAST_NAMES;
constexpr uint32_t kPreLines = LTPMG_OE_PRELOAD_LINES;
constexpr uint32_t kDeliveries = 10;  // entries of a Delivery_Query

inline uint32_t district_row(uint32_t w, uint32_t dd) { return w * kDistricts + dd; }
inline uint32_t customer_row(uint32_t w, uint32_t dd, uint32_t cc) { return district_row(w, dd) * kCustomers + cc; }
inline uint32_t stock_row(uint32_t w, uint32_t item) { return w * kStock + item; }
// preloaded order k of district (w, dd), and line l of a preloaded order
inline uint32_t order_row(uint32_t w, uint32_t dd, uint32_t k, uint32_t preload) {
    return district_row(w, dd) * preload + k;
}
inline uint32_t line_row(uint32_t order, uint32_t l) { return order * kPreLines + l; }

// The nine table descriptors and rows present at creation (the body of
// ltpmg_oe_schema).
int schema(uint32_t W, uint32_t cap, uint32_t preload, ltpmg_table_desc out[LTPMG_OE_TABLES],
           uint32_t n_present[LTPMG_OE_TABLES]);

// Last-name code of a customer row: the first little-endian u64 of its
// c_last cell, as ltpmg_fill_table(table_seed) writes it, mod kLastNames.
uint32_t last_name_code(uint64_t table_seed, uint32_t customer_row);
// The customer of district (w, dd) that a lookup by last-name code k finds:
// the lowest C_ID with that code, or C_ID 0 when none has it (the
// reference's index scan, cuda/src/gpuquery.cu:2641-2650).
uint32_t customer_by_last_name(uint64_t table_seed, uint32_t w, uint32_t dd, uint32_t k);

// Collects ops of one batch in SoA form.  begin(tid) marks the start of a
// transaction: within it, a second RMW_ADD of a cell is folded into the
// first one's delta (a query may name one stock row on two order lines, as
// the reference's generator allows, src/txn_generator.cpp:262-276), because
// the engine takes a written cell at most once per transaction.
struct Sink {
    std::vector<uint32_t> &rows;
    std::vector<uint16_t> &meta;
    std::vector<uint64_t> &args;
    uint64_t seed;        // payload seeds of written and inserted rows
    uint64_t table_seed;  // the tables' fill seed (last-name lookups)
    uint32_t preload;     // preloaded orders per district
    size_t txn_first = 0;
    uint32_t tid = 0;
    void begin(uint32_t t) {
        txn_first = rows.size();
        tid = t;
    }
    void put(uint32_t kind, uint32_t table, uint32_t col, uint32_t row, uint64_t arg) {
        rows.push_back(row);
        meta.push_back(LTPMG_META(kind, table, col));
        args.push_back(arg);
    }
    void read(uint32_t t, uint32_t col, uint32_t row) { put(LTPMG_OP_READ, t, col, row, 0); }
    void add(uint32_t t, uint32_t col, uint32_t row, int32_t delta) {
        const uint16_t m = LTPMG_META(LTPMG_OP_RMW_ADD, t, col);
        for (size_t k = txn_first; k < rows.size(); ++k)
            if (rows[k] == row && meta[k] == m) {
                args[k] = (uint64_t)(uint32_t)((uint32_t)args[k] + (uint32_t)delta);
                return;
            }
        put(LTPMG_OP_RMW_ADD, t, col, row, (uint64_t)(uint32_t)delta);
    }
    // a written cell gets a payload seeded by (seed, table, row, col)
    void set(uint32_t t, uint32_t col, uint32_t row) {
        put(LTPMG_OP_UPDATE, t, col, row, mixed(((uint64_t)row << 10) | ((uint64_t)t << 6) | col));
    }
    // an appended row gets payloads seeded by (seed, TID, table, line)
    void insert(uint32_t t, uint32_t line = 0) {
        put(LTPMG_OP_ROW, t, LTPMG_ROW_INSERT, 0, mixed(((uint64_t)tid << 12) | ((uint64_t)t << 8) | line | (1ull << 63)));
    }
    uint64_t mixed(uint64_t id) const {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (id + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

// The transaction's TID is the query's TID field.
void decompose(const Neworder_Query &q, Sink &out);
void decompose(const Payment_Query &q, Sink &out);
void decompose(const Orderstatus_Query &q, Sink &out);
void decompose(const Delivery_Query &q, Sink &out);
void decompose(const Stocklevel_Query &q, Sink &out);

// Query builders drawing from a transaction's own random stream.
struct Mix {
    uint32_t W, lines_min, lines_max, remote_bp;
    uint32_t preload;     // preloaded orders per district (read-only queries and deliveries pick among them)
    uint32_t byname_pct;  // Payment / Orderstatus: percent that select the customer by last name
};
Neworder_Query make_neworder(ltpmg_host::Rng &r, const Mix &m);
Payment_Query make_payment(ltpmg_host::Rng &r, const Mix &m);
Orderstatus_Query make_orderstatus(ltpmg_host::Rng &r, const Mix &m);
Delivery_Query make_delivery(ltpmg_host::Rng &r, const Mix &m);
Stocklevel_Query make_stocklevel(ltpmg_host::Rng &r, const Mix &m);

}  // namespace oe
