// oe_decompose.cpp -- the order-entry schema and query -> op decomposition
// (include/ltpmg_oe.h), shared by the generator in libltpmg_host.so and the
// reference-interface entry points in libltpmg_gpu.so.
#include <cstring>

#include "oe.h"
#include "rng.h"

namespace oe {

uint32_t last_name_code(uint64_t table_seed, uint32_t row) {
    const uint64_t id = ((uint64_t)row << 10) | ((uint64_t)CUSTOMER << 6) | c::LAST;
    const uint64_t cell = ltpmg_host::mix64(table_seed + ltpmg_host::kGolden * (id + 1));
    return (uint32_t)(ltpmg_host::mix64(cell + ltpmg_host::kGolden) % kLastNames);
}

uint32_t customer_by_last_name(uint64_t table_seed, uint32_t w, uint32_t dd, uint32_t k) {
    for (uint32_t cc = 0; cc < kCustomers; ++cc)
        if (last_name_code(table_seed, customer_row(w, dd, cc)) == k) return cc;
    return 0;
}

void decompose(const Neworder_Query &q, Sink &out) {
    out.begin();
    const uint32_t dr = district_row(q.W_ID, q.D_ID), cr = customer_row(q.W_ID, q.D_ID, q.C_ID);
    out.read(WAREHOUSE, w::TAX, q.W_ID);
    out.read(DISTRICT, d::TAX, dr);
    out.read(DISTRICT, d::NEXT_O_ID, dr);      // the order id this txn takes ...
    out.add(DISTRICT, d::NEXT_O_ID, dr, 1);    // ... and the increment after it
    out.read(CUSTOMER, c::DISCOUNT, cr);
    out.read(CUSTOMER, c::LAST, cr);
    out.read(CUSTOMER, c::CREDIT, cr);
    for (uint32_t col : {o::C_ID, o::OL_CNT, o::ALL_LOCAL, o::ENTRY_D}) out.insert(ORDER, col, q.O_ID);
    for (uint32_t col : {no::O_ID, no::D_ID, no::W_ID}) out.insert(NEWORDER, col, q.N_O_ID);
    for (uint32_t l = 0; l < q.O_OL_CNT && l < kMaxLines; ++l) {
        const auto &ln = q.INFO[l];
        const uint32_t sr = stock_row(ln.OL_SUPPLY_W_ID, ln.OL_I_ID);
        out.read(ITEM, i::PRICE, ln.OL_I_ID);
        out.read(ITEM, i::NAME, ln.OL_I_ID);
        out.add(STOCK, s::QUANTITY, sr, -(int32_t)ln.OL_QUANTITY);
        out.add(STOCK, s::YTD, sr, (int32_t)ln.OL_QUANTITY);
        out.add(STOCK, s::ORDER_CNT, sr, 1);
        if (ln.OL_SUPPLY_W_ID != q.W_ID) out.add(STOCK, s::REMOTE_CNT, sr, 1);
        for (uint32_t col : {ol::I_ID, ol::SUPPLY_W_ID, ol::QUANTITY, ol::AMOUNT, ol::DIST_INFO})
            out.insert(ORDERLINE, col, q.O_OL_ID + l);
    }
}

void decompose(const Payment_Query &q, Sink &out) {
    out.begin();
    const uint32_t cid = q.isName ? customer_by_last_name(out.table_seed, q.C_W_ID, q.C_D_ID, q.C_LAST) : q.C_ID;
    const uint32_t dr = district_row(q.W_ID, q.D_ID), cr = customer_row(q.C_W_ID, q.C_D_ID, cid);
    out.add(WAREHOUSE, w::YTD, q.W_ID, (int32_t)q.H_AMOUNT);
    out.read(WAREHOUSE, w::NAME, q.W_ID);
    out.add(DISTRICT, d::YTD, dr, (int32_t)q.H_AMOUNT);
    out.read(DISTRICT, d::NAME, dr);
    out.add(CUSTOMER, c::BALANCE, cr, -(int32_t)q.H_AMOUNT);
    out.add(CUSTOMER, c::YTD_PAYMENT, cr, (int32_t)q.H_AMOUNT);
    out.add(CUSTOMER, c::PAYMENT_CNT, cr, 1);
    out.read(CUSTOMER, c::FIRST, cr);
    out.read(CUSTOMER, c::LAST, cr);
    out.read(CUSTOMER, c::DATA, cr);
    for (uint32_t col : {h::C_ID, h::D_ID, h::W_ID, h::AMOUNT, h::DATA}) out.insert(HISTORY, col, q.H_ID);
}

void decompose(const Orderstatus_Query &q, Sink &out) {
    out.begin();
    const uint32_t cid = q.isName ? customer_by_last_name(out.table_seed, q.W_ID, q.D_ID, q.C_LAST) : q.C_ID;
    const uint32_t cr = customer_row(q.W_ID, q.D_ID, cid);
    out.read(CUSTOMER, c::BALANCE, cr);
    out.read(CUSTOMER, c::FIRST, cr);
    out.read(CUSTOMER, c::LAST, cr);
    out.read(ORDER, o::ID, q.O_ID);
    out.read(ORDER, o::CARRIER_ID, q.O_ID);
    out.read(ORDER, o::ENTRY_D, q.O_ID);
    for (uint32_t l = 0; l < 5; ++l) {
        out.read(ORDERLINE, ol::I_ID, q.OL_ID + l);
        out.read(ORDERLINE, ol::QUANTITY, q.OL_ID + l);
        out.read(ORDERLINE, ol::AMOUNT, q.OL_ID + l);
    }
}

void decompose(const Stocklevel_Query &q, Sink &out) {
    out.begin();
    out.read(DISTRICT, d::NEXT_O_ID, district_row(q.W_ID, q.D_ID));
    for (uint32_t k = 0; k < 10; ++k) {
        out.read(ORDERLINE, ol::I_ID, q.O_OL_ID[k]);
        out.read(STOCK, s::QUANTITY, stock_row(q.W_ID, q.I_ID[k]));
    }
}

static void table_desc(ltpmg_table_desc &t, uint32_t rows, std::initializer_list<uint32_t> widths) {
    std::memset(&t, 0, sizeof t);
    t.n_rows = rows;
    for (uint32_t wd : widths) t.col_bytes[t.n_cols++] = wd;
}

int schema(uint32_t W, uint32_t cap, ltpmg_table_desc out[LTPMG_OE_TABLES]) {
    if (!out || W == 0 || cap == 0 || (uint64_t)W * kStock > 0xFFFFFFFFull ||
        (uint64_t)cap * kMaxLines > 0xFFFFFFFFull)
        return LTPMG_EINVAL;
    const uint32_t S = 32;
    table_desc(out[WAREHOUSE], W, {4, 4, 4, S, S, S, S, S, S});
    table_desc(out[DISTRICT], W * kDistricts, {4, 4, 4, 4, 4, S, S, S, S, S, S});
    table_desc(out[CUSTOMER], W * kDistricts * kCustomers,
               {4, 4, 4, 4, 4, 4, 4, 4, 4, S, S, S, S, S, S, S, S, S, S, S, 256});
    table_desc(out[NEWORDER], cap, {4, 4, 4});
    table_desc(out[HISTORY], cap, {4, 4, 4, 4, 4, 4, S, S});
    table_desc(out[ORDER], cap, {4, 4, 4, 4, 4, 4, 4, S});
    table_desc(out[ORDERLINE], cap * kMaxLines, {4, 4, 4, 4, 4, 4, 4, 4, 4, 24});
    table_desc(out[STOCK], W * kStock, {4, 4, 4, 4, 4, 4, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, S});
    table_desc(out[ITEM], kItems, {4, 4, 4, 24, 52});
    return LTPMG_OK;
}


}  // namespace oe
