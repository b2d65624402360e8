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
    out.begin(q.TID);
    const uint32_t dr = district_row(q.W_ID, q.D_ID), cr = customer_row(q.W_ID, q.D_ID, q.C_ID);
    out.read(WAREHOUSE, w::TAX, q.W_ID);
    out.read(DISTRICT, d::TAX, dr);
    out.read(DISTRICT, d::NEXT_O_ID, dr);      // the order id this txn takes ...
    out.add(DISTRICT, d::NEXT_O_ID, dr, 1);    // ... and the increment after it
    out.read(CUSTOMER, c::DISCOUNT, cr);
    out.read(CUSTOMER, c::LAST, cr);
    out.read(CUSTOMER, c::CREDIT, cr);
    out.insert(ORDER);
    out.insert(NEWORDER);
    for (uint32_t l = 0; l < q.O_OL_CNT && l < kMaxLines; ++l) {
        const auto &ln = q.INFO[l];
        const uint32_t sr = stock_row(ln.OL_SUPPLY_W_ID, ln.OL_I_ID);
        out.read(ITEM, i::PRICE, ln.OL_I_ID);
        out.read(ITEM, i::NAME, ln.OL_I_ID);
        out.add(STOCK, s::QUANTITY, sr, -(int32_t)ln.OL_QUANTITY);
        out.add(STOCK, s::YTD, sr, (int32_t)ln.OL_QUANTITY);
        out.add(STOCK, s::ORDER_CNT, sr, 1);
        if (ln.OL_SUPPLY_W_ID != q.W_ID) out.add(STOCK, s::REMOTE_CNT, sr, 1);
        out.insert(ORDERLINE, l);
    }
}

void decompose(const Payment_Query &q, Sink &out) {
    out.begin(q.TID);
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
    out.insert(HISTORY);
}

void decompose(const Orderstatus_Query &q, Sink &out) {
    out.begin(q.TID);
    const uint32_t cid = q.isName ? customer_by_last_name(out.table_seed, q.W_ID, q.D_ID, q.C_LAST) : q.C_ID;
    const uint32_t cr = customer_row(q.W_ID, q.D_ID, cid);
    out.read(CUSTOMER, c::BALANCE, cr);
    out.read(CUSTOMER, c::FIRST, cr);
    out.read(CUSTOMER, c::LAST, cr);
    if (!out.preload) return;
    const uint32_t orow = order_row(q.W_ID, q.D_ID, q.O_ID % out.preload, out.preload);
    out.read(ORDER, o::ID, orow);
    out.read(ORDER, o::CARRIER_ID, orow);
    out.read(ORDER, o::ENTRY_D, orow);
    for (uint32_t l = 0; l < kPreLines; ++l) {
        out.read(ORDERLINE, ol::I_ID, line_row(orow, l));
        out.read(ORDERLINE, ol::QUANTITY, line_row(orow, l));
        out.read(ORDERLINE, ol::AMOUNT, line_row(orow, l));
    }
}

void decompose(const Delivery_Query &q, Sink &out) {
    out.begin(q.TID);
    if (!out.preload) return;
    uint32_t taken[kDeliveries];
    uint32_t n_taken = 0;
    for (uint32_t j = 0; j < kDeliveries; ++j) {
        const uint32_t orow = order_row(q.NO_W_ID[j], q.NO_D_ID[j], q.NO_O_ID[j] % out.preload, out.preload);
        bool dup = false;
        for (uint32_t k = 0; k < n_taken; ++k) dup |= taken[k] == orow;
        if (dup) continue;  // one DELETE of a row per txn
        taken[n_taken++] = orow;
        out.put(LTPMG_OP_ROW, NEWORDER, LTPMG_ROW_DELETE, orow, 0);
        out.set(ORDER, o::CARRIER_ID, orow);
        for (uint32_t l = 0; l < kPreLines; ++l) {
            out.read(ORDERLINE, ol::AMOUNT, line_row(orow, l));
            out.set(ORDERLINE, ol::DELIVERY_D, line_row(orow, l));
        }
        out.add(CUSTOMER, c::DELIVERY_CNT, customer_row(q.NO_W_ID[j], q.NO_D_ID[j], q.NO_C_ID[j] % kCustomers), 1);
    }
}

void decompose(const Stocklevel_Query &q, Sink &out) {
    out.begin(q.TID);
    out.read(DISTRICT, d::NEXT_O_ID, district_row(q.W_ID, q.D_ID));
    const uint32_t lines = out.preload * kPreLines;
    for (uint32_t k = 0; k < 10; ++k) {
        if (lines) out.read(ORDERLINE, ol::I_ID, district_row(q.W_ID, q.D_ID) * lines + q.O_OL_ID[k] % lines);
        out.read(STOCK, s::QUANTITY, stock_row(q.W_ID, q.I_ID[k]));
    }
}

static void table_desc(ltpmg_table_desc &t, uint32_t rows, std::initializer_list<uint32_t> widths) {
    std::memset(&t, 0, sizeof t);
    t.n_rows = rows;
    for (uint32_t wd : widths) t.col_bytes[t.n_cols++] = wd;
}

int schema(uint32_t W, uint32_t cap, uint32_t preload, ltpmg_table_desc out[LTPMG_OE_TABLES],
           uint32_t n_present[LTPMG_OE_TABLES]) {
    const uint64_t orders = (uint64_t)W * kDistricts * preload;
    if (!out || W == 0 || preload > LTPMG_OE_MAX_PRELOAD || (uint64_t)W * kStock > 0xFFFFFFFFull ||
        orders * kPreLines + (uint64_t)cap * kMaxLines > 0xFFFFFFFFull || orders + cap == 0 || cap == 0)
        return LTPMG_EINVAL;
    const uint32_t S = 32;
    table_desc(out[WAREHOUSE], W, {4, 4, 4, S, S, S, S, S, S});
    table_desc(out[DISTRICT], W * kDistricts, {4, 4, 4, 4, 4, S, S, S, S, S, S});
    table_desc(out[CUSTOMER], W * kDistricts * kCustomers,
               {4, 4, 4, 4, 4, 4, 4, 4, 4, S, S, S, S, S, S, S, S, S, S, S, 256});
    table_desc(out[NEWORDER], (uint32_t)(orders + cap), {4, 4, 4});
    table_desc(out[HISTORY], cap, {4, 4, 4, 4, 4, 4, S, S});
    table_desc(out[ORDER], (uint32_t)(orders + cap), {4, 4, 4, 4, 4, 4, 4, S});
    table_desc(out[ORDERLINE], (uint32_t)(orders * kPreLines + (uint64_t)cap * kMaxLines),
               {4, 4, 4, 4, 4, 4, 4, 4, 4, 24});
    table_desc(out[STOCK], W * kStock, {4, 4, 4, 4, 4, 4, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, S});
    table_desc(out[ITEM], kItems, {4, 4, 4, 24, 52});
    if (n_present) {
        for (uint32_t t = 0; t < LTPMG_OE_TABLES; ++t) n_present[t] = out[t].n_rows;
        n_present[NEWORDER] = n_present[ORDER] = (uint32_t)orders;
        n_present[HISTORY] = 0;
        n_present[ORDERLINE] = (uint32_t)(orders * kPreLines);
    }
    return LTPMG_OK;
}

}  // namespace oe
