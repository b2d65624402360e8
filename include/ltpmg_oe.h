/*
 * ltpmg_oe.h -- the order-entry workload adapter (libltpmg_host.so, C ABI).
 *
 * The reference's TPCC_* benchmarks: nine tables (data/TPCC.json; table ids
 * as src/param.cpp:25-100 orders them) and query structs decomposed into
 * single-row ops by parse<Neworder_Query> / parse<Payment_Query> / ...
 * (cuda/src/gpuquery.cu:2473-2927).  This adapter keeps the table ids, the
 * partition keys and the query structs, and turns each query into engine ops
 * (include/ltpmg.h) on these columns:
 *
 *   0 warehouse  W rows          ints w_id w_tax w_ytd; 6 strings of 32 B
 *   1 district   10 W            ints d_id d_w_id d_next_o_id d_tax d_ytd; 6 strings
 *   2 customer   30000 W         ints c_id c_d_id c_w_id c_payment_cnt c_delivery_cnt
 *                                c_discount c_credit_lim c_balance c_ytd_payment;
 *                                11 strings of 32 B; c_data 256 B   (644 B per row)
 *   3 neworder   append          ints no_o_id no_d_id no_w_id
 *   4 history    append          ints h_c_id h_c_d_id h_c_w_id h_d_id h_w_id h_amount;
 *                                h_date, h_data 32 B
 *   5 order      append          ints o_id o_d_id o_w_id o_c_id o_carrier_id o_ol_cnt
 *                                o_all_local; o_entry_d 32 B
 *   6 orderline  append          ints ol_o_id ol_d_id ol_w_id ol_number ol_i_id
 *                                ol_supply_w_id ol_delivery_d ol_quantity ol_amount;
 *                                ol_dist_info 24 B
 *   7 stock      100000 W        ints s_i_id s_w_id s_quantity s_ytd s_order_cnt
 *                                s_remote_cnt; 10 s_dist of 24 B; s_data 32 B (296 B)
 *   8 item       100000          ints i_id i_im_id i_price; i_name 24 B, i_data 52 B
 *
 * Money and rates are integer columns (cents, basis points) so that RMW_ADD
 * (a u32 add) is the arithmetic the transaction means.  Keys: district
 * W*10+D, customer (W*10+D)*3000+C, stock W*100000+I (the reference uses
 * W*10^6, gpuquery.cu:2559, outside its W*10^5-row table).
 *
 * Tables 3-6 are append tables (include/ltpmg.h, INSERT / DELETE).  At
 * creation each district holds `preload` delivered-or-not orders (the
 * reference sizes these tables W*10*3000 rows, src/param.cpp:74-77): order
 * and neworder row (W*10+D)*preload+k for k < preload, and LTPMG_OE_PRELOAD_LINES
 * orderlines per such order at rows order_row*LTPMG_OE_PRELOAD_LINES+l.
 * Capacities: order and neworder preload*10*W + cap, history cap, orderline
 * preload*10*W*LTPMG_OE_PRELOAD_LINES + 15*cap.  A transaction's INSERTs
 * take the rows the engine hands out (the result slot of an INSERT holds its
 * row id), so an aborted transaction leaves no hole and a full table fails
 * the batch (LTPMG_EFULL) instead of overwriting old rows.
 *
 * Query -> ops (a read of a cell may precede a write of it in one txn):
 *   Neworder (W, D, C, lines): READ w_tax; READ d_tax, d_next_o_id; RMW
 *     d_next_o_id +1; READ c_discount, c_last, c_credit; INSERT an order
 *     row and a neworder row; per line READ i_price, i_name, RMW s_quantity
 *     -q, s_ytd +q, s_order_cnt +1 (s_remote_cnt +1 when the supplying
 *     partition differs), INSERT an orderline row.
 *   Payment (W, D, C, amount): RMW w_ytd, d_ytd +amount; READ w_name,
 *     d_name; RMW c_balance -amount, c_ytd_payment +amount, c_payment_cnt
 *     +1; READ c_first, c_last, c_data; INSERT a history row.
 *   Orderstatus (read-only): customer balance/first/last; order O_ID of the
 *     district (row (W*10+D)*preload + O_ID mod preload): o_id, o_carrier_id,
 *     o_entry_d and its orderlines' ol_i_id, ol_quantity, ol_amount.
 *   Delivery (10 entries, the reference's parse<Delivery_Query>,
 *     gpuquery.cu:2794-2875): per entry, order k = NO_O_ID mod preload of
 *     district (NO_W_ID, NO_D_ID): DELETE its neworder row; UPDATE its
 *     o_carrier_id; per orderline READ ol_amount, UPDATE ol_delivery_d; RMW
 *     c_delivery_cnt +1 of customer NO_C_ID.  The ops of an entry whose
 *     neworder row is already gone still run (batches carry no data
 *     dependencies); the DELETE's status says whether it found the row.  An
 *     entry naming an order an earlier entry of the query names is dropped.
 *   Stocklevel (read-only): d_next_o_id; 10 orderlines of the district
 *     (line O_OL_ID[k] mod preload*LTPMG_OE_PRELOAD_LINES) and the
 *     s_quantity of items I_ID[k].
 * Customer by last name (the C_LAST index of cuda/src/gpudatabase.cu:201-230
 * and its scan in parse<Payment_Query>, gpuquery.cu:2641-2650): a Payment
 * or Orderstatus with isName set addresses the customer with the lowest C_ID
 * in its district whose last-name code equals C_LAST, or C_ID 0 when no
 * customer has it, as the reference's first-match scan does.  A customer's
 * code is the first little-endian u64 of its c_last cell, as
 * ltpmg_fill_table(table_seed) writes it, mod LTPMG_OE_LAST_NAMES (the
 * reference's 15-name vocabulary, src/database.cpp:312-316).  Names never
 * change, so the lookup is resolved when the query is decomposed into ops,
 * the step that does it in the reference.
 */
#ifndef LTPMG_OE_H
#define LTPMG_OE_H

#include <stdint.h>

#include "ltpmg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LTPMG_OE_TABLES 9
#define LTPMG_OE_LAST_NAMES 15
#define LTPMG_OE_APPEND_MASK 0x78u      /* tables 3 (neworder) .. 6 (orderline) */
#define LTPMG_OE_PRELOAD_LINES 5u       /* orderlines of each preloaded order */
#define LTPMG_OE_MAX_PRELOAD 3000u      /* the reference's orders per district */

typedef struct {
    uint32_t n_txns;
    uint32_t warehouses;      /* W partitions */
    uint32_t neworder_pct;    /* percent multi-line orders */
    uint32_t payment_pct;     /* percent payments */
    uint32_t delivery_pct;    /* percent deliveries; the rest read-only (half each kind) */
    uint32_t lines_min;       /* order lines per Neworder, uniform */
    uint32_t lines_max;
    uint32_t remote_bp;       /* per line: basis points of a remote supplying partition */
    uint32_t preload;         /* orders per district at table creation (ltpmg_oe_schema) */
    uint32_t byname_pct;      /* Payment / Orderstatus: percent by last name (the reference draws 15) */
    uint64_t seed;
    uint64_t table_seed;      /* the seed the tables were filled with (last-name lookups) */
} ltpmg_oe_spec;

/* Fills out[9] with the table descriptors for W partitions, `preload`
 * orders per district (<= LTPMG_OE_MAX_PRELOAD) and room for the rows cap
 * transactions append, and n_present[9] (if not NULL) with the rows present
 * at creation: every row of an ordinary table, the preloaded rows of an
 * append table (LTPMG_OE_APPEND_MASK; create those with
 * ltpmg_create_append_table). */
int ltpmg_oe_schema(uint32_t warehouses, uint32_t append_cap, uint32_t preload,
                    ltpmg_table_desc out[LTPMG_OE_TABLES], uint32_t n_present[LTPMG_OE_TABLES]);

/* first[(w * 10 + d) * LTPMG_OE_LAST_NAMES + k] = the C_ID a lookup of
 * last-name code k in district (w, d) finds, for tables filled from
 * table_seed. */
int ltpmg_oe_last_name_index(uint32_t warehouses, uint64_t table_seed, uint32_t *first);

/* Upper bound on device ops of a batch. */
uint64_t ltpmg_gen_oe_max_ops(const ltpmg_oe_spec *spec);

/* Generates a batch of the order-entry mix (seeded, thread-count
 * independent).  TIDs are positions in the batch; INSERT args are seeded by
 * (seed, TID, table, line). */
int ltpmg_gen_oe(ltpmg_oe_spec *spec, uint32_t *txn_off, uint32_t *rows, uint16_t *meta,
                 uint64_t *args, uint64_t max_ops, uint32_t *n_ops_out, uint32_t n_threads);

#ifdef __cplusplus
}
#endif
#endif
