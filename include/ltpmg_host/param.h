// param.h -- run parameters of the host API (the reference's include/param.h:
// same class, fields and methods, built from the flags in common.h).
//
// Benchmarks: YCSB_A..E (the reference's key-value mixes over its 10-column
// YCSB table); TPCC_PART (Neworder + Payment by neworder_percent) and OE
// (45% Neworder, 43% Payment, 12% read-only) over the order-entry schema of
// include/ltpmg_oe.h, table_size = partitions; plus this build's KV_C1
// (int32 columns, READ whole row / RMW +1) and KV (fields of kv_field_bytes,
// READ / UPDATE one field).  TEST and TPCC_ALL keep their sizes; TPCC_ALL's
// Delivery and the TEST queries are not built.
#pragma once
#include "common.h"
#include "db_structure.h"
#include "define.h"
#include "txn_structure.h"

// Result and Param keep the reference's members, types and order exactly
// (include/param.h:7-118), so the object a reference host allocates and
// passes through the extern entry points has the layout this build reads.
// This build's added settings are in run_options.h; commit counts come from
// the engine (ltpmg_stats), not from Result.
struct Result {
    double cost;          // device seconds of the last batch's CC + execute
    uint32_t epoch;
    uint32_t batch_size;
};

class Param {
public:
    std::string benchmark;
    uint32_t table_size;
    uint32_t batch_size;
    uint32_t epoch_tp;
    uint32_t epoch_sync;
    std::string deviceIDs;
    uint32_t neworder_percent;
    std::string filepath;
    uint32_t table_cnt;
    uint32_t device_cnt = 0;        // default member initialisers leave the layout unchanged
    uint32_t *device_IDs = nullptr;
    std::string data_distribution;
    double zipf_config;

    uint32_t test_1_size = 0;
    uint32_t test_2_size = 0;

    uint32_t warehouse_size = 0;
    uint32_t district_size = 0;
    uint32_t customer_size = 0;
    uint32_t neworder_size = 0;
    uint32_t history_size = 0;
    uint32_t order_size = 0;
    uint32_t orderline_size = 0;
    uint32_t stock_size = 0;
    uint32_t item_size = 100000;

    uint32_t ycsb_size = 0;
    uint32_t bitmap_row_cnt = 0;
    uint32_t bitmap_table_cnt = 0;

    uint32_t test_query_batch_size = 0;
    uint32_t test_query_2_batch_size = 0;
    uint32_t neworder_query_batch_size = 0;
    uint32_t payment_query_batch_size = 0;
    uint32_t orderstatus_query_batch_size = 0;
    uint32_t delivery_query_batch_size = 0;
    uint32_t stocklevel_query_batch_size = 0;
    uint32_t ycsb_a_query_batch_size = 0;
    uint32_t ycsb_b_query_batch_size = 0;
    uint32_t ycsb_c_query_batch_size = 0;
    uint32_t ycsb_d_query_batch_size = 0;
    uint32_t ycsb_e_query_batch_size = 0;

    uint32_t test_query_subtxn_cnt = 0;
    uint32_t test_query_2_subtxn_cnt = 0;
    uint32_t neworder_query_subtxn_cnt = 0;
    uint32_t payment_query_subtxn_cnt = 0;
    uint32_t orderstatus_query_subtxn_cnt = 0;
    uint32_t delivery_query_subtxn_cnt = 0;
    uint32_t stocklevel_query_subtxn_cnt = 0;
    uint32_t ycsb_a_query_subtxn_cnt = 0;
    uint32_t ycsb_b_query_subtxn_cnt = 0;
    uint32_t ycsb_c_query_subtxn_cnt = 0;
    uint32_t ycsb_d_query_subtxn_cnt = 0;
    uint32_t ycsb_e_query_subtxn_cnt = 0;

    uint32_t test_sub_txn_size = 0;
    uint32_t tpcc_all_sub_txn_size = 0;
    uint32_t tpcc_part_sub_txn_size = 0;
    uint32_t ycsb_a_sub_txn_size = 0;
    uint32_t ycsb_b_sub_txn_size = 0;
    uint32_t ycsb_c_sub_txn_size = 0;
    uint32_t ycsb_d_sub_txn_size = 0;
    uint32_t ycsb_e_sub_txn_size = 0;

    uint32_t test_query_subtxn_kinds = 0;
    uint32_t tpcc_all_query_subtxn_kinds = 0;
    uint32_t tpcc_part_query_subtxn_kinds = 0;
    uint32_t ycsb_a_query_subtxn_kinds = 0;
    uint32_t ycsb_b_query_subtxn_kinds = 0;
    uint32_t ycsb_c_query_subtxn_kinds = 0;
    uint32_t ycsb_d_query_subtxn_kinds = 0;
    uint32_t ycsb_e_query_subtxn_kinds = 0;

    uint32_t select_batch_size = 0;
    uint32_t insert_batch_size = 0;
    uint32_t update_batch_size = 0;
    uint32_t scan_batch_size = 0;
    uint32_t delete_batch_size = 0;

    Result result;

    Param();

    ~Param();

    void print();

    void analyse_deviceIDs();

    void free_deviceIDs();

    const std::string &get_filepath() { return filepath; }
    void set_table_cnt(uint32_t table_cnt) { this->table_cnt = table_cnt; }

    uint32_t get_subtxn_kinds();

    uint32_t get_sub_txn_size();

    uint32_t get_subtxn_cnt(const std::type_info &query_kind);

    uint32_t get_txn_batch_size(const std::type_info &query_kind);
};

#include "run_options.h"
