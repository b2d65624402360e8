// common.h -- flags and timers of the host API (include/common.h of the
// reference: the same flags and functions, plus this build's own flags).
#pragma once
#ifndef LTPMG_COMMON
#define LTPMG_COMMON
#include "define.h"

DECLARE_string(benchmark);
DECLARE_uint32(table_size);
DECLARE_uint32(batch_size);
DECLARE_uint32(epoch_tp);
DECLARE_uint32(epoch_sync);
DECLARE_string(deviceIDs);
DECLARE_uint32(neworder_percent);
DECLARE_string(log_path);
DECLARE_double(zipf_config);
DECLARE_string(data_distribution);

// Flags this build adds (no counterpart in the reference).
DECLARE_string(cc_mode);        // "O": ordered, every txn commits; "A": lowest TID wins
DECLARE_uint64(seed);           // seeds the table fill and the workload
DECLARE_uint32(kv_rows);        // KV benchmarks: rows
DECLARE_uint32(kv_fields);      // KV benchmarks: fields per row
DECLARE_uint32(kv_field_bytes); // KV benchmarks: bytes per field
DECLARE_uint32(kv_ops);         // KV benchmarks: ops per txn
DECLARE_uint32(read_percent);   // KV benchmarks: percent READ
DECLARE_string(dump_path);      // if set: write batch, results and image digest of epoch 0
DECLARE_uint32(append_cap);     // order-entry append tables, in transactions
DECLARE_uint32(oe_preload);     // order-entry orders per district at table creation
DECLARE_string(schema_dir);     // where <benchmark>.json schemas are looked up

void initialize_dependency(int argc, char **argv);

void free_dependency();

long long current_time();

float duration(long long start_t, long long end_t);

void sleep(float time_t);

#endif
