// run_options.h -- this build's run settings, kept beside Param.
//
// Param crosses the eight extern entry points by std::shared_ptr
// (include/database.h:66-70, include/query.h:63-70, include/scheduler.h:33-37
// of the reference), so its layout must stay exactly the reference's
// (include/param.h:13-118): a host built against the reference's own header
// then hands this build's GPU library an object it reads correctly.  The
// settings the reference has no field for live here instead, filled from the
// flags of common.h, and reach the GPU library through the environment
// (LTPMG_CC_MODE, LTPMG_SEED, LTPMG_APPEND_CAP, LTPMG_OE_PRELOAD), which the driver sets from
// the same flags.
#pragma once

#include <cstdint>

class Param;

namespace ltpmg_host {

struct RunOptions {
    uint32_t cc_mode = 1;               // LTPMG_MODE_O (1) or LTPMG_MODE_A (0)
    uint64_t seed = 0;                  // table fill and workload
    uint32_t kv_size = 0;               // rows of the KV / KV_C1 table
    uint32_t kv_fields = 10;
    uint32_t kv_field_bytes = 100;
    uint32_t kv_ops = 16;
    uint32_t read_percent = 50;
    uint32_t kv_sub_txn_size = 0;       // upper bound on device ops per batch
    uint32_t oe_append_cap = 1u << 20;  // order-entry append tables, in transactions
    uint32_t oe_preload = 3000;         // order-entry orders per district at table creation
};

// The settings of this run: taken from the flags on first use.
RunOptions &options();

// Benchmark families (by Param::benchmark, a field the reference has).
bool is_ycsb(const Param &p);
bool is_kv(const Param &p);
// order-entry schema (include/ltpmg_oe.h): TPCC_PART, TPCC_ALL and the OE mix
bool is_oe(const Param &p);

// Exports the settings the GPU library reads (LTPMG_CC_MODE, LTPMG_SEED,
// LTPMG_APPEND_CAP, LTPMG_OE_PRELOAD) into the environment.
void export_gpu_options();

}  // namespace ltpmg_host
