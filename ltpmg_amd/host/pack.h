// pack.h -- the reference's query structs -> the engine's SoA batch.
//
// The decomposition follows the reference's parse<T> kernels
// (cuda/src/gpuquery.cu:2473-3303); the rules are written out in pack.cpp.
// Shared by the reference-interface entry points (gpu_shim.cpp, in
// libltpmg_gpu.so) and the host code that records a run (scheduler.cpp), so
// both see the same batch.
#pragma once

#include <any>
#include <cstdint>
#include <string>
#include <vector>

#include "db_structure.h"
#include "param.h"
#include "txn_structure.h"

namespace ltpmg_pack {

// YCSB_E's table is an append table: rows [0, table_size) present and room
// for one insert per query and epoch.  Returns its capacity (0 when it does
// not fit 32-bit row ids).
uint32_t ycsb_e_capacity(const Param &p, uint32_t table_size);

// Packs the batch; info0 is table 0 of the loaded schema (the YCSB table's
// column count), seed the run's seed, preload the order-entry orders per
// district the tables were created with.  Returns false, with the query type
// in *why, for a query this engine does not decompose.
bool pack_queries(const std::vector<std::any> &batch, const Global_Table_Info &info0, uint64_t seed,
                  uint32_t preload, KV_Batch &out, std::string *why);

}  // namespace ltpmg_pack
