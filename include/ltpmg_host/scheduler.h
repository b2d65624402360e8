// scheduler.h -- host driver of the batch loop (the reference's
// include/scheduler.h: same class and the three extern entry points at
// scheduler.h:33-37).  This build adds fetch_results_from_GPU so a caller
// can see what the reference never returned: the commit flags and read
// results of the last batch.
#pragma once

#ifndef LTPMG_SCHEDULER
#define LTPMG_SCHEDULER

#include "common.h"
#include "database.h"
#include "param.h"
#include "query.h"

class Scheduler {
private:
    std::shared_ptr<Param> param = std::make_shared<Param>();

    std::shared_ptr<Database> db = std::make_shared<Database>(param->get_filepath());

    std::shared_ptr<Query> query = std::make_shared<Query>();

public:
    Scheduler(/* args */);
    ~Scheduler();
    void init_launcher(std::shared_ptr<Param> param);
    void execute(std::shared_ptr<Param> param);
    void free_launcher(std::shared_ptr<Param> param);
    void start();
    void end();
};

extern void initial_launcher_on_GPU(std::shared_ptr<Param> param);

extern void execute_on_GPU(std::shared_ptr<Param> param);

extern void free_launcher_on_GPU(std::shared_ptr<Param> param);

// Added: commit flags (one byte per txn) and read results (one slot of
// result_stride bytes per op) of the last executed batch, plus the SoA form
// of that batch as it was submitted.  Returns false if nothing has run.
struct Batch_Results {
    KV_Batch batch;
    std::vector<UINT8> commit;
    std::vector<UINT8> results;
    UINT32 result_stride = 0;
};
extern bool fetch_results_from_GPU(std::shared_ptr<Param> param, Batch_Results *out);
// Added: order-independent digest of a table on the device (ltpmg_table_digest).
extern UINT64 table_digest_on_GPU(std::shared_ptr<Param> param, UINT32 table_id);

#endif
