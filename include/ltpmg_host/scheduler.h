// scheduler.h -- host driver of the batch loop (the reference's
// include/scheduler.h: same class and the three extern entry points at
// scheduler.h:33-37).  What the reference never returned -- commit flags,
// read results, table digests of the last batch -- is read from the
// engine the entry points run on (ltpmg_shared_engine, include/ltpmg.h).
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

#endif
