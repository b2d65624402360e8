// database.h -- host image of the tables (the reference's include/database.h:
// same class, public methods and the two extern entry points at
// database.h:66-70).
//
// The reference reads the schema from data/<benchmark>.json with a vendored
// JSON parser (src/row_parser.cpp); this build fixes the schemas it runs
// (YCSB's 3 int + 3 double + 4 string columns, data/YCSB.json; the KV tables)
// and keeps the filepath argument for source compatibility.  Cells are filled
// from Param::seed with the engine's seeded cell definition (include/ltpmg.h),
// so the host image equals what ltpmg_fill_table would write.
#pragma once

#ifndef LTPMG_DATABASE
#define LTPMG_DATABASE

#include "db_structure.h"
#include "define.h"
#include "param.h"

class Database {
private:
    std::string filepath;

    Global_Table_Info *tables_for_GPU_info = nullptr;

    Global_Table *tables_for_GPU = nullptr;

    Global_Table_Index *index_for_GPU = nullptr;

    uint32_t n_tables = 0;

    std::vector<UINT32> string_to_uint(const std::string &input);

public:
    Database(const std::string &filepath);

    ~Database();

    void generate(std::shared_ptr<Param> param);

    void print() const;

    void print_global_row() const;

    void test();

    void malloc_global_row(std::shared_ptr<Param> param);

    void copy_to_global_row(std::shared_ptr<Param> param);

    void free_global_row(std::shared_ptr<Param> param);

    void transfer_to_GPU(std::shared_ptr<Param> param);

    void transfer_to_CPU(std::shared_ptr<Param> param);

    const Global_Table_Info *get_info() const { return tables_for_GPU_info; }
    const Global_Table *get_tables() const { return tables_for_GPU; }
};

extern void transfer_database_to_GPU(std::shared_ptr<Param> param,
                                     Global_Table_Info *table_for_gpu_info,
                                     Global_Table *table_for_gpu,
                                     Global_Table_Index *index_for_GPU);

extern void transfer_database_to_CPU(std::shared_ptr<Param> param,
                                     Global_Table_Info *table_for_gpu_info,
                                     Global_Table *table_for_gpu,
                                     Global_Table_Index *index_for_GPU);

#endif
