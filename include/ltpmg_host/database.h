// database.h -- host image of the tables (the reference's include/database.h:
// same class, public methods and the two extern entry points at
// database.h:66-70).
//
// As in the reference, the constructor parses the benchmark's schema file
// (Param's path, else the same name under --schema_dir, default the
// schemas/ directory beside the library) and the YCSB table's int / double /
// string column counts come from it; the order-entry schema's table list
// must match the adapter's nine tables; the KV benchmarks have no file.
// Cells are filled from Param::seed with the engine's seeded cell definition
// (include/ltpmg.h), so the host image equals what ltpmg_fill_table writes.
#pragma once

#ifndef LTPMG_DATABASE
#define LTPMG_DATABASE

#include "db_structure.h"
#include "define.h"
#include "param.h"
#include "row_parser.h"

class Database {
private:
    std::string filepath;

    Row_Parser parser;

    std::vector<Row> rows;

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
    const std::vector<Row> &get_rows() const { return rows; }
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
