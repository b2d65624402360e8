// db_structure.h -- table descriptors of the host API, the same PODs as the
// reference's include/db_structure.h:7-46 (a data format shared with callers).
// A table crosses to the engine as int columns (4 B), then double columns
// (8 B), then string columns (string_length u32 = 32 B), in that order.
#pragma once
#ifndef LTPMG_DB_STRUCTURE
#define LTPMG_DB_STRUCTURE

#include "define.h"

struct Global_Table_Info {
    UINT32 int_size = 0;
    UINT32 double_size = 0;
    UINT32 string_size = 0;
    UINT32 string_length = 8;  // u32 words per string (32 chars)
    UINT32 table_size = 0;
    UINT32 table_cnt = 0;
};

struct Global_Table {
    INT32 *int_data;       // row-major: [row * int_size + c]
    DOUBLE *double_data;   // row-major: [row * double_size + c]
    UINT32 *string_data;   // row-major: [(row * string_size + c) * string_length + k]
};

struct Global_Table_Index {
    UINT32 *index;
};

struct Global_Table_Meta {
    UINT32 row_start;
    UINT32 row_end;
    UINT32 table_slice_size;
    UINT32 table_size;
    UINT32 bitmap_row_slice_size;
};

struct Global_Table_Strategy {
    UINT32 *int_target_GPU;
    UINT32 *int_target_GPU_platform;
    UINT32 *string_target_GPU;
    UINT32 *string_target_GPU_platform;
    UINT32 *double_target_GPU;
    UINT32 *double_target_GPU_platform;
};

#endif
