// txn_structure.h -- query and op records of the host API.
//
// The query structs keep the reference's names and field layout
// (include/txn_structure.h:7-211) because callers fill them directly; they
// are a data format, so they stay alike.  What crosses to the GPU is no
// longer the 52-byte Global_Txn record but the SoA batch of include/ltpmg.h;
// KV_Batch below carries such a batch through the std::vector<std::any>
// interface unchanged.
#pragma once
#ifndef LTPMG_TXN_STRUCTURE
#define LTPMG_TXN_STRUCTURE

#include "define.h"

struct Global_Txn_Exec_Param {
    UINT32 global_txn_info_size;
    UINT32 global_sub_txn_size;
    UINT32 batch_size;
    UINT32 target_platform;
    UINT32 target_GPU;
    UINT32 start;
    UINT32 size;
};

struct Global_Txn_Sub {
    UINT32 TID = 0;
    UINT32 sub_txn_ID = 0;
    UINT32 ispopular = 0;
    UINT32 type = 'n';
    UINT32 benchmark = 0;
    UINT32 table_ID = 0;
    UINT32 dest_Row_1 = 0;
    UINT32 dest_Row_2 = 0;
    UINT32 data = 0;
    UINT32 dest_device = 0;
    UINT32 cur_mark = 0;
    UINT32 isCommitted = 0;
};

struct Global_Txn_Result {
    UINT32 TID;
    UINT32 int_attributes[16];
    UINT32 string_attributes[16 * 8];
    DOUBLE double_attributes[16];
};

struct Global_Txn_Info {
    UINT32 cur_subtxn_cnt = 0;
    UINT32 select_cnt = 0;
    UINT32 insert_cnt = 0;
    UINT32 update_cnt = 0;
    UINT32 scan_cnt = 0;
    UINT32 delete_cnt = 0;
};

struct Global_Txn {
    UINT32 global_txn_info_ID = 0;
    UINT32 dest_device = 0;
    Global_Txn_Sub *subtxn;
};

struct Test_Query {
    UINT32 TID;
    UINT32 Row_0, Row_1, Row_2, Row_3, Row_4, Row_5;
};

struct Test_Query_2 {
    UINT32 TID;
    UINT32 Row_0, Row_1, Row_2, Row_3, Row_4;
};

struct Neworder_Query {
    UINT32 TID;
    UINT32 W_ID;
    UINT32 D_ID;
    UINT32 C_ID;
    UINT32 O_ID;
    UINT32 N_O_ID;
    UINT32 O_OL_CNT;
    UINT32 O_OL_ID;
    struct Info {
        UINT32 OL_I_ID;
        UINT32 OL_SUPPLY_W_ID;
        UINT32 OL_QUANTITY;
    };
    Info INFO[15];
};

struct Payment_Query {
    UINT32 TID;
    UINT32 W_ID;
    UINT32 D_ID;
    UINT32 C_ID;
    UINT32 C_LAST;
    UINT32 isName;
    UINT32 C_D_ID;
    UINT32 C_W_ID;
    UINT32 H_AMOUNT;
    UINT32 H_ID;
};

struct Orderstatus_Query {
    UINT32 TID;
    UINT32 W_ID;
    UINT32 D_ID;
    UINT32 C_ID;
    UINT32 C_LAST;
    UINT32 O_ID;
    UINT32 OL_ID;
    UINT32 isName;
};

struct Delivery_Query {
    UINT32 TID;
    UINT32 NO_O_ID[10];
    UINT32 NO_W_ID[10];
    UINT32 NO_D_ID[10];
    UINT32 NO_C_ID[10];
};

struct Stocklevel_Query {
    UINT32 TID;
    UINT32 W_ID;
    UINT32 D_ID;
    UINT32 I_ID[10];
    UINT32 O_OL_ID[10];
};

struct YCSB_A_Query {
    UINT32 TID;
    UINT32 ROW_ID[10];
};
struct YCSB_B_Query {
    UINT32 TID;
    UINT32 ROW_ID[10];
};
struct YCSB_C_Query {
    UINT32 TID;
    UINT32 ROW_ID[10];
};
struct YCSB_D_Query {
    UINT32 TID;
    UINT32 ROW_ID[10];
};
struct YCSB_E_Query {
    UINT32 TID;
    UINT32 ROW_ID[10];
};

// A whole batch in the engine's SoA form (include/ltpmg.h: txn t owns ops
// [txn_off[t], txn_off[t+1]), op = row / meta / arg).  Used by the KV
// benchmarks, whose generator writes SoA directly.
struct KV_Batch {
    std::vector<UINT32> txn_off;
    std::vector<UINT32> rows;
    std::vector<UINT16> meta;
    std::vector<UINT64> args;
};

#endif
