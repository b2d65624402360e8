/*
 * ltpmg_fixture.h -- versioned binary batch fixtures (libltpmg_host.so).
 *
 * One file holds one batch in the engine's SoA form (include/ltpmg.h)
 * together with what is needed to replay it: the seed the tables were
 * filled from, the mode, and the table descriptors.  The reference keeps no
 * inputs (its generators draw from rand()/time, src/query.cpp:115-116,
 * src/random.cpp); these files let a CPU run and a GPU run, or two builds,
 * share the exact batch.  The ltpmg driver's --dump_path writes one.
 *
 * Layout (little-endian):
 *   magic "LTPMGFX\0" | u32 version | u32 header bytes
 *   u64 seed | u32 mode | u32 n_tables | u32 append_mask (version 2 on)
 *   per table: u32 n_rows | u32 n_cols | u32 col_bytes[n_cols]
 *              | u32 n_present (version 2 on, tables in append_mask only)
 *   u32 n_txns | u32 n_ops | u64 checksum of the arrays
 *   txn_off u32[n_txns + 1] | rows u32[n_ops] | meta u16[n_ops] | zeros
 *   until the meta bytes are a multiple of 8 | args u64[n_ops]
 * The checksum chains mix64 over the arrays' bytes as little-endian 64-bit
 * words (the last word zero padded), starting from 0x9E3779B97F4A7C15.
 * Version 1 files (no append tables) read with append_mask 0.  A reader
 * rejects another magic, a newer version, a short file or a checksum
 * mismatch with LTPMG_EINVAL.
 */
#ifndef LTPMG_FIXTURE_H
#define LTPMG_FIXTURE_H

#include <stdint.h>

#include "ltpmg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LTPMG_FIXTURE_VERSION 2
#define LTPMG_FIXTURE_NO_MODE 0xFFFFFFFFu

typedef struct {
    uint64_t seed;     /* ltpmg_fill_table seed of every table */
    uint32_t mode;     /* LTPMG_MODE_A, LTPMG_MODE_O or LTPMG_FIXTURE_NO_MODE */
    uint32_t n_tables; /* table ids 0 .. n_tables-1 */
    ltpmg_table_desc tables[LTPMG_MAX_TABLES];
    uint32_t append_mask;                 /* bit t: table t is an append table */
    uint32_t n_present[LTPMG_MAX_TABLES]; /* an append table's rows at creation */
} ltpmg_fixture_info;

int ltpmg_fixture_write(const char *path, const ltpmg_fixture_info *info, const ltpmg_batch *b);

/* Reads the header: the info and the batch sizes to allocate for. */
int ltpmg_fixture_probe(const char *path, ltpmg_fixture_info *info, uint32_t *n_txns, uint32_t *n_ops);

/* Reads the arrays into b (n_txns / n_ops as probed, arrays allocated by the
 * caller) and checks the checksum. */
int ltpmg_fixture_read(const char *path, ltpmg_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* LTPMG_FIXTURE_H */
