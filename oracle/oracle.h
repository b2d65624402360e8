/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of LTPMG's batched transaction-execution semantics, used
 * as the checker for the HIP engine.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker or the reported CPU baseline; the product never links it.
 *
 * What it restates (paths relative to the LTPMG tree):
 *   - per-row registration of an op's row            cuda/src/gpulauncher.cu:382-418
 *   - the commented "min_TID" commit rule
 *     (lowest TID wins, else redo)                    cuda/src/gpulauncher.cu:414, 2130-2142, 2177-2196
 *   - per-row application of ops in TID order         cuda/src/gpulauncher.cu:2439-2700
 *   - TID = position of the txn in the batch          src/query.cpp:284-292
 *
 * The reference's row operators are no-ops (gpulauncher.cu:64-165), so the
 * value semantics below (cell fill, UPDATE payload, RMW add) are this
 * project's own definitions, fixed here and restated independently in the
 * engine:
 *   mix64(z)        SplitMix64 finaliser (Steele, Lea, Flood 2014; constants
 *                   from Vigna's splitmix64.c).
 *   stream(s, j)    j-th output (j >= 0) of SplitMix64 started at state s:
 *                   mix64(s + (j+1) * 0x9E3779B97F4A7C15).
 *   cellseed        mix64(seed + 0x9E3779B97F4A7C15 * (1 + (row<<10 | table<<6 | col)))
 *   cell bytes      little-endian bytes of stream(cellseed, 0), stream(cellseed, 1), ...
 *                   truncated to the column width.
 *   UPDATE(arg)     cell <- little-endian bytes of stream(arg, 0..), truncated to width.
 *   RMW_ADD(arg)    every u32 word of the cell += (uint32_t)arg (wrapping).
 *   READ            result slot <- cell bytes; slot bytes past the width are 0.
 *   INSERT(arg)     (kind 3, column field 0; append tables only) appends a row
 *                   at the end of the batch: the committed INSERTs of a table
 *                   take rows cursor, cursor+1, ... in op order; column c of
 *                   the row <- bytes of stream(rowseed(arg, c), 0..), rowseed =
 *                   mix64(arg + (c+1) * 0xD1B54A32D192ED03); the result slot
 *                   <- the row id (u32 little endian, rest 0).
 *   DELETE          (kind 3, column field 1) clears the row's presence byte.
 *   An op on a row that is not present (an append table's deleted or
 *   unallocated row) has no effect; a READ of one gets a zeroed slot.
 *   found[op] = 1 when the op took effect (its txn committed and its row was
 *   present, or its INSERT got a row), else 0.
 *
 * Batch semantics (both modes equal a serial replay in TID order):
 *   Mode A (lowest-TID-wins): W[k] = min TID of any txn writing row k
 *     (UPDATE, RMW_ADD or DELETE; an INSERT writes no existing row);
 *     txn t commits iff W[k] >= t for every row k it touches.  The committed
 *     txns are replayed serially in ascending TID.  (Row ops: the reference's
 *     insert/delete executors, cuda/src/gpulauncher.cu:205-356.)
 *   Mode O (ordered): every txn commits; all txns are replayed serially in
 *     ascending TID (per-row TID order, the intent of filter_commit+commit).
 *
 * Parity pinning of each function (see tests/test_oracle.py):
 *   orc_mix64 / orc_stream   published SplitMix64 outputs for seed 0.
 *   orc_replay Mode A        closed forms (all-distinct, all-write-one-key,
 *                            readers+one-writer), pairwise brute force
 *                            (orc_commit_pairwise) on random tiny batches.
 *   orc_replay Mode O        closed form for RMW(+1): final = snapshot + #RMW,
 *                            READ = snapshot + #lower-TID RMWs; UPDATE/READ
 *                            last-writer closed form.
 *   orc_cpu_* (pipeline)     equal to orc_replay on random batches.
 */
#ifndef LTPMG_ORACLE_H
#define LTPMG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_COLS 64
#define ORC_KIND_READ 0u
#define ORC_KIND_UPDATE 1u
#define ORC_KIND_RMW_ADD 2u
#define ORC_KIND_ROW 3u
#define ORC_ROW_INSERT 0u
#define ORC_ROW_DELETE 1u

#define ORC_MODE_A 0u
#define ORC_MODE_O 1u

typedef struct {
    uint32_t n_rows;
    uint32_t n_cols;
    uint32_t width[ORC_MAX_COLS];  /* bytes, multiple of 4 */
    uint8_t *col[ORC_MAX_COLS];    /* packed: cell r at col[c] + r * width[c] */
    uint8_t *present;              /* append tables: 1 byte per row (NULL: an ordinary table) */
    uint32_t cursor;               /* append tables: rows allocated; INSERTs move it on */
} orc_table;

uint64_t orc_mix64(uint64_t z);
uint64_t orc_stream(uint64_t s, uint32_t j);
uint64_t orc_cellseed(uint64_t seed, uint32_t table, uint32_t row, uint32_t col);
void orc_bytes_from_stream(uint64_t s, uint32_t width, uint8_t *out);
void orc_fill_column(uint64_t seed, uint32_t table, uint32_t col, uint32_t width,
                     uint32_t row0, uint32_t nrows, uint8_t *out);

/* Fills only the listed rows of a packed column (sparse images). */
void orc_fill_rows(uint64_t seed, uint32_t table, uint32_t col, uint32_t width,
                   const uint32_t *rows, uint32_t n, uint8_t *col_base);

/* Serial replay.  commit[n_txns] (0/1), results[n_ops * res_stride] and,
 * if not NULL, found[n_ops] are written; result slots of ops that are
 * neither a READ nor an INSERT, and of aborted txns, are untouched.
 * Returns 0, a negative value for malformed input, or -5 when an append
 * table's INSERTs did not fit (the rows that fit were appended). */
int orc_replay(orc_table *tables, uint32_t n_tables, uint32_t n_txns,
               const uint32_t *txn_off, const uint32_t *rows, const uint16_t *meta,
               const uint64_t *args, uint32_t mode, uint8_t *commit,
               uint8_t *results, uint32_t res_stride, uint8_t *found);

/* Mode A commit set by the pairwise definition (txn t aborts iff some s < t
 * writes a row t touches).  O(B^2 L^2): tiny inputs only. */
int orc_commit_pairwise(uint32_t n_txns, const uint32_t *txn_off, const uint32_t *rows,
                        const uint16_t *meta, uint8_t *commit);

/* Threaded restatement of the reference's batch pipeline (CPU baseline).
 * Mode O: histogram -> exclusive scan -> scatter -> per-row TID sort ->
 * per-row apply (gpulauncher.cu:459-512, 1558-1589, 2375-2437, 2439-2523,
 * 2525-2700).  Mode A: reserve (atomic min per row) -> validate -> reads ->
 * writes.  Same outputs as orc_replay for batches of cell ops on ordinary
 * tables (row ops and append tables are rejected).  Returns elapsed seconds
 * of the CC + execute stages (steady clock), or a negative value on error. */
double orc_cpu_pipeline(orc_table *tables, uint32_t n_tables, uint32_t n_txns,
                        const uint32_t *txn_off, const uint32_t *rows, const uint16_t *meta,
                        const uint64_t *args, uint32_t mode, uint8_t *commit,
                        uint8_t *results, uint32_t res_stride, uint32_t n_threads);

#ifdef __cplusplus
}
#endif
#endif
