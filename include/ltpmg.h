/*
 * ltpmg.h -- C ABI of libltpmg_hip.so, the MI355X batched transaction engine.
 *
 * The reference exposes its GPU path as eight C++ functions taking
 * std::shared_ptr<Param>, std::vector<std::any> and table PODs
 * (include/database.h:66-70, include/query.h:63-70, include/scheduler.h:33-37,
 * defined in cuda/src/gpuscheduler.cu:3-47).  Only POD C types cross this
 * ABI; the C++ shim in ltpmg_amd/host/gpu_shim.cpp keeps those eight
 * signatures and calls the functions below.  Every function returns an int
 * status (LTPMG_OK or a negative LTPMG_E*), with a message available from
 * ltpmg_last_error().  Nothing here falls back to the CPU: without a usable
 * GPU, ltpmg_open() fails with LTPMG_EDEVICE.
 *
 * Batch model.  A batch is n_txns transactions in CSR form: txn t owns ops
 * [txn_off[t], txn_off[t+1]), 1..LTPMG_MAX_TXN_OPS of them, and its TID is t
 * (its position in the batch, as src/query.cpp:284-292 assigns it).  An op
 * is (row u32, meta u16, arg u64), meta = kind | col<<2 | table<<8, the SoA
 * replacement of the reference's 52-byte Global_Txn record
 * (cuda/include/txn_structure.cuh:21-41).
 *
 * Value semantics (the reference's row operators are no-ops,
 * cuda/src/gpulauncher.cu:8-165, so these are this engine's definitions):
 *   stream(s, j) = mix64(s + (j+1)*0x9E3779B97F4A7C15), mix64 = SplitMix64's
 *   finaliser.  A cell's initial bytes from ltpmg_fill_table(seed) are the
 *   little-endian bytes of stream(cellseed, 0..) with
 *   cellseed = mix64(seed + 0x9E3779B97F4A7C15 * (1 + (row<<10 | table<<6 | col))).
 *   READ      returns the cell (result slot bytes past the width are 0).
 *   UPDATE    sets the cell to the bytes of stream(arg, 0..).
 *   RMW_ADD   adds (uint32_t)arg to every little-endian u32 word of the cell.
 * Within one txn a cell that is written must not be accessed by another op.
 *
 * Append tables (ltpmg_create_append_table) hold rows [0, cursor), each
 * present or deleted, up to a capacity; two row-level op kinds act on them
 * (kind LTPMG_OP_ROW, the column field names which):
 *   INSERT    appends a row at the end of the batch: the committed INSERTs
 *             into a table take rows cursor, cursor+1, ... in op (TID) order,
 *             so aborted txns leave no holes; every column c gets the bytes of
 *             stream(rowseed(arg, c), 0..), rowseed(arg, c) =
 *             mix64(arg + (c+1) * 0xD1B54A32D192ED03), and the op's result
 *             slot the row id (u32, rest zero).  The row field must be 0.
 *   DELETE    marks row `row` deleted.
 * An op on a row that is not present (deleted, or not yet inserted when the
 * batch started: rows an INSERT appends become visible to the next batch)
 * has no effect and its status is 0; a READ of one returns zeros.  A txn
 * that deletes a row does not access that row with another op.  INSERT and
 * DELETE are only valid on append tables; every row of an ordinary table is
 * always present.  In the lowest-TID-wins mode a DELETE reserves its row like
 * a write; an INSERT touches no existing row and reserves nothing.
 *
 * Modes (both equal a serial replay in ascending TID of the committed txns):
 *   LTPMG_MODE_A  lowest TID wins: W[row] = min TID writing the row; a txn
 *                 commits iff no lower TID writes any row it touches (the
 *                 commented atomicMin rule of gpulauncher.cu:414, 2130-2142).
 *   LTPMG_MODE_O  ordered: every txn commits; each cell sees its ops in TID
 *                 order (what filter_commit + commit intend,
 *                 gpulauncher.cu:2439-2700).
 */
#ifndef LTPMG_H
#define LTPMG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LTPMG_MAX_TABLES 16
#define LTPMG_MAX_COLS 64
#define LTPMG_MAX_COL_BYTES 256
#define LTPMG_MAX_TXN_OPS 256
#define LTPMG_MAX_SLOTS 8

#define LTPMG_OK 0
#define LTPMG_EINVAL (-1)   /* bad argument or malformed batch */
#define LTPMG_ENOMEM (-2)   /* device or pinned allocation failed */
#define LTPMG_EBUDGET (-3)  /* request exceeds the configured HBM budget */
#define LTPMG_EHIP (-4)     /* HIP runtime error */
#define LTPMG_ESTATE (-5)   /* call out of order (no table / no batch) */
#define LTPMG_EDEVICE (-6)  /* no usable GPU */
#define LTPMG_EFULL (-7)    /* the INSERTs of a batch may not fit an append table's capacity */

#define LTPMG_OP_READ 0u
#define LTPMG_OP_UPDATE 1u
#define LTPMG_OP_RMW_ADD 2u
#define LTPMG_OP_ROW 3u    /* row-level op on an append table; the column field says which: */
#define LTPMG_ROW_INSERT 0u
#define LTPMG_ROW_DELETE 1u
#define LTPMG_META(kind, table, col) \
    ((uint16_t)((kind) | ((uint32_t)(col) << 2) | ((uint32_t)(table) << 8)))
#define LTPMG_META_INSERT(table) LTPMG_META(LTPMG_OP_ROW, table, LTPMG_ROW_INSERT)
#define LTPMG_META_DELETE(table) LTPMG_META(LTPMG_OP_ROW, table, LTPMG_ROW_DELETE)

#define LTPMG_MODE_A 0u
#define LTPMG_MODE_O 1u
#define LTPMG_EXEC_PROFILE 0x100u /* or-ed into the mode: record per-kernel HIP events */
/* or-ed into the mode: an ordered batch always runs the multi-kernel
 * pipeline, never the one-launch path for small batches (below) */
#define LTPMG_EXEC_PIPELINE 0x200u
/* Ordered batches of at most this many ops run as one launch of one
 * workgroup (sort, scan, cell moves and row ops in LDS) instead of the
 * multi-kernel pipeline, with the same results. */
#define LTPMG_SMALL_MAX_OPS 8192u

#define LTPMG_CFG_PROFILE 1u /* record per-kernel HIP events on every execute */

typedef struct ltpmg_engine ltpmg_engine;

typedef struct {
    int32_t device;      /* HIP device ordinal */
    uint32_t n_slots;    /* batch slots, 1..LTPMG_MAX_SLOTS */
    uint32_t max_txns;   /* per-slot capacity */
    uint32_t max_ops;    /* per-slot capacity */
    uint64_t hbm_budget; /* bytes the engine may allocate; 0 = 90% of free HBM */
    uint32_t flags;      /* LTPMG_CFG_* */
} ltpmg_cfg;

typedef struct {
    uint32_t n_rows;
    uint32_t n_cols;
    uint32_t col_bytes[LTPMG_MAX_COLS]; /* multiple of 4, 4..LTPMG_MAX_COL_BYTES */
} ltpmg_table_desc;

typedef struct {
    uint32_t n_txns;
    uint32_t n_ops;
    const uint32_t *txn_off; /* n_txns + 1 entries, txn_off[0] = 0 */
    const uint32_t *rows;    /* n_ops */
    const uint16_t *meta;    /* n_ops */
    const uint64_t *args;    /* n_ops */
} ltpmg_batch;

#define LTPMG_MAX_KSTATS 24
typedef struct {
    uint32_t n_txns;
    uint32_t n_ops;
    uint32_t committed;
    uint32_t aborted;
    uint32_t mode;
    float exec_ms;   /* device time from the first CC kernel to the end of execute */
    float upload_ms; /* device time of the host-to-device batch copies */
    uint32_t n_kernels;
    char kernel_name[LTPMG_MAX_KSTATS][32];
    float kernel_ms[LTPMG_MAX_KSTATS]; /* with LTPMG_CFG_PROFILE or LTPMG_EXEC_PROFILE */
    uint64_t hbm_allocated;
    uint32_t inserted;  /* rows the batch appended (committed INSERTs) */
} ltpmg_batch_stats;

/* Engine lifecycle.  Replaces GPUScheduler's global singleton
 * (cuda/include/gpuscheduler.cuh:54) with an explicit handle. */
int ltpmg_open(const ltpmg_cfg *cfg, ltpmg_engine **out);
int ltpmg_close(ltpmg_engine *e);
/* The process-wide engine of a device: opened with one slot (default
 * budget) on the first call, the same handle on every later call, until
 * ltpmg_close of that handle.  The reference's eight C++ entry points take no
 * handle (its GPUScheduler is a global singleton,
 * cuda/include/gpuscheduler.cuh:54); the library that implements them runs
 * on this engine, and host code beside it reads what they ran (commit
 * counts, results, digests) through the same handle. */
int ltpmg_shared_engine(int32_t device, ltpmg_engine **out);
const char *ltpmg_last_error(const ltpmg_engine *e);

/* Tables in HBM: columnar, each column a 256-byte-aligned array of slots of
 * col_bytes rounded up (4, 8, 16, then multiples of 32; a 100-byte field
 * takes a 128-byte slot).  Replaces GPUdatabase::malloc_global_row /
 * copy_to_global_row (cuda/src/gpudatabase.cu:9-238), whose row-major typed
 * arrays (int/double/8xu32 strings) map to 4-, 8- and 32-byte columns. */
int ltpmg_create_table(ltpmg_engine *e, uint32_t table_id, const ltpmg_table_desc *desc);
int ltpmg_fill_table(ltpmg_engine *e, uint32_t table_id, uint64_t seed);
int ltpmg_load_column(ltpmg_engine *e, uint32_t table_id, uint32_t col, uint32_t row0,
                      uint32_t n_rows, const void *src); /* packed, col_bytes per row */
int ltpmg_read_column(ltpmg_engine *e, uint32_t table_id, uint32_t col, uint32_t row0,
                      uint32_t n_rows, void *dst);
/* An append table (see INSERT / DELETE above): capacity desc->n_rows rows,
 * of which rows [0, n_present) exist at creation (filled by
 * ltpmg_fill_table / ltpmg_load_column like any table).  Replaces the
 * reference's fixed-size insert targets (insert_executor at host-chosen rows,
 * cuda/src/gpulauncher.cu:205-260, 52-79) for append-only tables. */
int ltpmg_create_append_table(ltpmg_engine *e, uint32_t table_id, const ltpmg_table_desc *desc,
                              uint32_t n_present);
/* The rows an append table has allocated so far (its cursor); n_rows of an
 * ordinary table. */
int ltpmg_table_rows(ltpmg_engine *e, uint32_t table_id, uint32_t *out);
/* Presence bytes (1 present, 0 deleted or unallocated) of rows [row0, row0 + n_rows). */
int ltpmg_read_presence(ltpmg_engine *e, uint32_t table_id, uint32_t row0, uint32_t n_rows, uint8_t *dst);
/* Sets an append table's cursor to n_rows (<= capacity) and clears the
 * presence of rows [n_rows, capacity): the rows appended after that point
 * are gone and the next INSERT takes row n_rows.  Waits for every slot.  No
 * counterpart in the reference (its tables are fixed-size). */
int ltpmg_truncate_table(ltpmg_engine *e, uint32_t table_id, uint32_t n_rows);
/* Order-independent digest of a table: sum over cells of a hash of
 * (table, row, col, cell bytes).  Replaces nothing in the reference (it has
 * no readback, gpudatabase.cu:240-339 frees without copying back). */
int ltpmg_table_digest(ltpmg_engine *e, uint32_t table_id, uint64_t *out);
/* Bytes per op in the result buffer: the largest slot over all tables. */
uint32_t ltpmg_result_stride(const ltpmg_engine *e);

/* Batches.  submit copies the SoA batch through pinned staging into the
 * slot's device buffers (replaces GPUquery::copy_global_txn + query_parse,
 * cuda/src/gpuquery.cu:593-2149); execute enqueues the CC + execute kernels
 * on the slot's stream (replaces GPUlauncher::txn_kernel_launcher,
 * cuda/src/gpulauncher.cu:755-1116); wait blocks until the slot is idle;
 * fetch copies the commit flags (n_txns bytes) and the read results
 * (n_ops * ltpmg_result_stride() bytes, slot i holds op i's READ: its
 * column's bytes, then zeros; an INSERT's: its row id, then zeros) back.  The
 * kernels write a slot only as far as the column's own slot size and the
 * fetch writes the zeros past it on the host, so the padding of wide result
 * strides never crosses HBM.  The reference has no results path
 * (Global_Txn_Result is never written). */
int ltpmg_submit_batch(ltpmg_engine *e, uint32_t slot, const ltpmg_batch *b);
/* The checks submit applies, without submitting (for a pool of requests). */
int ltpmg_check_batch(ltpmg_engine *e, const ltpmg_batch *b);
/* Zero-copy submit: stage returns pointers into the slot's pinned buffer
 * (waiting for the slot to be idle), the caller writes the batch there, and
 * submit_staged uploads it.  With LTPMG_SUBMIT_TRUSTED the host-side checks
 * are skipped (for batches cut from an already checked pool); the kernels
 * still never touch memory outside the tables, but a malformed op's result
 * is then undefined. */
typedef struct {
    uint32_t *txn_off; /* n_txns + 1 */
    uint32_t *rows;
    uint16_t *meta;
    uint64_t *args;
} ltpmg_staging;
#define LTPMG_SUBMIT_TRUSTED 1u
int ltpmg_stage_batch(ltpmg_engine *e, uint32_t slot, uint32_t n_txns, uint32_t n_ops,
                      ltpmg_staging *out);
int ltpmg_submit_staged(ltpmg_engine *e, uint32_t slot, uint32_t flags);
/* Submit from a pool without a host copy: after ltpmg_stage_batch, the
 * caller writes only staging.txn_off; the batch's ops are the runs, each
 * n_ops consecutive ops of pool (rows / meta / args; pool->txn_off is not
 * read) copied to ops [dst_op, dst_op + n_ops) of the batch.  The runs must
 * tile the batch's ops in order.  A batch of at most 8 runs out of a pool
 * in pinned memory (ltpmg_pinned_alloc) is read by one kernel straight from
 * host memory, with its offsets; otherwise runs of 4096 ops or more are
 * copied straight from the pool by DMA and shorter runs are gathered into
 * the staging buffer first.  Either way the ops do not pass through the
 * CPU when the pool is pinned.  The ops are trusted as
 * with LTPMG_SUBMIT_TRUSTED (check the pool once with ltpmg_check_batch);
 * the offsets are checked.  Replaces the reference's per-batch pinned
 * allocation and upload (cuda/src/gpuquery.cu:2098-2146) for the serving
 * form. */
typedef struct {
    uint32_t src_op; /* first op of the run in the pool */
    uint32_t dst_op; /* where it goes in the batch */
    uint32_t n_ops;
} ltpmg_op_run;
int ltpmg_submit_runs(ltpmg_engine *e, uint32_t slot, const ltpmg_batch *pool, const ltpmg_op_run *runs,
                      uint32_t n_runs);
/* Pinned (page-locked) host memory for pools and arrival rings; no
 * counterpart in the reference (it pins a new buffer per batch). */
int ltpmg_pinned_alloc(uint64_t bytes, void **out);
int ltpmg_pinned_free(void *p);
/* Batches take effect on the tables in the order of ltpmg_execute calls,
 * across slots: a batch's reads and writes of cells and lock words wait for
 * the previously executed batch.  In the ordered mode the key build, sort
 * and scan of a batch only touch its slot and overlap the previous batch;
 * an ordered batch of at most LTPMG_SMALL_MAX_OPS ops runs as one launch
 * (unless LTPMG_EXEC_PIPELINE is or-ed into the mode).
 * Replaces the reference's one-batch-at-a-time epoch loop
 * (src/scheduler.cpp:44-59) for the serving form. */
int ltpmg_execute(ltpmg_engine *e, uint32_t slot, uint32_t mode);
/* Non-blocking: LTPMG_OK when the slot is idle, 1 while work is queued. */
int ltpmg_query(ltpmg_engine *e, uint32_t slot);
int ltpmg_wait(ltpmg_engine *e, uint32_t slot);
int ltpmg_fetch_results(ltpmg_engine *e, uint32_t slot, uint8_t *commit, void *results,
                        uint64_t results_bytes);
int ltpmg_stats(ltpmg_engine *e, uint32_t slot, ltpmg_batch_stats *out);
/* Per op of the slot's last batch (n_ops bytes): 1 if the op took effect --
 * its txn committed and its row was present (an INSERT: its row was
 * appended) -- else 0.  No counterpart in the reference (its operators return
 * nothing). */
int ltpmg_fetch_status(ltpmg_engine *e, uint32_t slot, uint8_t *status);

/* Building-block self-tests, the counterparts of the reference's in-binary
 * LTPMG_GPULAUNCHER_TEST_MERGESORT and _PREFIXSUM checks
 * (cuda/src/gpulauncher.cu:1237-1321).  sort: stable sort of (key, val)
 * pairs by the low key_bits of key, in place on host arrays.  segscan:
 * segmented exclusive sum of u32 values, a segment starting wherever
 * head[i] != 0 (and at i = 0). */
int ltpmg_selftest_sort(ltpmg_engine *e, uint32_t *keys, uint32_t *vals, uint32_t n,
                        uint32_t key_bits);
int ltpmg_selftest_segscan(ltpmg_engine *e, const uint32_t *vals, const uint8_t *head,
                           uint32_t *out, uint32_t n);
/* Streaming copy of n_bytes (16-byte vectors) between two device buffers,
 * iters times; *gbps gets (read + write bytes) / average kernel time.  The
 * measured HBM copy peak that rooflines are also quoted against, and the
 * known-size kernel that calibrates hardware byte counters.  No counterpart
 * in the reference. */
int ltpmg_selftest_copy(ltpmg_engine *e, uint64_t n_bytes, uint32_t iters, double *gbps);

/* Allocation-free execute.  No counterpart in the reference: this is for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  It runs
 * the CC + execute kernels of ltpmg_execute on a batch that is already in
 * device memory (dev->txn_off / rows / meta / args are device pointers,
 * n_txns / n_ops host values) and writes dev_commit (n_txns bytes) and
 * dev_results (n_ops * ltpmg_result_stride() bytes, each READ's slot written
 * to its end, zeros past the column), both device buffers the caller owns,
 * on hip_stream (a hipStream_t; NULL = the null stream).
 * It allocates nothing: its sort, scan and tile scratch is one static
 * device array per GPU of about 5 MiB, held until the process ends, which
 * bounds a batch to LTPMG_STATIC_MAX_OPS ops (larger ones fail with
 * LTPMG_EINVAL).  It issues only kernel launches (no memset or copy nodes),
 * so it can be captured into a graph.  Every call on one GPU shares that
 * array: calls that overlap on a GPU race, and nothing orders a call
 * against batches run through slots -- the caller serialises them (e.g.
 * ltpmg_wait on every slot first).  The host cannot inspect device arrays,
 * so op fields are checked by the kernels (invalid ops are skipped), and a
 * malformed txn_off only leaves the affected commit flags undefined (every
 * access stays inside the batch and the tables); the batch must otherwise
 * follow the rules of ltpmg_submit_batch.  An engine with an append table
 * is refused (LTPMG_ESTATE): row ops need ltpmg_execute. */
#define LTPMG_STATIC_MAX_OPS (1u << 18)
int ltpmg_execute_static(ltpmg_engine *e, uint32_t mode, const ltpmg_batch *dev, uint8_t *dev_commit,
                         void *dev_results, void *hip_stream);

#ifdef __cplusplus
}
#endif
#endif
