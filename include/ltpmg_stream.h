/*
 * ltpmg_stream.h -- the batching form of libltpmg_hip.so: a stream of uneven
 * batches resident in HBM, run back to back.
 *
 * The reference runs one batch per GPUlauncher::txn_kernel_launcher call,
 * with about 3 + 6T launches and a device synchronisation after every stage
 * (cuda/src/gpulauncher.cu:755-1116), so a small batch costs mostly launch
 * and sync latency.  A stream keeps every batch of the stream on the device
 * and runs them under a plan:
 *
 *   ordered batches   consecutive ordered-mode batches are merged into one
 *                     launch group of at most merge_max_ops ops.  Ordered
 *                     mode is a serial replay in TID order, so a group of
 *                     batches b..c gives the same commit flags, read values
 *                     and final tables as running b, then b+1, ..., then c.
 *                     Rows that INSERTs append appear at the end of the
 *                     group, so a batch joins a group that INSERTs into an
 *                     append table only if its other ops on that table name
 *                     rows below the table's cursor at the start of the run.
 *                     Lowest-TID-wins batches with row ops run the
 *                     three-kernel pipeline plus the row-op kernels.
 *   lowest-TID-wins   each batch stays its own unit (its TIDs, its epoch).
 *                     A batch with at most fused_max_ops ops runs as one
 *                     persistent kernel (reserve | validate + read | write
 *                     behind two grid barriers, fused_tile_ops ops per
 *                     tile, fused_wgs_per_cu workgroups per CU); a larger
 *                     one runs the three-kernel pipeline.
 *   small groups      an ordered launch group of at most small_max_ops
 *                     ops (LTPMG_SMALL_MAX_OPS at most) runs as one launch
 *                     of one workgroup -- sort, scan, cell moves and row
 *                     ops fused in LDS -- instead of the multi-kernel
 *                     ordered pipeline.
 *   use_graph         the plan's whole launch sequence is captured once as
 *                     a hipGraph and replayed by each run.
 *
 * merge_max_ops, fused_max_ops and small_max_ops are the batch-cut
 * parameters, and fused_tile_ops / fused_wgs_per_cu the tile parameters,
 * that ltpmg_stream_tune picks by timing the stream itself.  With every
 * batch cut alone, no graph and a host wait after each batch
 * (LTPMG_PLAN_SYNC_EACH), a run is the plain per-batch path with its inputs
 * already in HBM (ltpmg_execute's path when small_max_ops is
 * LTPMG_SMALL_MAX_OPS).
 *
 * Every result equals that of submitting the batches one by one to
 * ltpmg_submit_batch + ltpmg_execute in stream order, whatever the plan.
 * Running a stream applies its writes to the tables; a second run applies
 * them again to the state the first one left.
 */
#ifndef LTPMG_STREAM_H
#define LTPMG_STREAM_H

#include "ltpmg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ltpmg_stream ltpmg_stream;

/* Keep the read results of every op of the stream (n_ops x stride bytes of
 * HBM).  Without it only the commit flags are kept, and results land in a
 * buffer that each launch group overwrites. */
#define LTPMG_STREAM_KEEP_RESULTS 0x1u
/* Every run first truncates (ltpmg_truncate_table) each append table the
 * stream INSERTs into back to its row count when the stream was created, so
 * repeated runs (timing, ltpmg_stream_tune) append the same rows instead of
 * filling the table.  Rows below that count keep their presence. */
#define LTPMG_STREAM_REWIND 0x2u

/* Plan flags */
#define LTPMG_PLAN_SYNC_EACH 0x1u /* host waits after every batch (plain form) */

typedef struct {
    uint32_t merge_max_ops;    /* ordered mode: ops per merged launch group (0 = no merging) */
    uint32_t fused_max_ops;    /* lowest-TID-wins: fused kernel for batches with <= this many ops */
    uint32_t fused_tile_ops;   /* 1024, 2048 or 4096 */
    uint32_t fused_wgs_per_cu; /* persistent grid = CUs x this, clamped to occupancy */
    uint32_t use_graph;        /* capture the launch sequence once and replay it */
    uint32_t flags;            /* LTPMG_PLAN_* */
    uint32_t small_max_ops;    /* ordered mode: one-launch kernel for groups with <= this many ops
                                  (at most LTPMG_SMALL_MAX_OPS count) */
} ltpmg_stream_plan;

/* Uploads a stream of n_batches batches.  all holds them back to back:
 * all->txn_off runs over every txn of the stream (all->txn_off[0] = 0), and
 * batch b is the next batch_txns[b] txns, whose TIDs restart at 0.
 * modes[b] is LTPMG_MODE_A or LTPMG_MODE_O.  Every batch is checked as
 * ltpmg_submit_batch checks it; the stream's ops may exceed 2^29 in total
 * but no batch may. */
int ltpmg_stream_create(ltpmg_engine *e, uint32_t n_batches, const uint32_t *batch_txns,
                        const uint8_t *modes, const ltpmg_batch *all, uint32_t flags,
                        ltpmg_stream **out);
int ltpmg_stream_destroy(ltpmg_engine *e, ltpmg_stream *s);

/* Runs every batch of the stream once, in order, under plan (NULL: the plan
 * the last ltpmg_stream_tune picked, else ltpmg_stream_default_plan).
 * Waits for the device; device_ms (may be NULL) gets the device time from
 * the first launch to the last. */
int ltpmg_stream_run(ltpmg_engine *e, ltpmg_stream *s, const ltpmg_stream_plan *plan,
                     double *device_ms);

/* Times candidate plans on the stream itself (each trial is a run, so the
 * tables receive the stream's writes once per trial) and returns the
 * fastest; later ltpmg_stream_run(NULL) calls use it.  trials_out (may be
 * NULL) gets the number of stream runs spent. */
int ltpmg_stream_tune(ltpmg_engine *e, ltpmg_stream *s, ltpmg_stream_plan *best,
                      uint32_t *trials_out);

void ltpmg_stream_default_plan(ltpmg_stream_plan *plan);

/* Batch b of the last run: commit flags (batch_txns[b] bytes) and, with
 * LTPMG_STREAM_KEEP_RESULTS, read results (its ops x ltpmg_result_stride
 * bytes, indexed by op within the batch).  Either pointer may be NULL. */
int ltpmg_stream_fetch(ltpmg_engine *e, ltpmg_stream *s, uint32_t batch, uint8_t *commit,
                       void *results, uint64_t results_bytes);

/* Committed txns per batch of the last run (n_batches entries). */
int ltpmg_stream_committed(ltpmg_engine *e, ltpmg_stream *s, uint32_t *per_batch);

/* Launch groups of a plan (the batch cut) for a run that starts from the
 * current append cursors (the cursors at creation with LTPMG_STREAM_REWIND):
 * writes up to cap first-batch indices, returns the number of groups (or a
 * negative status). */
int ltpmg_stream_groups(ltpmg_engine *e, ltpmg_stream *s, const ltpmg_stream_plan *plan,
                        uint32_t *group_first, uint32_t cap);

#ifdef __cplusplus
}
#endif

#endif /* LTPMG_STREAM_H */
