/*
 * ltpmg_serve.h -- the serving form (libltpmg_host.so, C ABI).
 *
 * The reference runs one fixed batch per epoch from a host loop
 * (src/scheduler.cpp:44-59) and has no arrival process or latency.  This
 * server replays an open-loop request stream against an engine: arrival i
 * (Poisson times at arrival_rate) is transaction (i mod pool_txns) of the
 * pool, taken by value.  Queued transactions are cut into batches and run on
 * the device while later ones keep arriving:
 *
 *   serving form (min_batch == 0): cut as soon as a slot is free and either
 *     max_batch transactions wait or the oldest waited max_wait_us; up to
 *     n_inflight batches are on the device at once (their sort/scan stages
 *     overlap the previous batch's apply, include/ltpmg.h), and the next
 *     batch is cut while the device works;
 *   plain form (min_batch > 0, n_inflight == 1): wait for min_batch
 *     transactions (or the end of the stream), one batch at a time.
 *
 * Batches take effect in cut order.  In LTPMG_MODE_A an aborted transaction
 * goes back to the head of the queue and so gets a low TID in the next
 * batch; TID 0 always commits, so every transaction eventually commits.
 * Latency of a transaction = completion of the batch it commits in minus
 * its scheduled arrival time.
 *
 * The pool's op arrays are copied once into pinned memory before the clock
 * starts; a cut batch is uploaded by DMA straight from there as runs of
 * consecutive pool ops (ltpmg_submit_runs), so the host writes only the
 * batch's offsets.  Admission is by HBM: slot buffers are allocated under
 * the engine's budget before the clock starts, and with hbm_inflight_bytes
 * set a batch is cut only as large as fits that many device bytes beside
 * the batches already in flight (ops 14 B + a result slot + 16 B of sort
 * keys per op, 5 B per transaction), otherwise it waits for one to finish.
 */
#ifndef LTPMG_SERVE_H
#define LTPMG_SERVE_H

#include <stdint.h>

#include "ltpmg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double arrival_rate;   /* txn/s; 0 = every transaction arrives at t = 0 */
    uint32_t n_arrivals;   /* transactions offered */
    uint32_t max_batch;    /* most transactions per batch */
    uint32_t min_batch;    /* plain form: cut at this size (0 = serving form) */
    double max_wait_us;    /* serving form: cut when the oldest waited this long */
    uint32_t n_inflight;   /* batches on the device at once, <= engine slots */
    uint32_t mode;         /* LTPMG_MODE_O or LTPMG_MODE_A */
    uint64_t seed;         /* arrival process */
    uint64_t hbm_inflight_bytes; /* admission budget of the batches in flight (0 = slots only) */
} ltpmg_serve_cfg;

typedef struct {
    uint64_t offered;
    uint64_t committed;
    uint64_t aborted_attempts; /* LTPMG_MODE_A: aborts that were retried */
    uint64_t batches;
    double wall_s;           /* first arrival to last completion */
    double committed_per_s;
    double p50_ms, p90_ms, p99_ms, p999_ms, max_ms; /* commit latency */
    double mean_batch_txns;
    double host_pack_s;      /* host time cutting and submitting batches (offsets, DMA issue) */
    double device_ms;        /* sum of the batches' device times */
    uint64_t peak_inflight_bytes; /* most device bytes of batches in flight at once */
    uint64_t budget_holds;   /* cuts made smaller, or put off, by hbm_inflight_bytes */
} ltpmg_serve_stats;

/* Optional record of what ran, for replay ("same alone and in the mix"). */
typedef struct {
    uint32_t *batch_off;    /* [cap_batches + 1]: entries of batch k */
    uint32_t *batch_txns;   /* [cap_entries]: arrival ids in TID order */
    uint8_t *batch_commit;  /* [cap_entries]: commit flag of each entry */
    uint32_t cap_batches;
    uint32_t cap_entries;
    uint32_t n_batches;     /* written */
    uint32_t n_entries;     /* written */
    uint32_t overflow;      /* written: 1 if a cap was hit and recording stopped */
    uint8_t *results;       /* optional, needs n_arrivals <= pool txns: READ result of
                               pool op i at results + i * result_stride, written when
                               its transaction commits */
} ltpmg_serve_record;

/* Runs the stream to completion.  The pool is checked once up front. */
int ltpmg_serve_run(ltpmg_engine *e, const ltpmg_serve_cfg *cfg, const ltpmg_batch *pool,
                    ltpmg_serve_stats *out, ltpmg_serve_record *rec);

#ifdef __cplusplus
}
#endif
#endif
