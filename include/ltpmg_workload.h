/*
 * ltpmg_workload.h -- seeded workload generators (libltpmg_host.so, C ABI).
 *
 * The reference generates YCSB batches with rand() from several threads that
 * share one Random object, seeds with time(), and shuffles with a
 * time-seeded engine (src/txn_generator.cpp:602-673, src/random.cpp:6,41,
 * src/query.cpp:282-283), so no two runs agree.  These generators are
 * counter based instead: transaction t draws from its own SplitMix64 stream
 * started at mix64(seed + (t+1) * 0x9E3779B97F4A7C15), so the output depends
 * only on the spec, never on the thread count.
 *
 * Keys: uniform (theta == 0), or Zipf(theta) by Gray et al., "Quickly
 * generating billion-record synthetic databases" (SIGMOD 1994) -- the method
 * of YCSB's ZipfianGenerator -- with key 0 the most popular, matching the
 * reference's "rows below 100 are hot" convention (src/param.cpp:105).  It
 * replaces Random::rand_zipf (src/random.cpp:105-161), which only shapes the
 * first 100 keys.  Keys are distinct within a txn: a repeated key is drawn
 * again, as txn_generator.cpp:627-641 does.
 *
 * Ops: each logical op is a READ with probability read_pct/100, else a write
 * of write_kind on one uniformly drawn column.  With read_all_cols a READ
 * reads the whole row and is emitted as n_cols single-cell READs, in column
 * order, so the row's result is n_cols consecutive result slots.  UPDATE args
 * are fresh 64-bit payload seeds; RMW_ADD args are rmw_delta.
 */
#ifndef LTPMG_WORKLOAD_H
#define LTPMG_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t n_txns;
    uint32_t ops_min;       /* logical ops per txn, uniform in [ops_min, ops_max] */
    uint32_t ops_max;
    uint32_t n_rows;        /* keys in [0, n_rows) */
    double zipf_theta;      /* 0 = uniform; else in (0, 1) */
    uint32_t n_cols;        /* columns to draw from */
    uint32_t read_pct;      /* 0..100 */
    uint32_t write_kind;    /* LTPMG_OP_UPDATE or LTPMG_OP_RMW_ADD */
    uint32_t read_all_cols; /* 1: READ = whole row */
    uint32_t table_id;
    int32_t rmw_delta;
    uint64_t seed;
} ltpmg_kv_spec;

/* Upper bound on the device ops the spec can produce. */
uint64_t ltpmg_gen_kv_max_ops(const ltpmg_kv_spec *spec);

/* Fills txn_off[n_txns+1], rows/meta/args[<= max_ops]; *n_ops_out gets the
 * op count.  n_threads = 0 uses every hardware thread.  Returns 0 or a
 * negative LTPMG_E* status. */
int ltpmg_gen_kv(const ltpmg_kv_spec *spec, uint32_t *txn_off, uint32_t *rows, uint16_t *meta,
                 uint64_t *args, uint64_t max_ops, uint32_t *n_ops_out, uint32_t n_threads);

/* zeta(n, theta) = sum_{i=1..n} i^-theta (the Zipf normaliser). */
double ltpmg_zeta(uint64_t n, double theta);

#ifdef __cplusplus
}
#endif
#endif
