// engine.hip -- implementation of the C ABI in include/ltpmg.h.
//
// One engine = one GPU.  It owns every device allocation (tables, lock
// words, per-slot batch buffers and scratch), one HIP stream per slot, and
// the pinned staging used to upload batches.  The reference's equivalents
// are GPUScheduler/GPUdatabase/GPUquery/GPUlauncher (cuda/src/gpuscheduler.cu,
// gpudatabase.cu, gpuquery.cu, gpulauncher.cu), which create and destroy
// streams per call, memset ~30 buffers per batch (gpuquery.cu:629-1157) and
// synchronise after every stage (gpulauncher.cu:798-1055).  Here per-batch
// state needs no clearing (epoch-tagged lock words, sizes carried in the
// launch), and the only host synchronisation is the caller's ltpmg_wait.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

using namespace ltpmg;

namespace {

uint32_t slot_bytes(uint32_t width) {
    if (width <= 4) return 4;
    if (width <= 8) return 8;
    if (width <= 16) return 16;
    return (width + 31u) & ~31u;
}

uint32_t ceil_log2(uint64_t x) {
    uint32_t b = 0;
    while ((1ull << b) < x) ++b;
    return b;
}

struct Stage {
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
};

struct Slot {
    hipStream_t stream = nullptr;
    uint32_t cap_txns = 0, cap_ops = 0, cap_res_stride = 0;
    uint32_t *d_txn_off = nullptr, *d_rows = nullptr, *d_tile_t0 = nullptr, *d_counters = nullptr;
    uint16_t *d_meta = nullptr;
    uint64_t *d_args = nullptr;
    uint8_t *d_commit = nullptr, *d_results = nullptr;
    OrderBufs ob{};
    uint8_t *h_stage = nullptr;
    size_t stage_bytes = 0;
    uint32_t n_txns = 0, n_ops = 0, mode = 0;
    uint32_t staged_txns = UINT32_MAX, staged_ops = 0;  // set by ltpmg_stage_batch
    bool has_batch = false, executed = false;
    hipEvent_t up0 = nullptr, up1 = nullptr, x0 = nullptr, x1 = nullptr;
    hipEvent_t done = nullptr;  // recorded after the batch's last table access
    std::vector<Stage> stages;
    uint32_t n_stages = 0;
};

struct Table {
    bool live = false;
    uint32_t n_rows = 0, n_cols = 0;
    uint32_t width[LTPMG_MAX_COLS] = {}, slot[LTPMG_MAX_COLS] = {};
    uint8_t *col[LTPMG_MAX_COLS] = {};
    unsigned long long *lock = nullptr;
};

}  // namespace

struct ltpmg_engine {
    int device = 0;
    uint32_t flags = 0;
    uint64_t budget = 0, allocated = 0;
    uint32_t default_txns = 0, default_ops = 0;
    uint32_t epoch = 0;
    int tail_slot = -1;  // slot of the most recently executed batch
    Table tables[LTPMG_MAX_TABLES];
    ColDesc *d_cols = nullptr;
    TabDesc *d_tabs = nullptr;
    unsigned long long *d_digest = nullptr;
    std::vector<Slot> slots;
    std::string err;
    std::mutex mu;
};

namespace {

int fail(ltpmg_engine *e, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (e) e->err = buf;
    return code;
}

#define HIP_TRY(e, call)                                                                 \
    do {                                                                                 \
        hipError_t _st = (call);                                                         \
        if (_st != hipSuccess)                                                           \
            return fail((e), LTPMG_EHIP, "%s failed: %s", #call, hipGetErrorString(_st)); \
    } while (0)

int dev_alloc(ltpmg_engine *e, void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return LTPMG_OK;
    if (e->allocated + bytes > e->budget)
        return fail(e, LTPMG_EBUDGET, "allocation of %zu bytes exceeds the HBM budget (%llu of %llu used)",
                    bytes, (unsigned long long)e->allocated, (unsigned long long)e->budget);
    hipError_t st = hipMalloc(p, bytes);
    if (st != hipSuccess) {
        *p = nullptr;
        return fail(e, LTPMG_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(st));
    }
    e->allocated += bytes;
    return LTPMG_OK;
}

template <class T>
int dev_alloc_n(ltpmg_engine *e, T **p, size_t n) {
    return dev_alloc(e, reinterpret_cast<void **>(p), n * sizeof(T));
}

void dev_free(ltpmg_engine *e, void *p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    e->allocated -= std::min<uint64_t>(e->allocated, bytes);
}

uint32_t max_slot_all(const ltpmg_engine *e) {
    uint32_t m = 4;
    for (const Table &t : e->tables)
        if (t.live)
            for (uint32_t c = 0; c < t.n_cols; ++c) m = std::max(m, t.slot[c]);
    return m;
}

uint32_t n_tabs_span(const ltpmg_engine *e) {
    uint32_t n = 0;
    for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t)
        if (e->tables[t].live) n = t + 1;
    return n;
}

// Writes the device descriptors and assigns ordered-mode global row bases.
int sync_descriptors(ltpmg_engine *e) {
    ColDesc cols[LTPMG_MAX_TABLES * LTPMG_MAX_COLS] = {};
    TabDesc tabs[LTPMG_MAX_TABLES] = {};
    uint64_t row_base = 0;
    for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t) {
        const Table &tb = e->tables[t];
        if (!tb.live) continue;
        tabs[t] = TabDesc{tb.lock, tb.n_rows, tb.n_cols, (uint32_t)row_base, 1u};
        row_base += tb.n_rows;
        for (uint32_t c = 0; c < tb.n_cols; ++c)
            cols[t * LTPMG_MAX_COLS + c] = ColDesc{tb.col[c], tb.slot[c], tb.width[c]};
    }
    if (row_base > 0xFFFFFFFFull) return fail(e, LTPMG_EINVAL, "more than 2^32 rows over all tables");
    HIP_TRY(e, hipMemcpy(e->d_cols, cols, sizeof cols, hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->d_tabs, tabs, sizeof tabs, hipMemcpyHostToDevice));
    return LTPMG_OK;
}

void free_slot_buffers(ltpmg_engine *e, Slot &s) {
    const size_t ops = s.cap_ops, txns = s.cap_txns;
    dev_free(e, s.d_txn_off, (txns + 1) * 4);
    dev_free(e, s.d_rows, ops * 4);
    dev_free(e, s.d_meta, ops * 2);
    dev_free(e, s.d_args, ops * 8);
    dev_free(e, s.d_commit, txns);
    dev_free(e, s.d_results, ops * (size_t)s.cap_res_stride);
    const size_t tiles = ops / kTileOps + 2, st = ops / kSortTile + 1, sg = seg_tiles((uint32_t)ops) + 1;
    dev_free(e, s.d_tile_t0, tiles * 4);
    dev_free(e, s.d_counters, 16);
    for (int i = 0; i < 2; ++i) {
        dev_free(e, s.ob.sort.keys[i], ops * 4);
        dev_free(e, s.ob.sort.vals[i], ops * 4);
    }
    dev_free(e, s.ob.sort.hist, 256 * st * 4);
    dev_free(e, s.ob.sort.partials, scan_partials_needed((uint32_t)(256 * st)) * 4);
    dev_free(e, s.ob.tile_flag, sg * 4 * 5);
    dev_free(e, s.ob.fin, ops * 12);
    dev_free(e, s.ob.counters, 16);
    if (s.h_stage) (void)hipHostFree(s.h_stage);
    s.d_txn_off = s.d_rows = s.d_tile_t0 = s.d_counters = nullptr;
    s.d_meta = nullptr;
    s.d_args = nullptr;
    s.d_commit = s.d_results = nullptr;
    s.ob = OrderBufs{};
    s.h_stage = nullptr;
    s.stage_bytes = 0;
    s.cap_txns = s.cap_ops = s.cap_res_stride = 0;
    s.has_batch = s.executed = false;
}

int ensure_slot_capacity(ltpmg_engine *e, Slot &s, uint32_t txns, uint32_t ops, uint32_t res_stride) {
    if (txns <= s.cap_txns && ops <= s.cap_ops && res_stride <= s.cap_res_stride && s.d_rows)
        return LTPMG_OK;
    HIP_TRY(e, hipStreamSynchronize(s.stream));
    free_slot_buffers(e, s);
    txns = std::max(txns, e->default_txns);
    ops = std::max(ops, e->default_ops);
    const size_t tiles = ops / kTileOps + 2, st = ops / kSortTile + 1, sg = seg_tiles(ops) + 1;
    int rc = LTPMG_OK;
#define ALLOC(ptr, n) if (rc == LTPMG_OK) rc = dev_alloc_n(e, &(ptr), (n))
    ALLOC(s.d_txn_off, (size_t)txns + 1);
    ALLOC(s.d_rows, ops);
    ALLOC(s.d_meta, ops);
    ALLOC(s.d_args, ops);
    ALLOC(s.d_commit, txns);
    ALLOC(s.d_results, (size_t)ops * res_stride);
    ALLOC(s.d_tile_t0, tiles);
    ALLOC(s.d_counters, 4);
    for (int i = 0; i < 2; ++i) {
        ALLOC(s.ob.sort.keys[i], ops);
        ALLOC(s.ob.sort.vals[i], ops);
    }
    ALLOC(s.ob.sort.hist, 256 * st);
    ALLOC(s.ob.sort.partials, scan_partials_needed((uint32_t)(256 * st)));
    ALLOC(s.ob.tile_flag, sg * 5);
    ALLOC(s.ob.fin, (size_t)ops * 3);
    ALLOC(s.ob.counters, 4);
#undef ALLOC
    s.cap_txns = txns;
    s.cap_ops = ops;
    s.cap_res_stride = res_stride;
    if (rc != LTPMG_OK) {
        free_slot_buffers(e, s);
        return rc;
    }
    s.ob.tile_set = s.ob.tile_flag + sg;
    s.ob.tile_delta = s.ob.tile_flag + 2 * sg;
    s.ob.carry_set = s.ob.tile_flag + 3 * sg;
    s.ob.carry_delta = s.ob.tile_flag + 4 * sg;
    s.stage_bytes = ((size_t)txns + 1) * 4 + (size_t)ops * (4 + 2 + 8) + 64;
    hipError_t hs = hipHostMalloc(reinterpret_cast<void **>(&s.h_stage), s.stage_bytes, hipHostMallocDefault);
    if (hs != hipSuccess) {
        s.h_stage = nullptr;
        free_slot_buffers(e, s);
        return fail(e, LTPMG_ENOMEM, "hipHostMalloc(%zu) failed: %s", s.stage_bytes, hipGetErrorString(hs));
    }
    return LTPMG_OK;
}

int validate_batch(ltpmg_engine *e, const ltpmg_batch *b) {
    if (!b || !b->txn_off || (b->n_ops && (!b->rows || !b->meta || !b->args)))
        return fail(e, LTPMG_EINVAL, "null batch array");
    if (b->n_ops >= (1u << 29)) return fail(e, LTPMG_EINVAL, "n_ops must be < 2^29");
    if (b->txn_off[0] != 0) return fail(e, LTPMG_EINVAL, "txn_off[0] must be 0");
    for (uint32_t t = 0; t < b->n_txns; ++t) {
        uint32_t len = b->txn_off[t + 1] - b->txn_off[t];
        if (b->txn_off[t + 1] < b->txn_off[t] || len < 1 || len > LTPMG_MAX_TXN_OPS)
            return fail(e, LTPMG_EINVAL, "txn %u has %d ops (allowed 1..%d)", t,
                        (int)(b->txn_off[t + 1] - b->txn_off[t]), LTPMG_MAX_TXN_OPS);
    }
    if (b->txn_off[b->n_txns] != b->n_ops)
        return fail(e, LTPMG_EINVAL, "txn_off[n_txns]=%u != n_ops=%u", b->txn_off[b->n_txns], b->n_ops);
    // per-op checks, split over host threads for large batches; the first
    // bad op (lowest index) is reported
    auto bad_op = [&](uint32_t i) -> bool {
        uint16_t m = b->meta[i];
        uint32_t kind = m & 3u, col = (m >> 2) & 63u;
        const Table &t = e->tables[(m >> 8) & 15u];
        return (m >> 12) || kind > LTPMG_OP_RMW_ADD || !t.live || col >= t.n_cols || b->rows[i] >= t.n_rows;
    };
    const uint32_t n = b->n_ops;
    const uint32_t n_threads = n < (1u << 20) ? 1u : std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<uint32_t> first(n_threads, UINT32_MAX);
    auto scan = [&](uint32_t w) {
        uint32_t a = (uint32_t)((uint64_t)n * w / n_threads), z = (uint32_t)((uint64_t)n * (w + 1) / n_threads);
        for (uint32_t i = a; i < z; ++i)
            if (bad_op(i)) {
                first[w] = i;
                return;
            }
    };
    if (n_threads == 1) {
        scan(0);
    } else {
        std::vector<std::thread> th;
        for (uint32_t w = 0; w < n_threads; ++w) th.emplace_back(scan, w);
        for (auto &t : th) t.join();
    }
    uint32_t i = *std::min_element(first.begin(), first.end());
    if (i != UINT32_MAX) {
        uint16_t m = b->meta[i];
        uint32_t kind = m & 3u, col = (m >> 2) & 63u, tab = (m >> 8) & 15u;
        const Table &t = e->tables[tab];
        if ((m >> 12) || kind > LTPMG_OP_RMW_ADD) return fail(e, LTPMG_EINVAL, "op %u: bad meta 0x%04x", i, m);
        if (!t.live) return fail(e, LTPMG_EINVAL, "op %u: table %u does not exist", i, tab);
        if (col >= t.n_cols) return fail(e, LTPMG_EINVAL, "op %u: column %u >= %u", i, col, t.n_cols);
        return fail(e, LTPMG_EINVAL, "op %u: row %u >= %u rows of table %u", i, b->rows[i], t.n_rows, tab);
    }
    return LTPMG_OK;
}

Slot *get_slot(ltpmg_engine *e, uint32_t slot) {
    if (slot >= e->slots.size()) return nullptr;
    return &e->slots[slot];
}

struct StageRec {
    ltpmg_engine *e;
    Slot &s;
    bool on;
    void begin(const char *name) {
        if (!on) return;
        if (s.n_stages == s.stages.size()) {
            Stage st;
            (void)hipEventCreate(&st.a);
            (void)hipEventCreate(&st.b);
            s.stages.push_back(st);
        }
        s.stages[s.n_stages].name = name;
        (void)hipEventRecord(s.stages[s.n_stages].a, s.stream);
    }
    void end() {
        if (!on) return;
        (void)hipEventRecord(s.stages[s.n_stages].b, s.stream);
        ++s.n_stages;
    }
};

}  // namespace

extern "C" {

int ltpmg_open(const ltpmg_cfg *cfg, ltpmg_engine **out) {
    if (!cfg || !out) return LTPMG_EINVAL;
    *out = nullptr;
    if (cfg->n_slots < 1 || cfg->n_slots > LTPMG_MAX_SLOTS) return LTPMG_EINVAL;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0 || cfg->device < 0 ||
        cfg->device >= n_dev)
        return LTPMG_EDEVICE;
    if (hipSetDevice(cfg->device) != hipSuccess) return LTPMG_EDEVICE;
    ltpmg_engine *e = new ltpmg_engine();
    e->device = cfg->device;
    e->flags = cfg->flags;
    e->default_txns = cfg->max_txns;
    e->default_ops = cfg->max_ops;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        delete e;
        return LTPMG_EDEVICE;
    }
    e->budget = cfg->hbm_budget ? cfg->hbm_budget : (uint64_t)(free_b * 0.9);
    int rc = dev_alloc_n(e, &e->d_cols, LTPMG_MAX_TABLES * LTPMG_MAX_COLS);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &e->d_tabs, LTPMG_MAX_TABLES);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &e->d_digest, 1);
    if (rc == LTPMG_OK) rc = sync_descriptors(e);
    e->slots.resize(cfg->n_slots);
    for (Slot &s : e->slots) {
        if (rc != LTPMG_OK) break;
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&s.up0) != hipSuccess || hipEventCreate(&s.up1) != hipSuccess ||
            hipEventCreate(&s.x0) != hipSuccess || hipEventCreate(&s.x1) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
            rc = LTPMG_EHIP;
    }
    if (rc != LTPMG_OK) {
        ltpmg_close(e);
        return rc;
    }
    *out = e;
    return LTPMG_OK;
}

int ltpmg_close(ltpmg_engine *e) {
    if (!e) return LTPMG_EINVAL;
    (void)hipSetDevice(e->device);
    for (Slot &s : e->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        free_slot_buffers(e, s);
        for (Stage &st : s.stages) {
            (void)hipEventDestroy(st.a);
            (void)hipEventDestroy(st.b);
        }
        for (hipEvent_t ev : {s.up0, s.up1, s.x0, s.x1, s.done})
            if (ev) (void)hipEventDestroy(ev);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    for (Table &t : e->tables) {
        for (uint32_t c = 0; c < t.n_cols; ++c) (void)hipFree(t.col[c]);
        if (t.lock) (void)hipFree(t.lock);
    }
    if (e->d_cols) (void)hipFree(e->d_cols);
    if (e->d_tabs) (void)hipFree(e->d_tabs);
    if (e->d_digest) (void)hipFree(e->d_digest);
    delete e;
    return LTPMG_OK;
}

const char *ltpmg_last_error(const ltpmg_engine *e) { return e ? e->err.c_str() : "no engine"; }

uint32_t ltpmg_result_stride(const ltpmg_engine *e) { return e ? max_slot_all(e) : 0; }

int ltpmg_create_table(ltpmg_engine *e, uint32_t table_id, const ltpmg_table_desc *d) {
    if (!e || !d) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (table_id >= LTPMG_MAX_TABLES) return fail(e, LTPMG_EINVAL, "table id %u >= %d", table_id, LTPMG_MAX_TABLES);
    if (e->tables[table_id].live) return fail(e, LTPMG_ESTATE, "table %u already exists", table_id);
    if (d->n_cols < 1 || d->n_cols > LTPMG_MAX_COLS) return fail(e, LTPMG_EINVAL, "n_cols %u", d->n_cols);
    if (d->n_rows < 1) return fail(e, LTPMG_EINVAL, "n_rows must be >= 1");
    for (uint32_t c = 0; c < d->n_cols; ++c)
        if (d->col_bytes[c] < 4 || d->col_bytes[c] > LTPMG_MAX_COL_BYTES || d->col_bytes[c] % 4)
            return fail(e, LTPMG_EINVAL, "column %u: width %u is not a multiple of 4 in 4..%d", c,
                        d->col_bytes[c], LTPMG_MAX_COL_BYTES);
    for (Slot &s : e->slots) HIP_TRY(e, hipStreamSynchronize(s.stream));
    Table t;
    t.n_rows = d->n_rows;
    t.n_cols = d->n_cols;
    int rc = LTPMG_OK;
    for (uint32_t c = 0; c < d->n_cols && rc == LTPMG_OK; ++c) {
        t.width[c] = d->col_bytes[c];
        t.slot[c] = slot_bytes(d->col_bytes[c]);
        rc = dev_alloc(e, reinterpret_cast<void **>(&t.col[c]), (size_t)t.slot[c] * t.n_rows);
        if (rc == LTPMG_OK && hipMemset(t.col[c], 0, (size_t)t.slot[c] * t.n_rows) != hipSuccess)
            rc = fail(e, LTPMG_EHIP, "hipMemset of column %u failed", c);
    }
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &t.lock, t.n_rows);
    if (rc == LTPMG_OK && hipMemset(t.lock, 0, sizeof(unsigned long long) * t.n_rows) != hipSuccess)
        rc = fail(e, LTPMG_EHIP, "hipMemset of lock words failed");
    if (rc != LTPMG_OK) {
        for (uint32_t c = 0; c < t.n_cols; ++c) dev_free(e, t.col[c], (size_t)t.slot[c] * t.n_rows);
        dev_free(e, t.lock, 8ull * t.n_rows);
        return rc;
    }
    t.live = true;
    e->tables[table_id] = t;
    return sync_descriptors(e);
}

int ltpmg_fill_table(ltpmg_engine *e, uint32_t table_id, uint64_t seed) {
    if (!e) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (table_id >= LTPMG_MAX_TABLES || !e->tables[table_id].live)
        return fail(e, LTPMG_ESTATE, "table %u does not exist", table_id);
    const Table &t = e->tables[table_id];
    hipStream_t s = e->slots[0].stream;
    for (uint32_t c = 0; c < t.n_cols; ++c)
        launch_fill_column(s, ColDesc{t.col[c], t.slot[c], t.width[c]}, table_id, c, t.n_rows, seed);
    HIP_TRY(e, hipGetLastError());
    HIP_TRY(e, hipStreamSynchronize(s));
    return LTPMG_OK;
}

static int column_io(ltpmg_engine *e, uint32_t table_id, uint32_t col, uint32_t row0,
                     uint32_t n_rows, void *host, bool to_device) {
    if (!e || (!host && n_rows)) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (table_id >= LTPMG_MAX_TABLES || !e->tables[table_id].live)
        return fail(e, LTPMG_ESTATE, "table %u does not exist", table_id);
    const Table &t = e->tables[table_id];
    if (col >= t.n_cols) return fail(e, LTPMG_EINVAL, "column %u >= %u", col, t.n_cols);
    if ((uint64_t)row0 + n_rows > t.n_rows)
        return fail(e, LTPMG_EINVAL, "rows [%u, %llu) outside the table", row0,
                    (unsigned long long)row0 + n_rows);
    if (n_rows == 0) return LTPMG_OK;
    for (Slot &s : e->slots) HIP_TRY(e, hipStreamSynchronize(s.stream));
    uint8_t *dev = t.col[col] + (size_t)row0 * t.slot[col];
    if (to_device)
        HIP_TRY(e, hipMemcpy2D(dev, t.slot[col], host, t.width[col], t.width[col], n_rows,
                               hipMemcpyHostToDevice));
    else
        HIP_TRY(e, hipMemcpy2D(host, t.width[col], dev, t.slot[col], t.width[col], n_rows,
                               hipMemcpyDeviceToHost));
    return LTPMG_OK;
}

int ltpmg_load_column(ltpmg_engine *e, uint32_t table_id, uint32_t col, uint32_t row0,
                      uint32_t n_rows, const void *src) {
    return column_io(e, table_id, col, row0, n_rows, const_cast<void *>(src), true);
}

int ltpmg_read_column(ltpmg_engine *e, uint32_t table_id, uint32_t col, uint32_t row0,
                      uint32_t n_rows, void *dst) {
    return column_io(e, table_id, col, row0, n_rows, dst, false);
}

int ltpmg_table_digest(ltpmg_engine *e, uint32_t table_id, uint64_t *out) {
    if (!e || !out) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (table_id >= LTPMG_MAX_TABLES || !e->tables[table_id].live)
        return fail(e, LTPMG_ESTATE, "table %u does not exist", table_id);
    for (Slot &s : e->slots) HIP_TRY(e, hipStreamSynchronize(s.stream));
    const Table &t = e->tables[table_id];
    hipStream_t s = e->slots[0].stream;
    HIP_TRY(e, hipMemsetAsync(e->d_digest, 0, sizeof(unsigned long long), s));
    for (uint32_t c = 0; c < t.n_cols; ++c)
        launch_digest_column(s, ColDesc{t.col[c], t.slot[c], t.width[c]}, table_id, c, t.n_rows,
                             e->d_digest);
    HIP_TRY(e, hipGetLastError());
    unsigned long long v = 0;
    HIP_TRY(e, hipMemcpyAsync(&v, e->d_digest, sizeof v, hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipStreamSynchronize(s));
    *out = v;
    return LTPMG_OK;
}

// Pinned staging layout of a slot: offsets | pad to 8 | args | rows | meta
// (keeps args 8-byte aligned).
static ltpmg_staging stage_ptrs(Slot &s, uint32_t n_txns, uint32_t n_ops) {
    const size_t off_b = ((size_t)n_txns + 1) * 4;
    uint8_t *p_args = s.h_stage + ((off_b + 7) & ~(size_t)7);
    uint8_t *p_rows = p_args + (size_t)n_ops * 8;
    uint8_t *p_meta = p_rows + (size_t)n_ops * 4;
    return ltpmg_staging{reinterpret_cast<uint32_t *>(s.h_stage), reinterpret_cast<uint32_t *>(p_rows),
                         reinterpret_cast<uint16_t *>(p_meta), reinterpret_cast<uint64_t *>(p_args)};
}

static int stage_locked(ltpmg_engine *e, Slot &s, uint32_t n_txns, uint32_t n_ops) {
    if (n_ops >= (1u << 29)) return fail(e, LTPMG_EINVAL, "n_ops must be < 2^29");
    int rc = ensure_slot_capacity(e, s, n_txns, n_ops, max_slot_all(e));
    if (rc != LTPMG_OK) return rc;
    // the staging buffer may still feed the slot's previous upload
    HIP_TRY(e, hipStreamSynchronize(s.stream));
    s.staged_txns = n_txns;
    s.staged_ops = n_ops;
    s.has_batch = s.executed = false;
    return LTPMG_OK;
}

static int upload_locked(ltpmg_engine *e, Slot &s, bool trusted) {
    const uint32_t n_txns = s.staged_txns, n_ops = s.staged_ops;
    ltpmg_staging st = stage_ptrs(s, n_txns, n_ops);
    if (!trusted) {
        ltpmg_batch b{n_txns, n_ops, st.txn_off, st.rows, st.meta, st.args};
        int rc = validate_batch(e, &b);
        if (rc != LTPMG_OK) return rc;
    }
    HIP_TRY(e, hipEventRecord(s.up0, s.stream));
    HIP_TRY(e, hipMemcpyAsync(s.d_txn_off, st.txn_off, ((size_t)n_txns + 1) * 4, hipMemcpyHostToDevice, s.stream));
    if (n_ops) {
        HIP_TRY(e, hipMemcpyAsync(s.d_args, st.args, (size_t)n_ops * 8, hipMemcpyHostToDevice, s.stream));
        HIP_TRY(e, hipMemcpyAsync(s.d_rows, st.rows, (size_t)n_ops * 4, hipMemcpyHostToDevice, s.stream));
        HIP_TRY(e, hipMemcpyAsync(s.d_meta, st.meta, (size_t)n_ops * 2, hipMemcpyHostToDevice, s.stream));
    }
    HIP_TRY(e, hipEventRecord(s.up1, s.stream));
    s.n_txns = n_txns;
    s.n_ops = n_ops;
    s.has_batch = true;
    s.executed = false;
    return LTPMG_OK;
}

int ltpmg_submit_batch(ltpmg_engine *e, uint32_t slot, const ltpmg_batch *b) {
    if (!e || !b) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    Slot *s = get_slot(e, slot);
    if (!s) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    int rc = validate_batch(e, b);
    if (rc != LTPMG_OK) return rc;
    rc = stage_locked(e, *s, b->n_txns, b->n_ops);
    if (rc != LTPMG_OK) return rc;
    ltpmg_staging st = stage_ptrs(*s, b->n_txns, b->n_ops);
    std::memcpy(st.txn_off, b->txn_off, ((size_t)b->n_txns + 1) * 4);
    if (b->n_ops) {
        std::memcpy(st.args, b->args, (size_t)b->n_ops * 8);
        std::memcpy(st.rows, b->rows, (size_t)b->n_ops * 4);
        std::memcpy(st.meta, b->meta, (size_t)b->n_ops * 2);
    }
    return upload_locked(e, *s, true);
}

int ltpmg_check_batch(ltpmg_engine *e, const ltpmg_batch *b) {
    if (!e || !b) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    return validate_batch(e, b);
}

int ltpmg_stage_batch(ltpmg_engine *e, uint32_t slot, uint32_t n_txns, uint32_t n_ops,
                      ltpmg_staging *out) {
    if (!e || !out) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    Slot *s = get_slot(e, slot);
    if (!s) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    int rc = stage_locked(e, *s, n_txns, n_ops);
    if (rc != LTPMG_OK) return rc;
    *out = stage_ptrs(*s, n_txns, n_ops);
    return LTPMG_OK;
}

int ltpmg_submit_staged(ltpmg_engine *e, uint32_t slot, uint32_t flags) {
    if (!e) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    Slot *s = get_slot(e, slot);
    if (!s) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    if (s->staged_txns == UINT32_MAX) return fail(e, LTPMG_ESTATE, "nothing staged in slot %u", slot);
    int rc = upload_locked(e, *s, (flags & LTPMG_SUBMIT_TRUSTED) != 0);
    s->staged_txns = UINT32_MAX;
    return rc;
}

int ltpmg_execute(ltpmg_engine *e, uint32_t slot, uint32_t mode) {
    if (!e) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    Slot *sp = get_slot(e, slot);
    if (!sp) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    Slot &s = *sp;
    if (!s.has_batch) return fail(e, LTPMG_ESTATE, "no batch submitted to slot %u", slot);
    const bool profile = (e->flags & LTPMG_CFG_PROFILE) || (mode & LTPMG_EXEC_PROFILE);
    mode &= ~LTPMG_EXEC_PROFILE;
    if (mode != LTPMG_MODE_A && mode != LTPMG_MODE_O) return fail(e, LTPMG_EINVAL, "mode %u", mode);
    const uint32_t rs = max_slot_all(e);
    if (rs > s.cap_res_stride) return fail(e, LTPMG_ESTATE, "tables changed since the batch was submitted");
    DevTables dt{e->d_cols, e->d_tabs};
    OpView ov{s.d_txn_off, s.d_rows, s.d_meta, s.d_args, s.n_txns, s.n_ops};
    StageRec rec{e, s, profile};
    s.n_stages = 0;
    s.mode = mode;
    HIP_TRY(e, hipEventRecord(s.x0, s.stream));
    if (mode == LTPMG_MODE_A) {
        const uint32_t epoch = ++e->epoch;
        if (epoch == 0) return fail(e, LTPMG_ESTATE, "epoch counter wrapped");
        const uint32_t n_tiles = std::max(1u, (s.n_ops + kTileOps - 1) / kTileOps);
        // lock words and table rows: after the previous batch, whatever its slot
        if (e->tail_slot >= 0 && e->tail_slot != (int)slot)
            HIP_TRY(e, hipStreamWaitEvent(s.stream, e->slots[e->tail_slot].done, 0));
        rec.begin("tile_bounds");
        launch_zero_u32(s.stream, s.d_counters, 4);
        launch_tile_bounds(s.stream, s.d_txn_off, s.n_txns, n_tiles, s.d_tile_t0);
        rec.end();
        rec.begin("cc_reserve");
        launch_cc_reserve(s.stream, dt, ov, s.d_tile_t0, n_tiles, epoch);
        rec.end();
        rec.begin("cc_validate_read");
        launch_cc_validate_read(s.stream, dt, ov, s.d_tile_t0, n_tiles, epoch, s.d_commit,
                                s.d_results, rs, s.d_counters);
        rec.end();
        rec.begin("cc_write");
        launch_cc_write(s.stream, dt, ov, s.d_tile_t0, n_tiles, s.d_commit, rs);
        rec.end();
    } else {
        uint64_t total_rows = 0;
        uint32_t max_cols = 1;
        for (const Table &t : e->tables)
            if (t.live) {
                total_rows += t.n_rows;
                max_cols = std::max(max_cols, t.n_cols);
            }
        const uint32_t col_bits = ceil_log2(max_cols), key_bits = ceil_log2(total_rows) + col_bits;
        if (key_bits > 31)
            return fail(e, LTPMG_EINVAL, "ordered mode needs <= 31 key bits (rows x columns), got %u", key_bits);
        const uint32_t n_tabs = n_tabs_span(e);
        rec.begin("mo_keys");
        launch_mo_keys(s.stream, dt, ov, col_bits, s.ob.sort.keys[0], s.ob.sort.vals[0]);
        rec.end();
        int sorted = 0;
        if (s.n_ops > 1) {
            for (uint32_t shift = 0; shift < key_bits; shift += 8) {
                rec.begin("rs_hist");
                launch_rs_hist(s.stream, s.ob.sort, sorted, s.n_ops, shift);
                rec.end();
                rec.begin("rs_scan");
                launch_rs_scan(s.stream, s.ob.sort, s.n_ops);
                rec.end();
                rec.begin("rs_scatter");
                launch_rs_scatter(s.stream, s.ob.sort, sorted, s.n_ops, shift);
                rec.end();
                sorted ^= 1;
            }
        }
        rec.begin("mo_seg_scan");
        launch_mo_seg_scan(s.stream, ov, s.ob, sorted);
        rec.end();
        // keys, sort and scan only touch this slot's buffers, so they overlap
        // the previous batch; reading and writing cells waits for it
        if (e->tail_slot >= 0 && e->tail_slot != (int)slot)
            HIP_TRY(e, hipStreamWaitEvent(s.stream, e->slots[e->tail_slot].done, 0));
        rec.begin("mo_apply_reads");
        launch_mo_apply_reads(s.stream, dt, ov, s.ob, sorted, col_bits, n_tabs, s.d_results, rs);
        rec.end();
        rec.begin("mo_final_writes");
        launch_mo_final_writes(s.stream, dt, ov, s.ob, col_bits, n_tabs, rs);
        launch_fill_u8(s.stream, s.d_commit, s.n_txns, 1);
        rec.end();
    }
    HIP_TRY(e, hipGetLastError());
    HIP_TRY(e, hipEventRecord(s.done, s.stream));
    e->tail_slot = (int)slot;
    HIP_TRY(e, hipEventRecord(s.x1, s.stream));
    s.executed = true;
    return LTPMG_OK;
}

int ltpmg_query(ltpmg_engine *e, uint32_t slot) {
    if (!e) return LTPMG_EINVAL;
    Slot *s = get_slot(e, slot);
    if (!s) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    hipError_t st = hipStreamQuery(s->stream);
    if (st == hipSuccess) return LTPMG_OK;
    if (st == hipErrorNotReady) return 1;
    return fail(e, LTPMG_EHIP, "hipStreamQuery failed: %s", hipGetErrorString(st));
}

int ltpmg_wait(ltpmg_engine *e, uint32_t slot) {
    if (!e) return LTPMG_EINVAL;
    HIP_TRY(e, hipSetDevice(e->device));
    Slot *s = get_slot(e, slot);
    if (!s) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    HIP_TRY(e, hipStreamSynchronize(s->stream));
    return LTPMG_OK;
}

int ltpmg_fetch_results(ltpmg_engine *e, uint32_t slot, uint8_t *commit, void *results,
                        uint64_t results_bytes) {
    if (!e) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    Slot *s = get_slot(e, slot);
    if (!s) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    if (!s->executed) return fail(e, LTPMG_ESTATE, "slot %u has not executed a batch", slot);
    HIP_TRY(e, hipStreamSynchronize(s->stream));
    if (commit && s->n_txns)
        HIP_TRY(e, hipMemcpy(commit, s->d_commit, s->n_txns, hipMemcpyDeviceToHost));
    if (results) {
        uint64_t need = (uint64_t)s->n_ops * max_slot_all(e);
        if (results_bytes < need)
            return fail(e, LTPMG_EINVAL, "results buffer holds %llu bytes, %llu needed",
                        (unsigned long long)results_bytes, (unsigned long long)need);
        if (need) HIP_TRY(e, hipMemcpy(results, s->d_results, need, hipMemcpyDeviceToHost));
    }
    return LTPMG_OK;
}

int ltpmg_stats(ltpmg_engine *e, uint32_t slot, ltpmg_batch_stats *out) {
    if (!e || !out) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    Slot *s = get_slot(e, slot);
    if (!s) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    std::memset(out, 0, sizeof *out);
    out->hbm_allocated = e->allocated;
    if (!s->executed) return fail(e, LTPMG_ESTATE, "slot %u has not executed a batch", slot);
    HIP_TRY(e, hipStreamSynchronize(s->stream));
    out->n_txns = s->n_txns;
    out->n_ops = s->n_ops;
    out->mode = s->mode;
    if (s->mode == LTPMG_MODE_A) {
        uint32_t c = 0;
        HIP_TRY(e, hipMemcpy(&c, s->d_counters, 4, hipMemcpyDeviceToHost));
        out->committed = c;
    } else {
        out->committed = s->n_txns;
    }
    out->aborted = s->n_txns - out->committed;
    HIP_TRY(e, hipEventElapsedTime(&out->exec_ms, s->x0, s->x1));
    HIP_TRY(e, hipEventElapsedTime(&out->upload_ms, s->up0, s->up1));
    out->n_kernels = std::min<uint32_t>(s->n_stages, LTPMG_MAX_KSTATS);
    for (uint32_t i = 0; i < out->n_kernels; ++i) {
        std::snprintf(out->kernel_name[i], sizeof out->kernel_name[i], "%s", s->stages[i].name);
        HIP_TRY(e, hipEventElapsedTime(&out->kernel_ms[i], s->stages[i].a, s->stages[i].b));
    }
    return LTPMG_OK;
}

int ltpmg_selftest_sort(ltpmg_engine *e, uint32_t *keys, uint32_t *vals, uint32_t n,
                        uint32_t key_bits) {
    if (!e || (n && (!keys || !vals)) || key_bits > 32) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    hipStream_t st = e->slots[0].stream;
    SortBufs b{};
    const size_t tiles = n / kSortTile + 1;
    int rc = LTPMG_OK;
    for (int i = 0; i < 2 && rc == LTPMG_OK; ++i) {
        rc = dev_alloc_n(e, &b.keys[i], std::max<size_t>(n, 1));
        if (rc == LTPMG_OK) rc = dev_alloc_n(e, &b.vals[i], std::max<size_t>(n, 1));
    }
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &b.hist, 256 * tiles);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &b.partials, scan_partials_needed((uint32_t)(256 * tiles)));
    if (rc == LTPMG_OK && n) {
        int cur = 0;
        if (hipMemcpy(b.keys[0], keys, 4ull * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b.vals[0], vals, 4ull * n, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(e, LTPMG_EHIP, "copy to device failed");
        if (rc == LTPMG_OK) cur = launch_radix_sort(st, b, n, key_bits);
        if (rc == LTPMG_OK && (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess))
            rc = fail(e, LTPMG_EHIP, "sort kernels failed");
        if (rc == LTPMG_OK &&
            (hipMemcpy(keys, b.keys[cur], 4ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(vals, b.vals[cur], 4ull * n, hipMemcpyDeviceToHost) != hipSuccess))
            rc = fail(e, LTPMG_EHIP, "copy to host failed");
    }
    for (int i = 0; i < 2; ++i) {
        dev_free(e, b.keys[i], 4ull * std::max<size_t>(n, 1));
        dev_free(e, b.vals[i], 4ull * std::max<size_t>(n, 1));
    }
    dev_free(e, b.hist, 1024 * tiles);
    dev_free(e, b.partials, 4 * scan_partials_needed((uint32_t)(256 * tiles)));
    return rc;
}

int ltpmg_selftest_segscan(ltpmg_engine *e, const uint32_t *vals, const uint8_t *head,
                           uint32_t *out, uint32_t n) {
    if (!e || (n && (!vals || !head || !out))) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    hipStream_t st = e->slots[0].stream;
    uint32_t *d_vals = nullptr, *d_out = nullptr, *d_scr = nullptr;
    uint8_t *d_head = nullptr;
    const size_t nn = std::max<size_t>(n, 1), scr = 5ull * (seg_tiles(n) + 1);
    int rc = dev_alloc_n(e, &d_vals, nn);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &d_out, nn);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &d_head, nn);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &d_scr, scr);
    if (rc == LTPMG_OK && n) {
        if (hipMemcpy(d_vals, vals, 4ull * n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_head, head, n, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(e, LTPMG_EHIP, "copy to device failed");
        if (rc == LTPMG_OK) launch_selftest_segscan(st, d_vals, d_head, d_out, n, d_scr);
        if (rc == LTPMG_OK && (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess))
            rc = fail(e, LTPMG_EHIP, "segscan kernels failed");
        if (rc == LTPMG_OK && hipMemcpy(out, d_out, 4ull * n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(e, LTPMG_EHIP, "copy to host failed");
    }
    dev_free(e, d_vals, 4 * nn);
    dev_free(e, d_out, 4 * nn);
    dev_free(e, d_head, nn);
    dev_free(e, d_scr, 4 * scr);
    return rc;
}

}  // extern "C"

extern "C" int ltpmg_selftest_copy(ltpmg_engine *e, uint64_t n_bytes, uint32_t iters, double *gbps) {
    if (!e || !gbps || iters == 0 || n_bytes < 16) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    hipStream_t st = e->slots[0].stream;
    const uint64_t n16 = n_bytes / 16, bytes = n16 * 16;
    void *a = nullptr, *b = nullptr;
    int rc = dev_alloc(e, &a, bytes);
    if (rc == LTPMG_OK) rc = dev_alloc(e, &b, bytes);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (rc == LTPMG_OK && (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess))
        rc = fail(e, LTPMG_EHIP, "hipEventCreate failed");
    if (rc == LTPMG_OK) {
        launch_zero_u32(st, static_cast<uint32_t *>(a), (uint32_t)std::min<uint64_t>(bytes / 4, 0xFFFFFFFFu));
        launch_stream_copy(st, a, b, n16);  // warm-up
        (void)hipEventRecord(t0, st);
        for (uint32_t i = 0; i < iters; ++i) launch_stream_copy(st, a, b, n16);
        (void)hipEventRecord(t1, st);
        float ms = 0.f;
        if (hipStreamSynchronize(st) != hipSuccess || hipEventElapsedTime(&ms, t0, t1) != hipSuccess)
            rc = fail(e, LTPMG_EHIP, "copy kernels failed");
        else
            *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    dev_free(e, a, bytes);
    dev_free(e, b, bytes);
    return rc;
}
