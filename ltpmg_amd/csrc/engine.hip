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
#include <array>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ltpmg_stream.h"
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
    // row-op buffers (allocated on the first batch that touches an append
    // table, for cap_ops ops) and what the current batch does with rows
    uint8_t *d_found = nullptr;
    uint16_t *d_meta_eff = nullptr;
    uint32_t *d_ins_cnt = nullptr, *d_ins_base = nullptr;
    bool has_row_bufs = false;
    RowBufs rb{};
    uint32_t inserts[LTPMG_MAX_TABLES] = {};  // INSERT ops of the batch per table
    uint8_t *h_stage = nullptr;
    uint8_t *stage_dev = nullptr;  // h_stage as the device sees it (null: not mapped)
    // device views of the last pool arrays (rows, meta, args) uploaded from
    // (null: not pinned host memory)
    const void *pool_host[3] = {}, *pool_dev[3] = {};
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
    // append tables: presence bytes, and an upper bound on the device cursor
    // (the cursor itself is d_cursor[table]: aborted INSERTs take no row, so
    // the host only knows a bound without reading it back)
    uint8_t *present = nullptr;
    uint64_t cursor_bound = 0;
};

}  // namespace

struct ltpmg_engine {
    int device = 0;
    uint32_t flags = 0;
    uint64_t budget = 0, allocated = 0;
    uint32_t default_txns = 0, default_ops = 0;
    int n_cu = 0;
    // device words: [0] lowest-TID-wins epoch (moved on by each such batch),
    // [1] grid-barrier count, [2] barrier generation, [3] barrier timeout flag
    uint32_t *d_sync = nullptr;
    int tail_slot = -1;  // slot of the most recently executed batch
    Table tables[LTPMG_MAX_TABLES];
    ColDesc *d_cols = nullptr;  // the sparse descriptor table, then the column list (kColList)
    uint32_t n_clist = 0;
    KeyDecode kdec{};  // ordered-mode key decode of the live tables
    uint32_t wide_lanes = 8;  // DevTables::wide_lanes
    TabDesc *d_tabs = nullptr;
    unsigned long long *d_digest = nullptr;
    uint8_t *static_scratch = nullptr;  // this device's copy of `scratch` (ltpmg_execute_static)
    uint32_t *d_cursor = nullptr;       // [LTPMG_MAX_TABLES] append cursors
    uint32_t n_append = 0;              // live append tables
    std::vector<Slot> slots;
    std::string err;
    std::mutex mu;
};

namespace {

// Scratch of ltpmg_execute_static: the sort, scan and tile buffers of one
// batch of up to LTPMG_STATIC_MAX_OPS ops (about 4.7 MB of the 5 MiB below
// at the default sort tile).  A file-scope device array, so each GPU holds
// one copy from module load to process end and the entry point never
// allocates.  Host code must not pass `scratch` itself to a kernel (in host
// code the name is a host-side stand-in, not a device address): the device
// address of each GPU's copy is looked up once per device in ltpmg_open and
// kept in g_static_ptr / ltpmg_engine::static_scratch.
constexpr size_t kStaticScratchBytes = 5u << 20;
__device__ __align__(256) uint8_t scratch[kStaticScratchBytes];
constexpr int kStaticDevices = 64;
// ltpmg_shared_engine's engines, by device
std::mutex g_shared_mu, g_shared_open_mu;
std::map<int32_t, ltpmg_engine *> g_shared;
std::mutex g_static_mu;
uint8_t *g_static_ptr[kStaticDevices] = {};

// Carves the static scratch into the buffers of one batch of up to
// LTPMG_STATIC_MAX_OPS ops; false if they do not fit (a build with larger
// sort tiles).
struct StaticBufs {
    OrderBufs ob{};
    uint32_t *tile_t0 = nullptr, *counters = nullptr;
};
bool carve_static(uint8_t *base, StaticBufs &out) {
    size_t at = 0;
    auto take = [&](size_t bytes) {
        uint8_t *p = base + at;
        at += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const size_t n = LTPMG_STATIC_MAX_OPS;
    const size_t sg = (seg_tiles((uint32_t)n) + 4) & ~(size_t)3;
    StaticBufs b;
    for (int i = 0; i < 2; ++i) {
        b.ob.sort.keys[i] = reinterpret_cast<uint32_t *>(take(n * 4));
        b.ob.sort.vals[i] = reinterpret_cast<uint32_t *>(take(n * 4));
    }
    b.ob.sort.hist = reinterpret_cast<uint32_t *>(take(sort_hist_words(n) * 4));
    b.ob.sort.partials = reinterpret_cast<uint32_t *>(take(sort_partials_needed() * 4));
    b.ob.tile_flag = reinterpret_cast<uint32_t *>(take((sg * 3 + 2) * 4));
    b.ob.tile_set = b.ob.tile_flag + sg;
    b.ob.tile_delta = b.ob.tile_flag + 2 * sg;
    b.ob.wtotal = b.ob.tile_flag + 3 * sg;
    b.ob.wlist = reinterpret_cast<uint4 *>(take(sg * 16));
    b.tile_t0 = reinterpret_cast<uint32_t *>(take((n / kTileOps + 2) * 4));
    b.counters = reinterpret_cast<uint32_t *>(take(16));
    if (at > kStaticScratchBytes) return false;
    out = b;
    return true;
}

// The device address of this device's `scratch`, looked up once per device
// (the device must be current).
hipError_t static_scratch_of(int device, uint8_t **out) {
    *out = nullptr;
    if (device < 0 || device >= kStaticDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_static_mu);
    if (!g_static_ptr[device]) {
        void *p = nullptr;
        hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess) return st;
        g_static_ptr[device] = static_cast<uint8_t *>(p);
    }
    *out = g_static_ptr[device];
    return hipSuccess;
}

}  // namespace

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

DevTables dev_tables(const ltpmg_engine *e) {
    return DevTables{e->d_cols, e->d_tabs, e->d_cols + LTPMG_MAX_TABLES * LTPMG_MAX_COLS, e->n_clist, e->kdec,
                     e->wide_lanes};
}

// Column stride of the ordered-mode cell keys (TabDesc::cstride): each
// table's column count rounded up to a power of two when the padded keys of
// all live tables fit below 2^32 - 1, else the column count itself.
#ifndef LTPMG_KEY_PAD
#define LTPMG_KEY_PAD 1
#endif
bool pad_key_columns(const ltpmg_engine *e) {
    if (!LTPMG_KEY_PAD) return false;
    uint64_t padded = 0;
    for (const Table &t : e->tables)
        if (t.live) padded += (uint64_t)t.n_rows << ceil_log2(t.n_cols);
    return padded <= 0xFFFFFFFFull;
}
uint32_t key_col_stride(const Table &t, bool pad) { return pad ? 1u << ceil_log2(t.n_cols) : t.n_cols; }

// A cell move takes G lanes, one 16-byte chunk each per step.  Past 128-byte
// result strides the chunks no longer fit one step of 8 lanes, and a wide
// lane group mostly idles on the narrow columns that schemas like order entry
// are made of (98 % of its 92 slots are at most 32 bytes, one is 256): 2
// lanes per move took its ordered batch from 3.59 to 3.12 ms against 8 (1
// lane: 3.74 ms, 4 lanes: 3.16 ms).  Tables of wide columns keep 8 (the
// apply kernel chooses per tile, from the tables its keys span), and so does
// the lowest-TID-wins mode (its W = 64 batch took 1.05 ms with 2 lanes
// against 1.01 ms with 8).
bool narrow_columns(const uint32_t *slots, uint32_t n) {
    uint32_t narrow = 0;
    for (uint32_t i = 0; i < n; ++i) narrow += slots[i] <= 32u;
    return n && narrow * 4 >= n * 3;
}

// Writes the device descriptors and assigns ordered-mode global row bases.
int sync_descriptors(ltpmg_engine *e) {
    ColDesc cols[LTPMG_MAX_TABLES * LTPMG_MAX_COLS + kColList] = {};
    ColDesc *clist = cols + LTPMG_MAX_TABLES * LTPMG_MAX_COLS;
    TabDesc tabs[LTPMG_MAX_TABLES] = {};
    uint64_t cell_base = 0;  // past 2^32 - 1 cells only the lowest-TID-wins mode runs
    uint32_t n_clist = 0;
    const bool pad = pad_key_columns(e);
    KeyDecode kd{};
    for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t) {
        const Table &tb = e->tables[t];
        if (!tb.live) continue;
        const uint32_t stride = key_col_stride(tb, pad);
        tabs[t] = TabDesc{tb.lock, tb.present, tb.n_rows, tb.n_cols,
                          (uint32_t)std::min<uint64_t>(cell_base, 0xFFFFFFFFu), 1u, stride};
        cell_base += (uint64_t)tb.n_rows * stride;
        kd.tab[kd.n] = t;
        kd.end[kd.n] = (uint32_t)std::min<uint64_t>(cell_base, 0xFFFFFFFFu);
        kd.stride[kd.n] = stride;
        kd.shift[kd.n] = (stride & (stride - 1u)) == 0u ? (uint32_t)__builtin_ctz(stride) : 32u;
        kd.cpos[kd.n] = kd.n_cols_all;
        kd.narrow |= (uint32_t)narrow_columns(tb.slot, tb.n_cols) << kd.n;
        kd.n_cols_all += tb.n_cols;
        kd.end_all = kd.end[kd.n];
        ++kd.n;
        for (uint32_t c = 0; c < tb.n_cols; ++c) {
            cols[t * LTPMG_MAX_COLS + c] = ColDesc{tb.col[c], tb.slot[c], tb.width[c]};
            if (n_clist < kColList) clist[n_clist++] = cols[t * LTPMG_MAX_COLS + c];
        }
    }
    HIP_TRY(e, hipMemcpy(e->d_cols, cols, sizeof cols, hipMemcpyHostToDevice));
    e->n_clist = n_clist;
    e->kdec = kd;
    std::vector<uint32_t> slots;
    for (const Table &tb : e->tables)
        if (tb.live) slots.insert(slots.end(), tb.slot, tb.slot + tb.n_cols);
    e->wide_lanes = narrow_columns(slots.data(), (uint32_t)slots.size()) ? 2u : 8u;
    HIP_TRY(e, hipMemcpy(e->d_tabs, tabs, sizeof tabs, hipMemcpyHostToDevice));
    return LTPMG_OK;
}

// Ordered-mode scratch for up to ops ops: sort buffers, per-tile scan state
// and the final-write list.
void free_order_bufs(ltpmg_engine *e, OrderBufs &ob, size_t ops) {
    const size_t sg = (seg_tiles((uint32_t)ops) + 4) & ~(size_t)3;  // a multiple of 4: 16-byte aligned arrays
    for (int i = 0; i < 2; ++i) {
        dev_free(e, ob.sort.keys[i], ops * 4);
        dev_free(e, ob.sort.vals[i], ops * 4);
    }
    dev_free(e, ob.sort.hist, sort_hist_words(ops) * 4);
    dev_free(e, ob.sort.partials, sort_partials_needed() * 4);
    dev_free(e, ob.tile_flag, (sg * 3 + 2) * 4);
    dev_free(e, ob.wlist, sg * 16);
    ob = OrderBufs{};
}

int alloc_order_bufs(ltpmg_engine *e, OrderBufs &ob, size_t ops) {
    const size_t sg = (seg_tiles((uint32_t)ops) + 4) & ~(size_t)3;  // a multiple of 4: 16-byte aligned arrays
    int rc = LTPMG_OK;
#define ALLOC(ptr, n) if (rc == LTPMG_OK) rc = dev_alloc_n(e, &(ptr), (n))
    for (int i = 0; i < 2; ++i) {
        ALLOC(ob.sort.keys[i], ops);
        ALLOC(ob.sort.vals[i], ops);
    }
    ALLOC(ob.sort.hist, sort_hist_words(ops));
    ALLOC(ob.sort.partials, sort_partials_needed());
    ALLOC(ob.tile_flag, sg * 3 + 2);
    ALLOC(ob.wlist, sg);
#undef ALLOC
    if (rc != LTPMG_OK) return rc;
    ob.tile_set = ob.tile_flag + sg;
    ob.tile_delta = ob.tile_flag + 2 * sg;
    ob.wtotal = ob.tile_flag + 3 * sg;
    return LTPMG_OK;
}

// Row-op buffers for batches of up to ops ops (found, meta_eff, insert
// counts per tile and table, the cursor at the batch start per table).
int alloc_row_bufs(ltpmg_engine *e, size_t ops, uint8_t *&found, uint16_t *&meta_eff, uint32_t *&ins_cnt,
                   uint32_t *&ins_base) {
    int rc = dev_alloc_n(e, &found, std::max<size_t>(ops, 1));
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &meta_eff, std::max<size_t>(ops, 1));
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &ins_cnt, (size_t)std::max(ins_tiles((uint32_t)ops), 1u) * LTPMG_MAX_TABLES);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &ins_base, LTPMG_MAX_TABLES);
    return rc;
}
void free_row_bufs(ltpmg_engine *e, size_t ops, uint8_t *&found, uint16_t *&meta_eff, uint32_t *&ins_cnt,
                   uint32_t *&ins_base) {
    dev_free(e, found, std::max<size_t>(ops, 1));
    dev_free(e, meta_eff, std::max<size_t>(ops, 1) * 2);
    dev_free(e, ins_cnt, (size_t)std::max(ins_tiles((uint32_t)ops), 1u) * LTPMG_MAX_TABLES * 4);
    dev_free(e, ins_base, LTPMG_MAX_TABLES * 4);
    found = nullptr;
    meta_eff = nullptr;
    ins_cnt = ins_base = nullptr;
}

// What a batch does with rows: any op on an append table, DELETEs, and the
// INSERTs per table (a pass over meta, split over host threads for large
// batches; nothing to do while no append table exists).
// row_end (if given) gets, per append table, one past the highest row any
// op other than an INSERT names (0 when none does).
void scan_row_ops(const ltpmg_engine *e, const ltpmg_batch *b, RowBufs &rb, uint32_t inserts[LTPMG_MAX_TABLES],
                  uint32_t *row_end = nullptr) {
    rb = RowBufs{};
    std::fill(inserts, inserts + LTPMG_MAX_TABLES, 0u);
    if (row_end) std::fill(row_end, row_end + LTPMG_MAX_TABLES, 0u);
    if (e->n_append == 0 || b->n_ops == 0) return;
    uint32_t append_mask = 0;
    for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t)
        if (e->tables[t].live && e->tables[t].present) append_mask |= 1u << t;
    const uint32_t n = b->n_ops;
    const uint32_t n_threads = n < (1u << 20) ? 1u : std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    struct Acc {
        uint32_t ins[LTPMG_MAX_TABLES] = {};
        uint32_t end[LTPMG_MAX_TABLES] = {};
        bool app = false, del = false;
    };
    std::vector<Acc> acc(n_threads);
    auto work = [&](uint32_t w) {
        Acc &a = acc[w];
        const uint32_t lo = (uint32_t)((uint64_t)n * w / n_threads), hi = (uint32_t)((uint64_t)n * (w + 1) / n_threads);
        for (uint32_t i = lo; i < hi; ++i) {
            const uint16_t m = b->meta[i];
            const uint32_t tb = (m >> 8) & 15u;
            if (!((append_mask >> tb) & 1u)) continue;
            a.app = true;
            if ((m & 3u) == LTPMG_OP_ROW && ((m >> 2) & 63u) == LTPMG_ROW_INSERT) {
                ++a.ins[tb];
                continue;
            }
            if ((m & 3u) == LTPMG_OP_ROW) a.del = true;
            a.end[tb] = std::max(a.end[tb], b->rows[i] + 1u);
        }
    };
    if (n_threads == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (uint32_t w = 0; w < n_threads; ++w) th.emplace_back(work, w);
        for (auto &t : th) t.join();
    }
    for (const Acc &a : acc) {
        rb.has_append |= a.app;
        rb.has_delete |= a.del;
        for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t) {
            inserts[t] += a.ins[t];
            if (row_end) row_end[t] = std::max(row_end[t], a.end[t]);
        }
    }
    for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t)
        if (inserts[t]) rb.insert_tables |= 1u << t;
    rb.has_insert = rb.insert_tables != 0;
}

// The INSERTs of a batch must fit: each append table's bound on its cursor
// plus the batch's INSERTs stays within its capacity.  When the bound says
// no, the exact cursor is read back (after every executed batch) and
// checked again.  Then the bounds move on by the batch's INSERTs.
int check_capacity(ltpmg_engine *e, const uint32_t inserts[LTPMG_MAX_TABLES]) {
    bool synced = false;
    for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t) {
        Table &tb = e->tables[t];
        if (!inserts[t] || tb.cursor_bound + inserts[t] <= tb.n_rows) continue;
        if (!synced) {
            for (Slot &sl : e->slots) HIP_TRY(e, hipStreamSynchronize(sl.stream));
            synced = true;
        }
        uint32_t cur = 0;
        HIP_TRY(e, hipMemcpy(&cur, e->d_cursor + t, 4, hipMemcpyDeviceToHost));
        tb.cursor_bound = cur;
        if ((uint64_t)cur + inserts[t] > tb.n_rows)
            return fail(e, LTPMG_EFULL, "table %u: the batch inserts up to %u rows, %u of %u rows are left", t,
                        inserts[t], tb.n_rows - cur, tb.n_rows);
    }
    for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t) e->tables[t].cursor_bound += inserts[t];
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
    dev_free(e, s.d_tile_t0, (ops / kTileOps + 2) * 4);
    dev_free(e, s.d_counters, 16);
    free_order_bufs(e, s.ob, ops);
    if (s.has_row_bufs) free_row_bufs(e, ops, s.d_found, s.d_meta_eff, s.d_ins_cnt, s.d_ins_base);
    s.has_row_bufs = false;
    if (s.h_stage) (void)hipHostFree(s.h_stage);
    s.d_txn_off = s.d_rows = s.d_tile_t0 = s.d_counters = nullptr;
    s.d_meta = nullptr;
    s.d_args = nullptr;
    s.d_commit = s.d_results = nullptr;
    s.h_stage = s.stage_dev = nullptr;
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
    const size_t tiles = ops / kTileOps + 2;
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
#undef ALLOC
    if (rc == LTPMG_OK) rc = alloc_order_bufs(e, s.ob, ops);
    s.cap_txns = txns;
    s.cap_ops = ops;
    s.cap_res_stride = res_stride;
    if (rc != LTPMG_OK) {
        free_slot_buffers(e, s);
        return rc;
    }
    s.stage_bytes = ((size_t)txns + 1) * 4 + (size_t)ops * (4 + 2 + 8) + 64;
    hipError_t hs = hipHostMalloc(reinterpret_cast<void **>(&s.h_stage), s.stage_bytes, hipHostMallocDefault);
    if (hs != hipSuccess) {
        s.h_stage = nullptr;
        free_slot_buffers(e, s);
        return fail(e, LTPMG_ENOMEM, "hipHostMalloc(%zu) failed: %s", s.stage_bytes, hipGetErrorString(hs));
    }
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&s.stage_dev), s.h_stage, 0) != hipSuccess) {
        (void)hipGetLastError();
        s.stage_dev = nullptr;
    }
    return LTPMG_OK;
}

// The CSR structure of a batch: txn_off starts at 0, every txn holds
// 1..LTPMG_MAX_TXN_OPS ops, and the last offset is n_ops.  The kernels derive
// every op range from txn_off, so this check runs on every submit, trusted
// or not.
int validate_structure(ltpmg_engine *e, const ltpmg_batch *b, uint64_t max_ops) {
    if (!b || !b->txn_off || (b->n_ops && (!b->rows || !b->meta || !b->args)))
        return fail(e, LTPMG_EINVAL, "null batch array");
    if (b->n_ops >= max_ops) return fail(e, LTPMG_EINVAL, "n_ops must be < %llu", (unsigned long long)max_ops);
    if (b->txn_off[0] != 0) return fail(e, LTPMG_EINVAL, "txn_off[0] must be 0");
    for (uint32_t t = 0; t < b->n_txns; ++t) {
        uint32_t len = b->txn_off[t + 1] - b->txn_off[t];
        if (b->txn_off[t + 1] < b->txn_off[t] || len < 1 || len > LTPMG_MAX_TXN_OPS)
            return fail(e, LTPMG_EINVAL, "txn %u has %d ops (allowed 1..%d)", t,
                        (int)(b->txn_off[t + 1] - b->txn_off[t]), LTPMG_MAX_TXN_OPS);
    }
    if (b->txn_off[b->n_txns] != b->n_ops)
        return fail(e, LTPMG_EINVAL, "txn_off[n_txns]=%u != n_ops=%u", b->txn_off[b->n_txns], b->n_ops);
    return LTPMG_OK;
}

// Runs fn(lo, hi) over [0, n) split across host threads (one thread below
// 2^20 items); returns the lowest index any range reported, or UINT32_MAX.
template <class Fn>
uint32_t first_bad_parallel(uint32_t n, Fn fn) {
    const uint32_t n_threads = n < (1u << 20) ? 1u : std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<uint32_t> first(n_threads, UINT32_MAX);
    auto scan = [&](uint32_t w) {
        first[w] = fn((uint32_t)((uint64_t)n * w / n_threads), (uint32_t)((uint64_t)n * (w + 1) / n_threads));
    };
    if (n_threads == 1) {
        scan(0);
    } else {
        std::vector<std::thread> th;
        for (uint32_t w = 0; w < n_threads; ++w) th.emplace_back(scan, w);
        for (auto &t : th) t.join();
    }
    return *std::min_element(first.begin(), first.end());
}

// Checks a batch as submitted; max_ops is 2^29 for one batch (ordered-mode
// op indices have 29 bits) and larger for a stream of batches.
int validate_batch(ltpmg_engine *e, const ltpmg_batch *b, uint64_t max_ops = 1ull << 29) {
    int rc = validate_structure(e, b, max_ops);
    if (rc != LTPMG_OK) return rc;
    // per-op checks, split over host threads for large batches; the first
    // bad op (lowest index) is reported
    auto bad_op = [&](uint32_t i) -> bool {
        uint16_t m = b->meta[i];
        uint32_t kind = m & 3u, col = (m >> 2) & 63u;
        const Table &t = e->tables[(m >> 8) & 15u];
        if ((m >> 12) || !t.live) return true;
        if (kind == LTPMG_OP_ROW)
            return !t.present || (col == LTPMG_ROW_DELETE ? b->rows[i] >= t.n_rows
                                                          : col != LTPMG_ROW_INSERT || b->rows[i] != 0);
        return col >= t.n_cols || b->rows[i] >= t.n_rows;
    };
    uint32_t i = first_bad_parallel(b->n_ops, [&](uint32_t lo, uint32_t hi) {
        for (uint32_t k = lo; k < hi; ++k)
            if (bad_op(k)) return k;
        return UINT32_MAX;
    });
    if (i != UINT32_MAX) {
        uint16_t m = b->meta[i];
        uint32_t kind = m & 3u, col = (m >> 2) & 63u, tab = (m >> 8) & 15u;
        const Table &t = e->tables[tab];
        if (m >> 12) return fail(e, LTPMG_EINVAL, "op %u: bad meta 0x%04x", i, m);
        if (!t.live) return fail(e, LTPMG_EINVAL, "op %u: table %u does not exist", i, tab);
        if (kind == LTPMG_OP_ROW) {
            if (!t.present) return fail(e, LTPMG_EINVAL, "op %u: INSERT / DELETE on table %u, not an append table", i, tab);
            if (col == LTPMG_ROW_INSERT) return fail(e, LTPMG_EINVAL, "op %u: an INSERT names row 0, not %u", i, b->rows[i]);
            if (col != LTPMG_ROW_DELETE) return fail(e, LTPMG_EINVAL, "op %u: bad row op %u", i, col);
            return fail(e, LTPMG_EINVAL, "op %u: DELETE of row %u >= %u rows of table %u", i, b->rows[i], t.n_rows, tab);
        }
        if (col >= t.n_cols) return fail(e, LTPMG_EINVAL, "op %u: column %u >= %u", i, col, t.n_cols);
        return fail(e, LTPMG_EINVAL, "op %u: row %u >= %u rows of table %u", i, b->rows[i], t.n_rows, tab);
    }
    // Within a txn, a cell that an op writes is not accessed by any later op
    // of the txn (a READ before the write is fine), and a row the txn
    // DELETEs is not accessed by any other op of the txn.  The kernels rely
    // on it: lowest-TID-wins reads see the batch-start snapshot and apply a
    // txn's writes to one cell in parallel.  Only pairs whose first op writes
    // (or either op deletes) are compared, so a 16-op txn with 8 writes costs
    // at most 8 x 15 compares.
    const uint32_t *off = b->txn_off;
    auto is_del = [&](uint32_t p) { return (b->meta[p] & 3u) == LTPMG_OP_ROW && ((b->meta[p] >> 2) & 63u) == LTPMG_ROW_DELETE; };
    auto is_row = [&](uint32_t p) { return (b->meta[p] & 3u) == LTPMG_OP_ROW; };
    // the first op (index) after p in txn [lo, hi) that conflicts with p, or hi
    auto clash = [&](uint32_t p, uint32_t lo, uint32_t hi) -> uint32_t {
        const uint32_t tab = (b->meta[p] >> 8) & 15u;
        if (is_del(p)) {  // every other op of the txn on that row
            for (uint32_t q = lo; q < hi; ++q)
                if (q != p && !(is_row(q) && !is_del(q)) && b->rows[q] == b->rows[p] && ((b->meta[q] >> 8) & 15u) == tab)
                    return q;
            return hi;
        }
        if ((b->meta[p] & 3u) == LTPMG_OP_READ || is_row(p)) return hi;
        const uint16_t cell = b->meta[p] & (uint16_t)~3u;
        for (uint32_t q = p + 1; q < hi; ++q)
            if (!is_row(q) && b->rows[q] == b->rows[p] && (b->meta[q] & (uint16_t)~3u) == cell) return q;
        return hi;
    };
    uint32_t bad_t = first_bad_parallel(b->n_txns, [&](uint32_t lo, uint32_t hi) {
        for (uint32_t t = lo; t < hi; ++t)
            for (uint32_t p = off[t]; p < off[t + 1]; ++p)
                if (clash(p, off[t], off[t + 1]) != off[t + 1]) return t;
        return UINT32_MAX;
    });
    if (bad_t != UINT32_MAX) {
        for (uint32_t p = off[bad_t]; p < off[bad_t + 1]; ++p) {
            const uint32_t q = clash(p, off[bad_t], off[bad_t + 1]);
            if (q == off[bad_t + 1]) continue;
            if (is_del(p))
                return fail(e, LTPMG_EINVAL, "txn %u: op %u accesses row %u of table %u, which op %u of the same txn "
                            "deletes", bad_t, q, b->rows[p], (b->meta[p] >> 8) & 15u, p);
            return fail(e, LTPMG_EINVAL,
                        "txn %u: op %u accesses the cell (table %u, row %u, column %u) that op %u of "
                        "the same txn writes",
                        bad_t, q, (b->meta[p] >> 8) & 15u, b->rows[p], (b->meta[p] >> 2) & 63u, p);
        }
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

// Lowest-TID-wins on the ops of ov with the three-kernel pipeline.
// wait_tables() runs before the first access to lock words or cells.
// rb: the batch's row-op buffers, or nullptr when it touches no append
// table (then the kernels never look at presence).
template <class Rec, class Wait>
int launch_mode_a(ltpmg_engine *e, hipStream_t st, const OpView &ov, uint32_t *tile_t0,
                  uint32_t *counters, uint8_t *commit, uint8_t *results, uint32_t rs, Rec &rec,
                  Wait wait_tables, RowBufs *rb = nullptr) {
    const uint32_t n_tiles = std::max(1u, (ov.n_ops + kTileOps - 1) / kTileOps);
    int rc = wait_tables();
    if (rc != LTPMG_OK) return rc;
    DevTables dt = dev_tables(e);
    uint8_t *found = rb ? rb->found : nullptr;
    rec.begin("tile_bounds");
    launch_zero_u32(st, counters, 4);
    launch_tile_bounds(st, ov.txn_off, ov.n_txns, n_tiles, tile_t0, e->d_sync);
    rec.end();
    rec.begin("cc_reserve");
    launch_cc_reserve(st, dt, ov, tile_t0, n_tiles, e->d_sync);
    rec.end();
    rec.begin("cc_validate_read");
    launch_cc_validate_read(st, dt, ov, tile_t0, n_tiles, e->d_sync, commit, results, rs, counters, found);
    rec.end();
    rec.begin("cc_write");
    launch_cc_write(st, dt, ov, tile_t0, n_tiles, commit, rs, found);
    rec.end();
    if (rb) {
        rec.begin("row_deletes");
        launch_row_deletes(st, dt, ov, *rb);
        rec.end();
        rec.begin("inserts");
        launch_inserts(st, dt, ov, *rb, results, rs);
        rec.end();
    }
    return LTPMG_OK;
}

// Key geometry of ordered mode: cells are numbered table by table, row
// major, each row taking TabDesc::cstride numbers, so a key needs
// ceil(log2(numbers)) bits; all-ones stays free as the kernels' end-of-list
// sentinel.
int order_key_bits(ltpmg_engine *e, uint32_t *key_bits) {
    uint64_t cells = 0;
    const bool pad = pad_key_columns(e);
    for (const Table &t : e->tables)
        if (t.live) cells += (uint64_t)t.n_rows * key_col_stride(t, pad);
    if (cells > 0xFFFFFFFFull)
        return fail(e, LTPMG_EINVAL, "ordered mode numbers at most 2^32 - 1 cells; the tables hold %llu",
                    (unsigned long long)cells);
    *key_bits = ceil_log2(cells);
    return LTPMG_OK;
}

// Ordered mode on the ops of ov (results indexed by op within ov).  Keys,
// sort and scan touch only the pipeline's own buffers; wait_tables() runs
// before the first cell access.
// rb as for launch_mode_a.  A batch that touches append tables reads the
// presence bytes before its sort, so it waits for the previous batch first
// (then its sort does not overlap that batch), and sorts by a private copy of
// the meta words in which ops on absent rows are invalid.
// small_max: a batch of 1 .. min(small_max, kSmallMaxOps) ops runs as the
// one-launch kernel instead (every stage, row ops included, in one
// workgroup).  pad: READ result slots are written to their end; otherwise
// only each READ's column slot, and a fetch zeroes the rest on the host
// (zero_result_padding).
template <class Rec, class Wait>
int launch_mode_o(ltpmg_engine *e, hipStream_t st, const OpView &ov_in, OrderBufs &ob,
                  uint8_t *commit, uint8_t *results, uint32_t rs, Rec &rec, Wait wait_tables,
                  RowBufs *rb = nullptr, uint32_t small_max = 0, bool pad = false) {
    uint32_t key_bits = 0;
    int rc = order_key_bits(e, &key_bits);
    if (rc != LTPMG_OK) return rc;
    DevTables dt = dev_tables(e);
    if (ov_in.n_ops >= 1 && ov_in.n_ops <= std::min(small_max, kSmallMaxOps)) {
        rc = wait_tables();
        if (rc != LTPMG_OK) return rc;
        rec.begin("mo_small");
        launch_mo_small(st, dt, ov_in, key_bits, results, rs, commit, rb);
        rec.end();
        return LTPMG_OK;
    }
    OpView ov = ov_in;
    if (rb) {
        rc = wait_tables();
        if (rc != LTPMG_OK) return rc;
        rec.begin("row_presence");
        launch_row_presence_o(st, dt, ov_in, *rb, e->d_sync, results, rs);
        rec.end();
        ov.meta = rb->meta_eff;
    }
    int sorted = 0;
    if (ov.n_ops > 1 && key_bits > 0) {
        // digit 0 straight from the ops: the unsorted keys are never stored
        const uint32_t bits = radix_digit_bits(key_bits);
        rec.begin("rs_hist_ops");
        launch_rs_hist_ops(st, dt, ov, ob.sort, bits);
        rec.end();
        rec.begin("rs_scan");
        launch_rs_scan(st, ob.sort, ov.n_ops, bits);
        rec.end();
        rec.begin("rs_scatter_ops");
        launch_rs_scatter_ops(st, dt, ov, ob.sort, bits);
        rec.end();
        sorted = 1;
        for (uint32_t shift = bits; shift < key_bits; shift += bits) {
            rec.begin("rs_hist");
            launch_rs_hist(st, ob.sort, sorted, ov.n_ops, shift, bits);
            rec.end();
            rec.begin("rs_scan");
            launch_rs_scan(st, ob.sort, ov.n_ops, bits);
            rec.end();
            rec.begin("rs_scatter");
            launch_rs_scatter(st, ob.sort, sorted, ov.n_ops, shift, bits);
            rec.end();
            sorted ^= 1;
        }
    } else {
        rec.begin("mo_keys");
        launch_mo_keys(st, dt, ov, ob.sort.keys[0], ob.sort.vals[0]);
        rec.end();
    }
    rec.begin("mo_seg_scan");
    launch_mo_seg_scan(st, ov, ob, sorted);
    rec.end();
    if (!rb) {
        rc = wait_tables();
        if (rc != LTPMG_OK) return rc;
    }
    rec.begin("mo_warm_seeds");
    launch_mo_warm_seeds(st, ov, ob);
    rec.end();
    rec.begin("mo_apply_reads");
    launch_mo_apply_reads(st, dt, ov, ob, sorted, results, rs, pad);
    rec.end();
    rec.begin("mo_final_writes");
    launch_mo_final_writes(st, dt, ov, ob, rs, commit);
    rec.end();
    if (rb) {
        rec.begin("row_deletes");
        launch_row_deletes(st, dt, ov_in, *rb);
        rec.end();
        rec.begin("inserts");
        launch_inserts(st, dt, ov_in, *rb, results, rs);
        rec.end();
    }
    return LTPMG_OK;
}

struct NoRec {
    void begin(const char *) {}
    void end() {}
};

// The ordered kernels write a READ's result slot only as far as its column's
// slot (an INSERT's, its first 16 bytes), so that the zeros past it never
// cross HBM; the fetch zeroes the rest of each slot here, on the host, from
// the batch's meta words, so callers get the same bytes as before.  (The
// lowest-TID-wins kernels write whole slots; zeroing again changes nothing.)
int zero_result_padding(ltpmg_engine *e, const uint16_t *d_meta, uint64_t n_ops, uint8_t *results,
                        uint32_t stride) {
    if (n_ops == 0) return LTPMG_OK;
    std::vector<uint16_t> meta(n_ops);
    HIP_TRY(e, hipMemcpy(meta.data(), d_meta, n_ops * 2, hipMemcpyDeviceToHost));
    uint32_t keep[LTPMG_MAX_TABLES][LTPMG_MAX_COLS + 1];  // defined bytes per (table, column)
    for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t)
        for (uint32_t c = 0; c <= LTPMG_MAX_COLS; ++c)
            keep[t][c] = e->tables[t].live && c < e->tables[t].n_cols ? std::min(e->tables[t].slot[c], stride) : stride;
    auto run = [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            const uint16_t m = meta[i];
            const uint32_t kind = m & 3u, col = (m >> 2) & 63u, tab = (m >> 8) & 15u;
            uint32_t from = stride;
            if (m >> 12) continue;  // not a valid op: no defined result
            if (kind == LTPMG_OP_READ) from = keep[tab][col];
            else if (kind == LTPMG_OP_ROW && col == LTPMG_ROW_INSERT) from = std::min(16u, stride);
            if (from < stride) std::memset(results + i * stride + from, 0, stride - from);
        }
    };
    const uint64_t kChunk = 1u << 20;
    if (n_ops <= kChunk) {
        run(0, n_ops);
    } else {
        std::vector<std::thread> th;
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        for (unsigned k = 0; k < nt; ++k)
            th.emplace_back(run, n_ops * k / nt, n_ops * (k + 1) / nt);
        for (auto &t : th) t.join();
    }
    return LTPMG_OK;
}

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
    int rc = dev_alloc_n(e, &e->d_cols, LTPMG_MAX_TABLES * LTPMG_MAX_COLS + kColList);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &e->d_tabs, LTPMG_MAX_TABLES);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &e->d_digest, 1);
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &e->d_sync, 4);
    if (rc == LTPMG_OK && hipMemset(e->d_sync, 0, 16) != hipSuccess) rc = LTPMG_EHIP;
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &e->d_cursor, LTPMG_MAX_TABLES);
    if (rc == LTPMG_OK && hipMemset(e->d_cursor, 0, 4 * LTPMG_MAX_TABLES) != hipSuccess) rc = LTPMG_EHIP;
    if (rc == LTPMG_OK &&
        hipDeviceGetAttribute(&e->n_cu, hipDeviceAttributeMultiprocessorCount, cfg->device) != hipSuccess)
        rc = LTPMG_EHIP;
    if (rc == LTPMG_OK) rc = sync_descriptors(e);
    if (rc == LTPMG_OK && cfg->device < kStaticDevices) {
        // the allocation-free entry point's scratch on this device (current)
        hipError_t st = static_scratch_of(cfg->device, &e->static_scratch);
        if (st != hipSuccess) rc = fail(e, LTPMG_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(st));
    }
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
    {
        std::lock_guard<std::mutex> g(g_shared_mu);
        for (auto it = g_shared.begin(); it != g_shared.end(); ++it)
            if (it->second == e) {
                g_shared.erase(it);
                break;
            }
    }
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
        if (t.present) (void)hipFree(t.present);
    }
    if (e->d_cursor) (void)hipFree(e->d_cursor);
    if (e->d_cols) (void)hipFree(e->d_cols);
    if (e->d_tabs) (void)hipFree(e->d_tabs);
    if (e->d_digest) (void)hipFree(e->d_digest);
    if (e->d_sync) (void)hipFree(e->d_sync);
    delete e;
    return LTPMG_OK;
}

int ltpmg_shared_engine(int32_t device, ltpmg_engine **out) {
    if (!out) return LTPMG_EINVAL;
    *out = nullptr;
    // one opener at a time; the map's own lock is not held across
    // ltpmg_open, whose failure path calls ltpmg_close, which takes it
    std::lock_guard<std::mutex> opening(g_shared_open_mu);
    {
        std::lock_guard<std::mutex> g(g_shared_mu);
        auto it = g_shared.find(device);
        if (it != g_shared.end()) {
            *out = it->second;
            return LTPMG_OK;
        }
    }
    ltpmg_cfg cfg{};
    cfg.device = device;
    cfg.n_slots = 1;
    int rc = ltpmg_open(&cfg, out);
    if (rc == LTPMG_OK) {
        std::lock_guard<std::mutex> g(g_shared_mu);
        g_shared[device] = *out;
    }
    return rc;
}

const char *ltpmg_last_error(const ltpmg_engine *e) { return e ? e->err.c_str() : "no engine"; }

uint32_t ltpmg_result_stride(const ltpmg_engine *e) { return e ? max_slot_all(e) : 0; }

}  // extern "C"

namespace {
// Creates table table_id; an append table (append set) also gets presence
// bytes, 1 for rows [0, n_present), and its cursor at n_present.
int create_table_locked(ltpmg_engine *e, uint32_t table_id, const ltpmg_table_desc *d, bool append,
                        uint32_t n_present) {
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
    if (append) {
        if (rc == LTPMG_OK) rc = dev_alloc_n(e, &t.present, t.n_rows);
        if (rc == LTPMG_OK && (hipMemset(t.present, 0, t.n_rows) != hipSuccess ||
                               (n_present && hipMemset(t.present, 1, n_present) != hipSuccess) ||
                               hipMemcpy(e->d_cursor + table_id, &n_present, 4, hipMemcpyHostToDevice) != hipSuccess))
            rc = fail(e, LTPMG_EHIP, "initialising the presence bytes failed");
        t.cursor_bound = n_present;
    }
    if (rc != LTPMG_OK) {
        for (uint32_t c = 0; c < t.n_cols; ++c) dev_free(e, t.col[c], (size_t)t.slot[c] * t.n_rows);
        dev_free(e, t.lock, 8ull * t.n_rows);
        dev_free(e, t.present, t.n_rows);
        return rc;
    }
    t.live = true;
    e->tables[table_id] = t;
    if (append) ++e->n_append;
    return sync_descriptors(e);
}
}  // namespace

extern "C" {

int ltpmg_create_table(ltpmg_engine *e, uint32_t table_id, const ltpmg_table_desc *d) {
    if (!e || !d) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    return create_table_locked(e, table_id, d, false, 0);
}

int ltpmg_create_append_table(ltpmg_engine *e, uint32_t table_id, const ltpmg_table_desc *d, uint32_t n_present) {
    if (!e || !d) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    if (n_present > d->n_rows) return fail(e, LTPMG_EINVAL, "%u present rows > capacity %u", n_present, d->n_rows);
    return create_table_locked(e, table_id, d, true, n_present);
}

int ltpmg_table_rows(ltpmg_engine *e, uint32_t table_id, uint32_t *out) {
    if (!e || !out) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (table_id >= LTPMG_MAX_TABLES || !e->tables[table_id].live)
        return fail(e, LTPMG_ESTATE, "table %u does not exist", table_id);
    Table &t = e->tables[table_id];
    if (!t.present) {
        *out = t.n_rows;
        return LTPMG_OK;
    }
    for (Slot &sl : e->slots) HIP_TRY(e, hipStreamSynchronize(sl.stream));
    HIP_TRY(e, hipMemcpy(out, e->d_cursor + table_id, 4, hipMemcpyDeviceToHost));
    t.cursor_bound = *out;
    return LTPMG_OK;
}

}  // extern "C"

namespace {
// Sets an append table's cursor to n_rows and clears the presence of every
// row from there on (caller holds e->mu, the slots are idle).
int truncate_locked(ltpmg_engine *e, uint32_t table_id, uint32_t n_rows) {
    Table &t = e->tables[table_id];
    if (n_rows < t.n_rows) HIP_TRY(e, hipMemset(t.present + n_rows, 0, t.n_rows - n_rows));
    HIP_TRY(e, hipMemcpy(e->d_cursor + table_id, &n_rows, 4, hipMemcpyHostToDevice));
    t.cursor_bound = n_rows;
    return LTPMG_OK;
}
}  // namespace

extern "C" {

int ltpmg_truncate_table(ltpmg_engine *e, uint32_t table_id, uint32_t n_rows) {
    if (!e) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (table_id >= LTPMG_MAX_TABLES || !e->tables[table_id].live)
        return fail(e, LTPMG_ESTATE, "table %u does not exist", table_id);
    Table &t = e->tables[table_id];
    if (!t.present) return fail(e, LTPMG_EINVAL, "table %u is not an append table", table_id);
    if (n_rows > t.n_rows) return fail(e, LTPMG_EINVAL, "%u rows > capacity %u", n_rows, t.n_rows);
    for (Slot &sl : e->slots) HIP_TRY(e, hipStreamSynchronize(sl.stream));
    return truncate_locked(e, table_id, n_rows);
}

int ltpmg_read_presence(ltpmg_engine *e, uint32_t table_id, uint32_t row0, uint32_t n_rows, uint8_t *dst) {
    if (!e || (!dst && n_rows)) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (table_id >= LTPMG_MAX_TABLES || !e->tables[table_id].live)
        return fail(e, LTPMG_ESTATE, "table %u does not exist", table_id);
    const Table &t = e->tables[table_id];
    if ((uint64_t)row0 + n_rows > t.n_rows)
        return fail(e, LTPMG_EINVAL, "rows [%u, %llu) outside the table", row0, (unsigned long long)row0 + n_rows);
    if (!n_rows) return LTPMG_OK;
    if (!t.present) {
        std::memset(dst, 1, n_rows);
        return LTPMG_OK;
    }
    for (Slot &sl : e->slots) HIP_TRY(e, hipStreamSynchronize(sl.stream));
    HIP_TRY(e, hipMemcpy(dst, t.present + row0, n_rows, hipMemcpyDeviceToHost));
    return LTPMG_OK;
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
    const size_t w = t.width[col], sl = t.slot[col];
    uint8_t *dev = t.col[col] + (size_t)row0 * sl;
    uint8_t *h = static_cast<uint8_t *>(host);
    if (w == sl) {
        HIP_TRY(e, hipMemcpy(to_device ? (void *)dev : (void *)h, to_device ? (void *)h : (void *)dev,
                             w * n_rows, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
        return LTPMG_OK;
    }
    // padded slots: repack on the host in chunks and copy each chunk flat (a
    // strided 2-D copy from pageable memory goes row by row)
    const size_t chunk = std::max<size_t>(1, (size_t(64) << 20) / sl);
    std::vector<uint8_t> buf(std::min<size_t>(chunk, n_rows) * sl, 0);
    for (size_t r0 = 0; r0 < n_rows; r0 += chunk) {
        const size_t nr = std::min<size_t>(chunk, n_rows - r0);
        if (to_device) {
            for (size_t r = 0; r < nr; ++r) std::memcpy(&buf[r * sl], h + (r0 + r) * w, w);
            HIP_TRY(e, hipMemcpy(dev + r0 * sl, buf.data(), nr * sl, hipMemcpyHostToDevice));
        } else {
            HIP_TRY(e, hipMemcpy(buf.data(), dev + r0 * sl, nr * sl, hipMemcpyDeviceToHost));
            for (size_t r = 0; r < nr; ++r) std::memcpy(h + (r0 + r) * w, &buf[r * sl], w);
        }
    }
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

// Row-op buffers of the slot, once its batch touches an append table.
static int slot_row_bufs(ltpmg_engine *e, Slot &s) {
    if (!s.rb.has_append) return LTPMG_OK;
    if (!s.has_row_bufs) {
        int rc = alloc_row_bufs(e, s.cap_ops, s.d_found, s.d_meta_eff, s.d_ins_cnt, s.d_ins_base);
        if (rc != LTPMG_OK) {
            free_row_bufs(e, s.cap_ops, s.d_found, s.d_meta_eff, s.d_ins_cnt, s.d_ins_base);
            return rc;
        }
        s.has_row_bufs = true;
    }
    s.rb.found = s.d_found;
    s.rb.meta_eff = s.d_meta_eff;
    s.rb.ins_cnt = s.d_ins_cnt;
    s.rb.ins_base = s.d_ins_base;
    s.rb.cursor = e->d_cursor;
    s.rb.inserted = s.d_counters + 1;
    return LTPMG_OK;
}

static void slot_uploaded(Slot &s, uint32_t n_txns, uint32_t n_ops) {
    s.n_txns = n_txns;
    s.n_ops = n_ops;
    s.has_batch = true;
    s.executed = false;
}

static int upload_locked(ltpmg_engine *e, Slot &s, bool trusted) {
    const uint32_t n_txns = s.staged_txns, n_ops = s.staged_ops;
    ltpmg_staging st = stage_ptrs(s, n_txns, n_ops);
    {
        // a trusted submit skips the per-op and per-txn checks, never the
        // structure the kernels derive their op ranges from
        ltpmg_batch b{n_txns, n_ops, st.txn_off, st.rows, st.meta, st.args};
        int rc = trusted ? validate_structure(e, &b, 1ull << 29) : validate_batch(e, &b);
        if (rc != LTPMG_OK) return rc;
        scan_row_ops(e, &b, s.rb, s.inserts);
        rc = slot_row_bufs(e, s);
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
    slot_uploaded(s, n_txns, n_ops);
    return LTPMG_OK;
}

// Runs of at least this many ops go to the device straight from the pool;
// shorter ones are gathered into the pinned staging buffer first, and each
// stretch of them goes up as one copy per array.
constexpr uint32_t kDirectRunOps = 4096;
#ifndef LTPMG_GATHER_MAX_OPS
#define LTPMG_GATHER_MAX_OPS ((1u << 29) - 1u)  // every batch (n_ops < 2^29)
#endif
constexpr uint32_t kGatherMaxOps = LTPMG_GATHER_MAX_OPS;

static int upload_runs_locked(ltpmg_engine *e, Slot &s, const ltpmg_batch *pool, const ltpmg_op_run *runs,
                              uint32_t n_runs) {
    const uint32_t n_txns = s.staged_txns, n_ops = s.staged_ops;
    ltpmg_staging st = stage_ptrs(s, n_txns, n_ops);
    // the runs tile the batch's ops in order and stay inside the pool
    uint64_t at = 0;
    for (uint32_t r = 0; r < n_runs; ++r) {
        if (runs[r].dst_op != at || runs[r].n_ops == 0 || (uint64_t)runs[r].src_op + runs[r].n_ops > pool->n_ops)
            return fail(e, LTPMG_EINVAL, "run %u: ops [%u, +%u) of the pool to op %u of the batch (expected op %llu)",
                        r, runs[r].src_op, runs[r].n_ops, runs[r].dst_op, (unsigned long long)at);
        at += runs[r].n_ops;
    }
    if (at != n_ops) return fail(e, LTPMG_EINVAL, "the runs hold %llu ops, %u staged", (unsigned long long)at, n_ops);
    {
        ltpmg_batch b{n_txns, n_ops, st.txn_off, st.rows, st.meta, st.args};  // only the offsets are read
        int rc = validate_structure(e, &b, 1ull << 29);
        if (rc != LTPMG_OK) return rc;
        s.rb = RowBufs{};
        std::fill(s.inserts, s.inserts + LTPMG_MAX_TABLES, 0u);
        for (uint32_t r = 0; r < n_runs && e->n_append; ++r) {
            const ltpmg_op_run &run = runs[r];
            ltpmg_batch v{0, run.n_ops, nullptr, pool->rows + run.src_op, pool->meta + run.src_op, nullptr};
            RowBufs one;
            uint32_t ins[LTPMG_MAX_TABLES];
            scan_row_ops(e, &v, one, ins);
            s.rb.has_append |= one.has_append;
            s.rb.has_delete |= one.has_delete;
            s.rb.insert_tables |= one.insert_tables;
            for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t) s.inserts[t] += ins[t];
        }
        s.rb.has_insert = s.rb.insert_tables != 0;
        rc = slot_row_bufs(e, s);
        if (rc != LTPMG_OK) return rc;
    }
    // A batch in a few runs out of a pool in pinned host memory is read by
    // one kernel straight from host memory, offsets included: no copy call.
    // The HIP runtime's host-to-device copy call had now and then blocked
    // the host for about 5 ms with the device idle (a few times per second
    // of serving, each putting that on the waiting requests' latency).  The
    // kernel also answers sooner (served at 10^8 txn/s: p50 0.41 against
    // 0.49 ms) and, reading 16 bytes per lane, carried 1.88 against 1.78 ·
    // 10^8 txn/s at 2 · 10^8 offered (p99 2.8 against 4.6 ms); only past
    // saturation (4 · 10^8 offered) do the copy engines' 53 GB/s carry
    // more (2.0-2.4 against 1.8 · 10^8 txn/s).
    const void *pv[3] = {pool->rows, pool->meta, pool->args};
    bool mapped = n_runs <= kGatherRuns && n_ops <= kGatherMaxOps && s.stage_dev != nullptr;
    for (int k = 0; k < 3 && mapped; ++k) {
        if (s.pool_host[k] != pv[k]) {
            void *d = nullptr;
            if (hipHostGetDevicePointer(&d, const_cast<void *>(pv[k]), 0) != hipSuccess) {
                (void)hipGetLastError();
                d = nullptr;
            }
            s.pool_host[k] = pv[k];
            s.pool_dev[k] = d;
        }
        mapped = s.pool_dev[k] != nullptr;
    }
    if (mapped) {
        GatherRuns g{static_cast<const uint32_t *>(s.pool_dev[0]), static_cast<const uint16_t *>(s.pool_dev[1]),
                     static_cast<const uint64_t *>(s.pool_dev[2]),
                     reinterpret_cast<const uint32_t *>(s.stage_dev + (reinterpret_cast<uint8_t *>(st.txn_off) - s.h_stage)),
                     {}, {}, n_runs, n_ops, n_txns};
        for (uint32_t r = 0; r < n_runs; ++r) {
            g.src[r] = runs[r].src_op;
            g.dst[r] = runs[r].dst_op;
        }
        HIP_TRY(e, hipEventRecord(s.up0, s.stream));
        launch_gather_runs(s.stream, g, s.d_txn_off, s.d_rows, s.d_meta, s.d_args);
        HIP_TRY(e, hipGetLastError());
        HIP_TRY(e, hipEventRecord(s.up1, s.stream));
        slot_uploaded(s, n_txns, n_ops);
        return LTPMG_OK;
    }
    HIP_TRY(e, hipEventRecord(s.up0, s.stream));
    HIP_TRY(e, hipMemcpyAsync(s.d_txn_off, st.txn_off, ((size_t)n_txns + 1) * 4, hipMemcpyHostToDevice, s.stream));
    auto up = [&](const uint32_t *rows, const uint16_t *meta, const uint64_t *args, uint32_t dst, uint32_t n) -> int {
        HIP_TRY(e, hipMemcpyAsync(s.d_args + dst, args, (size_t)n * 8, hipMemcpyHostToDevice, s.stream));
        HIP_TRY(e, hipMemcpyAsync(s.d_rows + dst, rows, (size_t)n * 4, hipMemcpyHostToDevice, s.stream));
        HIP_TRY(e, hipMemcpyAsync(s.d_meta + dst, meta, (size_t)n * 2, hipMemcpyHostToDevice, s.stream));
        return LTPMG_OK;
    };
    uint32_t packed0 = 0, packed1 = 0;  // the pending stretch of short runs in staging
    auto flush = [&]() -> int {
        int rc = LTPMG_OK;
        if (packed1 > packed0)
            rc = up(st.rows + packed0, st.meta + packed0, st.args + packed0, packed0, packed1 - packed0);
        packed0 = packed1;
        return rc;
    };
    for (uint32_t r = 0; r < n_runs; ++r) {
        const ltpmg_op_run &run = runs[r];
        int rc = LTPMG_OK;
        if (run.n_ops >= kDirectRunOps) {
            rc = flush();
            if (rc == LTPMG_OK)
                rc = up(pool->rows + run.src_op, pool->meta + run.src_op, pool->args + run.src_op, run.dst_op, run.n_ops);
            packed0 = packed1 = run.dst_op + run.n_ops;
        } else {
            std::memcpy(st.rows + run.dst_op, pool->rows + run.src_op, (size_t)run.n_ops * 4);
            std::memcpy(st.meta + run.dst_op, pool->meta + run.src_op, (size_t)run.n_ops * 2);
            std::memcpy(st.args + run.dst_op, pool->args + run.src_op, (size_t)run.n_ops * 8);
            packed1 = run.dst_op + run.n_ops;
        }
        if (rc != LTPMG_OK) return rc;
    }
    int rc = flush();
    if (rc != LTPMG_OK) return rc;
    HIP_TRY(e, hipEventRecord(s.up1, s.stream));
    slot_uploaded(s, n_txns, n_ops);
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

int ltpmg_submit_runs(ltpmg_engine *e, uint32_t slot, const ltpmg_batch *pool, const ltpmg_op_run *runs,
                      uint32_t n_runs) {
    if (!e || !pool || (n_runs && !runs) || (pool->n_ops && (!pool->rows || !pool->meta || !pool->args)))
        return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    Slot *s = get_slot(e, slot);
    if (!s) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    if (s->staged_txns == UINT32_MAX) return fail(e, LTPMG_ESTATE, "nothing staged in slot %u", slot);
    int rc = upload_runs_locked(e, *s, pool, runs, n_runs);
    s->staged_txns = UINT32_MAX;
    return rc;
}

int ltpmg_pinned_alloc(uint64_t bytes, void **out) {
    if (!out) return LTPMG_EINVAL;
    *out = nullptr;
    if (hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return LTPMG_ENOMEM;
    }
    return LTPMG_OK;
}

int ltpmg_pinned_free(void *p) {
    if (p && hipHostFree(p) != hipSuccess) return LTPMG_EHIP;
    return LTPMG_OK;
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
    const uint32_t small_max = (mode & LTPMG_EXEC_PIPELINE) ? 0u : kSmallMaxOps;
    mode &= ~(LTPMG_EXEC_PROFILE | LTPMG_EXEC_PIPELINE);
    if (mode != LTPMG_MODE_A && mode != LTPMG_MODE_O) return fail(e, LTPMG_EINVAL, "mode %u", mode);
    const uint32_t rs = max_slot_all(e);
    if (rs > s.cap_res_stride) return fail(e, LTPMG_ESTATE, "tables changed since the batch was submitted");
    OpView ov{s.d_txn_off, s.d_rows, s.d_meta, s.d_args, s.n_txns, s.n_ops};
    RowBufs *rb = s.rb.has_append ? &s.rb : nullptr;
    if (rb) {
        int rc = check_capacity(e, s.inserts);
        if (rc != LTPMG_OK) return rc;
    }
    StageRec rec{e, s, profile};
    s.n_stages = 0;
    s.mode = mode;
    HIP_TRY(e, hipEventRecord(s.x0, s.stream));
    // table rows and lock words: after the previous batch, whatever its slot
    auto wait_tail = [&]() -> int {
        if (e->tail_slot >= 0 && e->tail_slot != (int)slot)
            HIP_TRY(e, hipStreamWaitEvent(s.stream, e->slots[e->tail_slot].done, 0));
        return LTPMG_OK;
    };
    int rc = mode == LTPMG_MODE_A
                 ? launch_mode_a(e, s.stream, ov, s.d_tile_t0, s.d_counters, s.d_commit, s.d_results, rs,
                                 rec, wait_tail, rb)
                 : launch_mode_o(e, s.stream, ov, s.ob, s.d_commit, s.d_results, rs, rec, wait_tail, rb,
                                 small_max);
    if (rc != LTPMG_OK) return rc;
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
        int rc = zero_result_padding(e, s->d_meta, s->n_ops, static_cast<uint8_t *>(results), max_slot_all(e));
        if (rc != LTPMG_OK) return rc;
    }
    return LTPMG_OK;
}

int ltpmg_fetch_status(ltpmg_engine *e, uint32_t slot, uint8_t *status) {
    if (!e) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    Slot *s = get_slot(e, slot);
    if (!s) return fail(e, LTPMG_EINVAL, "slot %u out of range", slot);
    if (!s->executed) return fail(e, LTPMG_ESTATE, "slot %u has not executed a batch", slot);
    if (!s->n_ops) return LTPMG_OK;
    if (!status) return fail(e, LTPMG_EINVAL, "null status buffer");
    HIP_TRY(e, hipStreamSynchronize(s->stream));
    // the batch's offsets are still in the slot's staging buffer
    const uint32_t *off = stage_ptrs(*s, s->n_txns, s->n_ops).txn_off;
    std::vector<uint8_t> commit(s->n_txns);
    HIP_TRY(e, hipMemcpy(commit.data(), s->d_commit, s->n_txns, hipMemcpyDeviceToHost));
    if (s->rb.has_append) {
        HIP_TRY(e, hipMemcpy(status, s->d_found, s->n_ops, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < s->n_ops; ++i) status[i] = status[i] != 0;  // an INSERT's byte is 2 + its table
    } else {
        std::memset(status, 1, s->n_ops);
    }
    for (uint32_t t = 0; t < s->n_txns; ++t)
        if (!commit[t]) std::memset(status + off[t], 0, off[t + 1] - off[t]);
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
    if (s->rb.has_insert) HIP_TRY(e, hipMemcpy(&out->inserted, s->d_counters + 1, 4, hipMemcpyDeviceToHost));
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
    int rc = LTPMG_OK;
    for (int i = 0; i < 2 && rc == LTPMG_OK; ++i) {
        rc = dev_alloc_n(e, &b.keys[i], std::max<size_t>(n, 1));
        if (rc == LTPMG_OK) rc = dev_alloc_n(e, &b.vals[i], std::max<size_t>(n, 1));
    }
    if (rc == LTPMG_OK) rc = dev_alloc_n(e, &b.hist, sort_hist_words(n));
    if (rc == LTPMG_OK)
        rc = dev_alloc_n(e, &b.partials, sort_partials_needed());
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
    dev_free(e, b.hist, sort_hist_words(n) * 4);
    dev_free(e, b.partials, 4 * sort_partials_needed());
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
    const size_t nn = std::max<size_t>(n, 1), scr = 3ull * (seg_tiles(n) + 1);
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

// The allocation-free execute (include/ltpmg.h): the kernels of
// ltpmg_execute on a device-resident batch, their scratch carved from this
// device's static array.  Only kernel launches are issued (counters and the
// final-write list are cleared by kernels), no events are recorded and no
// memory is allocated or freed.
extern "C" int ltpmg_execute_static(ltpmg_engine *e, uint32_t mode, const ltpmg_batch *dev, uint8_t *dev_commit,
                                    void *dev_results, void *hip_stream) {
    if (!e || !dev) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (mode != LTPMG_MODE_A && mode != LTPMG_MODE_O) return fail(e, LTPMG_EINVAL, "mode %u", mode);
    if (dev->n_ops > LTPMG_STATIC_MAX_OPS)
        return fail(e, LTPMG_EINVAL, "batch of %u ops exceeds the %u ops the static scratch holds", dev->n_ops,
                    LTPMG_STATIC_MAX_OPS);
    if (dev->n_txns > dev->n_ops)
        return fail(e, LTPMG_EINVAL, "%u txns but %u ops (every txn has at least one op)", dev->n_txns, dev->n_ops);
    if (dev->n_ops == 0) return LTPMG_OK;
    if (!dev->txn_off || !dev->rows || !dev->meta || !dev->args || !dev_commit || !dev_results)
        return fail(e, LTPMG_EINVAL, "null device array");
    if (!e->static_scratch) return fail(e, LTPMG_ESTATE, "no static scratch on device %d", e->device);
    if (e->n_append)
        return fail(e, LTPMG_ESTATE, "append tables need ltpmg_execute (their row ops use per-slot buffers)");
    StaticBufs sb;
    if (!carve_static(e->static_scratch, sb))
        return fail(e, LTPMG_ESTATE, "the static scratch is too small for this build's sort tiles");
    const uint32_t rs = max_slot_all(e);
    OpView ov{dev->txn_off, dev->rows, dev->meta, dev->args, dev->n_txns, dev->n_ops};
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    NoRec rec;
    auto now = []() { return LTPMG_OK; };
    uint8_t *res = static_cast<uint8_t *>(dev_results);
    int rc = mode == LTPMG_MODE_A ? launch_mode_a(e, st, ov, sb.tile_t0, sb.counters, dev_commit, res, rs, rec, now)
                                  : launch_mode_o(e, st, ov, sb.ob, dev_commit, res, rs, rec, now, nullptr, 0u,
                                                  true);
    if (rc != LTPMG_OK) return rc;
    HIP_TRY(e, hipGetLastError());
    return LTPMG_OK;
}

// ---------------- batching form (include/ltpmg_stream.h) ----------------

struct ltpmg_stream {
    uint32_t n_batches = 0;
    uint32_t flags = 0;
    std::vector<uint32_t> txn_base, op_base;  // cumulative, n_batches + 1 entries
    std::vector<uint8_t> modes;
    uint32_t max_batch_ops = 0;
    uint32_t res_stride = 0;
    // the pool: batch b's own txn offsets (starting at 0) are at
    // d_txn_off + txn_base[b] + b; its ops at op_base[b]
    uint32_t *d_txn_off = nullptr, *d_rows = nullptr;
    uint16_t *d_meta = nullptr;
    uint64_t *d_args = nullptr;
    uint8_t *d_commit = nullptr;  // per txn of the stream
    // scratch of one launch group
    uint8_t *d_results = nullptr;
    uint64_t res_cap_ops = 0;
    uint32_t *d_tile_t0 = nullptr, *d_counters = nullptr;
    uint64_t tile_cap_ops = 0;
    OrderBufs ob{};
    uint64_t ob_cap_ops = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> gev;  // per-group marks of a timed (tuning) run
    hipGraphExec_t graph = nullptr;
    ltpmg_stream_plan graph_plan{};
    bool tuned = false, ran = false;
    ltpmg_stream_plan tuned_plan{};
    // row ops per batch (scan_row_ops), the INSERTs of the whole stream per
    // table, and the row-op scratch of one launch group
    std::vector<RowBufs> rows;
    uint32_t stream_inserts[LTPMG_MAX_TABLES] = {};
    uint32_t rewind_rows[LTPMG_MAX_TABLES] = {};  // LTPMG_STREAM_REWIND: cursors at creation
    // per batch and append table: one past the highest row its non-INSERT ops name
    std::vector<std::array<uint32_t, LTPMG_MAX_TABLES>> row_end;
    std::vector<uint64_t> graph_groups;  // the grouping the captured graph was made for
    uint8_t *d_found = nullptr;
    uint16_t *d_meta_eff = nullptr;
    uint32_t *d_ins_cnt = nullptr, *d_ins_base = nullptr;
    uint64_t row_cap_ops = 0;
};

namespace {

struct Group {
    uint32_t b0, b1, mode;
    bool fused;
    uint32_t grid;
    RowBufs rb;  // the row ops of batches [b0, b1) together
};

const uint32_t kTiles[3] = {1024, 2048, 4096};

bool plan_ok(const ltpmg_stream_plan &p) {
    bool tile = false;
    for (uint32_t t : kTiles) tile |= p.fused_tile_ops == t;
    return tile && p.fused_wgs_per_cu >= 1 && p.fused_wgs_per_cu <= 16 && p.merge_max_ops < (1u << 29) &&
           (p.flags & ~LTPMG_PLAN_SYNC_EACH) == 0;
}

// Workgroups of the fused kernel that fit on the device at once.
uint32_t fused_grid_cap(ltpmg_engine *e, uint32_t tile, uint32_t wgs_per_cu, uint32_t rs) {
    int per_cu = 0;
    const void *fn = cc_fused_kernel(tile, rs);
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0) != hipSuccess || per_cu < 1)
        return 0;
    return (uint32_t)e->n_cu * std::min<uint32_t>(wgs_per_cu, (uint32_t)per_cu);
}

// cursor: each append table's cursor when the run starts.
std::vector<Group> plan_groups(ltpmg_engine *e, const ltpmg_stream &s, const ltpmg_stream_plan &p,
                               const uint32_t cursor[LTPMG_MAX_TABLES]) {
    std::vector<Group> g;
    const uint32_t cap = fused_grid_cap(e, p.fused_tile_ops, p.fused_wgs_per_cu, s.res_stride);
    const bool each = (p.flags & LTPMG_PLAN_SYNC_EACH) != 0;
    for (uint32_t b = 0; b < s.n_batches; ++b) {
        const uint32_t ops = s.op_base[b + 1] - s.op_base[b];
        const RowBufs &rb = s.rows[b];
        if (s.modes[b] == LTPMG_MODE_O) {
            // A merged group runs as one ordered batch, which is the same as
            // its batches one by one -- except that rows an INSERT appends
            // become visible only after the group.  So a batch joins a group
            // with INSERTs into table t only if its other ops on t name rows
            // below t's cursor at the start of the run, which no INSERT of
            // the run creates.
            bool rows_ok = true;
            if (!g.empty())
                for (uint32_t m = g.back().rb.insert_tables; m; m &= m - 1)
                    rows_ok &= s.row_end[b][__builtin_ctz(m)] <= cursor[__builtin_ctz(m)];
            if (!each && !g.empty() && g.back().mode == LTPMG_MODE_O && rows_ok &&
                (uint64_t)s.op_base[b + 1] - s.op_base[g.back().b0] <= p.merge_max_ops) {
                Group &last = g.back();
                last.b1 = b + 1;
                last.rb.has_append |= rb.has_append;
                last.rb.has_delete |= rb.has_delete;
                last.rb.insert_tables |= rb.insert_tables;
                last.rb.has_insert |= rb.has_insert;
                continue;
            }
            g.push_back(Group{b, b + 1, LTPMG_MODE_O, false, 0, rb});
        } else {
            // the fused kernel has no row-op stages
            const uint32_t n_tiles = (ops + p.fused_tile_ops - 1) / p.fused_tile_ops;
            const bool fused = cap && !rb.has_append && ops <= p.fused_max_ops && n_tiles <= (uint64_t)kFusedMaxTiles * cap;
            g.push_back(Group{b, b + 1, LTPMG_MODE_A, fused, fused ? cc_fused_grid(ops, p.fused_tile_ops, cap) : 0, rb});
        }
    }
    return g;
}

void stream_free_scratch(ltpmg_engine *e, ltpmg_stream &s) {
    if (!(s.flags & LTPMG_STREAM_KEEP_RESULTS)) {
        dev_free(e, s.d_results, s.res_cap_ops * s.res_stride);
        s.d_results = nullptr;
        s.res_cap_ops = 0;
    }
    dev_free(e, s.d_tile_t0, (s.tile_cap_ops / kTileOps + 2) * 4);
    s.d_tile_t0 = nullptr;
    s.tile_cap_ops = 0;
    free_order_bufs(e, s.ob, s.ob_cap_ops);
    s.ob_cap_ops = 0;
    if (s.row_cap_ops) free_row_bufs(e, s.row_cap_ops, s.d_found, s.d_meta_eff, s.d_ins_cnt, s.d_ins_base);
    s.row_cap_ops = 0;
}

// Scratch for the largest group of the plan.
int stream_scratch(ltpmg_engine *e, ltpmg_stream &s, const std::vector<Group> &groups) {
    uint64_t max_o = 0, max_a = 0, max_any = 0, max_row = 0;
    for (const Group &g : groups) {
        const uint64_t ops = (uint64_t)s.op_base[g.b1] - s.op_base[g.b0];
        max_any = std::max(max_any, ops);
        if (g.mode == LTPMG_MODE_O) max_o = std::max(max_o, ops);
        else if (!g.fused) max_a = std::max(max_a, ops);
        if (g.rb.has_append) max_row = std::max(max_row, ops);
    }
    if (max_row > s.row_cap_ops) {
        if (s.row_cap_ops) free_row_bufs(e, s.row_cap_ops, s.d_found, s.d_meta_eff, s.d_ins_cnt, s.d_ins_base);
        s.row_cap_ops = 0;
        int rc = alloc_row_bufs(e, max_row, s.d_found, s.d_meta_eff, s.d_ins_cnt, s.d_ins_base);
        if (rc != LTPMG_OK) {
            free_row_bufs(e, max_row, s.d_found, s.d_meta_eff, s.d_ins_cnt, s.d_ins_base);
            return rc;
        }
        s.row_cap_ops = max_row;
    }
    if (!(s.flags & LTPMG_STREAM_KEEP_RESULTS) && max_any > s.res_cap_ops) {
        dev_free(e, s.d_results, s.res_cap_ops * s.res_stride);
        s.d_results = nullptr;
        s.res_cap_ops = 0;
        int rc = dev_alloc(e, reinterpret_cast<void **>(&s.d_results), std::max<uint64_t>(max_any, 1) * s.res_stride);
        if (rc != LTPMG_OK) return rc;
        s.res_cap_ops = std::max<uint64_t>(max_any, 1);
    }
    if (max_a > s.tile_cap_ops || !s.d_tile_t0) {
        dev_free(e, s.d_tile_t0, (s.tile_cap_ops / kTileOps + 2) * 4);
        s.d_tile_t0 = nullptr;
        s.tile_cap_ops = 0;
        int rc = dev_alloc_n(e, &s.d_tile_t0, max_a / kTileOps + 2);
        if (rc != LTPMG_OK) return rc;
        s.tile_cap_ops = max_a;
    }
    if (max_o > s.ob_cap_ops) {
        free_order_bufs(e, s.ob, s.ob_cap_ops);
        s.ob_cap_ops = 0;
        int rc = alloc_order_bufs(e, s.ob, max_o);
        if (rc != LTPMG_OK) return rc;
        s.ob_cap_ops = max_o;
    }
    return LTPMG_OK;
}

int issue_group(ltpmg_engine *e, ltpmg_stream &s, const Group &g, const ltpmg_stream_plan &p) {
    const uint32_t op0 = s.op_base[g.b0], t0 = s.txn_base[g.b0];
    const uint32_t ops = s.op_base[g.b1] - op0, txns = s.txn_base[g.b1] - t0;
    const uint32_t rs = s.res_stride;
    uint8_t *results = (s.flags & LTPMG_STREAM_KEEP_RESULTS) ? s.d_results + (uint64_t)op0 * rs : s.d_results;
    OpView ov{s.d_txn_off + t0 + g.b0, s.d_rows + op0, s.d_meta + op0, s.d_args + op0, txns, ops};
    NoRec rec;
    auto now = []() { return LTPMG_OK; };
    RowBufs rb = g.rb;
    rb.found = s.d_found;
    rb.meta_eff = s.d_meta_eff;
    rb.ins_cnt = s.d_ins_cnt;
    rb.ins_base = s.d_ins_base;
    rb.cursor = e->d_cursor;
    rb.inserted = s.d_counters + 1;
    RowBufs *rbp = g.rb.has_append ? &rb : nullptr;
    if (g.mode == LTPMG_MODE_O)
        return launch_mode_o(e, s.stream, ov, s.ob, s.d_commit + t0, results, rs, rec, now, rbp, p.small_max_ops);
    if (!g.fused)
        return launch_mode_a(e, s.stream, ov, s.d_tile_t0, s.d_counters, s.d_commit + t0, results, rs, rec, now, rbp);
    GridBar bar{e->d_sync + 1, e->d_sync + 2, e->d_sync + 3, 1u << 24};
    DevTables dt = dev_tables(e);
    HIP_TRY(e, launch_cc_fused(s.stream, dt, ov, p.fused_tile_ops, g.grid, e->d_sync, s.d_commit + t0, results,
                               rs, s.d_counters, bar));
    return LTPMG_OK;
}

// Reads (and clears) the grid-barrier timeout flag after the stream is idle.
int check_barrier(ltpmg_engine *e) {
    uint32_t w[4] = {};
    HIP_TRY(e, hipMemcpy(w, e->d_sync, sizeof w, hipMemcpyDeviceToHost));
    if (w[3] == 0) return LTPMG_OK;
    // The failed launch may have tagged lock words with epoch + 1 (its own)
    // without storing that epoch back; move past it so the next batch does
    // not read those tags as live reservations.
    w[0] += 2;
    w[1] = w[2] = w[3] = 0;
    HIP_TRY(e, hipMemcpy(e->d_sync, w, sizeof w, hipMemcpyHostToDevice));
    return fail(e, LTPMG_EHIP, "a fused kernel's grid barrier timed out (grid not co-resident); tables are undefined");
}

void stream_destroy_locked(ltpmg_engine *e, ltpmg_stream *s);

void drop_graph(ltpmg_stream &s) {
    if (s.graph) (void)hipGraphExecDestroy(s.graph);
    s.graph = nullptr;
}

bool same_plan(const ltpmg_stream_plan &a, const ltpmg_stream_plan &b) {
    return std::memcmp(&a, &b, sizeof a) == 0;
}

// One run of the stream; marks (if given) gets per-group device times.
int stream_run_locked(ltpmg_engine *e, ltpmg_stream &s, const ltpmg_stream_plan &p, double *ms,
                      std::vector<float> *marks) {
    if (!plan_ok(p)) return fail(e, LTPMG_EINVAL, "bad stream plan");
    if (max_slot_all(e) > s.res_stride) return fail(e, LTPMG_ESTATE, "tables changed since the stream was created");
    for (Slot &sl : e->slots) HIP_TRY(e, hipStreamSynchronize(sl.stream));
    e->tail_slot = -1;
    uint32_t cur[LTPMG_MAX_TABLES];  // the append cursors at the start of the run
    {  // every INSERT of the stream must fit (the run is synchronous: exact cursors)
        HIP_TRY(e, hipMemcpy(cur, e->d_cursor, sizeof cur, hipMemcpyDeviceToHost));
        for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t) {
            Table &tb = e->tables[t];
            if (!s.stream_inserts[t]) continue;
            if ((s.flags & LTPMG_STREAM_REWIND) && cur[t] != s.rewind_rows[t]) {
                if (s.rewind_rows[t] > tb.n_rows)
                    return fail(e, LTPMG_ESTATE, "table %u changed since the stream was created", t);
                int rc = truncate_locked(e, t, s.rewind_rows[t]);
                if (rc != LTPMG_OK) return rc;
                cur[t] = s.rewind_rows[t];
            }
            if ((uint64_t)cur[t] + s.stream_inserts[t] > tb.n_rows)
                return fail(e, LTPMG_EFULL, "table %u: the stream inserts up to %u rows, %u of %u rows are left", t,
                            s.stream_inserts[t], tb.n_rows - cur[t], tb.n_rows);
            tb.cursor_bound = (uint64_t)cur[t] + s.stream_inserts[t];
        }
    }
    std::vector<Group> groups = plan_groups(e, s, p, cur);
    int rc = stream_scratch(e, s, groups);
    if (rc != LTPMG_OK) return rc;
    const bool each = (p.flags & LTPMG_PLAN_SYNC_EACH) != 0;
    if (p.use_graph && !each && !marks) {
        std::vector<uint64_t> sig;  // the grouping can change with the cursors
        for (const Group &g : groups) sig.push_back((uint64_t)g.b0 << 1 | (g.fused ? 1u : 0u));
        if (!s.graph || !same_plan(s.graph_plan, p) || sig != s.graph_groups) {
            drop_graph(s);
            HIP_TRY(e, hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
            for (const Group &g : groups) {
                rc = issue_group(e, s, g, p);
                if (rc != LTPMG_OK) break;
            }
            hipGraph_t gr = nullptr;
            hipError_t st = hipStreamEndCapture(s.stream, &gr);
            if (rc != LTPMG_OK) {
                if (gr) (void)hipGraphDestroy(gr);
                return rc;
            }
            if (st != hipSuccess) return fail(e, LTPMG_EHIP, "graph capture failed: %s", hipGetErrorString(st));
            st = hipGraphInstantiate(&s.graph, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (st != hipSuccess) {
                s.graph = nullptr;
                return fail(e, LTPMG_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(st));
            }
            s.graph_plan = p;
            s.graph_groups = sig;
        }
        HIP_TRY(e, hipEventRecord(s.ev0, s.stream));
        HIP_TRY(e, hipGraphLaunch(s.graph, s.stream));
        HIP_TRY(e, hipEventRecord(s.ev1, s.stream));
    } else {
        if (marks)
            while (s.gev.size() < groups.size() + 1) {
                hipEvent_t ev = nullptr;
                HIP_TRY(e, hipEventCreate(&ev));
                s.gev.push_back(ev);
            }
        HIP_TRY(e, hipEventRecord(s.ev0, s.stream));
        for (size_t i = 0; i < groups.size(); ++i) {
            if (marks) HIP_TRY(e, hipEventRecord(s.gev[i], s.stream));
            rc = issue_group(e, s, groups[i], p);
            if (rc != LTPMG_OK) return rc;
            if (each) HIP_TRY(e, hipStreamSynchronize(s.stream));
        }
        if (marks) HIP_TRY(e, hipEventRecord(s.gev[groups.size()], s.stream));
        HIP_TRY(e, hipEventRecord(s.ev1, s.stream));
    }
    HIP_TRY(e, hipGetLastError());
    HIP_TRY(e, hipStreamSynchronize(s.stream));
    rc = check_barrier(e);
    if (rc != LTPMG_OK) return rc;
    s.ran = true;
    float f = 0.f;
    HIP_TRY(e, hipEventElapsedTime(&f, s.ev0, s.ev1));
    if (ms) *ms = f;
    if (marks) {
        marks->assign(s.n_batches, 0.f);
        for (size_t i = 0; i < groups.size(); ++i) {
            float g = 0.f;
            HIP_TRY(e, hipEventElapsedTime(&g, s.gev[i], s.gev[i + 1]));
            (*marks)[groups[i].b0] = g;  // tuning runs cut every batch alone
        }
    }
    return LTPMG_OK;
}

}  // namespace

extern "C" {

void ltpmg_stream_default_plan(ltpmg_stream_plan *p) {
    if (!p) return;
    *p = ltpmg_stream_plan{1u << 24, 1u << 18, 2048, 2, 1, 0, kSmallMaxOps};
}

int ltpmg_stream_create(ltpmg_engine *e, uint32_t n_batches, const uint32_t *batch_txns, const uint8_t *modes,
                        const ltpmg_batch *all, uint32_t flags, ltpmg_stream **out) {
    if (!e || !out || !batch_txns || !modes || !all || n_batches == 0) return LTPMG_EINVAL;
    *out = nullptr;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (flags & ~(LTPMG_STREAM_KEEP_RESULTS | LTPMG_STREAM_REWIND)) return fail(e, LTPMG_EINVAL, "stream flags 0x%x", flags);
    int rc = validate_batch(e, all, 0xFFFFFFFFull);
    if (rc != LTPMG_OK) return rc;
    ltpmg_stream *s = new ltpmg_stream();
    s->n_batches = n_batches;
    s->flags = flags;
    s->modes.assign(modes, modes + n_batches);
    s->txn_base.assign(n_batches + 1, 0);
    s->op_base.assign(n_batches + 1, 0);
    uint64_t t = 0;
    for (uint32_t b = 0; b < n_batches; ++b) {
        if (batch_txns[b] == 0 || modes[b] > LTPMG_MODE_O) {
            delete s;
            return fail(e, LTPMG_EINVAL, "batch %u: %u txns, mode %u", b, batch_txns[b], modes[b]);
        }
        t += batch_txns[b];
        if (t > all->n_txns) {
            delete s;
            return fail(e, LTPMG_EINVAL, "batch sizes add up to more than the %u txns given", all->n_txns);
        }
        s->txn_base[b + 1] = (uint32_t)t;
        s->op_base[b + 1] = all->txn_off[t];
        const uint32_t ops = s->op_base[b + 1] - s->op_base[b];
        if (ops >= (1u << 29)) {
            delete s;
            return fail(e, LTPMG_EINVAL, "batch %u has %u ops (must be < 2^29)", b, ops);
        }
        s->max_batch_ops = std::max(s->max_batch_ops, ops);
    }
    if (t != all->n_txns) {
        delete s;
        return fail(e, LTPMG_EINVAL, "batch sizes add up to %llu txns, %u given", (unsigned long long)t, all->n_txns);
    }
    s->res_stride = max_slot_all(e);
    s->rows.resize(n_batches);
    s->row_end.resize(n_batches);
    for (uint32_t b = 0; b < n_batches; ++b) {
        const uint32_t o0 = s->op_base[b];
        ltpmg_batch one{s->txn_base[b + 1] - s->txn_base[b], s->op_base[b + 1] - o0, nullptr,
                        all->rows + o0, all->meta + o0, all->args + o0};
        uint32_t ins[LTPMG_MAX_TABLES];
        scan_row_ops(e, &one, s->rows[b], ins, s->row_end[b].data());
        for (uint32_t t = 0; t < LTPMG_MAX_TABLES; ++t) s->stream_inserts[t] += ins[t];
    }
    if (flags & LTPMG_STREAM_REWIND) {  // the cursors every run starts from
        for (Slot &sl : e->slots) {
            if (hipStreamSynchronize(sl.stream) != hipSuccess) {
                delete s;
                return fail(e, LTPMG_EHIP, "hipStreamSynchronize failed");
            }
        }
        if (hipMemcpy(s->rewind_rows, e->d_cursor, sizeof s->rewind_rows, hipMemcpyDeviceToHost) != hipSuccess) {
            delete s;
            return fail(e, LTPMG_EHIP, "reading the append cursors failed");
        }
    }
    const uint64_t n_ops = all->n_ops, n_txns = all->n_txns;
    // per-batch offsets, each batch starting at 0
    std::vector<uint32_t> rel(n_txns + n_batches);
    for (uint32_t b = 0; b < n_batches; ++b) {
        const uint32_t base = all->txn_off[s->txn_base[b]];
        for (uint32_t k = s->txn_base[b]; k <= s->txn_base[b + 1]; ++k) rel[k + b] = all->txn_off[k] - base;
    }
#define ALLOC(ptr, n) if (rc == LTPMG_OK) rc = dev_alloc_n(e, &(ptr), (n))
    ALLOC(s->d_txn_off, rel.size());
    ALLOC(s->d_rows, std::max<uint64_t>(n_ops, 1));
    ALLOC(s->d_meta, std::max<uint64_t>(n_ops, 1));
    ALLOC(s->d_args, std::max<uint64_t>(n_ops, 1));
    ALLOC(s->d_commit, n_txns);
    ALLOC(s->d_counters, 4);
    if (flags & LTPMG_STREAM_KEEP_RESULTS) {
        ALLOC(s->d_results, std::max<uint64_t>(n_ops, 1) * s->res_stride);
        s->res_cap_ops = std::max<uint64_t>(n_ops, 1);
    }
#undef ALLOC
    if (rc == LTPMG_OK) {
        hipError_t st = hipMemcpy(s->d_txn_off, rel.data(), rel.size() * 4, hipMemcpyHostToDevice);
        if (st == hipSuccess && n_ops) st = hipMemcpy(s->d_rows, all->rows, n_ops * 4, hipMemcpyHostToDevice);
        if (st == hipSuccess && n_ops) st = hipMemcpy(s->d_meta, all->meta, n_ops * 2, hipMemcpyHostToDevice);
        if (st == hipSuccess && n_ops) st = hipMemcpy(s->d_args, all->args, n_ops * 8, hipMemcpyHostToDevice);
        if (st == hipSuccess) st = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (st == hipSuccess) st = hipEventCreate(&s->ev0);
        if (st == hipSuccess) st = hipEventCreate(&s->ev1);
        if (st != hipSuccess) rc = fail(e, LTPMG_EHIP, "stream upload failed: %s", hipGetErrorString(st));
    }
    if (rc != LTPMG_OK) {
        stream_destroy_locked(e, s);
        return rc;
    }
    *out = s;
    return LTPMG_OK;
}

int ltpmg_stream_destroy(ltpmg_engine *e, ltpmg_stream *s) {
    if (!e || !s) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    (void)hipSetDevice(e->device);
    stream_destroy_locked(e, s);
    return LTPMG_OK;
}

}  // extern "C"

namespace {
void stream_destroy_locked(ltpmg_engine *e, ltpmg_stream *s) {
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graph(*s);
    stream_free_scratch(e, *s);
    const uint64_t n_ops = s->op_base.empty() ? 0 : s->op_base.back();
    const uint64_t n_txns = s->txn_base.empty() ? 0 : s->txn_base.back();
    dev_free(e, s->d_txn_off, (n_txns + s->n_batches) * 4);
    dev_free(e, s->d_rows, std::max<uint64_t>(n_ops, 1) * 4);
    dev_free(e, s->d_meta, std::max<uint64_t>(n_ops, 1) * 2);
    dev_free(e, s->d_args, std::max<uint64_t>(n_ops, 1) * 8);
    dev_free(e, s->d_commit, n_txns);
    dev_free(e, s->d_counters, 16);
    if (s->flags & LTPMG_STREAM_KEEP_RESULTS) dev_free(e, s->d_results, s->res_cap_ops * s->res_stride);
    for (hipEvent_t ev : s->gev) (void)hipEventDestroy(ev);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}
}  // namespace

extern "C" {

int ltpmg_stream_run(ltpmg_engine *e, ltpmg_stream *s, const ltpmg_stream_plan *plan, double *device_ms) {
    if (!e || !s) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    ltpmg_stream_plan p;
    if (plan) p = *plan;
    else if (s->tuned) p = s->tuned_plan;
    else ltpmg_stream_default_plan(&p);
    return stream_run_locked(e, *s, p, device_ms, nullptr);
}

int ltpmg_stream_tune(ltpmg_engine *e, ltpmg_stream *s, ltpmg_stream_plan *best, uint32_t *trials_out) {
    if (!e || !s) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    ltpmg_stream_plan p;
    ltpmg_stream_default_plan(&p);
    uint32_t trials = 0;
    bool any_a = false, any_o = false;
    for (uint8_t m : s->modes) (m == LTPMG_MODE_A ? any_a : any_o) = true;
    int rc = LTPMG_OK;
    if (any_a) {
        // per-batch times of the three-kernel pipeline and of each fused
        // configuration; then the size cut that minimises their sum
        std::vector<float> multi, t;
        ltpmg_stream_plan q = p;
        q.fused_max_ops = 0;
        q.merge_max_ops = 0;
        q.use_graph = 0;
        rc = stream_run_locked(e, *s, q, nullptr, &multi);
        if (rc != LTPMG_OK) return rc;
        ++trials;
        std::vector<uint32_t> a_ops;
        for (uint32_t b = 0; b < s->n_batches; ++b)
            if (s->modes[b] == LTPMG_MODE_A) a_ops.push_back(s->op_base[b + 1] - s->op_base[b]);
        std::sort(a_ops.begin(), a_ops.end());
        double best_ms = 0;
        for (uint32_t b = 0; b < s->n_batches; ++b)
            if (s->modes[b] == LTPMG_MODE_A) best_ms += multi[b];
        p.fused_max_ops = 0;
        for (uint32_t tile : kTiles) {
            uint32_t last_cap = 0;
            for (uint32_t wgs : {1u, 2u, 4u, 8u}) {
                const uint32_t cap = fused_grid_cap(e, tile, wgs, s->res_stride);
                if (cap == 0 || cap == last_cap) continue;  // occupancy-clamped repeat
                last_cap = cap;
                q.fused_tile_ops = tile;
                q.fused_wgs_per_cu = wgs;
                q.fused_max_ops = 0xFFFFFFFFu;
                rc = stream_run_locked(e, *s, q, nullptr, &t);
                if (rc != LTPMG_OK) return rc;
                ++trials;
                for (uint32_t cut : a_ops) {
                    double sum = 0;
                    for (uint32_t b = 0; b < s->n_batches; ++b)
                        if (s->modes[b] == LTPMG_MODE_A)
                            sum += (s->op_base[b + 1] - s->op_base[b] <= cut) ? t[b] : multi[b];
                    if (sum < best_ms) {
                        best_ms = sum;
                        p.fused_max_ops = cut;
                        p.fused_tile_ops = tile;
                        p.fused_wgs_per_cu = wgs;
                    }
                }
            }
        }
    }
    if (any_o) {
        // the one-launch kernel for small ordered batches: per-batch times of
        // the pipeline and of the one-launch kernel, then the size cut that
        // minimises their sum
        std::vector<uint32_t> o_small;
        for (uint32_t b = 0; b < s->n_batches; ++b) {
            const uint32_t ops = s->op_base[b + 1] - s->op_base[b];
            if (s->modes[b] == LTPMG_MODE_O && ops <= kSmallMaxOps) o_small.push_back(ops);
        }
        p.small_max_ops = 0;
        if (!o_small.empty()) {
            ltpmg_stream_plan q = p;
            q.merge_max_ops = 0;
            q.use_graph = 0;
            std::vector<float> pipe, one;
            q.small_max_ops = 0;
            rc = stream_run_locked(e, *s, q, nullptr, &pipe);
            if (rc != LTPMG_OK) return rc;
            q.small_max_ops = kSmallMaxOps;
            rc = stream_run_locked(e, *s, q, nullptr, &one);
            if (rc != LTPMG_OK) return rc;
            trials += 2;
            std::sort(o_small.begin(), o_small.end());
            auto eligible = [&](uint32_t b) {
                return s->modes[b] == LTPMG_MODE_O && s->op_base[b + 1] - s->op_base[b] <= kSmallMaxOps;
            };
            double best_ms = 0;
            for (uint32_t b = 0; b < s->n_batches; ++b)
                if (eligible(b)) best_ms += pipe[b];
            for (uint32_t cut : o_small) {
                double sum = 0;
                for (uint32_t b = 0; b < s->n_batches; ++b)
                    if (eligible(b)) sum += (s->op_base[b + 1] - s->op_base[b] <= cut) ? one[b] : pipe[b];
                if (sum < best_ms) {
                    best_ms = sum;
                    p.small_max_ops = cut;
                }
            }
        }
        double best_ms = -1;
        uint32_t best_merge = 0;
        for (uint32_t merge : {0u, kSmallMaxOps, 1u << 20, 1u << 22, 1u << 24, 1u << 26, (1u << 28)}) {
            ltpmg_stream_plan q = p;
            q.merge_max_ops = merge;
            double ms = 0;
            rc = stream_run_locked(e, *s, q, &ms, nullptr);
            if (rc != LTPMG_OK) return rc;
            ++trials;
            if (best_ms < 0 || ms < best_ms) {
                best_ms = ms;
                best_merge = merge;
            }
        }
        p.merge_max_ops = best_merge;
    }
    p.use_graph = 1;
    s->tuned = true;
    s->tuned_plan = p;
    if (best) *best = p;
    if (trials_out) *trials_out = trials;
    return LTPMG_OK;
}

int ltpmg_stream_fetch(ltpmg_engine *e, ltpmg_stream *s, uint32_t batch, uint8_t *commit, void *results,
                       uint64_t results_bytes) {
    if (!e || !s) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (batch >= s->n_batches) return fail(e, LTPMG_EINVAL, "batch %u >= %u", batch, s->n_batches);
    if (!s->ran) return fail(e, LTPMG_ESTATE, "the stream has not run");
    const uint32_t t0 = s->txn_base[batch], nt = s->txn_base[batch + 1] - t0;
    const uint64_t op0 = s->op_base[batch], no = s->op_base[batch + 1] - op0;
    if (commit) HIP_TRY(e, hipMemcpy(commit, s->d_commit + t0, nt, hipMemcpyDeviceToHost));
    if (results) {
        if (!(s->flags & LTPMG_STREAM_KEEP_RESULTS))
            return fail(e, LTPMG_ESTATE, "the stream was created without LTPMG_STREAM_KEEP_RESULTS");
        const uint64_t need = no * s->res_stride;
        if (results_bytes < need)
            return fail(e, LTPMG_EINVAL, "results buffer holds %llu bytes, %llu needed",
                        (unsigned long long)results_bytes, (unsigned long long)need);
        if (need) HIP_TRY(e, hipMemcpy(results, s->d_results + op0 * s->res_stride, need, hipMemcpyDeviceToHost));
        int rc = zero_result_padding(e, s->d_meta + op0, no, static_cast<uint8_t *>(results), s->res_stride);
        if (rc != LTPMG_OK) return rc;
    }
    return LTPMG_OK;
}

int ltpmg_stream_committed(ltpmg_engine *e, ltpmg_stream *s, uint32_t *per_batch) {
    if (!e || !s || !per_batch) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    if (!s->ran) return fail(e, LTPMG_ESTATE, "the stream has not run");
    std::vector<uint8_t> c(s->txn_base.back());
    if (!c.empty()) HIP_TRY(e, hipMemcpy(c.data(), s->d_commit, c.size(), hipMemcpyDeviceToHost));
    for (uint32_t b = 0; b < s->n_batches; ++b) {
        uint32_t n = 0;
        for (uint32_t t = s->txn_base[b]; t < s->txn_base[b + 1]; ++t) n += c[t];
        per_batch[b] = n;
    }
    return LTPMG_OK;
}

int ltpmg_stream_groups(ltpmg_engine *e, ltpmg_stream *s, const ltpmg_stream_plan *plan, uint32_t *group_first,
                        uint32_t cap) {
    if (!e || !s) return LTPMG_EINVAL;
    std::lock_guard<std::mutex> g(e->mu);
    HIP_TRY(e, hipSetDevice(e->device));
    ltpmg_stream_plan p;
    if (plan) p = *plan;
    else if (s->tuned) p = s->tuned_plan;
    else ltpmg_stream_default_plan(&p);
    if (!plan_ok(p)) return fail(e, LTPMG_EINVAL, "bad stream plan");
    // the cursors the next run starts from
    uint32_t cur[LTPMG_MAX_TABLES];
    if (s->flags & LTPMG_STREAM_REWIND) {
        std::copy(s->rewind_rows, s->rewind_rows + LTPMG_MAX_TABLES, cur);
    } else {
        for (Slot &sl : e->slots) HIP_TRY(e, hipStreamSynchronize(sl.stream));
        HIP_TRY(e, hipStreamSynchronize(s->stream));
        HIP_TRY(e, hipMemcpy(cur, e->d_cursor, sizeof cur, hipMemcpyDeviceToHost));
    }
    std::vector<Group> gr = plan_groups(e, *s, p, cur);
    for (size_t i = 0; i < gr.size() && i < cap && group_first; ++i) group_first[i] = gr[i].b0;
    return (int)gr.size();
}

}  // extern "C"
