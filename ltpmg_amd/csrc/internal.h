// internal.h -- shared device-side definitions of the MI355X engine.
//
// Layout in HBM (one GPU):
//   column c of table t: an array of n_rows slots, slot = col_bytes rounded up
//   (4, 8, 16, then a multiple of 32), base 256-byte aligned; the bytes of a
//   slot past the column width are kept zero so whole slots can be moved with
//   16-byte vector accesses.  A 100-byte field uses a 128-byte slot: exactly
//   one 128-byte line per access.
//   lock word of row r of table t: u64 {epoch:32 | ~tid:32}; atomicMax keeps
//   the lowest TID of the current epoch, and words of older epochs read as
//   "no writer", so the array is never cleared between batches.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ltpmg.h"

namespace ltpmg {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint32_t kBlock = 256;       // threads per workgroup (4 waves of 64)
constexpr uint32_t kTileOps = 2048;    // nominal ops per CC tile
constexpr uint32_t kTileOpsMax = kTileOps + LTPMG_MAX_TXN_OPS;
// Keys per thread in a radix-sort tile (a build-time knob).  On MI355X, for
// 2^24 keys, 8 (2048-key tiles) measured fastest: 1.383 ms per C2 batch
// against 1.43 ms at 16, the smaller scatter tiles fitting more workgroups
// per CU (with the offset scan in 128 chunks, below).
#ifndef LTPMG_SORT_ITEMS
#define LTPMG_SORT_ITEMS 8
#endif
constexpr uint32_t kSortItems = LTPMG_SORT_ITEMS;
constexpr uint32_t kSortTile = kBlock * kSortItems;
// Cell moves in flight per lane group (independent loads issued before the
// stores of a round); build-time knobs for tuning sweeps.  Lowest-TID-wins
// moves (kMoveDepth) and ordered-mode moves (kMoveDepthO) are separate: the
// ordered apply kernel measured fastest at 7 (0.561 ms per C2 batch, against
// 0.578 at 5, 0.570 at 6 and 0.609 at 8, where its registers drop it to
// three waves per SIMD); the lowest-TID-wins kernels at 4 (0.117 ms for
// cc_validate_read against 0.136 at 6).
#ifndef LTPMG_MOVE_DEPTH
#define LTPMG_MOVE_DEPTH 4
#endif
#ifndef LTPMG_MOVE_DEPTH_O
#define LTPMG_MOVE_DEPTH_O 7
#endif
constexpr int kMoveDepth = LTPMG_MOVE_DEPTH;
constexpr int kMoveDepthO = LTPMG_MOVE_DEPTH_O;
// Tile order across the XCDs.  Workgroups are dealt to the 8 XCDs in turn
// (workgroup b to XCD b mod 8), and each XCD has its own L2.  xcd_tile maps
// workgroup b to a tile so each XCD takes one contiguous run of tiles:
// neighbouring tiles, whose outputs share cache lines, then meet in one L2.
// For kernels whose tiles carry equal work (the radix scatter); a kernel
// whose neighbouring tiles are all hot would pile them onto one XCD.  A
// bijection on [0, n) for any n, so correctness never depends on it.
#ifndef LTPMG_XCD_TILES
#define LTPMG_XCD_TILES 1
#endif
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t n) {
    if (!LTPMG_XCD_TILES) return b;
    const uint32_t x = b % kXcds, i = b / kXcds, q = n / kXcds, r = n % kXcds;
    return x * q + (x < r ? x : r) + i;
}
constexpr uint32_t kNoSet = 0xFFFFFFFFu;     // ordered-mode state: no write seen
constexpr uint32_t kRmwOnly = 0xFFFFFFFEu;   // ordered-mode state: only RMW_ADDs seen

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// j-th output of a SplitMix64 generator whose state starts at s.
__host__ __device__ inline uint64_t stream64(uint64_t s, uint32_t j) {
    return mix64(s + (uint64_t)(j + 1) * kGolden);
}
__host__ __device__ inline uint64_t cell_id(uint32_t table, uint32_t row, uint32_t col) {
    return ((uint64_t)row << 10) | ((uint64_t)table << 6) | (uint64_t)col;
}
__host__ __device__ inline uint64_t cellseed(uint64_t seed, uint32_t table, uint32_t row,
                                             uint32_t col) {
    return mix64(seed + kGolden * (cell_id(table, row, col) + 1));
}
// u32 word w of the byte string stream(s, 0), stream(s, 1), ... (little endian)
__host__ __device__ inline uint32_t stream_word(uint64_t s, uint32_t w) {
    return (uint32_t)(stream64(s, w >> 1) >> ((w & 1u) * 32));
}

struct ColDesc {
    uint8_t *base;
    uint32_t slot;   // bytes between consecutive rows
    uint32_t width;  // meaningful bytes of a cell
};

struct TabDesc {
    unsigned long long *lock;
    uint8_t *present;    // append tables: one byte per row, 1 = present (nullptr otherwise)
    uint32_t n_rows;
    uint32_t n_cols;
    uint32_t cell_base;  // ordered-mode key of (row 0, col 0): cells are numbered
                         // table by table, row-major, key = cell_base + row * cstride + col
    uint32_t live;
    uint32_t cstride;    // the keys' column stride: n_cols rounded up to a power of two
                         // (a key then decodes with a shift and a mask) when every live
                         // table's padded keys fit below 2^32 - 1, else n_cols
};

struct OpView {
    const uint32_t *txn_off;
    const uint32_t *rows;
    const uint16_t *meta;
    const uint64_t *args;
    uint32_t n_txns;
    uint32_t n_ops;
};

__device__ inline uint32_t op_kind(uint16_t m) { return m & 3u; }
__device__ inline uint32_t op_col(uint16_t m) { return (m >> 2) & 63u; }
__device__ inline uint32_t op_tab(uint16_t m) { return (m >> 8) & 15u; }

// A cell op the kernels may act on: reserved meta bits clear, a cell kind
// (READ / UPDATE / RMW_ADD), and row and column inside a live table.  The
// host checks every submitted op; this keeps an unchecked (trusted) submit
// from touching memory outside the tables.
__device__ inline bool op_valid(const TabDesc *tabs, uint16_t m, uint32_t row) {
    const TabDesc &td = tabs[op_tab(m)];
    return (m >> 12) == 0 && op_kind(m) <= LTPMG_OP_RMW_ADD && row < td.n_rows && op_col(m) < td.n_cols;
}
// Any op the kernels may act on, cell or row op, from the table's row and
// column counts (n_rows 0 for a dead table) and whether it is an append
// table: an INSERT names row 0, a DELETE a row of the table.
__device__ inline bool op_check(uint16_t m, uint32_t row, uint32_t n_rows, uint32_t n_cols, bool append) {
    if ((m >> 12) != 0) return false;
    if (op_kind(m) != LTPMG_OP_ROW) return row < n_rows && op_col(m) < n_cols;
    if (!append) return false;
    return op_col(m) == LTPMG_ROW_DELETE ? row < n_rows : (op_col(m) == LTPMG_ROW_INSERT && row == 0 && n_rows > 0);
}
__device__ inline bool is_insert(uint16_t m) { return op_kind(m) == LTPMG_OP_ROW && op_col(m) == LTPMG_ROW_INSERT; }
__device__ inline bool is_delete(uint16_t m) { return op_kind(m) == LTPMG_OP_ROW && op_col(m) == LTPMG_ROW_DELETE; }
// rowseed(arg, c): the payload seed of column c of a row an INSERT appends
constexpr uint64_t kRowGolden = 0xD1B54A32D192ED03ull;
__host__ __device__ inline uint64_t rowseed(uint64_t arg, uint32_t col) {
    return mix64(arg + (uint64_t)(col + 1) * kRowGolden);
}

// popcount of the bits of m below this lane (wave64)
__device__ inline uint32_t lanemask_lt_popc(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Vector helpers: 16-byte cell chunks.  Cells, result slots and lists are
// global memory, and these helpers say so (address space 1): a pointer read
// out of a descriptor is otherwise a generic one, its loads and stores
// compile to flat instructions, and a flat access counts as LDS traffic as
// well, so every wait for an LDS read would also wait for every cell load
// and store in flight.
#define LTPMG_GLOBAL __attribute__((address_space(1)))
struct u32x4 {
    uint32_t x, y, z, w;
};
typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));
__device__ inline u32x4 ld16(const uint8_t *p) {
    const v4u32 v = *(const LTPMG_GLOBAL v4u32 *)p;
    return {v.x, v.y, v.z, v.w};
}
__device__ inline void st16(uint8_t *p, u32x4 v) {
    *(LTPMG_GLOBAL v4u32 *)p = v4u32{v.x, v.y, v.z, v.w};
}
__device__ inline uint32_t ld4(const uint8_t *p) { return *(const LTPMG_GLOBAL uint32_t *)p; }
__device__ inline void st4(uint8_t *p, uint32_t x) { *(LTPMG_GLOBAL uint32_t *)p = x; }
// streaming store for write-once outputs (read results)
__device__ inline void st16_nt(uint8_t *p, u32x4 v) {
    v4u32 x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, (LTPMG_GLOBAL v4u32 *)p);
}
// zero the words of a 16-byte chunk at byte b that lie past width w
__device__ inline u32x4 clip16(u32x4 v, uint32_t b, uint32_t w) {
    if (b + 16 > w) {
        if (b + 4 > w) v.x = 0;
        if (b + 8 > w) v.y = 0;
        if (b + 12 > w) v.z = 0;
        v.w = 0;
    }
    return v;
}
// 16-byte chunk starting at byte b of the UPDATE payload of width w
__device__ inline u32x4 payload16(uint64_t arg, uint32_t b, uint32_t w) {
    uint64_t lo = stream64(arg, b >> 3), hi = stream64(arg, (b >> 3) + 1);
    return clip16(u32x4{(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32)}, b, w);
}
__device__ inline u32x4 add16(u32x4 v, uint32_t d, uint32_t b, uint32_t w) {
    if (b < w) v.x += d;
    if (b + 4 < w) v.y += d;
    if (b + 8 < w) v.z += d;
    if (b + 12 < w) v.w += d;
    return v;
}

// ---- launchers (defined in the kernel translation units) ----
// The live tables' columns in (table, column) order, at most kColList of
// them: a kernel copies this list into LDS so a cell's column descriptor
// costs no global load (KeyDecode gives the position in it).
constexpr uint32_t kColList = 256;
// Ordered-mode cell key -> (table, row, column), passed by value so every
// field reaches a kernel in its arguments (scalar loads, no per-cell
// global load).  The live tables own contiguous key ranges from 0, in table
// order (TabDesc::cell_base, TabDesc::cstride).
struct KeyDecode {
    uint32_t n;                         // live tables
    uint32_t end_all;                   // one past the last key
    uint32_t n_cols_all;                // columns of all live tables
    uint32_t end[LTPMG_MAX_TABLES];     // one past each live table's last key
    uint32_t tab[LTPMG_MAX_TABLES];     // its table id
    uint32_t shift[LTPMG_MAX_TABLES];   // log2 of its column stride, 32 if no power of two
    uint32_t stride[LTPMG_MAX_TABLES];  // its column stride
    uint32_t cpos[LTPMG_MAX_TABLES];    // position of its first column in the column list
    uint32_t narrow;                    // bit i: live table i's columns are narrow (narrow_columns)
    // the live table (index in key order) whose range holds the key
    __device__ uint32_t index(uint32_t key) const {
        uint32_t idx = 0;
        for (uint32_t i = 1; i < n; ++i) idx = key >= end[i - 1] ? i : idx;
        return idx;
    }
    // whether every live table whose range meets [lo, hi] is narrow
    __device__ bool narrow_range(uint32_t lo, uint32_t hi) const {
        const uint32_t i0 = index(lo), i1 = index(hi);
        const uint32_t m = (i1 >= 31u ? ~0u : (2u << i1) - 1u) & ~((1u << i0) - 1u);
        return (narrow & m) == m;
    }
    // cell key -> (table, row, col) and the column's position in the column
    // list (DevTables::clist order); false past the last key
    __device__ bool operator()(uint32_t key, uint32_t &t, uint32_t &row, uint32_t &col,
                               uint32_t &cp) const {
        if (key >= end_all) return false;
        uint32_t b = 0, i_t = tab[0], sh = shift[0], st = stride[0], c0 = cpos[0];
        // the table: the last one whose range starts at or below the key
        // (one live table, the usual case, skips the walk); every index is
        // uniform, so the fields are scalar loads
        for (uint32_t i = 1; i < n; ++i) {
            const bool in = key >= end[i - 1];
            b = in ? end[i - 1] : b;
            i_t = in ? tab[i] : i_t;
            sh = in ? shift[i] : sh;
            st = in ? stride[i] : st;
            c0 = in ? cpos[i] : c0;
        }
        const uint32_t off = key - b;
        if (sh < 32u) {
            row = off >> sh;
            col = off & (st - 1u);
        } else {
            row = off / st;
            col = off - row * st;
        }
        t = i_t;
        cp = c0 + col;
        return true;
    }
};
struct DevTables {
    const ColDesc *cols;   // [LTPMG_MAX_TABLES * LTPMG_MAX_COLS]
    const TabDesc *tabs;   // [LTPMG_MAX_TABLES]
    const ColDesc *clist;  // [kColList]
    uint32_t n_clist;      // entries of clist in use
    KeyDecode dec;
    // lanes per ordered-mode cell move when the result stride exceeds 128
    // bytes, in the kernels that take one lane count for the whole batch:
    // 2 when the live columns together are narrow (narrow_columns), else 8
    uint32_t wide_lanes;
};

// Whether columns with these slots are narrow: at least three in four
// slots of at most 32 bytes.
bool narrow_columns(const uint32_t *slots, uint32_t n);

void launch_fill_column(hipStream_t s, ColDesc cd, uint32_t table, uint32_t col,
                        uint32_t n_rows, uint64_t seed);
void launch_digest_column(hipStream_t s, ColDesc cd, uint32_t table, uint32_t col,
                          uint32_t n_rows, unsigned long long *acc);

// tile_t0[k] = first txn of CC tile k; also moves the device epoch on
void launch_tile_bounds(hipStream_t s, const uint32_t *txn_off, uint32_t n_txns,
                        uint32_t n_tiles, uint32_t *tile_t0, uint32_t *epoch_ctr);
void launch_zero_u32(hipStream_t s, uint32_t *p, uint32_t n);
void launch_stream_copy(hipStream_t s, const void *src, void *dst, uint64_t n16);
void launch_cc_reserve(hipStream_t s, DevTables dt, OpView ov, const uint32_t *tile_t0,
                       uint32_t n_tiles, const uint32_t *epoch_ctr);
// found (nullable): per op, 1 if its row is present at the batch start and
// its txn commits (validate writes it, write skips cell writes of absent
// rows)
void launch_cc_validate_read(hipStream_t s, DevTables dt, OpView ov, const uint32_t *tile_t0,
                             uint32_t n_tiles, const uint32_t *epoch_ctr, uint8_t *commit,
                             uint8_t *results, uint32_t res_stride, uint32_t *counters, uint8_t *found);
void launch_cc_write(hipStream_t s, DevTables dt, OpView ov, const uint32_t *tile_t0,
                     uint32_t n_tiles, const uint8_t *commit, uint32_t max_slot, const uint8_t *found);

// Grid barrier state in device memory (zero at start; reset by the last
// arriver, so it is reused by every launch).  spin_limit bounds each wait.
struct GridBar {
    uint32_t *count;
    uint32_t *gen;
    uint32_t *err;
    uint32_t spin_limit;
};
constexpr uint32_t kFusedMaxTiles = 64;  // tiles per workgroup in the fused CC kernel
// The persistent lowest-TID-wins kernel (reserve | validate + read | write in
// one launch).  tile_ops is 1024, 2048 or 4096; grid must be co-resident.
const void *cc_fused_kernel(uint32_t tile_ops, uint32_t res_stride);
uint32_t cc_fused_grid(uint32_t n_ops, uint32_t tile_ops, uint32_t grid);
hipError_t launch_cc_fused(hipStream_t s, DevTables dt, OpView ov, uint32_t tile_ops,
                           uint32_t grid, uint32_t *epoch_ctr, uint8_t *commit, uint8_t *results,
                           uint32_t res_stride, uint32_t *counters, GridBar bar);

// Per-batch buffers of a batch that touches append tables (INSERT, DELETE,
// or any op on an append table); absent (has_append false) otherwise.
struct RowBufs {
    bool has_append = false, has_delete = false, has_insert = false;
    uint32_t insert_tables = 0;  // bit t: the batch INSERTs into table t
    uint8_t *found = nullptr;    // [n_ops]: nonzero = the op takes effect: its txn commits and its
                                 // row is present (an INSERT: it gets a row; found holds 2 + its
                                 // table until the insert kernels have run)
    uint16_t *meta_eff = nullptr;  // [n_ops]: ordered mode's meta, ops on absent rows marked invalid
    uint32_t *ins_cnt = nullptr;   // [LTPMG_MAX_TABLES * ins_tiles(n_ops)] per table and tile, then offsets
    uint32_t *ins_base = nullptr;  // [LTPMG_MAX_TABLES] cursor of each table at the batch start
    uint32_t *cursor = nullptr;    // [LTPMG_MAX_TABLES] the engine's append cursors
    uint32_t *inserted = nullptr;  // [1] rows the batch appended
};
constexpr uint32_t kInsTile = 2048;  // ops per tile of the insert kernels
inline uint32_t ins_tiles(uint32_t n_ops) { return (n_ops + kInsTile - 1) / kInsTile; }
// ordered mode, before the sort: the lowest op that deletes each row (lock
// words, tagged with a new epoch), then found / meta_eff for every op (a READ
// of an absent row gets a zeroed result slot; every txn commits).  After the
// cell writes of either mode (found then holds presence AND commit):
// row_deletes clears the presence of rows whose DELETE found them; inserts
// appends the INSERTs found marks.
void launch_row_presence_o(hipStream_t s, const DevTables &dt, OpView ov, RowBufs &rb, uint32_t *epoch_ctr,
                           uint8_t *results, uint32_t res_stride);
void launch_row_deletes(hipStream_t s, const DevTables &dt, OpView ov, RowBufs &rb);
void launch_inserts(hipStream_t s, const DevTables &dt, OpView ov, RowBufs &rb, uint8_t *results,
                    uint32_t res_stride);

struct SortBufs {
    uint32_t *keys[2];
    uint32_t *vals[2];
    uint32_t *hist;      // sort_hist_words(n): per tile, one row of digit counts
    uint32_t *partials;  // sort_partials_needed() words
};
constexpr uint32_t kMaxDigitBits = 9;  // radix digits of 7, 8 or 9 bits
// words of SortBufs::hist for a sort of n keys
size_t sort_hist_words(size_t n);
// Sorts (keys[0], vals[0]) by the low key_bits bits, stably; returns which
// buffer index holds the result.
int launch_radix_sort(hipStream_t s, SortBufs &b, uint32_t n, uint32_t key_bits);
// digit width for key_bits: the narrowest of 7 or 8 bits that keeps the
// 8-bit pass count (9 bits, when that saves a pass, only if built with
// LTPMG_RADIX_MAX_BITS=9)
uint32_t radix_digit_bits(uint32_t key_bits);
// one LSD pass, split so each kernel can be timed: hist, scan, scatter
void launch_rs_hist(hipStream_t s, SortBufs &b, int cur, uint32_t n, uint32_t shift, uint32_t bits);
void launch_rs_scan(hipStream_t s, SortBufs &b, uint32_t n, uint32_t bits);
void launch_rs_scatter(hipStream_t s, SortBufs &b, int cur, uint32_t n, uint32_t shift, uint32_t bits);
// the ordered mode's first pass (digit 0) straight from the ops: histogram,
// then (after launch_rs_scan) the scatter into keys[1] / vals[1]
void launch_rs_hist_ops(hipStream_t s, DevTables dt, OpView ov, SortBufs &b, uint32_t bits);
void launch_rs_scatter_ops(hipStream_t s, DevTables dt, OpView ov, SortBufs &b, uint32_t bits);
// words of SortBufs::partials (per-chunk digit sums of one pass)
size_t sort_partials_needed();

struct OrderBufs {
    SortBufs sort;
    uint32_t *tile_flag;    // per seg-scan tile: has a segment head
    uint32_t *tile_set;     // per tile aggregate state
    uint32_t *tile_delta;
    uint4 *wlist;           // final writes {key, set, d, 0} of segments that span tiles, one per tile at most
    uint32_t *wtotal;       // entries in wlist (cleared by the seg scan), then one spare word
};
void launch_mo_keys(hipStream_t s, DevTables dt, OpView ov, uint32_t *keys,
                    uint32_t *vals);
// ordered-mode stages, run in this order after launch_mo_keys + launch_radix_sort
void launch_mo_seg_scan(hipStream_t s, OpView ov, OrderBufs &ob, int sorted);
void launch_mo_warm_seeds(hipStream_t s, OpView ov, OrderBufs &ob);
// pad: each READ's result slot is written to its end (zeros past the cell);
// else only the cell's slot bytes, and the host zeroes the rest on fetch
void launch_mo_apply_reads(hipStream_t s, const DevTables &dt, OpView ov, OrderBufs &ob, int sorted,
                           uint8_t *results, uint32_t res_stride, bool pad);
void launch_mo_final_writes(hipStream_t s, const DevTables &dt, OpView ov, OrderBufs &ob,
                            uint32_t res_stride, uint8_t *commit);  // also sets every commit flag
uint32_t seg_tiles(uint32_t n);
// An ordered batch of 1 .. kSmallMaxOps ops in one launch of one workgroup
// of kSmallBlock threads (sort, scan and moves in LDS); also sets every
// commit flag.  rb (nullable): the batch's row-op buffers when it touches an
// append table -- the kernel then also finds each op's presence (found),
// clears the rows DELETEs remove and appends the INSERTs' rows, as the row-op
// kernels do.  The caller checks the size and, for INSERTs, the capacity.
constexpr uint32_t kSmallBlock = 1024;
constexpr uint32_t kSmallMaxOps = LTPMG_SMALL_MAX_OPS;
void launch_mo_small(hipStream_t s, const DevTables &dt, OpView ov, uint32_t key_bits, uint8_t *results,
                     uint32_t res_stride, uint8_t *commit, const RowBufs *rb);
void launch_fill_u8(hipStream_t s, uint8_t *p, uint32_t n, uint8_t v);
// A batch's ops and offsets read by a kernel straight out of pinned host
// memory (device-visible pointers): up to kGatherRuns runs of pool ops,
// each to its place in the batch, and the staged offsets.
constexpr uint32_t kGatherRuns = 8;
struct GatherRuns {
    const uint32_t *rows;
    const uint16_t *meta;
    const uint64_t *args;
    const uint32_t *txn_off;
    uint32_t src[kGatherRuns], dst[kGatherRuns];
    uint32_t n_runs, n_ops, n_txns;
};
void launch_gather_runs(hipStream_t s, const GatherRuns &g, uint32_t *txn_off, uint32_t *rows, uint16_t *meta,
                        uint64_t *args);
void launch_selftest_segscan(hipStream_t s, const uint32_t *vals, const uint8_t *head,
                             uint32_t *out, uint32_t n, uint32_t *tile_scratch);

}  // namespace ltpmg
