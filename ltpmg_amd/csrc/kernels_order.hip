// kernels_order.hip -- ordered mode: every txn commits and each cell sees its
// ops in TID order.
//
// The reference builds per-row op lists (histogram, cnt_TID atomics,
// gpulauncher.cu:413-417; segment offsets, 1558-1589; scatter, 2106-2437),
// merge-sorts each row's list by TID with one warp per row (2439-2523) and
// applies the list serially (2525-2700).  Under zipf keys the hottest row's
// serial list is the tail.  Here the order is made data-parallel instead:
//
//   radix sort    stable LSD sort on key = cell number (table by table,
//                 row-major, a row's columns padded to a power of two when
//                 the keys still fit 32 bits, so a key decodes with a shift),
//                 val = op | kind<<30; the first pass builds both
//                 from the ops (mo_keys only for one-op batches); 7- or 8-bit
//                 digits (ops are already in TID order, so a stable sort
//                 yields per-cell TID order)
//   seg reduce /  segmented scan of the per-cell state monoid
//   apply           state = (last UPDATE op or none, sum of RMW deltas since),
//                 combine(a, b) = b has an UPDATE ? b : (a.set, a.d + b.d);
//                 a READ's value = (set ? payload(args[set]) : snapshot) + d,
//                 so no cell is ever walked serially
//   final writes  one write per written cell (the segment's inclusive
//                 state): in the apply kernel after its tile's reads when the
//                 segment lies in one tile, else in a later kernel, after
//                 every tile's reads.
#include "internal.h"

namespace ltpmg {

namespace {

constexpr uint32_t kSegItems = 8;
constexpr uint32_t kSegTile = kBlock * kSegItems;

struct St {
    uint32_t set, d;
};
struct Seg {
    uint32_t f, set, d;
};

__device__ inline St combine(St a, St b) {
    if (b.set < kRmwOnly) return b;
    uint32_t s = a.set < kRmwOnly ? a.set : (a.set < b.set ? a.set : b.set);
    return {s, a.d + b.d};
}
__device__ inline Seg seg_op(Seg a, Seg b) {
    if (b.f) return b;
    St c = combine({a.set, a.d}, {b.set, b.d});
    return {a.f, c.set, c.d};
}
// val = op index | invalid << 29 | kind << 30; an invalid op (only possible
// with an unchecked submit, or an op on an absent row) is an identity
// element and emits nothing.  An RMW_ADD whose delta fits 29 bits (signed)
// carries the delta itself instead of its op index, under kind 3 (row ops
// never reach the sorted list valid): its state then needs no random 8-byte
// gather of args (three per RMW in the scan and apply walks, most of an
// order-entry batch's scan time).
constexpr uint32_t kInvalid = 1u << 29;
constexpr uint32_t kOpMask = kInvalid - 1u;
constexpr uint32_t kInlineRmw = 3u << 30;
__device__ inline uint32_t rmw_val(uint32_t o, uint64_t arg) {
    const int32_t d = (int32_t)(uint32_t)arg;
    return d >= -(1 << 28) && d < (1 << 28) ? ((uint32_t)d & kOpMask) | kInlineRmw : o | (LTPMG_OP_RMW_ADD << 30);
}
__device__ inline uint32_t inline_delta(uint32_t val) { return (uint32_t)((int32_t)(val << 3) >> 3); }
__device__ inline St elem_state(uint32_t val, const uint64_t *args) {
    uint32_t kind = val >> 30, o = val & kOpMask;
    if (val & kInvalid) return {kNoSet, 0u};
    if (kind == LTPMG_OP_UPDATE) return {o, 0u};
    if (kind == LTPMG_OP_RMW_ADD) return {kRmwOnly, (uint32_t)args[o]};
    if (kind == 3u) return {kRmwOnly, inline_delta(val)};
    return {kNoSet, 0u};
}

// Exclusive segmented scan over the block's NT thread aggregates; ws is
// NT / 64 entries of LDS, free again when the scan returns.
template <uint32_t NT = kBlock>
__device__ inline Seg block_seg_exclusive(Seg x, Seg *total, Seg *ws) {
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    Seg inc = x;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        Seg y{(uint32_t)__shfl_up((int)inc.f, off), (uint32_t)__shfl_up((int)inc.set, off),
              (uint32_t)__shfl_up((int)inc.d, off)};
        if (lane >= off) inc = seg_op(y, inc);
    }
    Seg ex{(uint32_t)__shfl_up((int)inc.f, 1), (uint32_t)__shfl_up((int)inc.set, 1),
           (uint32_t)__shfl_up((int)inc.d, 1)};
    if (lane == 0) ex = Seg{0u, kNoSet, 0u};
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    Seg pre{0u, kNoSet, 0u};
    for (uint32_t i = 0; i < w; ++i) pre = seg_op(pre, ws[i]);
    Seg tot{0u, kNoSet, 0u};
    for (uint32_t i = 0; i < NT / 64; ++i) tot = seg_op(tot, ws[i]);
    __syncthreads();
    *total = tot;
    return seg_op(pre, ex);
}
__device__ inline Seg block_seg_exclusive(Seg x, Seg *total) {
    __shared__ Seg ws[kBlock / 64];
    return block_seg_exclusive(x, total, ws);
}

// Carry-in of scan tile `tile` from the tile aggregates, read newest first.
// A tile whose aggregate holds a segment head resets the state, so the fold
// only needs the tiles from the newest such one on; folding older tiles in
// as well changes nothing, because seg_op(a, b) = b when b has a head.  The
// calling wave reads 64 aggregates per step: most tiles stop at their
// neighbour, and a tile inside a hot cell's run reads back to where the run
// began.  Every lane of one wave calls it; every lane gets the result.
__device__ inline Seg tile_carry(const uint32_t *tflag, const uint32_t *tset, const uint32_t *tdelta,
                                 uint32_t tile) {
    const uint32_t lane = __lane_id();
    Seg acc{0u, kNoSet, 0u};  // fold of the newer tiles read so far
    for (uint32_t top = tile; top > 0;) {  // lane l reads tile top - 1 - l
        const bool in = lane < top;
        const uint32_t j = in ? top - 1 - lane : 0u;
        Seg x = in ? Seg{tflag[j] & 1u, tset[j], tdelta[j]} : Seg{0u, kNoSet, 0u};
        const bool found = __ballot(in && x.f != 0u) != 0ull;
        // fold lanes 63 .. 0 (oldest first): lane l ends with its own tile
        // and the older ones above it
        for (uint32_t off = 1; off < 64; off <<= 1) {
            Seg y{(uint32_t)__shfl_down((int)x.f, off), (uint32_t)__shfl_down((int)x.set, off),
                  (uint32_t)__shfl_down((int)x.d, off)};
            if (lane + off < 64) x = seg_op(y, x);
        }
        const Seg step{(uint32_t)__shfl((int)x.f, 0), (uint32_t)__shfl((int)x.set, 0),
                       (uint32_t)__shfl((int)x.d, 0)};
        acc = seg_op(step, acc);
        if (found) break;
        top = top > 64 ? top - 64 : 0u;
    }
    return acc;
}

// Exclusive sum over the block's NT threads; ws as for block_seg_exclusive.
template <uint32_t NT = kBlock>
__device__ inline uint32_t block_exclusive_sum(uint32_t x, uint32_t *total, uint32_t *ws) {
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t inc = x;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (uint32_t i = 0; i < NT / 64; ++i) {
        if (i < w) pre += ws[i];
        tot += ws[i];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - x;
}
__device__ inline uint32_t block_exclusive_sum(uint32_t x, uint32_t *total) {
    __shared__ uint32_t ws[kBlock / 64];
    return block_exclusive_sum(x, total, ws);
}

// Element sources for the segmented scan.
struct OrderSrc {
    const uint32_t *keys, *vals;
    const uint64_t *args;
    uint32_t n;
    __device__ bool head(uint32_t i) const { return i == 0 || keys[i] != keys[i - 1]; }
    __device__ St state(uint32_t i) const { return elem_state(vals[i], args); }
};
struct SelftestSrc {
    const uint32_t *vals;
    const uint8_t *heads;
    uint32_t n;
    __device__ bool head(uint32_t i) const { return i == 0 || heads[i] != 0; }
    __device__ St state(uint32_t i) const { return {kRmwOnly, vals[i]}; }
};

template <class Src>
__device__ inline Seg thread_aggregate(const Src &src, uint32_t first) {
    Seg a{0u, kNoSet, 0u};
    for (uint32_t j = 0; j < kSegItems; ++j) {
        uint32_t i = first + j;
        if (i >= src.n) break;
        St s = src.state(i);
        a = seg_op(a, Seg{src.head(i) ? 1u : 0u, s.set, s.d});
    }
    return a;
}

// The thread's 8 consecutive sorted elements (two 16-byte loads each for
// keys and vals when in range), and the key before them.  Positions past n
// read as all-ones keys / invalid vals; keys are < 2^31, so all-ones never
// equals a real key.
__device__ inline void load8(const OrderSrc &src, uint32_t first, uint32_t *key, uint32_t *val,
                             uint32_t &prev) {
    if (first + kSegItems <= src.n) {
        const uint4 *k4 = reinterpret_cast<const uint4 *>(src.keys + first);
        const uint4 *v4 = reinterpret_cast<const uint4 *>(src.vals + first);
        uint4 a = k4[0], b = k4[1], c = v4[0], d = v4[1];
        key[0] = a.x; key[1] = a.y; key[2] = a.z; key[3] = a.w;
        key[4] = b.x; key[5] = b.y; key[6] = b.z; key[7] = b.w;
        val[0] = c.x; val[1] = c.y; val[2] = c.z; val[3] = c.w;
        val[4] = d.x; val[5] = d.y; val[6] = d.z; val[7] = d.w;
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kSegItems; ++j) {
            key[j] = first + j < src.n ? src.keys[first + j] : 0xFFFFFFFFu;
            val[j] = first + j < src.n ? src.vals[first + j] : kInvalid;
        }
    }
    prev = (first > 0 && first - 1 < src.n) ? src.keys[first - 1] : 0xFFFFFFFFu;
}

__device__ inline Seg agg8(const OrderSrc &src, uint32_t first, const uint32_t *key,
                           const uint32_t *val, uint32_t prev) {
    Seg agg{0u, kNoSet, 0u};
#pragma unroll
    for (uint32_t j = 0; j < kSegItems; ++j) {
        if (first + j >= src.n) break;
        St s = elem_state(val[j], src.args);
        bool h = key[j] != (j ? key[j - 1] : prev);
        agg = seg_op(agg, Seg{h ? 1u : 0u, s.set, s.d});
    }
    return agg;
}

// Tile aggregates of the sorted (cell, op) list.
__global__ __launch_bounds__(kBlock) void k_seg_reduce(OrderSrc src, uint32_t *tflag,
                                                       uint32_t *tset, uint32_t *tdelta,
                                                       uint32_t *zero_word) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *zero_word = 0u;  // the apply kernel's counter
    const uint32_t first = blockIdx.x * kSegTile + threadIdx.x * kSegItems;
    uint32_t key[kSegItems], val[kSegItems], prev;
    load8(src, first, key, val, prev);
    // bit 1 of the tile's flag: the apply kernel has work here -- a valid
    // element, or the end of a segment (whose final write may fall to this
    // tile).  A tile of invalid elements inside one key's run (row ops and
    // ops on absent rows sort to key 0 together) has none.
    bool act = false;
#pragma unroll
    for (uint32_t j = 0; j < kSegItems; ++j) act |= !(val[j] & kInvalid);
    if (threadIdx.x == kBlock - 1) {
        const uint32_t end = blockIdx.x * kSegTile + kSegTile;  // one past the tile
        act |= end >= src.n || key[kSegItems - 1] != src.keys[end];
    }
    act = __syncthreads_or(act);
    Seg tot;
    block_seg_exclusive(agg8(src, first, key, val, prev), &tot);
    if (threadIdx.x == 0) {
        tflag[blockIdx.x] = tot.f | (act ? 2u : 0u);
        tset[blockIdx.x] = tot.set;
        tdelta[blockIdx.x] = tot.d;
    }
}

// The self-test's tile aggregates (explicit head flags).
__global__ __launch_bounds__(kBlock) void k_selftest_reduce(SelftestSrc src, uint32_t *tflag,
                                                            uint32_t *tset, uint32_t *tdelta) {
    uint32_t first = blockIdx.x * kSegTile + threadIdx.x * kSegItems;
    Seg tot;
    block_seg_exclusive(thread_aggregate(src, first), &tot);
    if (threadIdx.x == 0) {
        tflag[blockIdx.x] = tot.f;
        tset[blockIdx.x] = tot.set;
        tdelta[blockIdx.x] = tot.d;
    }
}

// Per-thread walk with the exclusive prefix of the thread's first element.
// The tile's carry comes from tile_carry in wave 0, through LDS (the scan's
// barriers order its store before the read).
template <class Src>
__device__ inline St thread_prefix(const Src &src, const uint32_t *tflag, const uint32_t *tset,
                                   const uint32_t *tdelta, uint32_t first) {
    __shared__ Seg carry;
    if (threadIdx.x < 64) {
        const Seg c = tile_carry(tflag, tset, tdelta, blockIdx.x);
        if (threadIdx.x == 0) carry = c;
    }
    Seg tot;
    Seg ex = block_seg_exclusive(thread_aggregate(src, first), &tot);
    Seg p = seg_op(Seg{0u, carry.set, carry.d}, ex);
    return {p.set, p.d};
}

__global__ __launch_bounds__(kBlock) void k_selftest_apply(SelftestSrc src, const uint32_t *tflag,
                                                           const uint32_t *tset, const uint32_t *tdelta,
                                                           uint32_t *out) {
    uint32_t first = blockIdx.x * kSegTile + threadIdx.x * kSegItems;
    St run = thread_prefix(src, tflag, tset, tdelta, first);
    for (uint32_t j = 0; j < kSegItems; ++j) {
        uint32_t i = first + j;
        if (i >= src.n) break;
        St s = src.state(i);
        bool h = src.head(i);
        St ex = h ? St{kNoSet, 0u} : run;
        out[i] = ex.d;
        run = h ? s : combine(run, s);
    }
}

// Pipelined cell moves.  A group of G lanes takes R entries per round and
// issues every independent source load (the payload seed of the UPDATE an
// entry sees, or the snapshot cell) before any store, so each lane keeps R
// HBM requests in flight instead of one.  fetch(k, key, st, op) gives entry
// k's cell key, state and destination: op's result slot (at results + op *
// res_span; the cell's slot bytes, and with res_fill = res_span zeros on to
// the end of the slot -- with res_fill = 0 the rest of the result slot is
// left alone, and the host zeroes it when it fetches the results) or, for
// op = kCellOp, the cell itself.  ONE: G * 16 bytes are the result span, so each lane moves
// at most one 16-byte chunk (lane * 16) of an entry.  LCOLS: every live
// column's descriptor is in lcols, the column list in LDS (else they come
// from dt.cols).  Per entry a round
// keeps only what its stores need (the cell address, slot and width packed
// in one word, the op), so seven entries fit the registers of four waves
// per SIMD.
constexpr uint32_t kCellOp = 0xFFFFFFFFu;
template <int G, bool VEC16, bool ONE, bool LCOLS, int R = kMoveDepthO, class Fetch>
__device__ inline void move_cells(uint32_t n, Fetch fetch, const DevTables &dt, const KeyDecode &dec,
                                  const uint64_t *args, uint8_t *results, uint32_t res_span,
                                  uint32_t res_fill, uint32_t g, uint32_t groups, const ColDesc *lcols) {
    const uint32_t lane = threadIdx.x % G, b0 = lane * 16;
    // stream64(s, j) = mix64(s + (j + 1) * golden): the lane's two words
    const uint64_t gold0 = (uint64_t)((b0 >> 3) + 1) * kGolden, gold1 = gold0 + kGolden;
    for (uint32_t base = g; base < n; base += groups * R) {
        St st[R];
        uint32_t op[R];
        uint32_t sw[R];  // slot | width << 16 (both <= LTPMG_MAX_COL_BYTES)
        const uint8_t *cell[R];
        uint64_t arg[R];
        u32x4 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t k = base + r * groups;
            cell[r] = nullptr;
            sw[r] = 0u;
            st[r] = St{kNoSet, 0u};
            op[r] = kCellOp;
            if (k < n) {
                uint32_t key;
                fetch(k, key, st[r], op[r]);
                uint32_t tab, row, col, cpos;
                if (dec(key, tab, row, col, cpos)) {
                    const ColDesc cd = LCOLS ? lcols[cpos] : dt.cols[tab * LTPMG_MAX_COLS + col];
                    cell[r] = cd.base + (uint64_t)row * cd.slot;
                    sw[r] = cd.slot | (cd.width << 16);
                }
            }
        }
        // the payload seeds depend only on the entries: issued after every
        // column-descriptor load (loads complete in order, so waiting for a
        // descriptor never waits for a seed) and before any descriptor is
        // used, so a seed's latency overlaps its descriptor's
#pragma unroll
        for (int r = 0; r < R; ++r) arg[r] = st[r].set < kRmwOnly ? args[st[r].set] : 0ull;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool has_set = st[r].set < kRmwOnly;
            const uint32_t slot = sw[r] & 0xFFFFu;
            v[r] = u32x4{0u, 0u, 0u, 0u};
            if (VEC16 && cell[r] && !has_set && slot >= 16 && b0 < slot) v[r] = ld16(cell[r] + b0);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!cell[r]) continue;
            const uint32_t slot = sw[r] & 0xFFFFu, width = sw[r] >> 16;
            const bool to_result = op[r] != kCellOp;
            uint8_t *dst = to_result ? results + (uint64_t)op[r] * res_span : const_cast<uint8_t *>(cell[r]);
            const uint32_t span = to_result && res_fill > slot ? res_fill : slot;
            const bool has_set = st[r].set < kRmwOnly;
            if (ONE && slot >= 16) {
                u32x4 x{0u, 0u, 0u, 0u};
                if (b0 < slot) {
                    if (has_set) {
                        const uint64_t lo = mix64(arg[r] + gold0), hi = mix64(arg[r] + gold1);
                        x = u32x4{(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32)};
                        x = clip16(x, b0, width);
                    } else {
                        x = v[r];
                    }
                    x = add16(x, st[r].d, b0, width);
                }
                // one streaming store for result slots and cells alike:
                // two store forms cost two instructions per move (0.566
                // against 0.530 ms for the apply kernel)
                if (b0 < span) st16_nt(dst + b0, x);
            } else if (VEC16 && slot >= 16) {
                for (uint32_t b = b0; b < span; b += G * 16) {
                    u32x4 x{0u, 0u, 0u, 0u};
                    if (b < slot) {
                        x = has_set ? payload16(arg[r], b, width) : (b == b0 ? v[r] : ld16(cell[r] + b));
                        x = add16(x, st[r].d, b, width);
                    }
                    // a plain store: here a result fills a few bytes of a
                    // wide slot, scattered over the whole result array,
                    // and L2 takes such partial lines better than the
                    // streaming path (1.03 against 1.14 ms for the apply
                    // kernel on the order-entry batch, 256-byte slots)
                    st16(dst + b, x);
                }
            } else {  // 4- and 8-byte slots: one word per lane step
                for (uint32_t b = lane * 4; b < span; b += G * 4) {
                    uint32_t x = 0;
                    if (b < width) {
                        x = has_set ? stream_word(arg[r], b >> 2) : ld4(cell[r] + b);
                        x += st[r].d;
                    }
                    st4(dst + b, x);
                }
            }
        }
    }
}

// One tile of the sorted (cell, op) list.  Every READ gets its value: the
// payload of the last UPDATE before it on the cell, else the snapshot, plus
// the deltas of the RMW_ADDs in between.  A segment (one cell) that starts
// and ends in this tile with a write has its final value written here,
// after the block's reads: no other tile reads that cell.  Only the tile's
// first segment can have begun earlier (its reads then span tiles); if it
// ends here with a write, its entry {key, set, d} goes to the batch's
// final-write list (at most one per tile, one global atomic each) for
// k_mo_final_writes, which runs after every tile's reads.
//
// TILEG (for G = 8 past 128-byte strides): a tile whose keys lie in narrow
// tables only (KeyDecode::narrow) moves its cells with 2 lanes each instead.
template <int G, bool VEC16, bool ONE, bool LCOLS, bool TILEG = false>
__global__ __launch_bounds__(kBlock) void k_mo_apply_reads(OrderSrc src, const uint32_t *tflag,
                                                           const uint32_t *tset,
                                                           const uint32_t *tdelta, DevTables dt,
                                                           uint8_t *results, uint32_t res_stride,
                                                           uint32_t res_fill, uint4 *wlist, uint32_t *wtotal) {
    // READs from the front {op, set, d, key}; in-tile final writes from the
    // back {key, set, d, 0}.  Such a write entry needs a write op of this
    // tile in its segment, so reads + write entries <= tile elements and the
    // two ends never meet.  With the column list beside it the kernel takes
    // 36 KiB of LDS, four workgroups per CU (the register count allows the
    // same four waves per SIMD); the two block scans borrow the list's first
    // bytes before any entry is written.
    __shared__ uint4 list[kSegTile];
    __shared__ ColDesc lcols[kColList];  // the column list: no global load per cell for its descriptor
    if (!(tflag[blockIdx.x] & 2u)) return;  // uniform: no work in this tile (k_seg_reduce)
    if (LCOLS)
        for (uint32_t i = threadIdx.x; i < dt.n_clist; i += kBlock) lcols[i] = dt.clist[i];
    const KeyDecode &dec = dt.dec;
    // tiles in launch order: the hottest cells' tiles are neighbours in the
    // sorted list, and dealing them round-robin spreads them over the XCDs
    // (a contiguous run per XCD measured 1.22 ms against 0.88 ms here)
    const uint32_t tile = blockIdx.x;
    Seg *seg_ws = reinterpret_cast<Seg *>(list);                   // bytes 0..47
    uint32_t *sum_ws = reinterpret_cast<uint32_t *>(list) + 16;    // bytes 64..79
    Seg *carry_ws = reinterpret_cast<Seg *>(reinterpret_cast<uint32_t *>(list) + 24);  // bytes 96..107
    const uint32_t first = tile * kSegTile + threadIdx.x * kSegItems;
    uint32_t key[kSegItems], val[kSegItems], prev;
    load8(src, first, key, val, prev);
    const uint32_t next = first + kSegItems < src.n ? src.keys[first + kSegItems] : 0xFFFFFFFFu;
    // the key of a segment continuing from the previous tile (all-ones: none)
    const uint32_t t0 = tile * kSegTile;
    const uint32_t cont_key = t0 > 0 && src.keys[t0 - 1] == src.keys[t0] ? src.keys[t0] : 0xFFFFFFFFu;
    // the tile's carry-in, after the key loads are issued so the two
    // latencies overlap; the scan's barriers order the LDS store before
    // the read below
    if (threadIdx.x < 64) {
        const Seg c = tile_carry(tflag, tset, tdelta, tile);
        if (threadIdx.x == 0) *carry_ws = c;
    }
    Seg tot;
    const Seg ex0 = block_seg_exclusive(agg8(src, first, key, val, prev), &tot, seg_ws);
    const Seg pre = seg_op(Seg{0u, carry_ws->set, carry_ws->d}, ex0);
    // item j of the thread: its state before (ex) and after (run) it, and
    // whether it is a READ (rd) or the end of a written segment (wr)
    auto walk = [&](auto &&visit) {
        St run{pre.set, pre.d};
#pragma unroll
        for (uint32_t j = 0; j < kSegItems; ++j) {
            if (first + j >= src.n) break;
            const St s = elem_state(val[j], src.args);
            const bool h = key[j] != (j ? key[j - 1] : prev);
            const St ex = h ? St{kNoSet, 0u} : run;
            run = h ? s : combine(run, s);
            const bool rd = (val[j] >> 30) == LTPMG_OP_READ && !(val[j] & kInvalid);
            const bool last = key[j] != (j + 1 < kSegItems ? key[j + 1] : next);
            visit(j, ex, run, rd, last && run.set != kNoSet);
        }
    };
    uint32_t n_rd_t = 0, n_wr_t = 0;
    walk([&](uint32_t j, St, St, bool rd, bool wr) {
        n_rd_t += rd;
        n_wr_t += wr && key[j] != cont_key;
    });
    // both counts <= kSegTile, so they share one word of the scan
    uint32_t both;
    const uint32_t pos = block_exclusive_sum(n_rd_t | (n_wr_t << 16), &both, sum_ws);
    const uint32_t n_rd = both & 0xFFFFu, n_wr = both >> 16;
    uint32_t rp = pos & 0xFFFFu, wp = pos >> 16;
    walk([&](uint32_t j, St ex, St run, bool rd, bool wr) {
        if (rd) list[rp++] = make_uint4(val[j] & kOpMask, ex.set, ex.d, key[j]);
        if (wr && key[j] != cont_key) list[kSegTile - 1 - wp++] = make_uint4(key[j], run.set, run.d, 0u);
        if (wr && key[j] == cont_key) wlist[atomicAdd(wtotal, 1u)] = make_uint4(key[j], run.set, run.d, 0u);
    });
    __syncthreads();
    auto moves = [&](auto lanes) {
        constexpr int L = decltype(lanes)::value;
        move_cells<L, VEC16, ONE && L == G, LCOLS>(n_rd, [&](uint32_t k, uint32_t &key_out, St &st, uint32_t &op) {
            uint4 e = list[k];
            key_out = e.w;
            st = St{e.y, e.z};
            op = e.x;
        }, dt, dec, src.args, results, res_stride, res_fill, threadIdx.x / L, kBlock / L, lcols);
        __syncthreads();  // every read of this tile's cells is done
        move_cells<L, VEC16, ONE && L == G, LCOLS>(n_wr, [&](uint32_t k, uint32_t &key_out, St &st, uint32_t &op) {
            uint4 e = list[kSegTile - 1 - k];
            key_out = e.x;
            st = St{e.y, e.z};
            op = kCellOp;  // the cell itself
        }, dt, dec, src.args, results, res_stride, res_fill, threadIdx.x / L, kBlock / L, lcols);
    };
    if constexpr (TILEG) {
        const uint32_t last = (src.n - t0 < kSegTile ? src.n : t0 + kSegTile) - 1u;
        if (dec.narrow_range(src.keys[t0], src.keys[last])) {  // uniform
            moves(std::integral_constant<int, 2>{});
            return;
        }
    }
    moves(std::integral_constant<int, G>{});
}

// The final writes of segments that span tiles, after every tile's reads:
// entries {key, set, d} from k_mo_apply_reads, *wtotal of them.  The same
// grid also sets the batch's commit flags (ordered mode commits every txn),
// which saves a launch.
template <int G, bool VEC16, bool ONE, bool LCOLS>  // LCOLS: unused, descriptors come from dt.cols
__global__ __launch_bounds__(kBlock) void k_mo_final_writes(const uint4 *wlist, const uint32_t *wtotal,
                                                            DevTables dt, const uint64_t *args,
                                                            uint8_t *commit, uint32_t n_txns) {
    {  // every txn of the batch commits: its flag is 1
        const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
        uint32_t head = 0;
        if (!(reinterpret_cast<uintptr_t>(commit) & 15)) {
            for (uint32_t i = tid; i < n_txns / 16; i += nt)
                reinterpret_cast<uint4 *>(commit)[i] = make_uint4(0x01010101u, 0x01010101u, 0x01010101u, 0x01010101u);
            head = n_txns / 16 * 16;
        }
        for (uint32_t i = head + tid; i < n_txns; i += nt) commit[i] = 1;
    }
    const uint32_t g = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const uint32_t groups = gridDim.x * blockDim.x / G;
    move_cells<G, VEC16, ONE, false>(*wtotal, [&](uint32_t k, uint32_t &key, St &st, uint32_t &op) {
        uint4 e = wlist[k];
        key = e.x;
        st = St{e.y, e.z};
        op = kCellOp;
    }, dt, dt.dec, args, nullptr, 0u, 0u, g, groups, nullptr);
}

// ---------------- radix sort ----------------

// Sort inputs.  SortedSrc: the key and value arrays of the previous pass.
// OpSrc: the first pass of the ordered mode reads the ops themselves and
// builds key = cell_base + row * cstride + col and val = op | invalid << 29 |
// kind << 30 on the fly, so the keys are never written out unsorted.
struct SortedSrc {
    const uint32_t *k, *v;
    struct Cache {};
    __device__ void prepare(Cache &) const {}
    __device__ uint32_t key(uint32_t i, const Cache &) const { return k[i]; }
    __device__ void kv(uint32_t i, uint32_t &key, uint32_t &val, const Cache &) const {
        key = k[i];
        val = v[i];
    }
    // kv in two steps: raw() only loads, cook() only computes, so a caller
    // can have every load of a tile in flight before the first result is
    // needed
    __device__ void raw(uint32_t i, uint32_t &a, uint32_t &b) const {
        a = k[i];
        b = v[i];
    }
    __device__ void cook(uint32_t, uint32_t &key, uint32_t &val, const Cache &) const {}
    // keys i .. i + 7 (i a multiple of 8, the array 16-byte aligned)
    __device__ void keys8(uint32_t i, uint32_t *key, const Cache &) const {
        const uint4 a = reinterpret_cast<const uint4 *>(k + i)[0], b = reinterpret_cast<const uint4 *>(k + i)[1];
        key[0] = a.x; key[1] = a.y; key[2] = a.z; key[3] = a.w;
        key[4] = b.x; key[5] = b.y; key[6] = b.z; key[7] = b.w;
    }
};
struct OpSrc {
    const TabDesc *tabs;
    const uint32_t *rows;
    const uint16_t *meta;
    const uint64_t *args;  // read for RMW_ADDs only (rmw_val)
    // the table descriptors in LDS, so a key costs no dependent global load
    // one 16-byte word per table {cell_base, n_rows (0 if not live), n_cols,
    // cstride}: a key costs one LDS read
    struct Cache {
        uint4 t[LTPMG_MAX_TABLES];
    };
    __device__ void prepare(Cache &c) const {
        if (threadIdx.x < LTPMG_MAX_TABLES) {
            const TabDesc &td = tabs[threadIdx.x];
            c.t[threadIdx.x] = make_uint4(td.cell_base, td.live ? td.n_rows : 0u, td.n_cols, td.cstride);
        }
    }
    __device__ void kv(uint32_t o, uint32_t &key, uint32_t &val, const Cache &c) const {
        raw(o, key, val);
        cook(o, key, val, c);
    }
    // raw: key = the op's meta word, val = its row; cook turns them into the
    // cell key and the op's sort value
    __device__ void raw(uint32_t o, uint32_t &a, uint32_t &b) const {
        a = meta[o];
        b = rows[o];
    }
    __device__ void cook(uint32_t o, uint32_t &key, uint32_t &val, const Cache &c) const {
        const uint16_t m = (uint16_t)key;
        const uint32_t row = val;
        const uint4 td = c.t[op_tab(m)];
        const bool ok = (m >> 12) == 0 && op_kind(m) <= LTPMG_OP_RMW_ADD && row < td.y && op_col(m) < td.z;
        key = ok ? td.x + row * td.w + op_col(m) : 0u;
        val = ok && op_kind(m) == LTPMG_OP_RMW_ADD ? rmw_val(o, args[o])
                                                   : o | (op_kind(m) << 30) | (ok ? 0u : kInvalid);
    }
    __device__ uint32_t key(uint32_t o, const Cache &c) const { return key_of(meta[o], rows[o], c); }
    __device__ uint32_t key_of(uint16_t m, uint32_t row, const Cache &c) const {
        const uint4 td = c.t[op_tab(m)];
        const bool ok = (m >> 12) == 0 && op_kind(m) <= LTPMG_OP_RMW_ADD && row < td.y && op_col(m) < td.z;
        return ok ? td.x + row * td.w + op_col(m) : 0u;
    }
    // keys of ops o .. o + 7 (o a multiple of 8, rows and meta 16-byte aligned)
    __device__ void keys8(uint32_t o, uint32_t *key, const Cache &c) const {
        const uint4 r0 = reinterpret_cast<const uint4 *>(rows + o)[0], r1 = reinterpret_cast<const uint4 *>(rows + o)[1];
        const uint4 m = reinterpret_cast<const uint4 *>(meta + o)[0];
        const uint32_t rw[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
        const uint32_t mw[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) key[j] = key_of((uint16_t)(mw[j >> 1] >> ((j & 1u) * 16)), rw[j], c);
    }
};

// Digit histogram of one sort tile: per-wave LDS counters, then the
// tile's row of D counts, written contiguously.
// VEC: each thread takes kSortItems consecutive keys, read 8 at a time
// with 16-byte loads (the source arrays are 16-byte aligned); the counts do
// not depend on which thread takes which key.
template <class Src, uint32_t BITS, bool VEC>
__global__ __launch_bounds__(kBlock) void k_rs_hist(Src src, uint32_t n, uint32_t shift,
                                                    uint32_t n_tiles, uint32_t *hist) {
    constexpr uint32_t D = 1u << BITS;
    __shared__ uint32_t h[kBlock / 64][D];
    __shared__ typename Src::Cache cache;
    for (uint32_t i = threadIdx.x; i < D * (kBlock / 64); i += kBlock) (&h[0][0])[i] = 0;
    src.prepare(cache);
    __syncthreads();
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t base = blockIdx.x * kSortTile;
    // every key of the thread first (independent loads in flight), then the
    // LDS counts
    uint32_t digit[kSortItems];
    const uint32_t vbase = base + threadIdx.x * kSortItems;
    if (VEC && vbase + kSortItems <= n) {
        static_assert(kSortItems % 8 == 0, "whole groups of 8 keys");
#pragma unroll
        for (uint32_t g = 0; g < kSortItems; g += 8) {
            uint32_t key[8];
            src.keys8(vbase + g, key, cache);
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j) digit[g + j] = (key[j] >> shift) & (D - 1u);
        }
    } else {
#pragma unroll
        for (uint32_t it = 0; it < kSortItems; ++it) {
            const uint32_t idx = VEC ? vbase + it : base + it * kBlock + threadIdx.x;
            const uint32_t key = idx < n ? src.key(idx, cache) : 0u;
            digit[it] = idx < n ? (key >> shift) & (D - 1u) : D;
        }
    }
    // equal neighbouring digits of the thread are merged first: skewed
    // digits (Zipf keys put most of them on one high digit) would otherwise
    // serialise on one LDS counter
    uint32_t run_d = digit[0], run_n = 0;
#pragma unroll
    for (uint32_t it = 0; it < kSortItems; ++it) {
        if (digit[it] != run_d) {
            if (run_d < D && run_n) atomicAdd(&h[w][run_d], run_n);
            run_d = digit[it];
            run_n = 0;
        }
        ++run_n;
    }
    if (run_d < D && run_n) atomicAdd(&h[w][run_d], run_n);
    __syncthreads();
    // one contiguous row per tile: hist[tile * D + d]
    for (uint32_t d = threadIdx.x; d < D; d += kBlock) {
        uint32_t c = 0;
        for (uint32_t i = 0; i < kBlock / 64; ++i) c += h[i][d];
        hist[(uint64_t)blockIdx.x * D + d] = c;
    }
}


// LDS of one scatter tile
template <uint32_t BITS>
struct ScatterLds {
    static constexpr uint32_t D = 1u << BITS;
    uint32_t cnt[kBlock / 64][D];  // per wave and digit: keys ranked so far, then the wave's offset
    uint32_t gbase[D];             // global position of the tile's first key of each digit
    uint32_t tile_off[D];          // position of each digit's run inside the tile
    uint32_t sk[kSortTile], sv[kSortTile];
};

// Ranks one tile: wave w ranks its kSortItems chunks of 64 keys in order;
// lanes holding the same digit are found with BITS ballots (wave64
// multisplit).  Leaves per-wave digit counts in L.cnt.
template <class Src, uint32_t BITS>
__device__ inline void rank_tile(const Src &src, const typename Src::Cache &cache, uint32_t tile,
                                 uint32_t n, uint32_t shift, ScatterLds<BITS> &L, uint32_t *k,
                                 uint32_t *v, uint32_t *r) {
    constexpr uint32_t D = 1u << BITS;
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    const uint32_t wbase = tile * kSortTile + w * (kSortTile / (kBlock / 64));
    // every chunk's loads first: the ranking below has a wave barrier per
    // chunk, which the compiler does not move loads across, so loading
    // inside it would cost one memory round trip per chunk
    // (lanes past n load element n - 1, which exists, so no load is
    // conditional, and then read as zero)
#pragma unroll
    for (uint32_t c = 0; c < kSortItems; ++c) {
        const uint32_t idx = wbase + c * 64 + lane;
        src.raw(idx < n ? idx : n - 1u, k[c], v[c]);
    }
#pragma unroll
    for (uint32_t c = 0; c < kSortItems; ++c) {
        const uint32_t idx = wbase + c * 64 + lane;
        src.cook(idx < n ? idx : n - 1u, k[c], v[c], cache);  // (cook may read args[op])
        k[c] = idx < n ? k[c] : 0u;
        v[c] = idx < n ? v[c] : 0u;
    }
#pragma unroll
    for (uint32_t c = 0; c < kSortItems; ++c) {
        uint32_t idx = wbase + c * 64 + lane;
        bool valid = idx < n;
        const uint32_t key = k[c];
        uint32_t d = (key >> shift) & (D - 1u);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < BITS; ++b) {
            uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t leader = peers ? (uint32_t)__builtin_ctzll(peers) : lane;
        uint32_t prior = 0;
        if (valid && lane == leader) {
            prior = L.cnt[w][d];
            L.cnt[w][d] = prior + (uint32_t)__popcll(peers);
        }
        prior = __shfl(prior, (int)leader);
        r[c] = prior + lanemask_lt_popc(peers);
        __builtin_amdgcn_wave_barrier();
    }
}

// Digits per thread when a thread owns PER consecutive digits (thread order
// is digit order); with fewer digits than threads, threads >= D own none.
template <uint32_t BITS>
struct DigitSplit {
    static constexpr uint32_t D = 1u << BITS;
    static constexpr uint32_t PER = D >= kBlock ? D / kBlock : 1u;
    static_assert(D % kBlock == 0 || kBlock % D == 0, "digits split evenly over threads");
    __device__ static bool owns() { return threadIdx.x * PER < D; }
};

// After rank_tile and a barrier: per digit, each wave's offset inside the
// digit's run (in L.cnt) and the run's start inside the tile (L.tile_off,
// an exclusive scan over digits in thread order).  runs[p] = the tile's
// count of digit threadIdx.x * PER + p (0 for a thread that owns none).
template <uint32_t BITS>
__device__ inline void digit_runs(ScatterLds<BITS> &L, uint32_t *runs) {
    using DS = DigitSplit<BITS>;
    constexpr uint32_t PER = DS::PER;
    const bool own = DS::owns();
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t p = 0; p < PER; ++p) {
        const uint32_t d = threadIdx.x * PER + p;
        uint32_t run = 0;
        if (own) {
            for (uint32_t i = 0; i < kBlock / 64; ++i) {
                uint32_t t = L.cnt[i][d];
                L.cnt[i][d] = run;
                run += t;
            }
        }
        runs[p] = run;
        mine += run;
    }
    uint32_t total;
    uint32_t ex = block_exclusive_sum(mine, &total);
#pragma unroll
    for (uint32_t p = 0; p < PER; ++p) {
        if (own) L.tile_off[threadIdx.x * PER + p] = ex;
        ex += runs[p];
    }
}

// After L.gbase and L.tile_off are set and a barrier: stages the tile in
// LDS in (digit, input) order, then writes each digit's run contiguously
// (neighbouring lanes store neighbouring addresses).
template <uint32_t BITS>
__device__ inline void stage_and_write(ScatterLds<BITS> &L, uint32_t tile, uint32_t n, uint32_t shift,
                                       const uint32_t *k, const uint32_t *v, const uint32_t *r,
                                       uint32_t *kout, uint32_t *vout) {
    constexpr uint32_t D = 1u << BITS;
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    const uint32_t wbase = tile * kSortTile + w * (kSortTile / (kBlock / 64));
#pragma unroll
    for (uint32_t c = 0; c < kSortItems; ++c) {
        uint32_t idx = wbase + c * 64 + lane;
        if (idx < n) {
            uint32_t d = (k[c] >> shift) & (D - 1u);
            uint32_t lp = L.tile_off[d] + L.cnt[w][d] + r[c];
            L.sk[lp] = k[c];
            L.sv[lp] = v[c];
        }
    }
    __syncthreads();
    const uint32_t tile0 = tile * kSortTile;
    const uint32_t tile_n = n - tile0 < kSortTile ? n - tile0 : kSortTile;
    // every LDS read first, into registers of its own, then every store:
    // reading the next key into the registers a store still reads from
    // would make each store wait for the one before it
    uint32_t kk[kSortItems], vv[kSortItems], pp[kSortItems];
#pragma unroll
    for (uint32_t c = 0; c < kSortItems; ++c) {
        const uint32_t j = threadIdx.x + c * kBlock;
        if (j < tile_n) {
            kk[c] = L.sk[j];
            const uint32_t d = (kk[c] >> shift) & (D - 1u);
            pp[c] = L.gbase[d] + (j - L.tile_off[d]);
            vv[c] = L.sv[j];
        }
    }
#pragma unroll
    for (uint32_t c = 0; c < kSortItems; ++c) {
        if (threadIdx.x + c * kBlock < tile_n) {
            kout[pp[c]] = kk[c];
            vout[pp[c]] = vv[c];
        }
    }
}

// Stable scatter of one tile (xcd_tile of the workgroup) with its digit
// offsets read from offs (one row of D per tile, from k_rs_offsets).
template <class Src, uint32_t BITS>
__global__ __launch_bounds__(kBlock) void k_rs_scatter(Src src, uint32_t *kout, uint32_t *vout,
                                                       uint32_t n, uint32_t shift,
                                                       uint32_t n_tiles, const uint32_t *offs) {
    constexpr uint32_t D = 1u << BITS, PER = DigitSplit<BITS>::PER;
    __shared__ ScatterLds<BITS> L;
    __shared__ typename Src::Cache cache;
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    for (uint32_t i = threadIdx.x; i < D * (kBlock / 64); i += kBlock) (&L.cnt[0][0])[i] = 0;
    for (uint32_t d = threadIdx.x; d < D; d += kBlock) L.gbase[d] = offs[(uint64_t)tile * D + d];
    src.prepare(cache);
    __syncthreads();
    uint32_t k[kSortItems], v[kSortItems], r[kSortItems];
    rank_tile<Src, BITS>(src, cache, tile, n, shift, L, k, v, r);
    __syncthreads();
    uint32_t runs[PER];
    digit_runs<BITS>(L, runs);
    __syncthreads();  // every tile_off[d] written before any lane reads it
    stage_and_write<BITS>(L, tile, n, shift, k, v, r, kout, vout);
}

// ---------------- digit offsets of a radix pass ----------------
//
// hist holds one row of D digit counts per tile.  The stable scatter needs,
// for tile t and digit d, offs = (keys of smaller digits) + (keys of digit d
// in tiles before t).  Tiles are grouped in chunks of up to kOffsetRows rows
// per pass of the walk: k_rs_chunk_sum sums each chunk's column per digit;
// k_rs_offsets gives each chunk its column prefix and the digit bases, then
// walks its tiles and overwrites the counts with offsets.  Every access
// reads or writes whole rows.


// Build-time knobs of the offset scan; 128 chunks with 32 row loads in
// flight measured fastest for 8192 tiles (0.064 ms for four passes, against
// 0.10 ms with 256 chunks and 16 loads).
#ifndef LTPMG_SORT_CHUNKS
#define LTPMG_SORT_CHUNKS 128
#endif
#ifndef LTPMG_OFFSET_ROWS
#define LTPMG_OFFSET_ROWS 32
#endif
constexpr uint32_t kMaxSortChunks = LTPMG_SORT_CHUNKS;  // tile groups of the offset scan
constexpr uint32_t kOffsetRows = LTPMG_OFFSET_ROWS;     // rows loaded together in the walk

__host__ __device__ inline uint32_t chunk_tiles(uint32_t n_tiles) {
    const uint32_t per = (n_tiles + kMaxSortChunks - 1) / kMaxSortChunks;
    return per < 32u ? 32u : per;
}

template <uint32_t BITS>
__global__ __launch_bounds__(kBlock) void k_rs_chunk_sum(const uint32_t *hist, uint32_t n_tiles,
                                                         uint32_t *chunks) {
    constexpr uint32_t D = 1u << BITS;
    const uint32_t ct = chunk_tiles(n_tiles), t0 = blockIdx.x * ct;
    const uint32_t t1 = t0 + ct < n_tiles ? t0 + ct : n_tiles;
    if constexpr (2 * D <= kBlock) {
        if (t1 - t0 <= 2 * kOffsetRows) {  // two threads per digit, half of the rows each
            __shared__ uint32_t part[D];
            const uint32_t d = threadIdx.x % D, h = threadIdx.x / D;
            const uint32_t half = (t1 - t0 + 1) / 2, r0 = t0 + (h ? half : 0u);
            const uint32_t nr = h == 0 ? half : h == 1 ? t1 - t0 - half : 0u;
            uint32_t x[kOffsetRows];
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r) x[r] = r < nr ? hist[(uint64_t)(r0 + r) * D + d] : 0u;
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r) sum += x[r];
            if (h == 1) part[d] = sum;
            __syncthreads();
            if (h == 0) chunks[blockIdx.x * D + d] = sum + part[d];
            return;
        }
    }
    for (uint32_t d = threadIdx.x; d < D; d += kBlock) {
        uint32_t sum = 0;
        for (uint32_t t = t0; t < t1; t += kOffsetRows) {  // kOffsetRows loads in flight
            uint32_t x[kOffsetRows];
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r) x[r] = t + r < t1 ? hist[(uint64_t)(t + r) * D + d] : 0u;
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r) sum += x[r];
        }
        chunks[blockIdx.x * D + d] = sum;
    }
}

template <uint32_t BITS>
__global__ __launch_bounds__(kBlock) void k_rs_offsets(uint32_t *hist, uint32_t n_tiles,
                                                       const uint32_t *chunks, uint32_t n_chunks) {
    using DS = DigitSplit<BITS>;
    constexpr uint32_t D = DS::D, PER = DS::PER;
    const bool own = DS::owns();
    const uint32_t ct = chunk_tiles(n_tiles), t0 = blockIdx.x * ct;
    const uint32_t t1 = t0 + ct < n_tiles ? t0 + ct : n_tiles;
    if constexpr (2 * D <= kBlock) {
        if (t1 - t0 <= 2 * kOffsetRows) {
            // Two threads per digit, each with half of the chunk's rows and
            // half of the earlier chunks' sums; the row loads go out first,
            // so both sets of loads are in flight together.
            __shared__ uint32_t sb[2][D], sa[2][D], sr[2][D], sbase[D];
            const uint32_t d = threadIdx.x % D, h = threadIdx.x / D;
            const bool act = h < 2;
            const uint32_t nrows = t1 - t0, half = (nrows + 1) / 2;
            const uint32_t r0 = t0 + (h ? half : 0u);
            const uint32_t nr = act ? (h ? nrows - half : half) : 0u;
            uint32_t cnt[kOffsetRows];
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r) cnt[r] = r < nr ? hist[(uint64_t)(r0 + r) * D + d] : 0u;
            uint32_t b = 0, a = 0;
            for (uint32_t c = h * kOffsetRows; act && c < n_chunks; c += 2 * kOffsetRows) {
                uint32_t x[kOffsetRows];
#pragma unroll
                for (uint32_t r = 0; r < kOffsetRows; ++r) x[r] = c + r < n_chunks ? chunks[(c + r) * D + d] : 0u;
#pragma unroll
                for (uint32_t r = 0; r < kOffsetRows; ++r) (c + r < blockIdx.x ? b : a) += x[r];
            }
            uint32_t rsum = 0;
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r) rsum += cnt[r];
            if (act) {
                sb[h][d] = b;
                sa[h][d] = a;
                sr[h][d] = rsum;
            }
            __syncthreads();
            const uint32_t before = act ? sb[0][d] + sb[1][d] : 0u;
            const uint32_t total = act ? before + sa[0][d] + sa[1][d] : 0u;
            uint32_t all;
            // threads 0 .. D - 1 hold the digits in order; the rest add 0
            const uint32_t base = block_exclusive_sum(h == 0 ? total : 0u, &all);
            if (h == 0) sbase[d] = base + before;
            __syncthreads();
            if (act) {
                uint32_t run = sbase[d] + (h ? sr[0][d] : 0u);
#pragma unroll
                for (uint32_t r = 0; r < kOffsetRows; ++r) {
                    if (r < nr) hist[(uint64_t)(r0 + r) * D + d] = run;
                    run += cnt[r];
                }
            }
            return;
        }
    }
    // thread owns digits threadIdx.x * PER + p, so thread order is digit order
    uint32_t before[PER], total[PER], mine = 0;
#pragma unroll
    for (uint32_t p = 0; p < PER; ++p) {
        const uint32_t d = threadIdx.x * PER + p;
        uint32_t b = 0, a = 0;
        for (uint32_t c = 0; own && c < n_chunks; c += kOffsetRows) {  // kOffsetRows loads in flight
            uint32_t x[kOffsetRows];
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r) x[r] = c + r < n_chunks ? chunks[(c + r) * D + d] : 0u;
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r) (c + r < blockIdx.x ? b : a) += x[r];
        }
        before[p] = b;
        total[p] = a + b;
        mine += a + b;
    }
    uint32_t all;
    uint32_t base = block_exclusive_sum(mine, &all);
    uint32_t run[PER];
#pragma unroll
    for (uint32_t p = 0; p < PER; ++p) {
        run[p] = base + before[p];
        base += total[p];
    }
    for (uint32_t t = t0; own && t < t1; t += kOffsetRows) {
        const uint32_t nr = t1 - t < kOffsetRows ? t1 - t : kOffsetRows;
#pragma unroll
        for (uint32_t p = 0; p < PER; ++p) {
            const uint32_t d = threadIdx.x * PER + p;
            uint32_t cnt[kOffsetRows];
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r)
                cnt[r] = r < nr ? hist[(uint64_t)(t + r) * D + d] : 0u;
#pragma unroll
            for (uint32_t r = 0; r < kOffsetRows; ++r) {
                if (r < nr) hist[(uint64_t)(t + r) * D + d] = run[p];
                run[p] += cnt[r];
            }
        }
    }
}

// ---------------- mode O keys ----------------

__global__ __launch_bounds__(kBlock) void k_mo_keys(DevTables dt, OpView ov, uint32_t *keys,
                                                    uint32_t *vals) {
    for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < ov.n_ops;
         o += gridDim.x * blockDim.x) {
        uint16_t m = ov.meta[o];
        const TabDesc &td = dt.tabs[op_tab(m)];
        uint32_t row = ov.rows[o];
        bool ok = (m >> 12) == 0 && op_kind(m) <= LTPMG_OP_RMW_ADD && row < td.n_rows &&
                  op_col(m) < td.n_cols;
        keys[o] = ok ? td.cell_base + row * td.cstride + op_col(m) : 0u;
        vals[o] = ok && op_kind(m) == LTPMG_OP_RMW_ADD ? rmw_val(o, ov.args[o])
                                                       : o | (op_kind(m) << 30) | (ok ? 0u : kInvalid);
    }
}

// ---------------- ordered batches in one workgroup ----------------
//
// A batch of up to kSmallMaxOps ops runs as one launch of one workgroup of
// kSmallBlock threads with every intermediate in LDS, so a small batch pays
// one launch instead of the pipeline's twenty (about 3 us of host time
// each; a 4,096-op batch took 134 us from execute to completion that way):
//
//   sort   LSD radix sort of the (cell key, op | kind) pairs, 8-bit
//          digits; wave w ranks its chunks of 64 consecutive elements in
//          order with one ballot per digit bit (wave64 multisplit) and
//          per-wave digit counters, so the sort is stable and the ops of a
//          cell stay in TID order
//   scan   the cell-state monoid over the sorted list, each thread folding
//          its run of consecutive elements, one block-wide segmented scan
//          between them
//   moves  the READs' moves, then (after a barrier) the final writes, from
//          one list in LDS, by the pipeline's lane groups (move_cells).
//          The whole batch is in the workgroup, so every segment starts
//          and ends in it: no carry between tiles, no cross-tile writes.
//
// The sorted pairs (64 KB) and the move list (16 B per entry, at most one
// per element: 128 KB) share one LDS region; with the column list the
// kernel takes about 133 KB of the CU's 160 KB.
constexpr uint32_t kSmallWaves = kSmallBlock / 64;
constexpr uint32_t kSmallItems = kSmallMaxOps / kSmallBlock;  // elements per thread
constexpr uint32_t kSmallBits = 8;
constexpr uint32_t kSmallD = 1u << kSmallBits;
static_assert(kSmallItems * kSmallBlock == kSmallMaxOps, "whole chunks per thread");

struct SmallSortLds {
    uint32_t sk[kSmallMaxOps], sv[kSmallMaxOps];
    uint32_t cnt[kSmallWaves][kSmallD];  // per wave and digit: ranked so far, then the wave's offset
    uint32_t dig_off[kSmallD];           // start of each digit's run
};
// Row ops: the first DELETE of each row, in an open-addressing hash over
// (table, row) with twice as many slots as a batch has ops, so a probe
// always ends; a slot holds key << 13 | op (op < kSmallMaxOps = 2^13).
constexpr uint32_t kDelSlots = 2 * kSmallMaxOps;
constexpr uint64_t kDelEmpty = ~0ull;
static_assert(kSmallMaxOps <= (1u << 13), "op indices fit 13 bits");
union SmallLds {
    SmallSortLds s;
    uint4 list[kSmallMaxOps];  // READs {op, set, d, key} from the front, final writes {key, set, d, 0} from the back
    unsigned long long del[kDelSlots];  // before the sort: the batch's DELETEs
};
static_assert(sizeof(SmallLds) == kSmallMaxOps * 16, "the move list sets the region's size");

__device__ inline uint32_t del_hash(uint64_t key) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 50) & (kDelSlots - 1u);  // 14 bits
}
// Keeps the lowest op deleting (table, row) = key (< 2^36).
__device__ inline void del_insert(unsigned long long *tab, uint64_t key, uint32_t op) {
    const unsigned long long want = (key << 13) | op;
    uint32_t h = del_hash(key);
    for (uint32_t probe = 0; probe < kDelSlots; ++probe) {
        unsigned long long cur = tab[h];
        if (cur == kDelEmpty) {
            cur = atomicCAS(tab + h, kDelEmpty, want);
            if (cur == kDelEmpty) return;
        }
        if ((cur >> 13) == key) {
            atomicMin(tab + h, want);
            return;
        }
        h = (h + 1u) & (kDelSlots - 1u);
    }
}
// The lowest op deleting (table, row) = key, or kSmallMaxOps for none.
__device__ inline uint32_t del_find(const unsigned long long *tab, uint64_t key) {
    uint32_t h = del_hash(key);
    for (uint32_t probe = 0; probe < kDelSlots; ++probe) {
        const unsigned long long cur = tab[h];
        if (cur == kDelEmpty) break;
        if ((cur >> 13) == key) return (uint32_t)(cur & 0x1FFFu);
        h = (h + 1u) & (kDelSlots - 1u);
    }
    return kSmallMaxOps;
}
// op_check from the LDS table cache ({cell_base, n_rows or 0 if not live,
// n_cols, cstride}) and the table's presence bytes (nullptr: not an append
// table)
__device__ inline bool small_op_ok(const uint4 &tc, const uint8_t *present, uint16_t m, uint32_t row) {
    return op_check(m, row, tc.y, tc.z, present != nullptr);
}

// Row-op state of a batch that touches append tables (on = 0 otherwise).
struct SmallRows {
    uint8_t *found;           // per op: takes effect (its row present; an INSERT: it gets a row)
    uint32_t *cursor;         // per table: the append cursor
    uint32_t *inserted;       // rows the batch appended
    uint32_t insert_tables;   // bit t: the batch INSERTs into table t
    uint32_t on;
};

template <int G, bool VEC16, bool ONE, bool LCOLS>
__global__ __launch_bounds__(kSmallBlock) void k_mo_small(DevTables dt, OpView ov, uint32_t key_bits,
                                                          uint8_t *results, uint32_t res_stride,
                                                          uint8_t *commit, SmallRows ra) {
    __shared__ SmallLds L;
    __shared__ ColDesc lcols[kColList];
    __shared__ OpSrc::Cache cache;
    __shared__ uint32_t sum_ws[kSmallWaves];
    __shared__ Seg seg_ws[kSmallWaves];
    const uint32_t n = ov.n_ops, lane = __lane_id(), w = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < ov.n_txns; i += kSmallBlock) commit[i] = 1;  // every txn commits
    if (LCOLS)
        for (uint32_t i = threadIdx.x; i < dt.n_clist; i += kSmallBlock) lcols[i] = dt.clist[i];
    __shared__ uint8_t *s_present[LTPMG_MAX_TABLES];  // each table's presence bytes (append tables)
    const OpSrc src{dt.tabs, ov.rows, ov.meta, ov.args};
    src.prepare(cache);
    if (ra.on) {
        for (uint32_t i = threadIdx.x; i < kDelSlots; i += kSmallBlock) L.del[i] = kDelEmpty;
        if (threadIdx.x < LTPMG_MAX_TABLES) s_present[threadIdx.x] = dt.tabs[threadIdx.x].present;
    }
    __syncthreads();
    // chunks of 64 elements per wave; element wbase + c * 64 + lane is the
    // lane's c-th, in op order before the sort
    const uint32_t nc = (n + kSmallBlock - 1) / kSmallBlock;
    const uint32_t wbase = w * 64 * nc;
    uint32_t k[kSmallItems], v[kSmallItems], r[kSmallItems];
    // first the raw meta word (k) and row (v) of each element
#pragma unroll
    for (uint32_t c = 0; c < kSmallItems; ++c) {
        const uint32_t i = wbase + c * 64 + lane;
        k[c] = 0xFFFFFFFFu;
        v[c] = kInvalid;
        if (c < nc && i < n) src.raw(i, k[c], v[c]);
    }
    // Row ops (the row-op kernels' semantics): an op on an append table
    // takes effect if its row is present at the batch start and no DELETE of
    // the row comes earlier in the batch; one that does not is an identity
    // in the sort and scan (a reserved meta bit), and a READ of it gets a
    // zeroed result slot.  Whether an op takes effect goes to found[], which
    // the same lane reads back after the cell writes.
    if (ra.on) {
        // r[c]: the presence byte of element c's row (1 if it has none to
        // read); every element's load is issued before any is used
#pragma unroll
        for (uint32_t c = 0; c < kSmallItems; ++c) {
            const uint32_t i = wbase + c * 64 + lane;
            r[c] = 1;
            if (c < nc && i < n) {
                const uint16_t m = (uint16_t)k[c];
                const uint8_t *pr = s_present[op_tab(m)];
                if (pr && !is_insert(m) && small_op_ok(cache.t[op_tab(m)], pr, m, v[c])) r[c] = pr[v[c]];
            }
        }
#pragma unroll
        for (uint32_t c = 0; c < kSmallItems; ++c) {
            const uint32_t i = wbase + c * 64 + lane;
            if (c < nc && i < n && is_delete((uint16_t)k[c])) {
                const uint16_t m = (uint16_t)k[c];
                if (small_op_ok(cache.t[op_tab(m)], s_present[op_tab(m)], m, v[c]))
                    del_insert(L.del, ((uint64_t)op_tab(m) << 32) | v[c], i);
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t c = 0; c < kSmallItems; ++c) {
            const uint32_t i = wbase + c * 64 + lane;
            if (c >= nc || i >= n) continue;
            const uint16_t m = (uint16_t)k[c];
            const uint32_t row = v[c];
            const uint8_t *pr = s_present[op_tab(m)];
            bool f = small_op_ok(cache.t[op_tab(m)], pr, m, row);
            if (f && pr && !is_insert(m)) {
                const uint32_t first_del = del_find(L.del, ((uint64_t)op_tab(m) << 32) | row);
                f = r[c] && !(first_del < i);
            }
            ra.found[i] = f ? 1 : 0;
            if (!f) {
                k[c] |= 0x8000u;
                if (op_kind(m) == LTPMG_OP_READ)
                    for (uint32_t b = 0; b < res_stride; b += 4) st4(results + (uint64_t)i * res_stride + b, 0u);
            }
        }
        __syncthreads();  // the hash's region is the sort's next
    }
#pragma unroll
    for (uint32_t c = 0; c < kSmallItems; ++c) {
        const uint32_t i = wbase + c * 64 + lane;
        if (c < nc && i < n) src.cook(i, k[c], v[c], cache);
    }
    const uint32_t passes = (key_bits + kSmallBits - 1) / kSmallBits;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = p * kSmallBits;
        for (uint32_t i = threadIdx.x; i < kSmallWaves * kSmallD; i += kSmallBlock) (&L.s.cnt[0][0])[i] = 0;
        __syncthreads();
#pragma unroll
        for (uint32_t c = 0; c < kSmallItems; ++c) {
            if (c >= nc) break;  // uniform
            const bool valid = wbase + c * 64 + lane < n;
            const uint32_t d = (k[c] >> shift) & (kSmallD - 1u);
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (uint32_t b = 0; b < kSmallBits; ++b) {
                const uint64_t bb = __ballot((d >> b) & 1u);
                peers &= ((d >> b) & 1u) ? bb : ~bb;
            }
            const uint32_t leader = peers ? (uint32_t)__builtin_ctzll(peers) : lane;
            uint32_t prior = 0;
            if (valid && lane == leader) {
                prior = L.s.cnt[w][d];
                L.s.cnt[w][d] = prior + (uint32_t)__popcll(peers);
            }
            prior = __shfl(prior, (int)leader);
            r[c] = prior + lanemask_lt_popc(peers);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        // thread d < kSmallD owns digit d: each wave's offset inside the
        // digit's run, then the runs' starts (an exclusive scan in thread
        // order, which is digit order)
        uint32_t run = 0;
        if (threadIdx.x < kSmallD)
            for (uint32_t q = 0; q < kSmallWaves; ++q) {
                const uint32_t t = L.s.cnt[q][threadIdx.x];
                L.s.cnt[q][threadIdx.x] = run;
                run += t;
            }
        uint32_t all;
        const uint32_t ex = block_exclusive_sum<kSmallBlock>(run, &all, sum_ws);
        if (threadIdx.x < kSmallD) L.s.dig_off[threadIdx.x] = ex;
        __syncthreads();
#pragma unroll
        for (uint32_t c = 0; c < kSmallItems; ++c) {
            if (c >= nc) break;
            if (wbase + c * 64 + lane < n) {
                const uint32_t d = (k[c] >> shift) & (kSmallD - 1u);
                const uint32_t pos = L.s.dig_off[d] + L.s.cnt[w][d] + r[c];
                L.s.sk[pos] = k[c];
                L.s.sv[pos] = v[c];
            }
        }
        __syncthreads();
        if (p + 1 < passes) {
#pragma unroll
            for (uint32_t c = 0; c < kSmallItems; ++c) {
                const uint32_t i = wbase + c * 64 + lane;
                if (c < nc && i < n) {
                    k[c] = L.s.sk[i];
                    v[c] = L.s.sv[i];
                }
            }
        }
    }
    if (passes == 0) {  // one cell: the op order is the sorted order
#pragma unroll
        for (uint32_t c = 0; c < kSmallItems; ++c) {
            const uint32_t i = wbase + c * 64 + lane;
            if (c < nc && i < n) {
                L.s.sk[i] = k[c];
                L.s.sv[i] = v[c];
            }
        }
        __syncthreads();
    }
    // the sorted list, thread t taking positions [t * nc, t * nc + nc);
    // positions past n read as all-ones keys (never a real key)
    const uint32_t first = threadIdx.x * nc;
#pragma unroll
    for (uint32_t j = 0; j < kSmallItems; ++j) {
        const bool in = j < nc && first + j < n;
        k[j] = in ? L.s.sk[first + j] : 0xFFFFFFFFu;
        v[j] = in ? L.s.sv[first + j] : kInvalid;
    }
    const uint32_t prev = first > 0 && first - 1 < n ? L.s.sk[first - 1] : 0xFFFFFFFFu;
    const uint32_t next = first + nc < n ? L.s.sk[first + nc] : 0xFFFFFFFFu;
    // each element's state, with the RMW deltas loaded once and all together
    // (loaded inside the walks below, one workgroup would wait for them one
    // round trip at a time)
    uint32_t dd[kSmallItems];
#pragma unroll
    for (uint32_t j = 0; j < kSmallItems; ++j) {
        const bool rmw = j < nc && first + j < n && (v[j] >> 30) == LTPMG_OP_RMW_ADD && !(v[j] & kInvalid);
        dd[j] = rmw ? (uint32_t)ov.args[v[j] & kOpMask] : 0u;
    }
    auto state = [&](uint32_t j) -> St {
        const uint32_t kind = v[j] >> 30;
        if (v[j] & kInvalid) return {kNoSet, 0u};
        if (kind == LTPMG_OP_UPDATE) return {v[j] & kOpMask, 0u};
        if (kind == LTPMG_OP_RMW_ADD) return {kRmwOnly, dd[j]};
        if (kind == 3u) return {kRmwOnly, inline_delta(v[j])};
        return {kNoSet, 0u};
    };
    Seg agg{0u, kNoSet, 0u};
#pragma unroll
    for (uint32_t j = 0; j < kSmallItems; ++j) {
        if (j >= nc || first + j >= n) break;
        const St s = state(j);
        const bool h = k[j] != (j ? k[j - 1] : prev);
        agg = seg_op(agg, Seg{h ? 1u : 0u, s.set, s.d});
    }
    Seg tot;
    // (its barriers also order every read of sk / sv above before the list
    // overwrites them)
    const Seg ex0 = block_seg_exclusive<kSmallBlock>(agg, &tot, seg_ws);
    auto walk = [&](auto &&visit) {
        St run{ex0.set, ex0.d};
#pragma unroll
        for (uint32_t j = 0; j < kSmallItems; ++j) {
            if (j >= nc || first + j >= n) break;
            const St s = state(j);
            const bool h = k[j] != (j ? k[j - 1] : prev);
            const St ex = h ? St{kNoSet, 0u} : run;
            run = h ? s : combine(run, s);
            const bool rd = (v[j] >> 30) == LTPMG_OP_READ && !(v[j] & kInvalid);
            const bool last = k[j] != (j + 1 < nc ? k[j + 1] : next);
            visit(j, ex, run, rd, last && run.set != kNoSet);
        }
    };
    uint32_t n_rd_t = 0, n_wr_t = 0;
    walk([&](uint32_t, St, St, bool rd, bool wr) {
        n_rd_t += rd;
        n_wr_t += wr;
    });
    // reads + write entries <= n <= kSmallMaxOps < 2^16: one word of the scan
    uint32_t both;
    const uint32_t pos = block_exclusive_sum<kSmallBlock>(n_rd_t | (n_wr_t << 16), &both, sum_ws);
    const uint32_t n_rd = both & 0xFFFFu, n_wr = both >> 16;
    uint32_t rp = pos & 0xFFFFu, wp = pos >> 16;
    walk([&](uint32_t j, St ex, St run, bool rd, bool wr) {
        if (rd) L.list[rp++] = make_uint4(v[j] & kOpMask, ex.set, ex.d, k[j]);
        if (wr) L.list[kSmallMaxOps - 1 - wp++] = make_uint4(k[j], run.set, run.d, 0u);
    });
    __syncthreads();
    // slots of several chunks per lane keep one move fewer in flight (the
    // deepest that stays within four waves per SIMD without spills)
    constexpr int RS = ONE ? kMoveDepthO : kMoveDepthO - 1;
    move_cells<G, VEC16, ONE, LCOLS, RS>(n_rd, [&](uint32_t q, uint32_t &key_out, St &st, uint32_t &op) {
        const uint4 e = L.list[q];
        key_out = e.w;
        st = St{e.y, e.z};
        op = e.x;
    }, dt, dt.dec, ov.args, results, res_stride, 0u, threadIdx.x / G, kSmallBlock / G, lcols);
    __syncthreads();  // every READ of the batch is done
    move_cells<G, VEC16, ONE, LCOLS, RS>(n_wr, [&](uint32_t q, uint32_t &key_out, St &st, uint32_t &op) {
        const uint4 e = L.list[kSmallMaxOps - 1 - q];
        key_out = e.x;
        st = St{e.y, e.z};
        op = kCellOp;
    }, dt, dt.dec, ov.args, results, res_stride, 0u, threadIdx.x / G, kSmallBlock / G, lcols);
    if (!ra.on) return;  // uniform
    // After the cell writes: a DELETE that took effect clears its row's
    // presence; the INSERTs that take effect get rows at the end of each
    // table in op order (rank over (wave, chunk, lane), which is op order).
    // The elements' meta words and rows are read again (their registers
    // held the sort).
    __shared__ uint32_t s_base;
    // k[c]: element c's meta word if the op takes effect, else 0xFFFF; v[c]:
    // its row (the registers are free again; every load is issued first)
#pragma unroll
    for (uint32_t c = 0; c < kSmallItems; ++c) {
        const uint32_t i = wbase + c * 64 + lane;
        k[c] = 0xFFFFu;
        v[c] = 0;
        r[c] = 0;
        if (c < nc && i < n) {
            k[c] = ov.meta[i];
            v[c] = ov.rows[i];
            r[c] = ra.found[i];
        }
    }
#pragma unroll
    for (uint32_t c = 0; c < kSmallItems; ++c) {
        if (!r[c]) k[c] = 0xFFFFu;  // kind 3, column 63: neither an INSERT nor a DELETE
        if (is_delete((uint16_t)k[c])) s_present[op_tab((uint16_t)k[c])][v[c]] = 0;
    }
    // Each table's INSERTs are listed in LDS in rank order (op index and
    // payload seed), then the whole workgroup writes the rows: presence and
    // result slots, then column by column with consecutive threads on
    // consecutive rows and the table's column descriptors in LDS (each lane
    // writing its own rows column by column took 125 of 239 us for a
    // 128-txn order-entry batch).
    uint32_t *ins_op = reinterpret_cast<uint32_t *>(L.list);                       // [kSmallMaxOps]
    unsigned long long *ins_arg = reinterpret_cast<unsigned long long *>(L.list) + kSmallMaxOps / 2;  // [kSmallMaxOps]
    __shared__ ColDesc tcols[LTPMG_MAX_COLS];
    uint32_t n_ins = 0;
    for (uint32_t mask = ra.insert_tables; mask; mask &= mask - 1) {
        const uint32_t t = __builtin_ctz(mask);  // uniform
        // whether element c of the lane is an INSERT into t that takes effect
        auto inserts = [&](uint32_t c) { return is_insert((uint16_t)k[c]) && op_tab((uint16_t)k[c]) == t; };
        uint32_t wave_n = 0;  // the wave's INSERTs into t
        for (uint32_t c = 0; c < nc; ++c) wave_n += (uint32_t)__popcll(__ballot(inserts(c)));
        if (lane == 0) sum_ws[w] = wave_n;
        if (threadIdx.x == 0) s_base = ra.cursor[t];
        if (threadIdx.x < LTPMG_MAX_COLS) tcols[threadIdx.x] = dt.cols[t * LTPMG_MAX_COLS + threadIdx.x];
        __syncthreads();  // (the first table's: every cell move, which reads the list, is done)
        uint32_t before = 0, total = 0;  // INSERTs of earlier waves; of the batch
        for (uint32_t q = 0; q < kSmallWaves; ++q) {
            before += q < w ? sum_ws[q] : 0u;
            total += sum_ws[q];
        }
        const TabDesc td = dt.tabs[t];
        const uint32_t base = s_base;
        for (uint32_t c = 0; c < nc; ++c) {
            const bool ins = inserts(c);
            const uint64_t bal = __ballot(ins);
            const uint32_t rank = before + lanemask_lt_popc(bal);
            before += (uint32_t)__popcll(bal);
            if (ins) {
                const uint32_t i = wbase + c * 64 + lane;
                ins_op[rank] = i;
                ins_arg[rank] = ov.args[i];
            }
        }
        __syncthreads();  // every rank listed; sum_ws and s_base read
        // rows past the capacity never occur with the host's capacity check
        const uint32_t fit = base >= td.n_rows ? 0u : (td.n_rows - base < total ? td.n_rows - base : total);
        if (threadIdx.x == 0) ra.cursor[t] = base + fit;
        n_ins += total;
        for (uint32_t q = threadIdx.x; q < total; q += kSmallBlock) {
            const uint32_t i = ins_op[q];
            if (q >= fit) {
                ra.found[i] = 0;
                continue;
            }
            td.present[base + q] = 1;
            // the row id and zeros to 16 bytes; the host zeroes the rest
            uint8_t *res = results + (uint64_t)i * res_stride;
            for (uint32_t b = 0; b < res_stride && b < 16; b += 4) st4(res + b, b == 0 ? base + q : 0u);
        }
        // one (row, column) cell per lane, column by column
        const uint32_t items = fit * td.n_cols;
        for (uint32_t i = threadIdx.x; i < items; i += kSmallBlock) {
            const uint32_t col = i / fit, q = i - col * fit;
            const ColDesc cd = tcols[col];
            uint8_t *cell = cd.base + (uint64_t)(base + q) * cd.slot;
            const uint64_t seed = rowseed(ins_arg[q], col);
            if (cd.slot >= 16) {
                for (uint32_t b = 0; b < cd.slot; b += 16) st16(cell + b, payload16(seed, b, cd.width));
            } else {
                for (uint32_t b = 0; b < cd.slot; b += 4) st4(cell + b, b < cd.width ? stream_word(seed, b >> 2) : 0u);
            }
        }
        __syncthreads();  // the list and tcols are read before the next table rewrites them
    }
    if (ra.insert_tables && threadIdx.x == 0) *ra.inserted = n_ins;
}

}  // namespace

size_t sort_partials_needed() { return (size_t)kMaxSortChunks << kMaxDigitBits; }

size_t sort_hist_words(size_t n) { return (n / kSortTile + 1) << kMaxDigitBits; }

#define LTPMG_RS_BITS(BITS_VAR, KERNEL, SRC, ...)                                            \
    do {                                                                                   \
        if (BITS_VAR == 9)                                                                 \
            hipLaunchKernelGGL((KERNEL<SRC, 9>), dim3(n_tiles), dim3(kBlock), 0, s, __VA_ARGS__); \
        else if (BITS_VAR == 7)                                                            \
            hipLaunchKernelGGL((KERNEL<SRC, 7>), dim3(n_tiles), dim3(kBlock), 0, s, __VA_ARGS__); \
        else                                                                               \
            hipLaunchKernelGGL((KERNEL<SRC, 8>), dim3(n_tiles), dim3(kBlock), 0, s, __VA_ARGS__); \
    } while (0)

#define LTPMG_RS_HIST(BITS_VAR, VEC, SRC, ...)                                                       \
    do {                                                                                            \
        if (BITS_VAR == 9)                                                                          \
            hipLaunchKernelGGL((k_rs_hist<SRC, 9, VEC>), dim3(n_tiles), dim3(kBlock), 0, s, __VA_ARGS__); \
        else if (BITS_VAR == 7)                                                                     \
            hipLaunchKernelGGL((k_rs_hist<SRC, 7, VEC>), dim3(n_tiles), dim3(kBlock), 0, s, __VA_ARGS__); \
        else                                                                                        \
            hipLaunchKernelGGL((k_rs_hist<SRC, 8, VEC>), dim3(n_tiles), dim3(kBlock), 0, s, __VA_ARGS__); \
    } while (0)

static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

void launch_rs_hist(hipStream_t s, SortBufs &b, int cur, uint32_t n, uint32_t shift, uint32_t bits) {
    const uint32_t n_tiles = (n + kSortTile - 1) / kSortTile;
    const SortedSrc src{b.keys[cur], b.vals[cur]};
    if (aligned16(src.k)) LTPMG_RS_HIST(bits, true, SortedSrc, src, n, shift, n_tiles, b.hist);
    else LTPMG_RS_HIST(bits, false, SortedSrc, src, n, shift, n_tiles, b.hist);
}

void launch_rs_scan(hipStream_t s, SortBufs &b, uint32_t n, uint32_t bits) {
    const uint32_t n_tiles = (n + kSortTile - 1) / kSortTile;
    const uint32_t n_chunks = (n_tiles + chunk_tiles(n_tiles) - 1) / chunk_tiles(n_tiles);
    if (bits == 9) {
        hipLaunchKernelGGL(k_rs_chunk_sum<9>, dim3(n_chunks), dim3(kBlock), 0, s, b.hist, n_tiles, b.partials);
        hipLaunchKernelGGL(k_rs_offsets<9>, dim3(n_chunks), dim3(kBlock), 0, s, b.hist, n_tiles, b.partials,
                           n_chunks);
    } else if (bits == 7) {
        hipLaunchKernelGGL(k_rs_chunk_sum<7>, dim3(n_chunks), dim3(kBlock), 0, s, b.hist, n_tiles, b.partials);
        hipLaunchKernelGGL(k_rs_offsets<7>, dim3(n_chunks), dim3(kBlock), 0, s, b.hist, n_tiles, b.partials,
                           n_chunks);
    } else {
        hipLaunchKernelGGL(k_rs_chunk_sum<8>, dim3(n_chunks), dim3(kBlock), 0, s, b.hist, n_tiles, b.partials);
        hipLaunchKernelGGL(k_rs_offsets<8>, dim3(n_chunks), dim3(kBlock), 0, s, b.hist, n_tiles, b.partials,
                           n_chunks);
    }
}

void launch_rs_scatter(hipStream_t s, SortBufs &b, int cur, uint32_t n, uint32_t shift, uint32_t bits) {
    const uint32_t n_tiles = (n + kSortTile - 1) / kSortTile;
    LTPMG_RS_BITS(bits, k_rs_scatter, SortedSrc, SortedSrc{b.keys[cur], b.vals[cur]}, b.keys[cur ^ 1],
                  b.vals[cur ^ 1], n, shift, n_tiles, b.hist);
}

void launch_rs_hist_ops(hipStream_t s, DevTables dt, OpView ov, SortBufs &b, uint32_t bits) {
    const uint32_t n = ov.n_ops, n_tiles = (n + kSortTile - 1) / kSortTile;
    const OpSrc src{dt.tabs, ov.rows, ov.meta, ov.args};
    if (aligned16(ov.rows) && aligned16(ov.meta)) LTPMG_RS_HIST(bits, true, OpSrc, src, n, 0u, n_tiles, b.hist);
    else LTPMG_RS_HIST(bits, false, OpSrc, src, n, 0u, n_tiles, b.hist);
}

void launch_rs_scatter_ops(hipStream_t s, DevTables dt, OpView ov, SortBufs &b, uint32_t bits) {
    const uint32_t n = ov.n_ops, n_tiles = (n + kSortTile - 1) / kSortTile;
    LTPMG_RS_BITS(bits, k_rs_scatter, OpSrc, OpSrc{dt.tabs, ov.rows, ov.meta, ov.args}, b.keys[1], b.vals[1], n,
                  0u, n_tiles, b.hist);
}

// The pass count of 8-bit digits, with the narrowest digit (7 or 8 bits)
// that keeps it: half as many digits make each digit's run in the scatter
// twice as long, so fewer of its lines are written in part (measured 0.05 ms
// per C2 batch faster than 8 bits).  With LTPMG_RADIX_MAX_BITS=9, 9-bit digits are taken when
// they save a pass; on MI355X that measured slower for the 27-bit unpadded
// keys of a 10M x 10 table (three 9-bit passes 1.28 ms per batch, four 7-bit passes
// 1.22 ms): a 9-bit scatter pass takes 0.10 ms against 0.07 ms, and its
// digit-offset scan reads four times the counts.
#ifndef LTPMG_RADIX_MAX_BITS
#define LTPMG_RADIX_MAX_BITS 8
#endif
#ifndef LTPMG_RADIX_MIN_BITS
#define LTPMG_RADIX_MIN_BITS 7
#endif
uint32_t radix_digit_bits(uint32_t key_bits) {
    if (LTPMG_RADIX_MAX_BITS >= 9 && (key_bits + 8) / 9 < (key_bits + 7) / 8) return 9u;
    const uint32_t passes = (key_bits + 7) / 8;
    return LTPMG_RADIX_MIN_BITS <= 7 && passes * 7 >= key_bits ? 7u : 8u;
}

int launch_radix_sort(hipStream_t s, SortBufs &b, uint32_t n, uint32_t key_bits) {
    int cur = 0;
    if (n <= 1) return cur;
    const uint32_t bits = radix_digit_bits(key_bits);
    for (uint32_t shift = 0; shift < key_bits; shift += bits) {
        launch_rs_hist(s, b, cur, n, shift, bits);
        launch_rs_scan(s, b, n, bits);
        launch_rs_scatter(s, b, cur, n, shift, bits);
        cur ^= 1;
    }
    return cur;
}

void launch_mo_keys(hipStream_t s, DevTables dt, OpView ov, uint32_t *keys, uint32_t *vals) {
    uint32_t blocks = (ov.n_ops + kBlock - 1) / kBlock;
    if (blocks > 16384) blocks = 16384;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_mo_keys, dim3(blocks), dim3(kBlock), 0, s, dt, ov, keys, vals);
}

// G lanes per cell move (a 16-byte chunk each) from the result stride, and
// past 128 bytes from dt.wide_lanes; ONE: one chunk per lane covers a slot.
// (Lane groups that always cover the stride, G = 16 for 256-byte slots,
// measured slower on order-entry batches: 0.356 against 0.250 ms for a
// 4,863-op batch through the pipeline, half as many moves in flight.)
#define LTPMG_MO_LAUNCH_B(KERNEL, G_, LC, GRID, BLOCK, ...)                                         \
    do {                                                                                           \
        if (res_stride == G_ * 16)                                                                 \
            hipLaunchKernelGGL((KERNEL<G_, true, true, LC>), GRID, BLOCK, 0, s, __VA_ARGS__);         \
        else                                                                                       \
            hipLaunchKernelGGL((KERNEL<G_, true, false, LC>), GRID, BLOCK, 0, s, __VA_ARGS__);        \
    } while (0)
#define LTPMG_MO_DISPATCH_B(KERNEL, LC, GRID, BLOCK, ...)                                       \
    do {                                                                                        \
        if (res_stride > 128 && dt.wide_lanes == 2)                                             \
            LTPMG_MO_LAUNCH_B(KERNEL, 2, LC, GRID, BLOCK, __VA_ARGS__);                         \
        else if (res_stride >= 128)                                                             \
            LTPMG_MO_LAUNCH_B(KERNEL, 8, LC, GRID, BLOCK, __VA_ARGS__);                         \
        else if (res_stride >= 64)                                                              \
            LTPMG_MO_LAUNCH_B(KERNEL, 4, LC, GRID, BLOCK, __VA_ARGS__);                         \
        else if (res_stride >= 32)                                                              \
            LTPMG_MO_LAUNCH_B(KERNEL, 2, LC, GRID, BLOCK, __VA_ARGS__);                         \
        else if (res_stride >= 16)                                                              \
            LTPMG_MO_LAUNCH_B(KERNEL, 1, LC, GRID, BLOCK, __VA_ARGS__);                         \
        else                                                                                    \
            hipLaunchKernelGGL((KERNEL<1, false, false, LC>), GRID, BLOCK, 0, s, __VA_ARGS__);  \
    } while (0)
#define LTPMG_MO_DISPATCH(KERNEL, LC, GRID, ...) LTPMG_MO_DISPATCH_B(KERNEL, LC, GRID, dim3(kBlock), __VA_ARGS__)

void launch_mo_seg_scan(hipStream_t s, OpView ov, OrderBufs &ob, int sorted) {
    const uint32_t n = ov.n_ops;
    if (n == 0) return;
    const uint32_t n_tiles = (n + kSegTile - 1) / kSegTile;
    OrderSrc src{ob.sort.keys[sorted], ob.sort.vals[sorted], ov.args, n};
    hipLaunchKernelGGL(k_seg_reduce, dim3(n_tiles), dim3(kBlock), 0, s, src, ob.tile_flag,
                       ob.tile_set, ob.tile_delta, ob.wtotal);
}

// Streams the batch's payload seeds (8 bytes per op, in op order) through
// the memory-side cache just before the apply kernel, whose seed reads come
// in sorted-cell order, one random 8-byte read each.  The loads' values are
// folded into a word that is stored only if it equals a constant, so the
// loads stay.
__global__ __launch_bounds__(kBlock) void k_mo_warm_seeds(const uint4 *p, uint32_t n, uint32_t *sink) {
    uint32_t acc = 0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint4 v = p[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9E3779B9u) *sink = acc;
}

void launch_mo_warm_seeds(hipStream_t s, OpView ov, OrderBufs &ob) {
    // the MI355X memory-side cache holds 256 MB; a larger seed array would
    // only push itself out
    constexpr uint64_t kWarmMax = 192ull << 20;
    if ((uint64_t)ov.n_ops * 8 > kWarmMax || (reinterpret_cast<uintptr_t>(ov.args) & 15)) return;
    const uint32_t n4 = ov.n_ops / 2;  // whole 16-byte pairs of seeds
    if (n4 == 0) return;
    const uint32_t blocks = std::min<uint32_t>(4096u, (n4 + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_mo_warm_seeds, dim3(blocks), dim3(kBlock), 0, s,
                       reinterpret_cast<const uint4 *>(ov.args), n4, ob.wtotal + 1);
}

void launch_mo_apply_reads(hipStream_t s, const DevTables &dt, OpView ov, OrderBufs &ob, int sorted,
                           uint8_t *results, uint32_t res_stride, bool pad) {
    const uint32_t res_fill = pad ? res_stride : 0u;
    const uint32_t n = ov.n_ops;
    if (n == 0) return;
    const uint32_t n_tiles = (n + kSegTile - 1) / kSegTile;
    OrderSrc src{ob.sort.keys[sorted], ob.sort.vals[sorted], ov.args, n};
    // the column list in LDS when it holds every live column; past 128-byte
    // strides with narrow tables, lanes per move chosen per tile
    if (res_stride > 128 && dt.dec.narrow != 0) {
        if (dt.n_clist >= dt.dec.n_cols_all)
            hipLaunchKernelGGL((k_mo_apply_reads<8, true, false, true, true>), dim3(n_tiles), dim3(kBlock), 0, s, src,
                               ob.tile_flag, ob.tile_set, ob.tile_delta, dt, results, res_stride, res_fill, ob.wlist,
                               ob.wtotal);
        else
            hipLaunchKernelGGL((k_mo_apply_reads<8, true, false, false, true>), dim3(n_tiles), dim3(kBlock), 0, s,
                               src, ob.tile_flag, ob.tile_set, ob.tile_delta, dt, results, res_stride, res_fill,
                               ob.wlist, ob.wtotal);
    } else if (dt.n_clist >= dt.dec.n_cols_all)
        LTPMG_MO_DISPATCH(k_mo_apply_reads, true, dim3(n_tiles), src, ob.tile_flag, ob.tile_set, ob.tile_delta,
                          dt, results, res_stride, res_fill, ob.wlist, ob.wtotal);
    else
        LTPMG_MO_DISPATCH(k_mo_apply_reads, false, dim3(n_tiles), src, ob.tile_flag, ob.tile_set, ob.tile_delta,
                          dt, results, res_stride, res_fill, ob.wlist, ob.wtotal);
}

void launch_mo_final_writes(hipStream_t s, const DevTables &dt, OpView ov, OrderBufs &ob,
                            uint32_t res_stride, uint8_t *commit) {
    if (ov.n_ops == 0) {
        launch_fill_u8(s, commit, ov.n_txns, 1);
        return;
    }
    // at most one entry per scan tile; one lane group per entry
    const uint32_t n_tiles = (ov.n_ops + kSegTile - 1) / kSegTile;
    const uint32_t blocks = (n_tiles * 8 + kBlock - 1) / kBlock;
    LTPMG_MO_DISPATCH(k_mo_final_writes, false, dim3(blocks), ob.wlist, ob.wtotal, dt, ov.args, commit,
                      ov.n_txns);
}

void launch_mo_small(hipStream_t s, const DevTables &dt, OpView ov, uint32_t key_bits, uint8_t *results,
                     uint32_t res_stride, uint8_t *commit, const RowBufs *rb) {
    if (ov.n_ops == 0 || ov.n_ops > kSmallMaxOps || key_bits > 32) return;
    SmallRows ra{};
    if (rb && rb->has_append) ra = SmallRows{rb->found, rb->cursor, rb->inserted, rb->insert_tables, 1u};
    if (dt.n_clist >= dt.dec.n_cols_all)
        LTPMG_MO_DISPATCH_B(k_mo_small, true, dim3(1), dim3(kSmallBlock), dt, ov, key_bits, results, res_stride,
                            commit, ra);
    else
        LTPMG_MO_DISPATCH_B(k_mo_small, false, dim3(1), dim3(kSmallBlock), dt, ov, key_bits, results, res_stride,
                            commit, ra);
}

void launch_selftest_segscan(hipStream_t s, const uint32_t *vals, const uint8_t *head,
                             uint32_t *out, uint32_t n, uint32_t *scratch) {
    if (n == 0) return;
    const uint32_t n_tiles = (n + kSegTile - 1) / kSegTile;
    uint32_t *tf = scratch, *ts = scratch + n_tiles, *td = scratch + 2 * n_tiles;
    SelftestSrc src{vals, head, n};
    hipLaunchKernelGGL(k_selftest_reduce, dim3(n_tiles), dim3(kBlock), 0, s, src, tf, ts, td);
    hipLaunchKernelGGL(k_selftest_apply, dim3(n_tiles), dim3(kBlock), 0, s, src, tf, ts, td, out);
}

uint32_t seg_tiles(uint32_t n) { return (n + kSegTile - 1) / kSegTile; }

}  // namespace ltpmg
