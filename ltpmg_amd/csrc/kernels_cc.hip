// kernels_cc.hip -- lowest-TID-wins concurrency control and execution.
//
// Replaces the reference's registration + apply pipeline
// (cuda/src/gpulauncher.cu:459-512 txn_executor/register_cc, 1558-2700
// prefix/analyze/filter_commit/commit) for the mode in which txns abort.
// The reference only sketches that rule in comments (atomicMin on a per-row
// min_TID, commit iff min_TID == tid, gpulauncher.cu:414, 2130-2142); here
// it is three passes over the SoA batch:
//
//   cc_reserve        write op of txn t on row k: atomicMax(lock[k], epoch<<32 | ~t)
//                     (a tile's ops on one row meet in LDS first; skipped when
//                     the lock already shows a lower TID)
//   cc_validate_read  txn t commits iff no op's lock holds a lower TID of this
//                     epoch; committed READs copy the cell to the result slot
//   cc_write          committed writes (UPDATE payload / RMW_ADD)
//
// Reads finish before any write starts (kernel boundary): every committed
// reader of a row written by committed txn w has TID <= w, so reads see the
// batch-start snapshot, which is what a serial replay in TID order gives.
//
// Work split: a workgroup owns the txns whose first op falls in its
// 2048-op tile, so it validates whole txns without cross-workgroup
// traffic; the op -> txn map of the tile is rebuilt in LDS from txn_off
// (4 bytes per txn) instead of shipping a TID per op.
//
// The epoch lives in device memory: tile_bounds (or the fused kernel) moves
// it on once per batch, so a captured launch sequence replays correctly.
//
// cc_fused runs the three passes in one persistent launch for small and
// mid-size batches: each workgroup takes a contiguous run of tiles, finds
// their txn bounds itself, and passes two grid barriers (reserve | validate
// + read | write); with one tile per workgroup the op -> txn map and the
// commit flags stay in LDS across all three passes.
#include "internal.h"

namespace ltpmg {

namespace {

struct Tile {
    uint32_t t0, nt, op0, nops;
};

// Txns [t0, t1) of a tile; fills owner[o - op0] = txn index in the tile.
// max_nt / max_ops are the LDS capacities of owner[] and the caller's
// per-txn arrays.  The host checks txn_off before a batch runs, but a batch
// handed over on the device (or submitted as trusted) is only checked here:
// a tile whose offsets are not nondecreasing, or that would overrun LDS or
// the batch, becomes empty, so a malformed batch never makes a kernel
// access memory outside its buffers (its txns then simply do not commit).
__device__ inline Tile fill_owner(const OpView &ov, uint32_t t0, uint32_t t1, uint16_t *owner,
                                  uint32_t max_nt = kTileOps, uint32_t max_ops = kTileOpsMax) {
    __shared__ uint32_t s_bad;
    __syncthreads();  // the previous user of owner[] and s_bad is done
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    bool bad = t1 < t0 || t1 > ov.n_txns || t1 - t0 > max_nt;
    const uint32_t op0 = bad ? 0u : ov.txn_off[t0];
    const uint32_t end = bad ? 0u : ov.txn_off[t1];
    bad = bad || end < op0 || end > ov.n_ops || end - op0 > max_ops;
    Tile t{t0, bad ? 0u : t1 - t0, op0, bad ? 0u : end - op0};
    for (uint32_t j = threadIdx.x; j < t.nt; j += blockDim.x)
        if (ov.txn_off[t.t0 + j] > ov.txn_off[t.t0 + j + 1]) s_bad = 1;
    __syncthreads();
    if (s_bad) t.nt = t.nops = 0;  // uniform: every thread read the same flag
    for (uint32_t j = threadIdx.x; j < t.nt; j += blockDim.x) {
        uint32_t a = ov.txn_off[t.t0 + j] - t.op0, b = ov.txn_off[t.t0 + j + 1] - t.op0;
        for (uint32_t o = a; o < b; ++o) owner[o] = (uint16_t)j;
    }
    __syncthreads();
    return t;
}

// Appends v to an LDS list with one LDS atomic per wave.
__device__ inline void wave_append(bool pred, uint32_t v, uint32_t *list, uint32_t *count) {
    uint64_t m = __ballot(pred);
    if (m == 0) return;
    uint32_t leader = __builtin_ctzll(m);
    uint32_t base = 0;
    if (__lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    if (pred) list[base + lanemask_lt_popc(m)] = v;
}

// Moves the cells of the ops in an LDS list, G lanes per op and R ops per
// round, all independent loads of a round issued before its stores.
// READS: cell -> result slot (res_stride bytes, zero past the slot).
// Otherwise: UPDATE writes the payload of args[o], RMW_ADD adds (u32)args[o].
template <int G, bool VEC16, bool READS>
__device__ inline void move_ops(uint32_t n, const uint32_t *list, const OpView &ov, DevTables dt,
                                uint8_t *results, uint32_t res_stride) {
    constexpr int R = kMoveDepth;
    constexpr uint32_t groups = kBlock / G;
    const uint32_t g = threadIdx.x / G, lane = threadIdx.x % G, b0 = lane * 16;
    for (uint32_t base = g; base < n; base += groups * R) {
        uint32_t o[R];
        ColDesc cd[R];
        uint8_t *cell[R];
        uint64_t arg[R];
        bool upd[R];
        u32x4 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t k = base + r * groups;
            cell[r] = nullptr;
            cd[r] = ColDesc{nullptr, 0u, 0u};
            upd[r] = false;
            o[r] = 0;
            if (k < n) {
                o[r] = list[k];
                const uint16_t m = ov.meta[o[r]];
                cd[r] = dt.cols[op_tab(m) * LTPMG_MAX_COLS + op_col(m)];
                cell[r] = cd[r].base + (uint64_t)ov.rows[o[r]] * cd[r].slot;
                upd[r] = op_kind(m) == LTPMG_OP_UPDATE;
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            arg[r] = (!READS && cell[r]) ? ov.args[o[r]] : 0ull;
            v[r] = u32x4{0u, 0u, 0u, 0u};
            if (VEC16 && cell[r] && !upd[r] && cd[r].slot >= 16 && b0 < cd[r].slot) v[r] = ld16(cell[r] + b0);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!cell[r]) continue;
            if (READS) {
                uint8_t *dst = results + (uint64_t)o[r] * res_stride;
                if (VEC16 && cd[r].slot >= 16) {
                    for (uint32_t b = b0; b < res_stride; b += G * 16) {
                        u32x4 x = b < cd[r].slot ? (b == b0 ? v[r] : ld16(cell[r] + b)) : u32x4{0u, 0u, 0u, 0u};
                        st16_nt(dst + b, x);
                    }
                } else {  // 4- or 8-byte slots (also a narrow column beside wider ones)
                    for (uint32_t b = lane * 4; b < res_stride; b += G * 4) {
                        uint32_t x = b < cd[r].slot ? ld4(cell[r] + b) : 0u;
                        st4(dst + b, x);
                    }
                }
            } else if (VEC16 && cd[r].slot >= 16) {
                for (uint32_t b = b0; b < cd[r].slot; b += G * 16) {
                    u32x4 x = upd[r] ? payload16(arg[r], b, cd[r].width)
                                     : add16(b == b0 ? v[r] : ld16(cell[r] + b), (uint32_t)arg[r], b, cd[r].width);
                    st16(cell[r] + b, x);
                }
            } else {
                for (uint32_t b = lane * 4; b < cd[r].slot; b += G * 4) {
                    uint8_t *c = cell[r] + b;
                    if (b < cd[r].width) st4(c, upd[r] ? stream_word(arg[r], b >> 2) : ld4(c) + (uint32_t)arg[r]);
                }
            }
        }
    }
}


// The lock word an invalid op reads in the validation pass (never written,
// so it holds no tag of any epoch).
__device__ unsigned long long g_no_lock = 0ull;

// Pass 1 on one tile: every valid write op offers its TID for its row.  The
// tile's ops on one row first meet in an LDS hash (key = table << 32 | row,
// value = the lowest TID), so under Zipf a hot row costs the tile one
// global check and at most one atomicMax instead of one per op.  An op
// whose slot holds another row goes to memory by itself.  keys / tids are
// n_slots (a power of two) words of LDS scratch; the check reads the lock
// word at device scope, so it skips the atomic when a lower TID's tag (or
// ours) is already there.
__device__ inline void tile_reserve(const DevTables &dt, const OpView &ov, const Tile &t,
                                    const uint16_t *owner, uint32_t epoch, unsigned long long *keys,
                                    uint32_t *tids, uint32_t n_slots) {
    constexpr unsigned long long kEmpty = ~0ull;
    const unsigned long long ep = (unsigned long long)epoch << 32;
    // {rows (0 for a dead table), columns, lock words} per table, in LDS,
    // and a bit per append table
    __shared__ uint4 tc[LTPMG_MAX_TABLES];
    __shared__ uint32_t s_append;
    if (threadIdx.x < LTPMG_MAX_TABLES) {
        const TabDesc &td = dt.tabs[threadIdx.x];
        const uint64_t lp = reinterpret_cast<uint64_t>(td.lock);
        tc[threadIdx.x] = make_uint4(td.live ? td.n_rows : 0u, td.n_cols, (uint32_t)lp, (uint32_t)(lp >> 32));
    }
    if (threadIdx.x < 64) {  // wave 0
        const uint64_t b = __ballot(threadIdx.x < LTPMG_MAX_TABLES && dt.tabs[threadIdx.x].present != nullptr);
        if (threadIdx.x == 0) s_append = (uint32_t)b;
    }
    auto offer = [&](uint32_t tab, uint32_t row, uint32_t tid) {
        const uint4 c = tc[tab];
        const unsigned long long tag = ep | (uint32_t)~tid;
        // global-address-space pointer: global (not flat) load and atomic
        LTPMG_GLOBAL unsigned long long *p =
            (LTPMG_GLOBAL unsigned long long *)(((uint64_t)c.w << 32) | c.z) + row;
        const unsigned long long cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur < tag) __hip_atomic_fetch_max(p, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    for (uint32_t i = threadIdx.x; i < n_slots; i += blockDim.x) {
        keys[i] = kEmpty;
        tids[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    // the ops' own loads go out eight per thread at a time; the lock words
    // are still read one row at a time below (reading them together measured
    // slower: a later read sees more lower-TID tags and skips its atomic)
    constexpr uint32_t kBatch = 8;
    for (uint32_t base = 0; base < t.nops; base += kBatch * blockDim.x) {
        uint32_t mm[kBatch], rr[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const uint32_t i = base + j * blockDim.x + threadIdx.x;
            const uint32_t o = t.op0 + (i < t.nops ? i : t.nops - 1u);
            mm[j] = ov.meta[o];
            rr[j] = ov.rows[o];
        }
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const uint32_t i = base + j * blockDim.x + threadIdx.x;
            const uint16_t m = (uint16_t)mm[j];
            const uint32_t row = rr[j];
            // writes (UPDATE, RMW_ADD, DELETE) reserve their row; an INSERT
            // appends a new row and reserves nothing
            if (i >= t.nops || op_kind(m) == LTPMG_OP_READ || is_insert(m)) continue;
            const uint4 c = tc[op_tab(m)];
            // never touch memory outside the tables
            if (!op_check(m, row, c.x, c.y, (s_append >> op_tab(m)) & 1u)) continue;
            const unsigned long long key = ((unsigned long long)op_tab(m) << 32) | row;
            const uint32_t slot = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 40) & (n_slots - 1);
            const unsigned long long prev = atomicCAS(&keys[slot], kEmpty, key);
            if (prev == kEmpty || prev == key) atomicMin(&tids[slot], t.t0 + owner[i]);
            else offer(op_tab(m), row, t.t0 + owner[i]);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n_slots; i += blockDim.x) {
        const unsigned long long key = keys[i];
        if (key != kEmpty) offer((uint32_t)(key >> 32), (uint32_t)key, tids[i]);
    }
}

// Pass 2 on one tile: commit flags into ok[] and commit[], committed READs
// copied to their result slots.  Returns the tile's committed count.  With
// found (a batch that touches append tables), found[o] gets whether op o's
// row is present at the batch start -- what a committed op sees, since a
// committed DELETE excludes every higher TID from its row -- and a committed
// READ of an absent row gets a zeroed slot instead of the cell.
template <int G, bool VEC16>
__device__ inline uint32_t tile_validate_read(const DevTables &dt, const OpView &ov, const Tile &t,
                                              const uint16_t *owner, uint32_t epoch, uint8_t *ok,
                                              uint32_t *list, uint8_t *commit, uint8_t *results,
                                              uint32_t res_stride, uint8_t *found) {
    __shared__ uint32_t n_list, n_commit;
    // {rows (0 for a dead table), columns, lock words} per table, and the
    // presence bytes of append tables
    __shared__ uint4 tc[LTPMG_MAX_TABLES];
    __shared__ const uint8_t *s_pres[LTPMG_MAX_TABLES];
    __syncthreads();  // the previous pass's readers of ok[] and list[] are done
    for (uint32_t j = threadIdx.x; j < t.nt; j += blockDim.x) ok[j] = 1;
    if (threadIdx.x == 0) n_list = n_commit = 0;
    if (threadIdx.x < LTPMG_MAX_TABLES) {
        const TabDesc &td = dt.tabs[threadIdx.x];
        const uint64_t lp = reinterpret_cast<uint64_t>(td.lock);
        tc[threadIdx.x] = make_uint4(td.live ? td.n_rows : 0u, td.n_cols, (uint32_t)lp, (uint32_t)(lp >> 32));
        s_pres[threadIdx.x] = td.present;
    }
    __syncthreads();
    // kBatch ops per thread at a time, every load of a batch issued before
    // any result is used: first the ops, then their lock words (an invalid
    // op reads a dummy word instead, so no load is conditional)
    constexpr uint32_t kBatch = 8;
    for (uint32_t base = 0; base < t.nops; base += kBatch * blockDim.x) {
        uint32_t mm[kBatch], rr[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const uint32_t i = base + j * blockDim.x + threadIdx.x;
            const uint32_t o = t.op0 + (i < t.nops ? i : t.nops - 1u);
            mm[j] = ov.meta[o];
            rr[j] = ov.rows[o];
        }
        unsigned long long lk[kBatch];
        bool valid[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const uint16_t m = (uint16_t)mm[j];
            const uint4 c = tc[op_tab(m)];
            valid[j] = op_check(m, rr[j], c.x, c.y, s_pres[op_tab(m)] != nullptr);
            // an INSERT touches no existing row: no lock to check
            const unsigned long long *p =
                valid[j] && !is_insert(m)
                    ? reinterpret_cast<const unsigned long long *>(((uint64_t)c.w << 32) | c.z) + rr[j]
                    : &g_no_lock;
            lk[j] = *(const LTPMG_GLOBAL unsigned long long *)p;
        }
        if (found) {  // presence at the batch start (INSERTs: decided when appended)
#pragma unroll
            for (uint32_t j = 0; j < kBatch; ++j) {
                const uint32_t i = base + j * blockDim.x + threadIdx.x;
                if (i >= t.nops) continue;
                const uint16_t m = (uint16_t)mm[j];
                const uint8_t *pres = s_pres[op_tab(m)];
                found[t.op0 + i] = (valid[j] && pres && !is_insert(m)) ? pres[rr[j]] : (uint8_t)1;
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const uint32_t i = base + j * blockDim.x + threadIdx.x;
            if (i >= t.nops) continue;
            if (!valid[j]) {  // an invalid op aborts its txn
                ok[owner[i]] = 0;
                continue;
            }
            const uint32_t tid = t.t0 + owner[i];
            if ((uint32_t)(lk[j] >> 32) == epoch && (uint32_t)~(uint32_t)lk[j] < tid) ok[owner[i]] = 0;
        }
    }
    __syncthreads();
    uint32_t mine = 0;
    for (uint32_t j = threadIdx.x; j < t.nt; j += blockDim.x) {
        commit[t.t0 + j] = ok[j];
        mine += ok[j];
    }
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
    if (__lane_id() == 0 && mine) atomicAdd(&n_commit, mine);
    // committed READs of this tile -> LDS list (loop bound uniform per wave);
    // a READ of an absent row gets zeros.  found[] of op i was written by this
    // same thread above (both loops give op i to thread i % blockDim.x); an
    // op of an aborted txn takes no effect, so found[] becomes presence AND
    // commit, all the row-op kernels after this one need.  An INSERT of a
    // committed txn is valid (an invalid op aborts its txn): its byte becomes
    // 2 + its table, which the insert kernels read back.
    for (uint32_t base = 0; base < t.nops; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool take = false;
        if (i < t.nops) {
            const bool c = ok[owner[i]] != 0;
            const uint16_t m = ov.meta[t.op0 + i];
            take = op_kind(m) == LTPMG_OP_READ && c;
            if (found && !c) found[t.op0 + i] = 0;
            else if (found && is_insert(m)) found[t.op0 + i] = (uint8_t)(2u + op_tab(m));
        }
        if (take && found && !found[t.op0 + i]) {
            for (uint32_t b = 0; b < res_stride; b += 4) st4(results + (uint64_t)(t.op0 + i) * res_stride + b, 0u);
            take = false;
        }
        wave_append(take, t.op0 + i, list, &n_list);
    }
    __syncthreads();
    const uint32_t committed = n_commit;
    move_ops<G, VEC16, true>(n_list, list, ov, dt, results, res_stride);
    return committed;
}

// Pass 3 on one tile: committed writes.  ok[] holds the tile's commit
// flags, or is filled here from commit[] when fill_ok is set.
// With found, a write of a row absent at the batch start is skipped (row
// ops are applied by their own kernels afterwards).
template <int G, bool VEC16>
__device__ inline void tile_write(const DevTables &dt, const OpView &ov, const Tile &t,
                                  const uint16_t *owner, uint8_t *ok, bool fill_ok,
                                  const uint8_t *commit, uint32_t *list, const uint8_t *found) {
    __shared__ uint32_t n_list;
    __syncthreads();  // the previous pass's readers of ok[] and list[] are done
    if (fill_ok)
        for (uint32_t j = threadIdx.x; j < t.nt; j += blockDim.x) ok[j] = commit[t.t0 + j];
    if (threadIdx.x == 0) n_list = 0;
    __syncthreads();
    for (uint32_t base = 0; base < t.nops; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool take = false;
        if (i < t.nops && ok[owner[i]]) {
            uint32_t o = t.op0 + i;
            take = op_kind(ov.meta[o]) != LTPMG_OP_READ && op_valid(dt.tabs, ov.meta[o], ov.rows[o]) &&
                   (!found || found[o]);
        }
        wave_append(take, t.op0 + i, list, &n_list);
    }
    __syncthreads();
    move_ops<G, VEC16, false>(n_list, list, ov, dt, nullptr, 0u);
}

// Smallest t in [0, n] with a[t] >= target, given a[n] >= target and a
// nondecreasing.  256 samples per round: a 1M-txn batch takes 3 rounds of
// one load per thread.
__device__ inline uint32_t block_lower_bound(const uint32_t *a, uint32_t n, uint64_t target) {
    __shared__ uint32_t s_lo, s_hi, s_first;
    __syncthreads();  // earlier readers of s_lo / s_hi are done
    if (threadIdx.x == 0) {
        s_lo = 0;
        s_hi = n;
    }
    __syncthreads();
    while (true) {
        const uint32_t lo = s_lo, hi = s_hi;
        if (lo >= hi) break;  // uniform: every thread read the same values
        const uint32_t p = lo + (uint32_t)(((uint64_t)(hi - lo) * threadIdx.x) / (kBlock - 1));
        const bool ge = (uint64_t)a[p] >= target;
        if (threadIdx.x == 0) s_first = kBlock - 1;  // a[hi] >= target
        __syncthreads();  // s_first reset, s_lo / s_hi read by all
        if (ge) atomicMin(&s_first, threadIdx.x);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t j = s_first;
            if (j == 0) {
                s_hi = lo;
            } else {
                s_lo = lo + (uint32_t)(((uint64_t)(hi - lo) * (j - 1)) / (kBlock - 1)) + 1;
                s_hi = lo + (uint32_t)(((uint64_t)(hi - lo) * j) / (kBlock - 1));
            }
        }
        __syncthreads();
    }
    return s_lo;
}

// Reusable grid barrier (count + generation).  The last workgroup to arrive
// resets the count and moves the generation on; the others wait for the
// generation with a bounded spin, so every wave leaves even if the grid were
// not co-resident (then err is set and the kernel stops).
__device__ inline bool grid_sync(const GridBar &b) {
    __shared__ uint32_t s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = __hip_atomic_load(b.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();  // this workgroup's writes before its arrival
        uint32_t ok = 1;
        if (atomicAdd(b.count, 1u) == gridDim.x - 1) {
            __hip_atomic_store(b.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            atomicAdd(b.gen, 1u);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(b.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > b.spin_limit) {
                    atomicOr(b.err, 1u);
                    ok = 0;
                    break;
                }
            }
        }
        __threadfence();  // the other workgroups' writes after leaving
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

}  // namespace

// Every txn has >= 1 op, so a tile owns <= kTileOps txns and
// < kTileOps + LTPMG_MAX_TXN_OPS ops.
__global__ void k_tile_bounds(const uint32_t *txn_off, uint32_t n_txns, uint32_t n_tiles,
                              uint32_t *tile_t0, uint32_t *epoch_ctr) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *epoch_ctr += 1;  // this batch's epoch, read by the next kernels
    if (k > n_tiles) return;
    uint64_t target = (uint64_t)k * kTileOps;
    uint32_t lo = 0, hi = n_txns;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)txn_off[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    tile_t0[k] = lo;
}

constexpr uint32_t kReserveSlots = 2048;  // LDS hash of a tile's written rows

__global__ __launch_bounds__(kBlock) void k_cc_reserve(DevTables dt, OpView ov,
                                                       const uint32_t *tile_t0,
                                                       const uint32_t *epoch_ctr) {
    __shared__ uint16_t owner[kTileOpsMax];
    __shared__ unsigned long long keys[kReserveSlots];
    __shared__ uint32_t tids[kReserveSlots];
    const uint32_t epoch = *epoch_ctr;
    Tile t = fill_owner(ov, tile_t0[blockIdx.x], tile_t0[blockIdx.x + 1], owner);
    tile_reserve(dt, ov, t, owner, epoch, keys, tids, kReserveSlots);
}

// G lanes cooperate on one op; each moves 16 bytes (VEC16) or a 4-byte word.
template <int G, bool VEC16>
__global__ __launch_bounds__(kBlock) void k_cc_validate_read(DevTables dt, OpView ov,
                                                             const uint32_t *tile_t0,
                                                             const uint32_t *epoch_ctr,
                                                             uint8_t *commit, uint8_t *results,
                                                             uint32_t res_stride,
                                                             uint32_t *counters, uint8_t *found) {
    __shared__ uint16_t owner[kTileOpsMax];
    __shared__ uint8_t ok[kTileOps];
    __shared__ uint32_t list[kTileOpsMax];
    const uint32_t epoch = *epoch_ctr;
    Tile t = fill_owner(ov, tile_t0[blockIdx.x], tile_t0[blockIdx.x + 1], owner);
    uint32_t c = tile_validate_read<G, VEC16>(dt, ov, t, owner, epoch, ok, list, commit, results,
                                              res_stride, found);
    if (threadIdx.x == 0 && c) atomicAdd(&counters[0], c);
}

template <int G, bool VEC16>
__global__ __launch_bounds__(kBlock) void k_cc_write(DevTables dt, OpView ov,
                                                     const uint32_t *tile_t0,
                                                     const uint8_t *commit, const uint8_t *found) {
    __shared__ uint16_t owner[kTileOpsMax];
    __shared__ uint8_t ok[kTileOps];
    __shared__ uint32_t list[kTileOpsMax];
    Tile t = fill_owner(ov, tile_t0[blockIdx.x], tile_t0[blockIdx.x + 1], owner);
    tile_write<G, VEC16>(dt, ov, t, owner, ok, true, commit, list, found);
}

// All three passes in one persistent launch.  Workgroup b owns tiles
// [b * tpb, (b + 1) * tpb) of TILE nominal ops; the host sizes the grid so
// that it is co-resident (the occupancy of this kernel times the CUs).
template <int G, bool VEC16, uint32_t TILE>
__global__ __launch_bounds__(kBlock) void k_cc_fused(DevTables dt, OpView ov, uint32_t n_tiles,
                                                     uint32_t tpb, uint32_t *epoch_ctr,
                                                     uint8_t *commit, uint8_t *results,
                                                     uint32_t res_stride, uint32_t *counters,
                                                     GridBar bar) {
    __shared__ uint16_t owner[TILE + LTPMG_MAX_TXN_OPS];
    __shared__ uint8_t ok[TILE];
    // list: the validate and write passes' op list; the reserve pass, which
    // runs before them, uses it as its LDS hash (12 bytes per slot)
    __shared__ __align__(16) uint32_t list[TILE + LTPMG_MAX_TXN_OPS];
    constexpr uint32_t kSlots = TILE / 4;  // 12 * TILE / 4 bytes <= sizeof list
    static_assert(12 * kSlots <= 4 * (TILE + LTPMG_MAX_TXN_OPS), "hash fits the list");
    __shared__ uint32_t s_t0[kFusedMaxTiles + 1];
    const uint32_t epoch = *epoch_ctr + 1;  // every workgroup reads it before the first barrier
    const uint32_t k0 = min(n_tiles, blockIdx.x * tpb), k1 = min(n_tiles, k0 + tpb);
    const uint32_t nk = k1 - k0;
    if (nk) {
        // s_t0[j] = first txn of tile k0 + j: a search for the first, then
        // one coalesced sweep over the run's txns for the rest
        const uint32_t first = block_lower_bound(ov.txn_off, ov.n_txns, (uint64_t)k0 * TILE);
        for (uint32_t j = threadIdx.x; j <= nk; j += blockDim.x) s_t0[j] = j ? ov.n_txns : first;
        __syncthreads();
        const uint64_t end = (uint64_t)k1 * TILE;
        for (uint32_t t = first + 1 + threadIdx.x;; t += blockDim.x) {
            bool more = false;
            if (t <= ov.n_txns) {
                const uint64_t lo = ov.txn_off[t - 1], hi = ov.txn_off[t];
                // tiles k with lo < k * TILE <= hi start at txn t
                for (uint64_t k = lo / TILE + 1; k * TILE <= hi; ++k)
                    if (k > k0 && k <= k1) s_t0[k - k0] = (uint32_t)t;
                more = hi < end;
            }
            if (!__syncthreads_or(more)) break;
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) counters[0] = 0;
    Tile t{0u, 0u, 0u, 0u};
    for (uint32_t j = 0; j < nk; ++j) {
        t = fill_owner(ov, s_t0[j], s_t0[j + 1], owner, TILE, TILE + LTPMG_MAX_TXN_OPS);
        tile_reserve(dt, ov, t, owner, epoch, reinterpret_cast<unsigned long long *>(list),
                     list + 2 * kSlots, kSlots);
    }
    if (!grid_sync(bar)) return;
    uint32_t committed = 0;
    for (uint32_t j = 0; j < nk; ++j) {
        if (nk > 1) t = fill_owner(ov, s_t0[j], s_t0[j + 1], owner, TILE, TILE + LTPMG_MAX_TXN_OPS);
        committed += tile_validate_read<G, VEC16>(dt, ov, t, owner, epoch, ok, list, commit,
                                                  results, res_stride, nullptr);
    }
    if (threadIdx.x == 0 && committed) atomicAdd(&counters[0], committed);
    if (!grid_sync(bar)) return;
    for (uint32_t j = 0; j < nk; ++j) {
        if (nk > 1) t = fill_owner(ov, s_t0[j], s_t0[j + 1], owner, TILE, TILE + LTPMG_MAX_TXN_OPS);
        tile_write<G, VEC16>(dt, ov, t, owner, ok, nk > 1, commit, list, nullptr);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *epoch_ctr = epoch;
}

void launch_tile_bounds(hipStream_t s, const uint32_t *txn_off, uint32_t n_txns,
                        uint32_t n_tiles, uint32_t *tile_t0, uint32_t *epoch_ctr) {
    uint32_t blocks = (n_tiles + 1 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_tile_bounds, dim3(blocks), dim3(kBlock), 0, s, txn_off, n_txns,
                       n_tiles, tile_t0, epoch_ctr);
}

void launch_cc_reserve(hipStream_t s, DevTables dt, OpView ov, const uint32_t *tile_t0,
                       uint32_t n_tiles, const uint32_t *epoch_ctr) {
    hipLaunchKernelGGL(k_cc_reserve, dim3(n_tiles), dim3(kBlock), 0, s, dt, ov, tile_t0,
                       epoch_ctr);
}

// Lanes per op: one 16-byte chunk per lane, up to 8 lanes (128-byte slots).
#define LTPMG_CC_DISPATCH(KERNEL, GRID, ...)                                               \
    do {                                                                                   \
        if (max_slot >= 128)                                                               \
            hipLaunchKernelGGL((KERNEL<8, true>), GRID, dim3(kBlock), 0, s, __VA_ARGS__);  \
        else if (max_slot >= 64)                                                           \
            hipLaunchKernelGGL((KERNEL<4, true>), GRID, dim3(kBlock), 0, s, __VA_ARGS__);  \
        else if (max_slot >= 32)                                                           \
            hipLaunchKernelGGL((KERNEL<2, true>), GRID, dim3(kBlock), 0, s, __VA_ARGS__);  \
        else if (max_slot >= 16)                                                           \
            hipLaunchKernelGGL((KERNEL<1, true>), GRID, dim3(kBlock), 0, s, __VA_ARGS__);  \
        else                                                                               \
            hipLaunchKernelGGL((KERNEL<1, false>), GRID, dim3(kBlock), 0, s, __VA_ARGS__); \
    } while (0)

void launch_cc_validate_read(hipStream_t s, DevTables dt, OpView ov, const uint32_t *tile_t0,
                             uint32_t n_tiles, const uint32_t *epoch_ctr, uint8_t *commit,
                             uint8_t *results, uint32_t res_stride, uint32_t *counters, uint8_t *found) {
    // the result stride is the largest slot, so it selects the vector width
    uint32_t max_slot = res_stride;
    LTPMG_CC_DISPATCH(k_cc_validate_read, dim3(n_tiles), dt, ov, tile_t0, epoch_ctr, commit,
                      results, res_stride, counters, found);
}

void launch_cc_write(hipStream_t s, DevTables dt, OpView ov, const uint32_t *tile_t0,
                     uint32_t n_tiles, const uint8_t *commit, uint32_t max_slot, const uint8_t *found) {
    LTPMG_CC_DISPATCH(k_cc_write, dim3(n_tiles), dt, ov, tile_t0, commit, found);
}

namespace {
template <int G, bool VEC16>
const void *fused_fn(uint32_t tile) {
    if (tile == 1024) return reinterpret_cast<const void *>(&k_cc_fused<G, VEC16, 1024>);
    if (tile == 2048) return reinterpret_cast<const void *>(&k_cc_fused<G, VEC16, 2048>);
    if (tile == 4096) return reinterpret_cast<const void *>(&k_cc_fused<G, VEC16, 4096>);
    return nullptr;
}
}  // namespace

const void *cc_fused_kernel(uint32_t tile_ops, uint32_t res_stride) {
    if (res_stride >= 128) return fused_fn<8, true>(tile_ops);
    if (res_stride >= 64) return fused_fn<4, true>(tile_ops);
    if (res_stride >= 32) return fused_fn<2, true>(tile_ops);
    if (res_stride >= 16) return fused_fn<1, true>(tile_ops);
    return fused_fn<1, false>(tile_ops);
}

uint32_t cc_fused_grid(uint32_t n_ops, uint32_t tile_ops, uint32_t grid) {
    const uint32_t n_tiles = (n_ops + tile_ops - 1) / tile_ops;
    if (n_tiles == 0 || grid == 0) return 0;
    const uint32_t tpb = (n_tiles + grid - 1) / grid;
    return (n_tiles + tpb - 1) / tpb;
}

hipError_t launch_cc_fused(hipStream_t s, DevTables dt, OpView ov, uint32_t tile_ops,
                           uint32_t grid, uint32_t *epoch_ctr, uint8_t *commit, uint8_t *results,
                           uint32_t res_stride, uint32_t *counters, GridBar bar) {
    const void *fn = cc_fused_kernel(tile_ops, res_stride);
    if (!fn || ov.n_txns == 0 || ov.n_ops == 0 || grid == 0) return hipErrorInvalidValue;
    uint32_t n_tiles = (ov.n_ops + tile_ops - 1) / tile_ops;
    uint32_t tpb = (n_tiles + grid - 1) / grid;
    if (tpb > kFusedMaxTiles) return hipErrorInvalidValue;
    grid = (n_tiles + tpb - 1) / tpb;
    void *args[] = {&dt, &ov, &n_tiles, &tpb, &epoch_ctr, &commit, &results, &res_stride,
                    &counters, &bar};
    return hipLaunchKernel(fn, dim3(grid), dim3(kBlock), args, 0, s);
}

}  // namespace ltpmg
