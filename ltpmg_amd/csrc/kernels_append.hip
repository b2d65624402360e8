// kernels_append.hip -- row-level ops on append tables: INSERT and DELETE.
//
// The reference's insert/delete executors write at host-chosen rows through
// operators whose stores are commented out (cuda/src/gpulauncher.cu:205-356,
// 52-128); its NewOrder "inserts" target random row ids
// (src/txn_generator.cpp:253-258) and its Delivery deletes them
// (cuda/src/gpuquery.cu:2794-2807).  Here an append table has a cursor and
// one presence byte per row:
//
//   INSERT   rows are handed out at the end of the batch, in op (= TID)
//            order over the committed INSERTs of each table: count per tile,
//            scan over tiles (which also moves the cursor), then each INSERT
//            writes its row.  An aborted txn takes no row, so the table has no
//            holes.  Which INSERTs commit is in found[] (the lowest-TID-wins
//            validate kernel folds the commit flags into it; in the ordered
//            mode every txn commits), so no kernel here maps ops to txns.
//   DELETE   clears the row's presence byte, after the batch's cell writes.
//
// An op sees a row as present if its presence byte is set at the batch start
// and no DELETE of it comes earlier in the batch.  In the lowest-TID-wins
// mode a committed DELETE excludes every higher TID from its row, so the
// snapshot is what committed ops see (kernels_cc.hip reads it).  In the
// ordered mode every txn commits, so the lowest op that deletes each row is
// found first (atomicMax of {epoch, ~op} on the row's lock word, which that
// mode does not otherwise use), and ops after it see the row as deleted:
// they are marked invalid in a private copy of the meta words, which makes
// them identities in the sort and scan.
#include "internal.h"

namespace ltpmg {

namespace {

__device__ inline bool op_ok(const TabDesc &td, uint16_t m, uint32_t row) {
    return op_check(m, row, td.live ? td.n_rows : 0u, td.n_cols, td.present != nullptr);
}

__global__ void k_next_epoch(uint32_t *epoch_ctr) { *epoch_ctr += 1; }

// Ordered mode: the lowest op deleting each row, in the row's lock word.
__global__ __launch_bounds__(kBlock) void k_row_first_delete(DevTables dt, OpView ov, const uint32_t *epoch_ctr) {
    const unsigned long long ep = (unsigned long long)*epoch_ctr << 32;
    for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < ov.n_ops; o += gridDim.x * blockDim.x) {
        const uint16_t m = ov.meta[o];
        if (!is_delete(m)) continue;
        const TabDesc &td = dt.tabs[op_tab(m)];
        const uint32_t row = ov.rows[o];
        if (!op_ok(td, m, row)) continue;
        __hip_atomic_fetch_max(td.lock + row, ep | (uint32_t)~o, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The tables' fields these kernels need, in LDS (a TabDesc gathered from
// global memory per op, by lanes of different tables, made the row-op
// kernels latency-bound).
struct RowTab {
    const uint8_t *present;
    const unsigned long long *lock;
    uint32_t n_rows;  // 0 for a table that is not live
    uint32_t n_cols;
};
__device__ inline void load_row_tabs(const DevTables &dt, RowTab *rt) {
    if (threadIdx.x < LTPMG_MAX_TABLES) {
        const TabDesc &td = dt.tabs[threadIdx.x];
        rt[threadIdx.x] = RowTab{td.present, td.lock, td.live ? td.n_rows : 0u, td.n_cols};
    }
}
constexpr uint32_t kPresPerThread = 8;  // ops per thread per round, their loads issued together

// Ordered mode: found[o] and meta_eff[o] for every op; a READ of an absent
// row gets a zeroed result slot (the apply kernel skips invalid ops).  Each
// thread takes kPresPerThread ops a round (consecutive threads on
// consecutive ops), issues every meta and row load, then every lock-word and
// presence load, before it uses any.
// Without a DELETE in the batch (has_delete false) no lock word can carry
// this batch's epoch, so none is read.
__global__ __launch_bounds__(kBlock) void k_row_presence(DevTables dt, OpView ov, const uint32_t *epoch_ctr,
                                                          uint8_t *found, uint16_t *meta_eff, uint8_t *results,
                                                          uint32_t res_stride, bool has_delete) {
    __shared__ RowTab rt[LTPMG_MAX_TABLES];
    load_row_tabs(dt, rt);
    __syncthreads();
    const uint32_t epoch = *epoch_ctr;
    constexpr uint32_t P = kPresPerThread;
    for (uint32_t o0 = blockIdx.x * blockDim.x * P; o0 < ov.n_ops; o0 += gridDim.x * blockDim.x * P) {
        uint32_t m[P], row[P], pres[P];
        unsigned long long lk[P];
#pragma unroll
        for (uint32_t j = 0; j < P; ++j) {
            const uint32_t o = o0 + j * blockDim.x + threadIdx.x;
            m[j] = o < ov.n_ops ? ov.meta[o] : 0xF000u;  // reserved bits: not a valid op
            row[j] = o < ov.n_ops ? ov.rows[o] : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < P; ++j) {
            const uint16_t mj = (uint16_t)m[j];
            const RowTab &t = rt[op_tab(mj)];
            const bool ok = op_check(mj, row[j], t.n_rows, t.n_cols, t.present != nullptr);
            const bool look = ok && t.present && !is_insert(mj);
            pres[j] = look ? t.present[row[j]] : 1u;
            lk[j] = look && has_delete ? t.lock[row[j]] : 0ull;
            m[j] |= (ok ? 0u : 0x10000u) | (look ? 0x20000u : 0u);  // bit 16: op_check failed; 17: looked up
        }
#pragma unroll
        for (uint32_t j = 0; j < P; ++j) {
            const uint32_t o = o0 + j * blockDim.x + threadIdx.x;
            if (o >= ov.n_ops) continue;
            const uint16_t mj = (uint16_t)m[j];
            const bool deleted_before =
                (m[j] & 0x20000u) && (uint32_t)(lk[j] >> 32) == epoch && (uint32_t)~(uint32_t)lk[j] < o;
            const bool f = !(m[j] & 0x10000u) && pres[j] && !deleted_before;
            // 2 + table for an INSERT that takes effect (taken_inserts)
            found[o] = f ? (is_insert(mj) ? (uint8_t)(2u + op_tab(mj)) : (uint8_t)1) : (uint8_t)0;
            meta_eff[o] = f ? mj : (uint16_t)(mj | 0x8000u);  // a reserved bit: an identity in the sort
            if (!f && op_kind(mj) == LTPMG_OP_READ)
                for (uint32_t b = 0; b < res_stride; b += 4) st4(results + (uint64_t)o * res_stride + b, 0u);
        }
    }
}

// Clears the presence of every row whose DELETE took effect (found: the row
// was present and the txn commits).
__global__ __launch_bounds__(kBlock) void k_row_deletes(DevTables dt, OpView ov, const uint8_t *found) {
    for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < ov.n_ops; o += gridDim.x * blockDim.x) {
        const uint16_t m = ov.meta[o];
        if (!is_delete(m) || !found[o]) continue;
        const TabDesc &td = dt.tabs[op_tab(m)];
        const uint32_t row = ov.rows[o];
        if (!op_ok(td, m, row)) continue;
        td.present[row] = 0;
    }
}

constexpr uint32_t kInsPerThread = kInsTile / kBlock;
static_assert(kInsPerThread == 8, "a thread's found bytes are one 8-byte load");
// The tables of ops o0 .. o0 + 7 that are INSERTs taking effect, else 0xFF.
// found[] holds 2 + table for exactly those (k_row_presence, and
// tile_validate_read in the lowest-TID-wins mode: a valid INSERT of a
// committed txn), so one 8-byte load of found answers for all eight (the
// meta, row and found loads per op this replaced had made the insert
// kernels latency-bound).
__device__ inline void taken_inserts(uint32_t n_ops, const uint8_t *found, uint32_t o0, uint32_t *tb) {
    uint32_t f[kInsPerThread];
    if (o0 + kInsPerThread <= n_ops && (reinterpret_cast<uintptr_t>(found + o0) & 7u) == 0) {
        const uint2 w = *reinterpret_cast<const uint2 *>(found + o0);
#pragma unroll
        for (uint32_t j = 0; j < kInsPerThread; ++j) f[j] = ((j < 4 ? w.x : w.y) >> (8 * (j & 3u))) & 0xFFu;
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kInsPerThread; ++j) f[j] = o0 + j < n_ops ? found[o0 + j] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < kInsPerThread; ++j) tb[j] = f[j] >= 2u ? f[j] - 2u : 0xFFu;
}

// INSERTs per table in each kInsTile-op tile: cnt[table * n_tiles + tile].
__global__ __launch_bounds__(kBlock) void k_ins_count(OpView ov, const uint8_t *found, uint32_t *cnt,
                                                       uint32_t n_tiles) {
    __shared__ uint32_t c[LTPMG_MAX_TABLES];
    if (threadIdx.x < LTPMG_MAX_TABLES) c[threadIdx.x] = 0;
    __syncthreads();
    uint32_t tb[kInsPerThread];
    taken_inserts(ov.n_ops, found, blockIdx.x * kInsTile + threadIdx.x * kInsPerThread, tb);
#pragma unroll
    for (uint32_t j = 0; j < kInsPerThread; ++j)
        if (tb[j] != 0xFFu) atomicAdd(&c[tb[j]], 1u);
    __syncthreads();
    if (threadIdx.x < LTPMG_MAX_TABLES) cnt[(uint64_t)threadIdx.x * n_tiles + blockIdx.x] = c[threadIdx.x];
}

constexpr uint32_t kScanThreads = 1024;

// One workgroup per table (only the tables in table_mask do work): each
// thread sums a contiguous chunk of the table's tile counts, one block scan
// gives every chunk its start, and the thread rewrites its chunk's counts as
// exclusive offsets.  base[t] gets the cursor at the batch start, and the
// cursor moves past the rows handed out (the host has checked that they fit
// the table).
__global__ __launch_bounds__(kScanThreads) void k_ins_scan(DevTables dt, uint32_t *cnt, uint32_t n_tiles,
                                                           uint32_t table_mask, uint32_t *cursor, uint32_t *base,
                                                           uint32_t *inserted) {
    __shared__ uint32_t ws[kScanThreads / 64];
    const uint32_t t = blockIdx.x;
    if (!((table_mask >> t) & 1u)) return;  // uniform per workgroup
    uint32_t *c = cnt + (uint64_t)t * n_tiles;
    const uint32_t chunk = (n_tiles + kScanThreads - 1) / kScanThreads;
    const uint32_t k0 = threadIdx.x * chunk, k1 = k0 + chunk < n_tiles ? k0 + chunk : n_tiles;
    uint32_t sum = 0;
    for (uint32_t k = k0; k < k1; ++k) sum += c[k];
    // block exclusive sum of the chunk sums
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t inc = sum;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t pre = 0, total = 0;
    for (uint32_t i = 0; i < kScanThreads / 64; ++i) {
        if (i < w) pre += ws[i];
        total += ws[i];
    }
    uint32_t run = pre + inc - sum;
    for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t v = c[k];
        c[k] = run;
        run += v;
    }
    if (threadIdx.x == 0) {
        const uint32_t b = cursor[t], cap = dt.tabs[t].present ? dt.tabs[t].n_rows : b;
        base[t] = b;
        cursor[t] = (uint64_t)b + total < cap ? b + total : cap;
        atomicAdd(inserted, total);
    }
}

constexpr uint32_t kInsGroup = 4;  // insert tables per pass of k_ins_write

// Exclusive sum over the block's threads of kInsGroup 16-bit counts packed
// in 64 bits (no field carries: a tile has kInsTile < 2^16 ops); *total
// gets the packed totals.
__device__ inline uint64_t block_excl_packed(uint64_t x, unsigned long long *ws, uint64_t *total) {
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint64_t inc = x;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up((unsigned long long)inc, off);
        if (lane >= off) inc += y;
    }
    __syncthreads();  // earlier readers of ws are done
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint64_t pre = 0, tot = 0;
    for (uint32_t i = 0; i < kBlock / 64; ++i) {
        if (i < w) pre += ws[i];
        tot += ws[i];
    }
    *total = tot;
    return pre + inc - x;
}

// Every INSERT that takes effect writes its row: row = base + tile offset +
// rank in the tile (op order; a thread takes kInsPerThread consecutive ops,
// so thread order then item order is op order).  The tile's insert tables
// are taken kInsGroup at a time: one block scan of packed counts ranks the
// INSERTs of all of them, which are listed in LDS table by table in rank
// order, with each table's fields and column descriptors (loaded while the
// scan runs).  Then the whole workgroup walks the group's INSERTs (presence
// byte; the result slot's first 16 bytes get the row id and zeros) and its
// cells, one (row, column) per lane, column by column within a table, so
// each column's stores are contiguous.  Every column gets the payload of
// rowseed(arg, column).  (A loop per table, with its own scan and loads,
// and per column had made this kernel latency-bound: most tiles hold a few
// rows of each table.)
__global__ __launch_bounds__(kBlock) void k_ins_write(DevTables dt, OpView ov, uint32_t table_mask,
                                                       const uint32_t *cnt, uint32_t n_tiles, const uint32_t *base,
                                                       uint8_t *results, uint32_t res_stride, uint8_t *found) {
    struct GroupTab {
        uint8_t *present;
        uint32_t first;   // the tile's first row in the table
        uint32_t n_rows;
        uint32_t n_cols;
    };
    __shared__ unsigned long long ws[kBlock / 64];
    __shared__ uint32_t list[kInsTile];            // op index of each INSERT, table by table in rank order
    __shared__ unsigned long long seeds[kInsTile];  // its arg (the row's payload seed)
    __shared__ ColDesc tcols[kInsGroup][LTPMG_MAX_COLS];
    __shared__ GroupTab gt[kInsGroup];
    __shared__ uint32_t busy;
    // A tile none of whose INSERTs takes effect writes nothing.  The counts
    // are offsets by now, so a tile's count per table is the next tile's
    // offset less its own; the last tile always takes the walk.
    if (threadIdx.x == 0) busy = blockIdx.x + 1 == n_tiles ? 1u : 0u;
    __syncthreads();
    if (threadIdx.x < LTPMG_MAX_TABLES && ((table_mask >> threadIdx.x) & 1u) && blockIdx.x + 1 < n_tiles) {
        const uint32_t *c = cnt + (uint64_t)threadIdx.x * n_tiles + blockIdx.x;
        if (c[1] != c[0]) busy = 1u;
    }
    __syncthreads();
    if (!busy) return;  // uniform
    const uint32_t o0 = blockIdx.x * kInsTile + threadIdx.x * kInsPerThread;
    uint32_t o[kInsPerThread], tb[kInsPerThread];
#pragma unroll
    for (uint32_t j = 0; j < kInsPerThread; ++j) o[j] = o0 + j;
    taken_inserts(ov.n_ops, found, o0, tb);
    for (uint32_t tm = table_mask; tm;) {  // uniform
        uint32_t gtab[kInsGroup], ng = 0;
        for (; tm && ng < kInsGroup; tm &= tm - 1) gtab[ng++] = __builtin_ctz(tm);
        __syncthreads();  // the previous group's readers of gt, tcols, list and seeds are done
        for (uint32_t x = threadIdx.x; x < ng * LTPMG_MAX_COLS; x += kBlock) {
            const uint32_t g = x / LTPMG_MAX_COLS, c = x % LTPMG_MAX_COLS;
            tcols[g][c] = dt.cols[gtab[g] * LTPMG_MAX_COLS + c];
        }
        if (threadIdx.x < ng) {
            const uint32_t t = gtab[threadIdx.x];
            const TabDesc &td = dt.tabs[t];
            gt[threadIdx.x] = GroupTab{td.present, base[t] + cnt[(uint64_t)t * n_tiles + blockIdx.x], td.n_rows,
                                       td.n_cols};
        }
        uint32_t grp[kInsPerThread];  // the INSERT's place in the group, or kInsGroup
        uint64_t mine = 0;
#pragma unroll
        for (uint32_t j = 0; j < kInsPerThread; ++j) {
            grp[j] = kInsGroup;
            for (uint32_t g = 0; g < ng; ++g)
                if (tb[j] == gtab[g]) grp[j] = g;
            if (grp[j] < kInsGroup) mine += 1ull << (16 * grp[j]);
        }
        uint64_t tot;
        const uint64_t ex = block_excl_packed(mine, ws, &tot);  // its barriers publish gt and tcols
        if (tot == 0) continue;  // uniform
        uint32_t at[kInsGroup], n[kInsGroup], fit[kInsGroup], item_at[kInsGroup + 1];
        uint32_t run = 0, items = 0;
#pragma unroll
        for (uint32_t g = 0; g < kInsGroup; ++g) {
            n[g] = (uint32_t)(tot >> (16 * g)) & 0xFFFFu;
            at[g] = run;
            run += n[g];
            const uint32_t first = g < ng ? gt[g].first : 0u, rows = g < ng ? gt[g].n_rows : 0u;
            // rows past the capacity never occur with the host's capacity check
            fit[g] = first >= rows ? 0u : (rows - first < n[g] ? rows - first : n[g]);
            item_at[g] = items;
            items += g < ng ? fit[g] * gt[g].n_cols : 0u;
        }
        item_at[kInsGroup] = items;
        {
            uint64_t rank = ex;
#pragma unroll
            for (uint32_t j = 0; j < kInsPerThread; ++j) {
                if (grp[j] == kInsGroup) continue;
                const uint32_t g = grp[j];
                const uint32_t pos = at[g] + ((uint32_t)(rank >> (16 * g)) & 0xFFFFu);
                rank += 1ull << (16 * g);
                seeds[pos] = ov.args[o[j]];
                list[pos] = o[j];
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < run; i += kBlock) {
            uint32_t g = 0;
            while (g + 1 < kInsGroup && i >= at[g + 1]) ++g;
            const uint32_t k = i - at[g], op = list[i];
            if (k >= fit[g]) {
                found[op] = 0;
                continue;
            }
            const uint32_t row = gt[g].first + k;
            gt[g].present[row] = 1;
            // the row id and zeros to 16 bytes (the host zeroes the rest of
            // the result slot when it fetches the results)
            uint8_t *res = results + (uint64_t)op * res_stride;
            if ((res_stride & 15u) == 0) {
                st16(res, u32x4{row, 0u, 0u, 0u});
            } else {
                st4(res, row);
                for (uint32_t b = 4; b < res_stride && b < 16; b += 4) st4(res + b, 0u);
            }
        }
        for (uint32_t i = threadIdx.x; i < items; i += kBlock) {
            uint32_t g = 0;
            while (g + 1 < kInsGroup && i >= item_at[g + 1]) ++g;
            const uint32_t local = i - item_at[g], c = local / fit[g], k = local - c * fit[g];
            const ColDesc cd = tcols[g][c];
            uint8_t *cell = cd.base + (uint64_t)(gt[g].first + k) * cd.slot;
            const uint64_t seed = rowseed(seeds[at[g] + k], c);
            if (cd.slot >= 16) {
                for (uint32_t b = 0; b < cd.slot; b += 16) st16(cell + b, payload16(seed, b, cd.width));
            } else {
                for (uint32_t b = 0; b < cd.slot; b += 4) st4(cell + b, b < cd.width ? stream_word(seed, b >> 2) : 0u);
            }
        }
    }
}

uint32_t grid_for(uint32_t n) {
    uint32_t blocks = (n + kBlock - 1) / kBlock;
    return blocks > 4096 ? 4096 : (blocks ? blocks : 1);
}

}  // namespace

void launch_row_presence_o(hipStream_t s, const DevTables &dt, OpView ov, RowBufs &rb, uint32_t *epoch_ctr,
                           uint8_t *results, uint32_t res_stride) {
    if (ov.n_ops == 0) return;
    hipLaunchKernelGGL(k_next_epoch, dim3(1), dim3(1), 0, s, epoch_ctr);
    if (rb.has_delete)
        hipLaunchKernelGGL(k_row_first_delete, dim3(grid_for(ov.n_ops)), dim3(kBlock), 0, s, dt, ov, epoch_ctr);
    hipLaunchKernelGGL(k_row_presence, dim3(grid_for((ov.n_ops + kPresPerThread - 1) / kPresPerThread)),
                       dim3(kBlock), 0, s, dt, ov, epoch_ctr, rb.found, rb.meta_eff, results, res_stride,
                       rb.has_delete);
}

void launch_row_deletes(hipStream_t s, const DevTables &dt, OpView ov, RowBufs &rb) {
    if (ov.n_ops == 0 || !rb.has_delete) return;
    hipLaunchKernelGGL(k_row_deletes, dim3(grid_for(ov.n_ops)), dim3(kBlock), 0, s, dt, ov, rb.found);
}

void launch_inserts(hipStream_t s, const DevTables &dt, OpView ov, RowBufs &rb, uint8_t *results,
                    uint32_t res_stride) {
    if (ov.n_ops == 0 || !rb.has_insert) return;
    const uint32_t tiles = ins_tiles(ov.n_ops);
    hipLaunchKernelGGL(k_ins_count, dim3(tiles), dim3(kBlock), 0, s, ov, rb.found, rb.ins_cnt, tiles);
    launch_zero_u32(s, rb.inserted, 1);
    hipLaunchKernelGGL(k_ins_scan, dim3(LTPMG_MAX_TABLES), dim3(kScanThreads), 0, s, dt, rb.ins_cnt, tiles,
                       rb.insert_tables, rb.cursor, rb.ins_base, rb.inserted);
    hipLaunchKernelGGL(k_ins_write, dim3(tiles), dim3(kBlock), 0, s, dt, ov, rb.insert_tables, rb.ins_cnt, tiles,
                       rb.ins_base, results, res_stride, rb.found);
}

}  // namespace ltpmg
