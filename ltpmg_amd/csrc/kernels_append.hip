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
//            holes.
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

// The txn that owns op o: the largest t < n_txns with txn_off[t] <= o.
__device__ inline uint32_t txn_of(const OpView &ov, uint32_t o) {
    uint32_t lo = 0, hi = ov.n_txns;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ov.txn_off[mid] <= o) lo = mid;
        else hi = mid;
    }
    return lo;
}

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

// Ordered mode: found[o] and meta_eff[o] for every op; a READ of an absent
// row gets a zeroed result slot (the apply kernel skips invalid ops).
__global__ __launch_bounds__(kBlock) void k_row_presence(DevTables dt, OpView ov, const uint32_t *epoch_ctr,
                                                          uint8_t *found, uint16_t *meta_eff, uint8_t *results,
                                                          uint32_t res_stride) {
    const uint32_t epoch = *epoch_ctr;
    for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < ov.n_ops; o += gridDim.x * blockDim.x) {
        const uint16_t m = ov.meta[o];
        const uint32_t row = ov.rows[o];
        const TabDesc &td = dt.tabs[op_tab(m)];
        bool f = op_ok(td, m, row);
        if (f && td.present && !is_insert(m)) {
            const unsigned long long lk = td.lock[row];
            const bool deleted_before = (uint32_t)(lk >> 32) == epoch && (uint32_t)~(uint32_t)lk < o;
            f = td.present[row] && !deleted_before;
        }
        found[o] = f ? 1 : 0;
        meta_eff[o] = f ? m : (uint16_t)(m | 0x8000u);  // a reserved bit: an identity in the sort
        if (!f && op_kind(m) == LTPMG_OP_READ)
            for (uint32_t b = 0; b < res_stride; b += 4) st4(results + (uint64_t)o * res_stride + b, 0u);
    }
}

// Clears the presence of every row whose DELETE found it (in a committed txn).
__global__ __launch_bounds__(kBlock) void k_row_deletes(DevTables dt, OpView ov, const uint8_t *found,
                                                         const uint8_t *commit) {
    for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < ov.n_ops; o += gridDim.x * blockDim.x) {
        const uint16_t m = ov.meta[o];
        if (!is_delete(m) || !found[o]) continue;
        const TabDesc &td = dt.tabs[op_tab(m)];
        const uint32_t row = ov.rows[o];
        if (!op_ok(td, m, row) || !commit[txn_of(ov, o)]) continue;
        td.present[row] = 0;
    }
}

// The table of op o if it is a valid INSERT of a committed txn, else 0xFF.
__device__ inline uint32_t committed_insert(const DevTables &dt, const OpView &ov, const uint8_t *commit,
                                            uint32_t o) {
    if (o >= ov.n_ops) return 0xFFu;
    const uint16_t m = ov.meta[o];
    if (!is_insert(m) || !op_ok(dt.tabs[op_tab(m)], m, ov.rows[o]) || !commit[txn_of(ov, o)]) return 0xFFu;
    return op_tab(m);
}

// Committed INSERTs per table in each kInsTile-op tile.
__global__ __launch_bounds__(kBlock) void k_ins_count(DevTables dt, OpView ov, const uint8_t *commit,
                                                       uint32_t *cnt) {
    __shared__ uint32_t c[LTPMG_MAX_TABLES];
    if (threadIdx.x < LTPMG_MAX_TABLES) c[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t o0 = blockIdx.x * kInsTile;
    for (uint32_t i = threadIdx.x; i < kInsTile; i += blockDim.x) {
        const uint32_t tb = committed_insert(dt, ov, commit, o0 + i);
        if (tb != 0xFFu) atomicAdd(&c[tb], 1u);
    }
    __syncthreads();
    if (threadIdx.x < LTPMG_MAX_TABLES) cnt[(uint64_t)blockIdx.x * LTPMG_MAX_TABLES + threadIdx.x] = c[threadIdx.x];
}

// One block of LTPMG_MAX_TABLES waves, wave w for table w: the counts become
// each tile's exclusive offset, base[w] gets the cursor at the batch start,
// and the cursor moves past the rows handed out (the host has checked that
// they fit the table).
__global__ __launch_bounds__(64 * LTPMG_MAX_TABLES) void k_ins_scan(DevTables dt, uint32_t *cnt, uint32_t n_tiles,
                                                                    uint32_t *cursor, uint32_t *base,
                                                                    uint32_t *inserted) {
    __shared__ uint32_t total[LTPMG_MAX_TABLES];
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    uint32_t run = 0;
    for (uint32_t k0 = 0; k0 < n_tiles; k0 += 64) {
        const uint32_t k = k0 + lane;
        const uint32_t v = k < n_tiles ? cnt[(uint64_t)k * LTPMG_MAX_TABLES + w] : 0u;
        uint32_t inc = v;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(inc, off);
            if (lane >= off) inc += y;
        }
        if (k < n_tiles) cnt[(uint64_t)k * LTPMG_MAX_TABLES + w] = run + inc - v;
        run += __shfl(inc, 63);
    }
    if (lane == 0) {
        const uint32_t b = cursor[w], cap = dt.tabs[w].present ? dt.tabs[w].n_rows : b;
        base[w] = b;
        cursor[w] = (uint64_t)b + run < cap ? b + run : cap;
        total[w] = run;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t i = 0; i < LTPMG_MAX_TABLES; ++i) t += total[i];
        *inserted = t;
    }
}

// Exclusive sum over the block's threads.
__device__ inline uint32_t block_excl(uint32_t x, uint32_t *ws) {
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t inc = x;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
    }
    __syncthreads();  // earlier readers of ws are done
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t i = 0; i < w; ++i) pre += ws[i];
    return pre + inc - x;
}

constexpr uint32_t kInsPerThread = kInsTile / kBlock;

// Every committed INSERT writes its row: row = base + tile offset + rank in
// the tile (op order; a thread takes kInsPerThread consecutive ops, so
// thread order then item order is op order).  All columns get the payload of
// rowseed(arg, column), the presence byte is set, the result slot gets the
// row id.
__global__ __launch_bounds__(kBlock) void k_ins_write(DevTables dt, OpView ov, const uint8_t *commit,
                                                       uint32_t table_mask, const uint32_t *cnt, const uint32_t *base,
                                                       uint8_t *results, uint32_t res_stride, uint8_t *found) {
    __shared__ uint32_t ws[kBlock / 64];
    const uint32_t o0 = blockIdx.x * kInsTile + threadIdx.x * kInsPerThread;
    uint32_t tb[kInsPerThread];
#pragma unroll
    for (uint32_t j = 0; j < kInsPerThread; ++j) tb[j] = committed_insert(dt, ov, commit, o0 + j);
    for (uint32_t mask = table_mask; mask; mask &= mask - 1) {
        const uint32_t t = __builtin_ctz(mask);  // uniform
        uint32_t mine = 0;
#pragma unroll
        for (uint32_t j = 0; j < kInsPerThread; ++j) mine += tb[j] == t;
        uint32_t rank = block_excl(mine, ws);
        if (!mine) continue;
        const TabDesc td = dt.tabs[t];
        const uint32_t first = base[t] + cnt[(uint64_t)blockIdx.x * LTPMG_MAX_TABLES + t];
        for (uint32_t j = 0; j < kInsPerThread; ++j) {
            if (tb[j] != t) continue;
            const uint32_t o = o0 + j, row = first + rank++;
            uint8_t *res = results + (uint64_t)o * res_stride;
            if (row >= td.n_rows) {  // never with the host's capacity check
                found[o] = 0;
                continue;
            }
            const uint64_t arg = ov.args[o];
            for (uint32_t c = 0; c < td.n_cols; ++c) {
                const ColDesc cd = dt.cols[t * LTPMG_MAX_COLS + c];
                uint8_t *cell = cd.base + (uint64_t)row * cd.slot;
                const uint64_t seed = rowseed(arg, c);
                if (cd.slot >= 16) {
                    for (uint32_t b = 0; b < cd.slot; b += 16) st16(cell + b, payload16(seed, b, cd.width));
                } else {
                    for (uint32_t b = 0; b < cd.slot; b += 4) st4(cell + b, b < cd.width ? stream_word(seed, b >> 2) : 0u);
                }
            }
            td.present[row] = 1;
            found[o] = 1;
            st4(res, row);
            for (uint32_t b = 4; b < res_stride; b += 4) st4(res + b, 0u);
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
    hipLaunchKernelGGL(k_row_presence, dim3(grid_for(ov.n_ops)), dim3(kBlock), 0, s, dt, ov, epoch_ctr, rb.found,
                       rb.meta_eff, results, res_stride);
}

void launch_row_deletes(hipStream_t s, const DevTables &dt, OpView ov, RowBufs &rb, const uint8_t *commit) {
    if (ov.n_ops == 0 || !rb.has_delete) return;
    hipLaunchKernelGGL(k_row_deletes, dim3(grid_for(ov.n_ops)), dim3(kBlock), 0, s, dt, ov, rb.found, commit);
}

void launch_inserts(hipStream_t s, const DevTables &dt, OpView ov, RowBufs &rb, const uint8_t *commit,
                    uint8_t *results, uint32_t res_stride) {
    if (ov.n_ops == 0 || !rb.has_insert) return;
    const uint32_t tiles = ins_tiles(ov.n_ops);
    hipLaunchKernelGGL(k_ins_count, dim3(tiles), dim3(kBlock), 0, s, dt, ov, commit, rb.ins_cnt);
    hipLaunchKernelGGL(k_ins_scan, dim3(1), dim3(64 * LTPMG_MAX_TABLES), 0, s, dt, rb.ins_cnt, tiles, rb.cursor,
                       rb.ins_base, rb.inserted);
    hipLaunchKernelGGL(k_ins_write, dim3(tiles), dim3(kBlock), 0, s, dt, ov, commit, rb.insert_tables, rb.ins_cnt,
                       rb.ins_base, results, res_stride, rb.found);
}

}  // namespace ltpmg
