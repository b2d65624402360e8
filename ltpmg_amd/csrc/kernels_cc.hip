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
//                     (skipped when a plain load already shows a lower TID)
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
#include "internal.h"

namespace ltpmg {

namespace {

struct Tile {
    uint32_t t0, nt, op0, nops;
};

// Loads the tile's txn range and fills owner[o - op0] = txn index in tile.
__device__ inline Tile load_tile(const OpView &ov, const uint32_t *tile_t0, uint16_t *owner) {
    __shared__ uint32_t s_t[2], s_o[2];
    if (threadIdx.x == 0) {
        uint32_t t0 = tile_t0[blockIdx.x], t1 = tile_t0[blockIdx.x + 1];
        s_t[0] = t0;
        s_t[1] = t1;
        s_o[0] = ov.txn_off[t0];
        s_o[1] = ov.txn_off[t1];
    }
    __syncthreads();
    Tile t{s_t[0], s_t[1] - s_t[0], s_o[0], s_o[1] - s_o[0]};
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
    constexpr int R = 4;
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
                        uint32_t x = b < cd[r].slot ? *reinterpret_cast<const uint32_t *>(cell[r] + b) : 0u;
                        *reinterpret_cast<uint32_t *>(dst + b) = x;
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
                    uint32_t *c = reinterpret_cast<uint32_t *>(cell[r] + b);
                    if (b < cd[r].width) *c = upd[r] ? stream_word(arg[r], b >> 2) : *c + (uint32_t)arg[r];
                }
            }
        }
    }
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_cc_reserve(DevTables dt, OpView ov,
                                                       const uint32_t *tile_t0, uint32_t epoch) {
    __shared__ uint16_t owner[kTileOpsMax];
    Tile t = load_tile(ov, tile_t0, owner);
    const unsigned long long ep = (unsigned long long)epoch << 32;
    for (uint32_t i = threadIdx.x; i < t.nops; i += blockDim.x) {
        uint32_t o = t.op0 + i;
        uint16_t m = ov.meta[o];
        if (op_kind(m) == LTPMG_OP_READ) continue;
        uint32_t row = ov.rows[o];
        if (!op_valid(dt.tabs, m, row)) continue;  // never touch memory outside the tables
        const TabDesc &td = dt.tabs[op_tab(m)];
        unsigned long long tag = ep | (uint32_t)~(t.t0 + owner[i]);
        unsigned long long *p = td.lock + row;
        unsigned long long cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur < tag) atomicMax(p, tag);
    }
}

// G lanes cooperate on one op; each moves 16 bytes (VEC16) or a 4-byte word.
template <int G, bool VEC16>
__global__ __launch_bounds__(kBlock) void k_cc_validate_read(DevTables dt, OpView ov,
                                                             const uint32_t *tile_t0,
                                                             uint32_t epoch, uint8_t *commit,
                                                             uint8_t *results,
                                                             uint32_t res_stride,
                                                             uint32_t *counters) {
    __shared__ uint16_t owner[kTileOpsMax];
    __shared__ uint8_t ok[kTileOps];
    __shared__ uint32_t list[kTileOpsMax];
    __shared__ uint32_t n_list, n_commit;
    Tile t = load_tile(ov, tile_t0, owner);
    for (uint32_t j = threadIdx.x; j < t.nt; j += blockDim.x) ok[j] = 1;
    if (threadIdx.x == 0) n_list = n_commit = 0;
    __syncthreads();

    for (uint32_t i = threadIdx.x; i < t.nops; i += blockDim.x) {
        uint32_t o = t.op0 + i;
        uint16_t m = ov.meta[o];
        uint32_t row = ov.rows[o];
        if (!op_valid(dt.tabs, m, row)) {  // an invalid op aborts its txn
            ok[owner[i]] = 0;
            continue;
        }
        const TabDesc &td = dt.tabs[op_tab(m)];
        unsigned long long lk = td.lock[row];
        uint32_t tid = t.t0 + owner[i];
        if ((uint32_t)(lk >> 32) == epoch && (uint32_t)~(uint32_t)lk < tid) ok[owner[i]] = 0;
    }
    __syncthreads();

    uint32_t mine = 0;
    for (uint32_t j = threadIdx.x; j < t.nt; j += blockDim.x) {
        commit[t.t0 + j] = ok[j];
        mine += ok[j];
    }
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
    if (__lane_id() == 0 && mine) atomicAdd(&n_commit, mine);

    // committed READs of this tile -> LDS list (loop bound uniform per wave)
    for (uint32_t base = 0; base < t.nops; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool take = false;
        if (i < t.nops) take = op_kind(ov.meta[t.op0 + i]) == LTPMG_OP_READ && ok[owner[i]];
        wave_append(take, t.op0 + i, list, &n_list);
    }
    __syncthreads();
    if (threadIdx.x == 0 && n_commit) atomicAdd(&counters[0], n_commit);

    move_ops<G, VEC16, true>(n_list, list, ov, dt, results, res_stride);
}

template <int G, bool VEC16>
__global__ __launch_bounds__(kBlock) void k_cc_write(DevTables dt, OpView ov,
                                                     const uint32_t *tile_t0,
                                                     const uint8_t *commit) {
    __shared__ uint16_t owner[kTileOpsMax];
    __shared__ uint8_t ok[kTileOps];
    __shared__ uint32_t list[kTileOpsMax];
    __shared__ uint32_t n_list;
    Tile t = load_tile(ov, tile_t0, owner);
    for (uint32_t j = threadIdx.x; j < t.nt; j += blockDim.x) ok[j] = commit[t.t0 + j];
    if (threadIdx.x == 0) n_list = 0;
    __syncthreads();
    for (uint32_t base = 0; base < t.nops; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool take = false;
        if (i < t.nops && ok[owner[i]]) {
            uint32_t o = t.op0 + i;
            take = op_kind(ov.meta[o]) != LTPMG_OP_READ && op_valid(dt.tabs, ov.meta[o], ov.rows[o]);
        }
        wave_append(take, t.op0 + i, list, &n_list);
    }
    __syncthreads();

    move_ops<G, VEC16, false>(n_list, list, ov, dt, nullptr, 0u);
}

void launch_cc_reserve(hipStream_t s, DevTables dt, OpView ov, const uint32_t *tile_t0,
                       uint32_t n_tiles, uint32_t epoch) {
    hipLaunchKernelGGL(k_cc_reserve, dim3(n_tiles), dim3(kBlock), 0, s, dt, ov, tile_t0, epoch);
}

// Lanes per op: one 16-byte chunk per lane, up to 8 lanes (128-byte slots).
#define LTPMG_CC_DISPATCH(KERNEL, ...)                                                     \
    do {                                                                                   \
        if (max_slot >= 128)                                                               \
            hipLaunchKernelGGL((KERNEL<8, true>), dim3(n_tiles), dim3(kBlock), 0, s,       \
                               __VA_ARGS__);                                               \
        else if (max_slot >= 64)                                                           \
            hipLaunchKernelGGL((KERNEL<4, true>), dim3(n_tiles), dim3(kBlock), 0, s,       \
                               __VA_ARGS__);                                               \
        else if (max_slot >= 32)                                                           \
            hipLaunchKernelGGL((KERNEL<2, true>), dim3(n_tiles), dim3(kBlock), 0, s,       \
                               __VA_ARGS__);                                               \
        else if (max_slot >= 16)                                                           \
            hipLaunchKernelGGL((KERNEL<1, true>), dim3(n_tiles), dim3(kBlock), 0, s,       \
                               __VA_ARGS__);                                               \
        else                                                                               \
            hipLaunchKernelGGL((KERNEL<1, false>), dim3(n_tiles), dim3(kBlock), 0, s,      \
                               __VA_ARGS__);                                               \
    } while (0)

void launch_cc_validate_read(hipStream_t s, DevTables dt, OpView ov, const uint32_t *tile_t0,
                             uint32_t n_tiles, uint32_t epoch, uint8_t *commit,
                             uint8_t *results, uint32_t res_stride, uint32_t *counters) {
    // the result stride is the largest slot, so it selects the vector width
    uint32_t max_slot = res_stride;
    LTPMG_CC_DISPATCH(k_cc_validate_read, dt, ov, tile_t0, epoch, commit, results, res_stride,
                      counters);
}

void launch_cc_write(hipStream_t s, DevTables dt, OpView ov, const uint32_t *tile_t0,
                     uint32_t n_tiles, const uint8_t *commit, uint32_t max_slot) {
    LTPMG_CC_DISPATCH(k_cc_write, dt, ov, tile_t0, commit);
}

}  // namespace ltpmg
