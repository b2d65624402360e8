// kernels_table.hip -- seeded table fill, table digest and small utilities.
//
// The reference fills tables on the host with rand() and copies them over
// synchronously (src/database.cpp:176-395, cuda/src/gpudatabase.cu:183-198).
// Here a table of any size is filled in place in HBM from a seed, so a 10 GB
// table costs one write pass instead of a host generation plus a PCIe copy.
#include "internal.h"

namespace ltpmg {

// One thread per 16-byte chunk of a slot; chunks past the width are zeroed.
__global__ __launch_bounds__(kBlock) void k_fill_column(ColDesc cd, uint32_t table,
                                                        uint32_t col, uint32_t n_rows,
                                                        uint64_t seed) {
    const uint32_t chunks = cd.slot >= 16 ? cd.slot / 16 : 1;
    const uint64_t total = (uint64_t)n_rows * chunks;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t row = (uint32_t)(i / chunks), c = (uint32_t)(i % chunks);
        uint64_t cs = cellseed(seed, table, row, col);
        uint8_t *cell = cd.base + (uint64_t)row * cd.slot;
        if (cd.slot >= 16) {
            st16(cell + c * 16, payload16(cs, c * 16, cd.width));
        } else {
            for (uint32_t w = 0; w < cd.slot / 4; ++w)
                reinterpret_cast<uint32_t *>(cell)[w] = (w * 4 < cd.width) ? stream_word(cs, w) : 0u;
        }
    }
}

void launch_fill_column(hipStream_t s, ColDesc cd, uint32_t table, uint32_t col,
                        uint32_t n_rows, uint64_t seed) {
    uint64_t chunks = (uint64_t)n_rows * (cd.slot >= 16 ? cd.slot / 16 : 1);
    uint64_t blocks = (chunks + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_fill_column, dim3((uint32_t)blocks), dim3(kBlock), 0, s, cd, table,
                       col, n_rows, seed);
}

// Digest of one cell: h0 = mix64(golden * (1 + cell_id)), then for each
// 8-byte little-endian word of the width (last one zero padded)
// h = mix64(h ^ word).  Table digest = sum of cell digests mod 2^64.
__global__ __launch_bounds__(kBlock) void k_digest_column(ColDesc cd, uint32_t table,
                                                          uint32_t col, uint32_t n_rows,
                                                          unsigned long long *acc) {
    __shared__ unsigned long long part[kBlock / 64];
    unsigned long long sum = 0;
    for (uint32_t row = blockIdx.x * blockDim.x + threadIdx.x; row < n_rows;
         row += gridDim.x * blockDim.x) {
        const uint8_t *cell = cd.base + (uint64_t)row * cd.slot;
        uint64_t h = mix64(kGolden * (cell_id(table, row, col) + 1));
        for (uint32_t b = 0; b < cd.width; b += 8) {
            uint64_t lo = *reinterpret_cast<const uint32_t *>(cell + b);
            uint64_t hi = (b + 4 < cd.width) ? *reinterpret_cast<const uint32_t *>(cell + b + 4) : 0;
            h = mix64(h ^ (lo | (hi << 32)));
        }
        sum += h;
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < kBlock / 64; ++w) t += part[w];
        atomicAdd(acc, t);
    }
}

void launch_digest_column(hipStream_t s, ColDesc cd, uint32_t table, uint32_t col,
                          uint32_t n_rows, unsigned long long *acc) {
    uint32_t blocks = (n_rows + kBlock - 1) / kBlock;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_digest_column, dim3(blocks), dim3(kBlock), 0, s, cd, table, col,
                       n_rows, acc);
}

// Plain streaming copy (the bandwidth reference): 16 B per lane, grid-stride.
__global__ __launch_bounds__(kBlock) void k_stream_copy(const uint4 *__restrict__ src,
                                                        uint4 *__restrict__ dst, uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_stream_copy(hipStream_t s, const void *src, void *dst, uint64_t n16) {
    if (n16 == 0) return;
    hipLaunchKernelGGL(k_stream_copy, dim3(256 * 64), dim3(kBlock), 0, s,
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void k_fill_u8(uint8_t *p, uint32_t n, uint8_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

void launch_fill_u8(hipStream_t s, uint8_t *p, uint32_t n, uint8_t v) {
    uint32_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_fill_u8, dim3(blocks), dim3(kBlock), 0, s, p, n, v);
}

// Four ops per lane and step, all their loads issued before any store;
// neighbouring lanes read neighbouring pool ops, so each wave's reads of
// host memory are contiguous.  The grid stays small (kGatherBlocks): the
// kernel mostly waits on host memory, and the other slot's kernels need the
// CUs.
#ifndef LTPMG_GATHER16
#define LTPMG_GATHER16 1
#endif
#ifndef LTPMG_GATHER_BLOCKS
#define LTPMG_GATHER_BLOCKS 128  // 64: 1.72, 256: 1.72-1.86 against 1.88 · 10^8 txn/s served at 2 · 10^8
#endif
constexpr uint32_t kGatherBlocks = LTPMG_GATHER_BLOCKS, kGatherPer = 4;
__global__ __launch_bounds__(kBlock) void k_gather_runs(GatherRuns g, uint32_t *txn_off, uint32_t *rows,
                                                        uint16_t *meta, uint64_t *args) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < g.n_ops; i0 += stride * kGatherPer) {
        uint64_t a[kGatherPer];
        uint32_t w[kGatherPer];
        uint16_t m[kGatherPer];
#pragma unroll
        for (uint32_t j = 0; j < kGatherPer; ++j) {
            const uint32_t i = i0 + j * stride;
            if (i >= g.n_ops) continue;
            uint32_t r = 0;
            for (uint32_t k = 1; k < g.n_runs; ++k) r = i >= g.dst[k] ? k : r;
            const uint32_t s = g.src[r] + (i - g.dst[r]);
            a[j] = g.args[s];
            w[j] = g.rows[s];
            m[j] = g.meta[s];
        }
#pragma unroll
        for (uint32_t j = 0; j < kGatherPer; ++j) {
            const uint32_t i = i0 + j * stride;
            if (i >= g.n_ops) continue;
            args[i] = a[j];
            rows[i] = w[j];
            meta[i] = m[j];
        }
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t <= g.n_txns; t += stride) txn_off[t] = g.txn_off[t];
}

// The same with 16-byte reads of host memory (the pool arrays 16-byte
// aligned): a wave takes 512 consecutive pool ops of a run at a time, from
// an op index that is a multiple of 8, and every one of its load
// instructions reads 1 KiB in one piece -- args as four, rows as two, meta
// as one.  Stores go to HBM one element at a time.  The ops of a run before
// its first multiple of 8 and after its last whole group of 8 are copied one
// by one.  (Reads of 16 bytes at a time raise the rate at which a kernel
// pulls host memory over PCIe.)
constexpr uint32_t kChunkOps = 512;
__global__ __launch_bounds__(kBlock) void k_gather_runs16(GatherRuns g, uint32_t *txn_off, uint32_t *rows,
                                                          uint16_t *meta, uint64_t *args) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, n_waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (uint32_t r = 0; r < g.n_runs; ++r) {  // uniform
        const uint32_t s0 = g.src[r], n = (r + 1 < g.n_runs ? g.dst[r + 1] : g.n_ops) - g.dst[r], d0 = g.dst[r];
        const uint32_t b0 = s0 + n < ((s0 + 7u) & ~7u) ? s0 + n : ((s0 + 7u) & ~7u);  // body [b0, b1): groups of 8
        const uint32_t b1 = b0 + ((s0 + n - b0) & ~7u);
        // head and tail, one op per thread
        for (uint32_t k = tid; k < (b0 - s0) + (s0 + n - b1); k += nt) {
            const uint32_t s = k < b0 - s0 ? s0 + k : b1 + (k - (b0 - s0));
            const uint32_t d = d0 + (s - s0);
            args[d] = g.args[s];
            rows[d] = g.rows[s];
            meta[d] = g.meta[s];
        }
        for (uint32_t c0 = b0 + wave * kChunkOps; c0 < b1; c0 += n_waves * kChunkOps) {
            u32x4 va[4], vr[2], vm;
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t o = c0 + 128u * j + 2u * lane;
                if (o + 2u <= b1) va[j] = *reinterpret_cast<const u32x4 *>(g.args + o);
            }
#pragma unroll
            for (uint32_t j = 0; j < 2; ++j) {
                const uint32_t o = c0 + 256u * j + 4u * lane;
                if (o + 4u <= b1) vr[j] = *reinterpret_cast<const u32x4 *>(g.rows + o);
            }
            {
                const uint32_t o = c0 + 8u * lane;
                if (o + 8u <= b1) vm = *reinterpret_cast<const u32x4 *>(g.meta + o);
            }
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t o = c0 + 128u * j + 2u * lane;
                if (o + 2u <= b1) {
                    uint64_t *d = args + d0 + (o - s0);
                    d[0] = (uint64_t)va[j].x | (uint64_t)va[j].y << 32;
                    d[1] = (uint64_t)va[j].z | (uint64_t)va[j].w << 32;
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < 2; ++j) {
                const uint32_t o = c0 + 256u * j + 4u * lane;
                if (o + 4u <= b1) {
                    uint32_t *d = rows + d0 + (o - s0);
                    d[0] = vr[j].x;
                    d[1] = vr[j].y;
                    d[2] = vr[j].z;
                    d[3] = vr[j].w;
                }
            }
            {
                const uint32_t o = c0 + 8u * lane;
                if (o + 8u <= b1) {
                    uint16_t *d = meta + d0 + (o - s0);
                    const uint32_t wv[4] = {vm.x, vm.y, vm.z, vm.w};
#pragma unroll
                    for (uint32_t q = 0; q < 4; ++q) {
                        d[2 * q] = (uint16_t)wv[q];
                        d[2 * q + 1] = (uint16_t)(wv[q] >> 16);
                    }
                }
            }
        }
    }
    for (uint32_t t = tid; t <= g.n_txns; t += nt) txn_off[t] = g.txn_off[t];
}

void launch_gather_runs(hipStream_t s, const GatherRuns &g, uint32_t *txn_off, uint32_t *rows, uint16_t *meta,
                        uint64_t *args) {
    const bool al = ((reinterpret_cast<uintptr_t>(g.rows) | reinterpret_cast<uintptr_t>(g.meta) |
                      reinterpret_cast<uintptr_t>(g.args)) & 15u) == 0;
    uint32_t blocks = (std::max(g.n_ops, g.n_txns + 1) + kBlock * kGatherPer - 1) / (kBlock * kGatherPer);
    if (blocks > kGatherBlocks) blocks = kGatherBlocks;
    if (al && LTPMG_GATHER16)
        hipLaunchKernelGGL(k_gather_runs16, dim3(blocks), dim3(kBlock), 0, s, g, txn_off, rows, meta, args);
    else
        hipLaunchKernelGGL(k_gather_runs, dim3(blocks), dim3(kBlock), 0, s, g, txn_off, rows, meta, args);
}

void launch_zero_u32(hipStream_t s, uint32_t *p, uint32_t n) {
    uint32_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks), dim3(kBlock), 0, s, p, n);
}

}  // namespace ltpmg
