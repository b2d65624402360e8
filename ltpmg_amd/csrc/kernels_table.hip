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

void launch_zero_u32(hipStream_t s, uint32_t *p, uint32_t n) {
    uint32_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks), dim3(kBlock), 0, s, p, n);
}

}  // namespace ltpmg
