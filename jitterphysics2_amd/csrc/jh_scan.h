// Device-count exclusive scan / stream compaction helpers (counts live in device memory).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jh {
namespace scan {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kBlock / 64];
    int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // wave inclusive scan with shuffles
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        uint32_t s = wsum[w];
        if (w < wave) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

// Pass 1: per-block sums of flags[0..n).
__global__ void __launch_bounds__(kBlock) block_sums(const uint32_t* __restrict__ flags, const int* __restrict__ n_ptr,
                                                     uint32_t* __restrict__ sums) {
    int n = *n_ptr;
    int i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t v = i < n ? flags[i] : 0u;
    uint32_t tot;
    block_excl_scan(v, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
// Pass 2: exclusive scan of block sums (one workgroup of 1024, any count), total to *out_total.
__global__ void __launch_bounds__(1024) scan_sums(uint32_t* __restrict__ sums, int nblocks, int* __restrict__ out_total) {
    __shared__ uint32_t part[1024];
    int tid = threadIdx.x;
    int per = (nblocks + 1023) / 1024;
    int beg = tid * per, end = min(beg + per, nblocks);
    uint32_t s = 0;
    for (int i = beg; i < end; ++i) s += sums[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = part[tid] - s;
    for (int i = beg; i < end; ++i) {
        uint32_t v = sums[i];
        sums[i] = run;
        run += v;
    }
    if (tid == 1023 && out_total) *out_total = int(part[1023]);
}
// Pass 3: exclusive scan of flags into offsets.
__global__ void __launch_bounds__(kBlock) block_scan(const uint32_t* __restrict__ flags, const int* __restrict__ n_ptr,
                                                     const uint32_t* __restrict__ sums, uint32_t* __restrict__ offs) {
    int n = *n_ptr;
    int i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t v = i < n ? flags[i] : 0u;
    uint32_t tot;
    uint32_t e = block_excl_scan(v, &tot);
    if (i < n) offs[i] = sums[blockIdx.x] + e;
}

// offs[i] = sum(flags[0..i)), *out_total = sum(flags[0..n)); capacity `cap` sizes the grid.
inline void exclusive_scan(const uint32_t* flags, uint32_t* offs, const int* n_ptr, int cap, uint32_t* sums,
                           int* out_total, hipStream_t st) {
    int nb = (cap + kBlock - 1) / kBlock;
    if (nb < 1) nb = 1;
    block_sums<<<nb, kBlock, 0, st>>>(flags, n_ptr, sums);
    scan_sums<<<1, 1024, 0, st>>>(sums, nb, out_total);
    block_scan<<<nb, kBlock, 0, st>>>(flags, n_ptr, sums, offs);
}

}  // namespace scan
}  // namespace jh
