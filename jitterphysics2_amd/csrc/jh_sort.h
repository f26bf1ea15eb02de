// Stable LSD radix sort of 64-bit keys with optional 32-bit values, 8-bit digits per pass.
// The element count lives in device memory (it is produced by an earlier kernel), so no host
// synchronisation is needed; grids are sized by the buffer capacity and tiles past the count exit.
// Per pass: tile histograms (LDS atomics) -> one exclusive scan over digit-major counts -> stable
// scatter where each wave ranks equal digits with eight ballots.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jh {
namespace rsort {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;

__global__ void __launch_bounds__(kThreads) upsweep(const uint64_t* __restrict__ keys, const int* __restrict__ n_ptr,
                                                    int shift, uint32_t* __restrict__ hist, int num_tiles) {
    __shared__ uint32_t h[256];
    int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    int n = *n_ptr;
    int base = blockIdx.x * kTile;
    if (base < n) {
        for (int c = 0; c < kItems; ++c) {
            int idx = base + c * kThreads + tid;
            if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    hist[tid * num_tiles + blockIdx.x] = h[tid];
}

// Exclusive scan of `count` uint32 in place by one workgroup of 1024 threads.
__global__ void __launch_bounds__(1024) scan_exclusive(uint32_t* __restrict__ data, int count) {
    __shared__ uint32_t part[1024];
    int tid = threadIdx.x;
    int per = (count + 1023) / 1024;
    int beg = tid * per;
    int end = min(beg + per, count);
    uint32_t s = 0;
    for (int i = beg; i < end; ++i) s += data[i];
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
        uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

__global__ void __launch_bounds__(kThreads) scatter(const uint64_t* __restrict__ kin, uint64_t* __restrict__ kout,
                                                    const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
                                                    const int* __restrict__ n_ptr, int shift,
                                                    const uint32_t* __restrict__ hist, int num_tiles) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcnt[kThreads / 64][256];
    int tid = threadIdx.x;
    int lane = tid & 63, wave = tid >> 6;
    int n = *n_ptr;
    int base = blockIdx.x * kTile;
    if (base >= n) return;
    running[tid] = hist[tid * num_tiles + blockIdx.x];
    uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int c = 0; c < kItems; ++c) {
        for (int w = 0; w < kThreads / 64; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        int idx = base + c * kThreads + tid;
        bool valid = idx < n;
        uint64_t key = valid ? kin[idx] : 0ull;
        uint32_t d = uint32_t(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            uint64_t m = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pre = 0;
            for (int w = 0; w < wave; ++w) pre += wcnt[w][d];
            uint32_t pos = running[d] + pre + rank;
            kout[pos] = key;
            if (vin) vout[pos] = vin[idx];
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < kThreads / 64; ++w) tot += wcnt[w][tid];
        running[tid] += tot;
        __syncthreads();
    }
}

// Sorts keys[0..*n_ptr) (capacity `cap`) on the 8-bit digits starting at shifts[0..nshifts) (least
// significant first; digits that are zero for every key may be left out).  Ping-pongs between
// (k0, v0) and (k1, v1); returns true when the result ended in (k1, v1).
inline bool sort_pairs(uint64_t* k0, uint64_t* k1, uint32_t* v0, uint32_t* v1, const int* n_ptr, int cap,
                       const int* shifts, int nshifts, uint32_t* hist, hipStream_t st) {
    int tiles = (cap + kTile - 1) / kTile;
    if (tiles < 1) tiles = 1;
    bool flip = false;
    for (int p = 0; p < nshifts; ++p) {
        int shift = shifts[p];
        const uint64_t* ki = flip ? k1 : k0;
        uint64_t* ko = flip ? k0 : k1;
        const uint32_t* vi = flip ? v1 : v0;
        uint32_t* vo = flip ? v0 : v1;
        upsweep<<<tiles, kThreads, 0, st>>>(ki, n_ptr, shift, hist, tiles);
        scan_exclusive<<<1, 1024, 0, st>>>(hist, tiles * 256);
        scatter<<<tiles, kThreads, 0, st>>>(ki, ko, vi, vo, n_ptr, shift, hist, tiles);
        flip = !flip;
    }
    return flip;
}

}  // namespace rsort
}  // namespace jh
