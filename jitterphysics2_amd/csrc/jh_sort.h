// Stable LSD radix sort of 64-bit keys with optional 32-bit values, 8-bit digits per pass.
// The element count lives in device memory (it is produced by an earlier kernel), so no host
// synchronisation is needed; grids are sized by the buffer capacity and tiles past the count exit.
// Per pass: tile histograms (LDS atomics) -> one exclusive scan over digit-major counts -> stable
// scatter where each wave ranks equal digits with eight ballots.  The histogram holds
// 256 * ceil(n / kTile) entries for the current count n (kTile = 2048 keys), so the scan costs what
// the data needs, not what the buffer could hold; above kMultiScanEntries it runs over many workgroups.
// The histogram buffer must hold hist_entries(cap) entries.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jh {
namespace rsort {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;

__device__ __forceinline__ int tiles_for(int n) { return (n + kTile - 1) / kTile; }

// Tile histograms, digit-major: hist[d * tiles + tile] with tiles = ceil(n / kTile) from the device count.
__global__ void __launch_bounds__(kThreads) upsweep(const uint64_t* __restrict__ keys, const int* __restrict__ n_ptr,
                                                    int shift, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    int tid = threadIdx.x;
    int n = *n_ptr;
    int base = blockIdx.x * kTile;
    if (base >= n) return;
    h[tid] = 0;
    __syncthreads();
    for (int c = 0; c < kItems; ++c) {
        int idx = base + c * kThreads + tid;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[tid * tiles_for(n) + blockIdx.x] = h[tid];
}

// Exclusive scan of one chunk of up to kScanChunk entries of data[0..count) in place, plus `carry`;
// one workgroup of 1024 threads, four consecutive entries per thread (one 16-byte load), wave scans
// with shuffles, then the 16 wave totals.  Returns the chunk's total (same value on every thread).
constexpr int kScanChunk = 4096;
__device__ __forceinline__ uint32_t scan_chunk(uint32_t* __restrict__ data, int base, int count, uint32_t carry) {
    __shared__ uint32_t wsum[16];
    int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int idx = base + tid * 4;
    uint4 v = idx < count ? *reinterpret_cast<const uint4*>(data + idx) : make_uint4(0, 0, 0, 0);
    uint32_t t = v.x + v.y + v.z + v.w;
    uint32_t inc = t;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t o = __shfl_up(inc, off, 64);
        if (lane >= off) inc += o;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t wpre = 0, total = 0;
    for (int w = 0; w < 16; ++w) {
        uint32_t x = wsum[w];
        if (w < wave) wpre += x;
        total += x;
    }
    uint32_t run = carry + wpre + inc - t;
    if (idx < count) {
        uint4 o;
        o.x = run; run += v.x;
        o.y = run; run += v.y;
        o.z = run; run += v.z;
        o.w = run;
        *reinterpret_cast<uint4*>(data + idx) = o;
    }
    __syncthreads();
    return total;
}

// Exclusive scan in place of the 256 * ceil(n / kTile) histogram entries by one workgroup.
__global__ void __launch_bounds__(1024) scan_exclusive(uint32_t* __restrict__ data, const int* __restrict__ n_ptr) {
    int count = tiles_for(*n_ptr) * 256;
    uint32_t carry = 0;
    for (int base = 0; base < count; base += kScanChunk) carry += scan_chunk(data, base, count, carry);
}

// The same scan over many workgroups for large histograms: chunk sums, an exclusive scan of the
// sums (one workgroup), then every chunk scanned with its carry.
__global__ void __launch_bounds__(1024) scan_chunk_sums(const uint32_t* __restrict__ data, const int* __restrict__ n_ptr,
                                                        uint32_t* __restrict__ csum) {
    __shared__ uint32_t wsum[16];
    int count = tiles_for(*n_ptr) * 256;
    int base = blockIdx.x * kScanChunk;
    if (base >= count) return;
    int idx = base + threadIdx.x * 4;
    uint4 v = idx < count ? *reinterpret_cast<const uint4*>(data + idx) : make_uint4(0, 0, 0, 0);
    uint32_t t = v.x + v.y + v.z + v.w;
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < 16; ++w) s += wsum[w];
        csum[blockIdx.x] = s;
    }
}
__global__ void __launch_bounds__(1024) scan_chunk_offsets(uint32_t* __restrict__ csum, const int* __restrict__ n_ptr) {
    int chunks = (tiles_for(*n_ptr) * 256 + kScanChunk - 1) / kScanChunk;
    uint32_t carry = 0;
    for (int base = 0; base < chunks; base += kScanChunk) carry += scan_chunk(csum, base, chunks, carry);
}
__global__ void __launch_bounds__(1024) scan_chunk_apply(uint32_t* __restrict__ data, const int* __restrict__ n_ptr,
                                                         const uint32_t* __restrict__ csum) {
    int count = tiles_for(*n_ptr) * 256;
    int base = blockIdx.x * kScanChunk;
    if (base >= count) return;
    scan_chunk(data, base, count, csum[blockIdx.x]);
}

// Histogram entries above which the scan runs over many workgroups (one workgroup scans 4,096
// entries per round, ~0.7 us a round; three short multi-workgroup kernels take ~8 us in all).
constexpr int kMultiScanEntries = 1 << 15;
// uint32 entries a histogram buffer needs for `cap` keys: the histogram, then the chunk sums.
inline size_t hist_entries(size_t cap) {
    size_t tiles = (cap + kTile - 1) / kTile + 1;
    return tiles * 256 + (tiles * 256 + kScanChunk - 1) / kScanChunk + 64;
}

__global__ void __launch_bounds__(kThreads) scatter(const uint64_t* __restrict__ kin, uint64_t* __restrict__ kout,
                                                    const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
                                                    const int* __restrict__ n_ptr, int shift,
                                                    const uint32_t* __restrict__ hist) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcnt[kThreads / 64][256];
    int tid = threadIdx.x;
    int lane = tid & 63, wave = tid >> 6;
    int n = *n_ptr;
    int base = blockIdx.x * kTile;
    if (base >= n) return;
    running[tid] = hist[tid * tiles_for(n) + blockIdx.x];
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
    size_t entries = size_t(tiles) * 256;
    int chunks = int((entries + kScanChunk - 1) / kScanChunk);
    uint32_t* csum = hist + (size_t(tiles) + 1) * 256;  // the chunk sums follow the histogram (hist_entries)
    bool flip = false;
    for (int p = 0; p < nshifts; ++p) {
        int shift = shifts[p];
        const uint64_t* ki = flip ? k1 : k0;
        uint64_t* ko = flip ? k0 : k1;
        const uint32_t* vi = flip ? v1 : v0;
        uint32_t* vo = flip ? v0 : v1;
        upsweep<<<tiles, kThreads, 0, st>>>(ki, n_ptr, shift, hist);
        if (entries > size_t(kMultiScanEntries)) {
            scan_chunk_sums<<<chunks, 1024, 0, st>>>(hist, n_ptr, csum);
            scan_chunk_offsets<<<1, 1024, 0, st>>>(csum, n_ptr);
            scan_chunk_apply<<<chunks, 1024, 0, st>>>(hist, n_ptr, csum);
        } else {
            scan_exclusive<<<1, 1024, 0, st>>>(hist, n_ptr);
        }
        scatter<<<tiles, kThreads, 0, st>>>(ki, ko, vi, vo, n_ptr, shift, hist);
        flip = !flip;
    }
    return flip;
}

}  // namespace rsort
}  // namespace jh
