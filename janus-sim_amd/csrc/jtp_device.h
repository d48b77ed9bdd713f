// Device-side helpers shared by the kernels (wavefront scans and masks).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace jtp {

__device__ inline int lane_id() { return threadIdx.x & 63; }

__device__ inline uint64_t lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive prefix sum across the 64 lanes of a wavefront.
template <typename T>
__device__ inline T wave_incl_scan(T v) {
    const int l = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T o = __shfl_up(v, off, 64);
        if (l >= off) v += o;
    }
    return v;
}

// Inclusive prefix sum of an int across the 64 lanes with DPP row shifts and row broadcasts (no
// LDS round trips): within each row of 16 lanes by shifts of 1, 2, 4, 8, then the last lane of row
// 0 (2) is added to rows 1 and 3 (row_bcast:15) and lane 31 to rows 2 and 3 (row_bcast:31).
__device__ inline int wave_incl_scan_dpp(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return x;
}

// Inclusive running maximum of a uint32 across the 64 lanes (same DPP pattern as the sum; 0 is
// the identity that lanes outside a shift or broadcast receive).
__device__ inline uint32_t wave_incl_max_dpp(uint32_t x) {
    int v = (int)x;
    v = (int)max((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));
    v = (int)max((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));
    v = (int)max((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));
    v = (int)max((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));
    v = (int)max((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
    v = (int)max((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));
    return (uint32_t)v;
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}
__device__ inline uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}

// Appends the lanes with `pred` to a global list using one atomic per wavefront; returns the
// lane's slot (or -1).  Every lane of the wavefront must call it.
__device__ inline int wave_append(bool pred, int32_t* counter) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return -1;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, (int)__popcll(m));
    base = __shfl(base, leader, 64);
    return pred ? base + (int)__popcll(m & lanemask_lt()) : -1;
}

// Reserves `count` consecutive slots per lane of a global list (one atomic per wavefront); returns
// the lane's first slot (meaningless where count is 0).  Every lane of the wavefront must call it.
__device__ inline int wave_reserve(int count, int32_t* counter) {
    int incl = count;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane_id() >= off) incl += v;
    }
    const int total = __shfl(incl, 63, 64);
    if (total == 0) return -1;
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(counter, total);
    base = __shfl(base, 0, 64);
    return base + incl - count;
}

// Appends the threads with `pred` to a global list using one atomic per 256-thread block (the
// throughput of same-address atomics, not their count per wavefront, limits frontier appends).
// Every thread of the block must call it; `lds` holds 6 int32 of scratch.
__device__ inline int block_append(bool pred, int32_t* counter, int32_t* lds) {
    const int w = threadIdx.x >> 6;
    const uint64_t m = __ballot(pred);
    if (lane_id() == 0) lds[w] = (int32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int32_t tot = lds[0] + lds[1] + lds[2] + lds[3];
        lds[4] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    int off = lds[4];
    for (int i = 0; i < w; ++i) off += lds[i];
    const int r = pred ? off + (int)__popcll(m & lanemask_lt()) : -1;
    __syncthreads();
    return r;
}

// Exclusive prefix sum over a 256-thread block, one value per thread.
// `lds` must hold 4 uint32 scratch entries; `total` receives the block sum.
__device__ inline uint32_t block256_excl_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    const int w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint32_t t = lds[i];
        if (i < w) off += t;
        tot += t;
    }
    __syncthreads();
    if (total) *total = tot;
    return off + inc - v;
}

}  // namespace jtp
