// Fixed-point conversion, Morton keys and a stable LSD radix sort of
// (63-bit Morton key, int32 particle index) pairs for gfx950.
//
// Reference behaviour replaced: compute_morton_all + radix_histogram /
// radix_prefix_sum / radix_scatter + reorder_by_idx_* (src/nbody_gpu_twopass.rs:
// 1447-1658).  The reference sort is a stable LSD radix sort, so its output is
// the unique stable argsort of the keys; this one is too, with a parallel
// digit-major scan instead of the reference's serial per-bucket prefix loop and
// an LDS-staged tile reorder so that the scatter writes contiguous runs.
#include "jtp_internal.h"
#include "jtp_device.h"

namespace jtp {

// ------------------------------------------------------------------ fixed point
__device__ inline uint32_t to_fixed(double x, double L) {
    double t = x / L + 0.5;
    t -= floor(t);
    uint64_t w = (uint64_t)(t * 4294967296.0);
    if (w >= 4294967296ULL) w -= 4294967296ULL;
    return (uint32_t)w;
}

__global__ void k_fixed_from_f64(const double* __restrict__ pos, uint32_t* __restrict__ out, int64_t n3,
                                 double L) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) out[i] = to_fixed(pos[i], L);
}
__global__ void k_fixed_from_f32(const float* __restrict__ pos, uint32_t* __restrict__ out, int64_t n3,
                                 double L) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) out[i] = to_fixed((double)pos[i], L);
}

hipError_t launch_fixed_from_f64(const double* pos, uint32_t* out, int64_t n3, double L, hipStream_t s) {
    if (n3 > 0)
        k_fixed_from_f64<<<dim3((unsigned)((n3 + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(pos, out, n3, L);
    return hipGetLastError();
}
hipError_t launch_fixed_from_f32(const float* pos, uint32_t* out, int64_t n3, double L, hipStream_t s) {
    if (n3 > 0)
        k_fixed_from_f32<<<dim3((unsigned)((n3 + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(pos, out, n3, L);
    return hipGetLastError();
}

// ------------------------------------------------------------------ Morton keys
// Native keys: the top 21 bits of each fixed-point axis, interleaved (x in bit 0), with the low
// kKeyLowBits bits cleared (the sort skips their passes).  Each particle
// is also packed into one 16-byte record (x, y, z, species) in input order, so that the gather
// after the sort reads one aligned 16-byte word per particle instead of 12 bytes that may straddle
// two segments plus a byte in a third.
__global__ void k_morton_keys(const uint32_t* __restrict__ pos, const int8_t* __restrict__ species,
                              uint64_t* __restrict__ keys, int32_t* __restrict__ vals, uint4* __restrict__ rec,
                              int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    keys[i] = morton_fixed(x, y, z) & kKeySortMask;
    vals[i] = (int32_t)i;
    rec[i] = make_uint4(x, y, z, (uint32_t)(int32_t)species[i]);
}

// Reference-compatible keys from f32 positions (compute_morton_all, nbody_gpu_twopass.rs:1458-1484):
// ix = (unsigned) fminf(fmaxf((x + box_half) * inv_cell, 0), 2^21 - 1), same interleave.
__global__ void k_morton_keys_f32(const float* __restrict__ pos, uint64_t* __restrict__ keys, int64_t n,
                                  float box_half, float inv_cell) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float g = __fmul_rn(__fadd_rn(pos[3 * i + a], box_half), inv_cell);
        c[a] = (uint32_t)fminf(fmaxf(g, 0.0f), 2097151.0f);
    }
    keys[i] = spread3_21(c[0]) | (spread3_21(c[1]) << 1) | (spread3_21(c[2]) << 2);
}

// compute_morton_f32 (nbody_gpu_twopass.rs:302-326): 10 bits per axis of (x + box_half) * 1024 / L in f32,
// the float-to-u32 conversion saturating (negative and NaN -> 0) and clamped to 1023, key = mc << 32 | i.
__device__ inline uint32_t cell10(float x) {
    return !(x > 0.0f) ? 0u : (x >= 1024.0f ? 1023u : (uint32_t)x);
}
__global__ void k_morton30_keys_f32(const float* __restrict__ pos, uint64_t* __restrict__ keys,
                                    int32_t* __restrict__ idx, int64_t n, float box_half) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float scale = __fdiv_rn(1024.0f, __fmul_rn(2.0f, box_half));
    const uint32_t cx = cell10(__fmul_rn(__fadd_rn(pos[3 * i], box_half), scale));
    const uint32_t cy = cell10(__fmul_rn(__fadd_rn(pos[3 * i + 1], box_half), scale));
    const uint32_t cz = cell10(__fmul_rn(__fadd_rn(pos[3 * i + 2], box_half), scale));
    // a 10-bit value sits in the low bits of the 21-bit spread, so spread3_21 gives the 30-bit interleave
    const uint64_t mc = spread3_21(cx) | (spread3_21(cy) << 1) | (spread3_21(cz) << 2);
    keys[i] = (mc << 32) | (uint64_t)(uint32_t)i;
    idx[i] = (int32_t)i;
}

hipError_t launch_morton30_keys_f32(const float* pos, uint64_t* keys, int32_t* idx, int64_t n, float box_half,
                                    hipStream_t s) {
    if (n > 0)
        k_morton30_keys_f32<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(pos, keys, idx, n,
                                                                                                box_half);
    return hipGetLastError();
}

hipError_t launch_morton_keys(const uint32_t* pos, const int8_t* species, uint64_t* keys, int32_t* vals,
                              uint4* rec, int64_t n, hipStream_t s) {
    if (n > 0)
        k_morton_keys<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(pos, species, keys, vals,
                                                                                         rec, n);
    return hipGetLastError();
}
hipError_t launch_morton_keys_f32(const float* pos, uint64_t* keys, int64_t n, float box_half, float inv_cell,
                                  hipStream_t s) {
    if (n > 0)
        k_morton_keys_f32<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(pos, keys, n,
                                                                                              box_half, inv_cell);
    return hipGetLastError();
}

// ------------------------------------------------------------------ radix sort
// Pass structure (8-bit digits from first_bit up, stable; 6 passes for the particle keys):
//   k_radix_hist    per-tile digit counts, stored digit-major: hist[d * nblk + b]
//   k_scan_chunks   exclusive scan of 2048-entry chunks of hist, chunk totals -> part[]
//   k_scan_part     exclusive scan of part[] (one workgroup)
//   k_radix_scatter stable in-tile ranks (wavefront bit-match + per-wave LDS counters),
//                   LDS reorder of the tile by digit, contiguous global writes.
// Digit-major exclusive prefix of the counts is exactly the global start of
// digit d's run for tile b.

constexpr int kScanChunk = 2048;

__global__ __launch_bounds__(256) void k_radix_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                    uint32_t* __restrict__ hist, int nblk) {
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortIPT; ++r) {
        int64_t idx = base + r * 256 + tid;
        if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)tid * nblk + blockIdx.x] = h[tid];
}

__global__ __launch_bounds__(256) void k_scan_chunks(uint32_t* __restrict__ data, int64_t m,
                                                     uint32_t* __restrict__ part) {
    __shared__ uint32_t lds[4];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + tid * 8;
    uint32_t v[8], s = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        v[i] = (base + i < m) ? data[base + i] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t off = block256_excl_scan(s, lds, &tot);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (base + i < m) data[base + i] = off;
        off += v[i];
    }
    if (tid == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_part(uint32_t* __restrict__ part, int64_t np) {
    __shared__ uint32_t lds[4];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < np; b0 += 256) {
        int64_t i = b0 + threadIdx.x;
        uint32_t v = (i < np) ? part[i] : 0u;
        uint32_t tot;
        uint32_t off = block256_excl_scan(v, lds, &tot);
        uint32_t c = carry;
        if (i < np) part[i] = c + off;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
}

// Wavefront w of the block owns the contiguous slice [512 w, 512 w + 512) of the tile (lane l
// holds the keys at 512 w + 64 r + l, r = 0..7), so the ranks follow the input order with no
// block barrier between rounds: a round's rank is the wave's running count of the digit (its own
// LDS counters, read and then bumped by the digit's first lane, in program order within the wave)
// plus the number of earlier lanes holding the digit.  One block step then turns the per-wave
// counts into each wave's start within the tile's run of the digit.
__global__ __launch_bounds__(256) void k_radix_scatter(const uint64_t* __restrict__ kin,
                                                       const int32_t* __restrict__ vin,
                                                       uint64_t* __restrict__ kout, int32_t* __restrict__ vout,
                                                       int64_t n, int shift, const uint32_t* __restrict__ hist,
                                                       const uint32_t* __restrict__ part, int nblk) {
    static_assert(kSortTile == 4 * 64 * kSortIPT, "four wavefronts of kSortIPT keys per lane");
    __shared__ uint32_t wcnt[4][256];  // per-wave digit counts, then each wave's start in the tile
    __shared__ int64_t gbase[256];
    __shared__ uint32_t scan_lds[4];
    __shared__ uint64_t skeys[kSortTile];
    __shared__ int32_t svals[kSortTile];

    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int b = blockIdx.x;
    const int64_t base = (int64_t)b * kSortTile;
    const int64_t wbase = base + (int64_t)w * 64 * kSortIPT + lane;
    const uint64_t lt = lanemask_lt();

    uint64_t key[kSortIPT];
    int32_t val[kSortIPT];
    uint32_t rk[kSortIPT];
#pragma unroll
    for (int r = 0; r < kSortIPT; ++r) {
        const int64_t idx = wbase + r * 64;
        if (idx < n) {
            key[r] = kin[idx];
            val[r] = vin[idx];
        } else {
            key[r] = 0;
            val[r] = 0;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) wcnt[i][tid] = 0;
    __syncthreads();

    uint32_t* mine = wcnt[w];
#pragma unroll
    for (int r = 0; r < kSortIPT; ++r) {
        const bool valid = wbase + r * 64 < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool bs = (d >> bit) & 1u;
            const uint64_t bb = __ballot(bs);
            peers &= bs ? bb : ~bb;
        }
        const uint32_t before = mine[d];
        rk[r] = before + (uint32_t)__popcll(peers & lt);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // every lane's read before the bump
        if (valid && (peers & lt) == 0) mine[d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    // digit tid: the tile's run starts at toff; wave ww's keys of the digit follow those of waves < ww
    {
        const uint32_t c0 = wcnt[0][tid], c1 = wcnt[1][tid], c2 = wcnt[2][tid], c3 = wcnt[3][tid];
        const uint32_t toff = block256_excl_scan(c0 + c1 + c2 + c3, scan_lds, nullptr);
        wcnt[0][tid] = toff;
        wcnt[1][tid] = toff + c0;
        wcnt[2][tid] = toff + c0 + c1;
        wcnt[3][tid] = toff + c0 + c1 + c2;
        const int64_t hidx = (int64_t)tid * nblk + b;
        gbase[tid] = (int64_t)hist[hidx] + (int64_t)part[hidx / kScanChunk] - (int64_t)toff;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortIPT; ++r) {
        if (wbase + r * 64 < n) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
            const uint32_t l = mine[d] + rk[r];
            skeys[l] = key[r];
            svals[l] = val[r];
        }
    }
    __syncthreads();
    const int64_t cnt = (n - base) < kSortTile ? (n - base) : kSortTile;
    for (int l = tid; l < cnt; l += 256) {
        const uint64_t k = skeys[l];
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        const int64_t o = gbase[d] + l;
        kout[o] = k;
        vout[o] = svals[l];
    }
}

hipError_t radix_sort_pairs(Ctx& c, int64_t n, hipStream_t s, int first_bit) {
    if (n <= 1) return hipGetLastError();
    if (first_bit < 0 || first_bit > 48 || first_bit % 16 != 0) return hipErrorInvalidValue;
    const int nblk = (int)((n + kSortTile - 1) / kSortTile);
    const int64_t m = (int64_t)nblk * 256;
    const int64_t nchunk = (m + kScanChunk - 1) / kScanChunk;
    uint64_t* ka = c.keys0;
    uint64_t* kb = c.keys1;
    int32_t* va = c.vals0;
    int32_t* vb = c.vals1;
    for (int shift = first_bit; shift < 64; shift += 8) {
        k_radix_hist<<<dim3(nblk), dim3(256), 0, s>>>(ka, n, shift, c.hist, nblk);
        k_scan_chunks<<<dim3((unsigned)nchunk), dim3(256), 0, s>>>(c.hist, m, c.hist_blk);
        k_scan_part<<<dim3(1), dim3(256), 0, s>>>(c.hist_blk, nchunk);
        k_radix_scatter<<<dim3(nblk), dim3(256), 0, s>>>(ka, va, kb, vb, n, shift, c.hist, c.hist_blk, nblk);
        uint64_t* tk = ka; ka = kb; kb = tk;
        int32_t* tv = va; va = vb; vb = tv;
    }
    // an even number of passes: the result is back in keys0/vals0
    return hipGetLastError();
}

// ------------------------------------------------------------------ device copy
// Streaming copy (16-byte words when both pointers and the size allow it).  Also the
// reference copy kernel for calibrating HBM byte counters.
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t device_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    const bool aligned = (((uintptr_t)dst | (uintptr_t)src | bytes) & 15) == 0;
    if (!aligned) return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
    const int64_t n16 = (int64_t)(bytes / 16);
    int64_t blocks = (n16 + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    k_copy16<<<dim3((unsigned)blocks), dim3(kBlock), 0, s>>>((const uint4*)src, (uint4*)dst, n16);
    return hipGetLastError();
}

// ------------------------------------------------------------------ particle gather
// parts[i] = (x, y, z, species) of the i-th particle in key order (reorder_by_idx_*,
// nbody_gpu_twopass.rs:1490-1514, fused into one 16-byte record).  With ts, also the start of every
// Morton-aligned tile of 8^3 PM cells (lt = log2 G - 3 tile bits per axis; the tile index is a
// prefix of the key): the particle that begins a run writes the starts of the tiles up to its own,
// so every entry is written once, ts[ntile] = n.  Lanes read the previous particle's tile from
// their neighbour; lane 0 gathers it again.
__global__ void k_gather_parts(const uint4* __restrict__ rec, const int32_t* __restrict__ perm,
                               uint4* __restrict__ parts, int64_t n, int32_t* __restrict__ ts, int lt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < n;
    uint4 r = make_uint4(0u, 0u, 0u, 0u);
    if (in) {
        r = rec[perm[i]];
        parts[i] = r;
    }
    if (!ts) return;
    const int sh = 32 - lt;
    const int32_t ntile = 1 << (3 * lt);
    auto tile_of = [&](const uint4& q) {
        return (int32_t)(spread3_21(q.x >> sh) | (spread3_21(q.y >> sh) << 1) | (spread3_21(q.z >> sh) << 2));
    };
    const int32_t tc = in ? tile_of(r) : ntile;
    int32_t tp = __shfl_up(tc, 1, 64);
    if ((threadIdx.x & 63) == 0) tp = (i == 0 || !in) ? -1 : tile_of(rec[perm[i - 1]]);
    if (!in) return;
    for (int32_t t = tp + 1; t <= tc; ++t) ts[t] = (int32_t)i;
    if (i == n - 1)
        for (int32_t t = tc + 1; t <= ntile; ++t) ts[t] = (int32_t)n;
}

hipError_t launch_gather_parts(const uint4* rec, const int32_t* perm, uint4* parts, int64_t n, int32_t* ts, int logG,
                               hipStream_t s) {
    if (n > 0) {
        k_gather_parts<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(rec, perm, parts, n, ts,
                                                                                         logG - 3);
    } else if (ts) {
        const int64_t ntile = (int64_t)1 << (3 * (logG - 3));
        return launch_fill_i32(ts, ntile + 1, 0, s);
    }
    return hipGetLastError();
}

}  // namespace jtp
