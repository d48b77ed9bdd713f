// Long-range (PM) part: deterministic CIC deposit of rho+ - rho-, 3-D real FFT
// Poisson solve with the split Green function, gradient + CIC interpolation.
//
// Reference behaviour replaced:
//   cic_scatter (FP64 CAS atomics into two grids), nbody_gpu_twopass.rs:85-158,
//     and PmGrid::assign_mass, src/treepm/pm_grid.rs:63-103 -> one net grid rho+ - rho- from the
//     Morton-sorted particles, no float atomics: FP32 k_dep_tile (each particle read once, integer
//     LDS sums per 8^3-cell tile) + k_dep_assemble (each tile's own node blocks written as whole rows,
//     the lower neighbours' overlaps added in a fixed order); FP64 k_deposit(_both) (each node
//     gathers its source cells in a fixed order).
//   cufft_solve_device_filtered (cuFFT D2Z / apply_green_kernel / Z2D / normalize),
//     cuda/cufft_wrapper.cu:148-208,290-295,350-398, and the rustfft passes of
//     PmGrid::solve_poisson_filtered, pm_grid.rs:132-271 -> five LDS Stockham passes
//     (R2C x, y, z with the Green multiply fused between forward and inverse z, y, C2R x).
//   cic_gather (nbody_gpu_twopass.rs:162-257) and PmGrid::interpolate_force
//     (pm_grid.rs:275-359) -> on the single net potential (Phi- = -Phi+): FP32 k_interp_light /
//     k_interp_heavy per 8^3-cell tile from LDS; FP64 k_grad + k_interp through gradient grids.
//
// Grid layout: real[z][y][x] (x fastest; index i + G(j + G k) as in pm_grid.rs:97),
// half spectrum complex[z][y][kx], kx in [0, G/2].
#include "jtp_internal.h"
#include "jtp_device.h"

#include <cmath>
#include <cstdio>
#include <mutex>
#include <vector>

namespace jtp {

template <typename R>
__device__ inline R frac_of(uint32_t u, int logG) {
    const uint32_t fb = 32 - logG;
    const uint32_t m = (fb >= 32) ? 0xffffffffu : ((1u << fb) - 1u);
    return (R)(u & m) * (R)(1.0 / (double)(1ull << fb));
}

// ------------------------------------------------------------------ deposit
// FP64 form: deterministic CIC deposit ("sort + segmented reduction").  The particles are
// in Morton order and G is a power of two, so every PM cell is a contiguous run
// (cell_start).  Cells holding more than kDenseCell particles (halo cores hold up
// to ~1e5) are first reduced by wavefronts into their 8 corner sums, in a
// fixed lane/tree order; then every grid node gathers, in a fixed order, the
// corner sums of dense source cells and the particles of the others.
template <typename R>
__device__ inline void cic_corners(const uint4 r, int logG, R q, R w[8]) {
    const R fx = frac_of<R>(r.x, logG), fy = frac_of<R>(r.y, logG), fz = frac_of<R>(r.z, logG);
    const R wx[2] = {(R)1 - fx, fx}, wy[2] = {(R)1 - fy, fy}, wz[2] = {(R)1 - fz, fz};
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = q * (wx[c & 1] * wy[(c >> 1) & 1] * wz[c >> 2]);
}

// A dense cell is reduced in chunks of kDenseChunk particles (one wavefront each, so a halo core
// of ~1e4 particles is not one long serial loop); slot[c] is its first chunk, list[] maps each
// chunk back to its cell, and k_dense_combine adds a cell's chunk sums, in chunk order, into its
// first chunk, where the consumers read them.
__device__ inline int dense_chunks(int32_t cnt) { return (cnt + kDenseChunk - 1) / kDenseChunk; }

__global__ void k_dense_mark(const int32_t* __restrict__ cs, int64_t ncells, int32_t* __restrict__ slot,
                             int32_t* __restrict__ list, int32_t* __restrict__ count) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t cnt = c < ncells ? cs[c + 1] - cs[c] : 0;
    const int nch = cnt > kDenseCell ? dense_chunks(cnt) : 0;
    const int k = wave_reserve(nch, count);
    if (nch > 0) {
        slot[c] = k;
        for (int i = 0; i < nch; ++i) list[k + i] = (int32_t)c;
    }
}

// Thread per (chunk, value): the first chunk of a cell sums the value over the cell's chunks in order.
template <typename R>
__global__ __launch_bounds__(256) void k_dense_combine(const int32_t* __restrict__ cs,
                                                       const int32_t* __restrict__ list,
                                                       const int32_t* __restrict__ slot,
                                                       const int32_t* __restrict__ count, int nv,
                                                       R* __restrict__ vals) {
    const int32_t nd = *count;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)nd * nv;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = t / nv;
        const int q = (int)(t - k * nv);
        const int32_t c = list[k];
        if (slot[c] != k) continue;
        const int nch = dense_chunks(cs[c + 1] - cs[c]);
        if (nch == 1) continue;
        R v = vals[k * nv + q];
        for (int i = 1; i < nch; ++i) v += vals[(k + i) * nv + q];
        vals[k * nv + q] = v;
    }
}

// One wavefront per dense-cell chunk: lane l takes particles s + l, s + l + 64, ...; then a
// fixed butterfly reduction across lanes.  Unshifted corner index c = dx + 2 dy + 4 dz
// (offset of the node from the cell's lower corner).  With interlacing, also the 27 sums
// for the shifted grid, whose nodes sit at the centres of cells c-1, c, c+1 per axis:
// index (ex+1) + 3 (ey+1) + 9 (ez+1) for the node at the centre of cell c + e.
template <typename R>
__device__ inline void shifted_weights(R f, R w[3]) {
    const R h = (R)0.5;
    w[0] = f < h ? h - f : (R)0;        // node at the centre of the cell below
    w[1] = f < h ? h + f : (R)1.5 - f;  // 1 - |f - 1/2|
    w[2] = f < h ? (R)0 : f - h;        // node at the centre of the cell above
}

template <typename R>
__global__ __launch_bounds__(256) void k_dense_reduce(const uint4* __restrict__ parts, const int32_t* __restrict__ cs,
                                                      const int32_t* __restrict__ list,
                                                      const int32_t* __restrict__ slot,
                                                      const int32_t* __restrict__ count, int logG, R qp, R qm,
                                                      R* __restrict__ corner) {
    const int lane = threadIdx.x & 63;
    const int32_t nd = *count;
    for (int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nd;
         k += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        const int32_t c = list[k];
        const int32_t s = cs[c] + (int32_t)(k - slot[c]) * kDenseChunk, e = min(cs[c + 1], s + kDenseChunk);
        R acc[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = 0;
#pragma unroll 4
        for (int32_t p = s + lane; p < e; p += 64) {
            const uint4 r = parts[p];
            const R q = (int32_t)r.w > 0 ? qp : qm;
            R w[8];
            cic_corners<R>(r, logG, q, w);
#pragma unroll
            for (int q2 = 0; q2 < 8; ++q2) acc[q2] += w[q2];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            R v = acc[q];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            acc[q] = v;
        }
        if (lane < 8) {
            R v = acc[0];
#pragma unroll
            for (int q = 1; q < 8; ++q) v = lane == q ? acc[q] : v;
            corner[8 * (int64_t)k + lane] = v;
        }
    }
}

// Both grids of an interlaced deposit in one pass over a dense cell's particles: the 8 node-grid
// corner sums (as k_dense_reduce) and the 27 sums for the shifted grid's nodes at the centres of
// cells c-1, c, c+1 per axis, index (ex+1) + 3 (ey+1) + 9 (ez+1) for the centre of cell c + e.
template <typename R>
__global__ __launch_bounds__(256) void k_dense_reduce_both(const uint4* __restrict__ parts,
                                                           const int32_t* __restrict__ cs,
                                                           const int32_t* __restrict__ list,
                                                           const int32_t* __restrict__ slot,
                                                           const int32_t* __restrict__ count, int logG, R qp, R qm,
                                                           R* __restrict__ corner8, R* __restrict__ corner27) {
    const int lane = threadIdx.x & 63;
    const int32_t nd = *count;
    for (int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nd;
         k += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        const int32_t c = list[k];
        const int32_t s = cs[c] + (int32_t)(k - slot[c]) * kDenseChunk, e = min(cs[c + 1], s + kDenseChunk);
        R a8[8], a27[27];
#pragma unroll
        for (int q = 0; q < 8; ++q) a8[q] = 0;
#pragma unroll
        for (int q = 0; q < 27; ++q) a27[q] = 0;
#pragma unroll 4
        for (int32_t p = s + lane; p < e; p += 64) {
            const uint4 r = parts[p];
            const R q = (int32_t)r.w > 0 ? qp : qm;
            R w[8];
            cic_corners<R>(r, logG, q, w);
#pragma unroll
            for (int q2 = 0; q2 < 8; ++q2) a8[q2] += w[q2];
            R wx[3], wy[3], wz[3];
            shifted_weights<R>(frac_of<R>(r.x, logG), wx);
            shifted_weights<R>(frac_of<R>(r.y, logG), wy);
            shifted_weights<R>(frac_of<R>(r.z, logG), wz);
#pragma unroll
            for (int cz = 0; cz < 3; ++cz)
#pragma unroll
                for (int cy = 0; cy < 3; ++cy)
#pragma unroll
                    for (int cx = 0; cx < 3; ++cx) a27[cx + 3 * cy + 9 * cz] += q * (wx[cx] * wy[cy] * wz[cz]);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            R v = a8[q];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            a8[q] = v;
        }
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            R v = a27[q];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            a27[q] = v;
        }
        if (lane < 8) {
            R v = a8[0];
#pragma unroll
            for (int q = 1; q < 8; ++q) v = lane == q ? a8[q] : v;
            corner8[8 * (int64_t)k + lane] = v;
        }
        if (lane < 27) {
            R v = a27[0];
#pragma unroll
            for (int q = 1; q < 27; ++q) v = lane == q ? a27[q] : v;
            corner27[27 * (int64_t)k + lane] = v;
        }
    }
}

template <typename R> struct Vec4T;
template <> struct Vec4T<float> { using T = float4; };
template <> struct Vec4T<double> { using T = double4; };

// Both grids in one gather: the thread of cell m = (i, j, k) reads the particles of cells m-1..m+1
// per axis once and sums (a) the node-grid node (i, j, k) from cells i-1, i per axis, in the order
// of k_deposit, and (b) the interlaced node at the centre of cell m, written at (i+1, j+1, k+1): per
// axis, with d the cell offset from m and f the fraction in that cell, weight max(0, f - 1/2) for
// d = -1, 1 - |f - 1/2| for d = 0 and max(0, 1/2 - f) for d = +1 (shifted_weights).
template <typename R>
__global__ void k_deposit_both(const uint4* __restrict__ parts, const int32_t* __restrict__ cs,
                               const int32_t* __restrict__ slot, const R* __restrict__ corner8,
                               const R* __restrict__ corner27, int logG, R qp, R qm, R* __restrict__ rho,
                               R* __restrict__ rho2) {
    const int64_t G = 1ll << logG;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= G * G * G) return;
    const int32_t i = (int32_t)compact3_21((uint64_t)m);
    const int32_t j = (int32_t)compact3_21((uint64_t)m >> 1);
    const int32_t k = (int32_t)compact3_21((uint64_t)m >> 2);
    const int32_t msk = (int32_t)(G - 1);
    R acc1 = 0, acc2 = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const bool both = dx <= 0 && dy <= 0 && dz <= 0;  // also a source cell of node (i, j, k)
                const uint32_t ci = (uint32_t)((i + dx) & msk), cj = (uint32_t)((j + dy) & msk),
                               ck = (uint32_t)((k + dz) & msk);
                const uint64_t cm = spread3_21(ci) | (spread3_21(cj) << 1) | (spread3_21(ck) << 2);
                const int32_t s = cs[cm], e = cs[cm + 1];
                if (e - s > kDenseCell) {
                    const int64_t sl = slot[cm];
                    acc2 += corner27[27 * sl + (1 - dx) + 3 * (1 - dy) + 9 * (1 - dz)];
                    if (both) acc1 += corner8[8 * sl + (-dx) + 2 * (-dy) + 4 * (-dz)];
                    continue;
                }
                for (int32_t p = s; p < e; ++p) {
                    const uint4 r = parts[p];
                    const R fx = frac_of<R>(r.x, logG), fy = frac_of<R>(r.y, logG), fz = frac_of<R>(r.z, logG);
                    const R q = (int32_t)r.w > 0 ? qp : qm;
                    R wx[3], wy[3], wz[3];
                    shifted_weights<R>(fx, wx);
                    shifted_weights<R>(fy, wy);
                    shifted_weights<R>(fz, wz);
                    acc2 += q * (wx[1 - dx] * wy[1 - dy] * wz[1 - dz]);
                    if (both) {
                        const R ux = dx == 0 ? (R)1 - fx : fx;
                        const R uy = dy == 0 ? (R)1 - fy : fy;
                        const R uz = dz == 0 ? (R)1 - fz : fz;
                        acc1 += q * (ux * uy * uz);
                    }
                }
            }
        }
    }
    rho[((int64_t)k * G + j) * G + i] = acc1;
    const int64_t ii = (i + 1) & msk, jj = (j + 1) & msk, kk = (k + 1) & msk;
    rho2[(kk * G + jj) * G + ii] = acc2;
}

constexpr int kTile = 8;     // deposit and interpolation tiles: Morton-aligned blocks of 8^3 cells
constexpr int kFixBits = 22;  // FP32 deposit: |q| 2^k <= 2^kFixBits, so one contribution fits an int32

// Tiles are classed by particle count: light tiles (1 .. kLightMax particles, kLightMax / 256 per thread) are
// listed for k_interp_light; every chunk of kInterpChunk particles of a heavier tile (a halo core
// holds 1e4-1e5) is listed, as (tile, first particle), for a k_interp_heavy workgroup.  The list
// order varies between runs; each particle is still computed by one thread with the same
// arithmetic, so the results do not.
//
// Tile lists of the FP32 deposit and interpolation, from the tile starts ts[t] (first sorted particle
// of 8^3-cell tile t, ts[ntile] = n): light tiles (1 .. kLightMax particles) and the kInterpChunk-particle
// chunks of the others for the interpolation; for the deposit, the chunks of tiles holding more than
// kInterpChunk particles, each tile's chunks in consecutive entries in chunk order.
// count: light tiles, interpolation chunks, deposit chunks.
__global__ __launch_bounds__(256) void k_tile_mark(const int32_t* __restrict__ ts, int32_t ntile,
                                                   int32_t* __restrict__ light, int2* __restrict__ heavy,
                                                   int2* __restrict__ dchunks, int32_t* __restrict__ count) {
    const int32_t t = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    int32_t beg = 0, np = 0;
    if (t < ntile) {
        beg = ts[t];
        np = ts[t + 1] - beg;
    }
    const bool lt = np > 0 && np <= kLightMax;
    const int32_t nh = np > kLightMax ? (np + kInterpChunk - 1) / kInterpChunk : 0;
    const int at_l = wave_reserve(lt ? 1 : 0, count);
    if (lt) light[at_l] = t;
    const int at_h = wave_reserve(nh, count + 1);
    for (int32_t k = 0; k < nh; ++k) heavy[at_h + k] = make_int2(t, beg + k * kInterpChunk);
    const int32_t nd = np > kInterpChunk ? nh : 0;
    const int at_d = wave_reserve(nd, count + 2);
    for (int32_t k = 0; k < nd; ++k) dchunks[at_d + k] = make_int2(t, beg + k * kInterpChunk);
}

// ---- FP32 deposit, each particle read once.
// A workgroup takes the contiguous Morton run of one 8^3-cell tile's particles (a tile of more
// than kInterpChunk particles is cut into chunks of that many, one workgroup each) and adds every
// CIC contribution into LDS as 64-bit fixed point (charges pre-scaled by 2^k, |q| 2^k <= 2^22,
// each contribution rounded to an int32 once): the node-grid nodes o .. o + 8 per axis its cells
// touch and the interlaced nodes o .. o + 9 (LDS coordinate L is interlaced node o + L, the centre
// of cell o + L - 1).  Integer addition is associative, so the sums do not depend on the order in
// which the LDS atomics land.  Each tile owns the sector-aligned block of nodes o .. o + 7 per axis
// of each grid; its LDS nodes beyond that block (coordinate 8 on some axis, interlaced 8 or 9) are
// own nodes of its upper neighbours.  k_dep_tile converts all its sums to float into the tile's
// record (own blocks, then the beyond nodes); k_dep_assemble writes every tile's own blocks as whole
// rows, adding to each face node (coordinate 0 on some axis, interlaced 0 or 1) the beyond values
// of the lower neighbours that reach it, in a fixed order.  So every 32-byte sector of the grids is
// written once and whole.  A chunked tile's chunks write their whole LDS grid as int64 and
// k_dep_merge adds them exactly, in chunk order, into the tile's record.  The result is
// bit-reproducible and needs no cell table, no zeroing pass and no float atomics.
// Each thread takes a contiguous share of the run; the particles of one (cell, half-cell octant) all
// reach the same 8 + 8 nodes, so a thread sums consecutive ones in registers and adds one LDS atomic
// per node and run (dep_range).
constexpr int kDepPer = kInterpChunk / 256;        // at most 16 particles per thread
static_assert((int64_t)kDepPer << kFixBits <= (int64_t)1 << 30, "a thread's run sum fits an int32");
constexpr int kD1 = kTile + 1;                     // node-grid LDS coordinates 0..8 per axis
constexpr int kD2 = kTile + 2;                     // interlaced LDS coordinates 0..9 per axis
constexpr int kDN1 = kD1 * kD1 * kD1;              // 729
constexpr int kDN2 = kD2 * kD2 * kD2;              // 1000
constexpr int kDepOwn = kTile * kTile * kTile;     // 512 own nodes per grid
constexpr int kDepH1 = kDN1 - kDepOwn;             // 217 beyond nodes, node grid
constexpr int kDepH2 = kDN2 - kDepOwn;             // 488, interlaced grid
static_assert(kTile == 8, "own blocks are indexed by the bits of a 9-bit index");

// A tile's record (floats): own node-grid block (x fastest), own interlaced block, node-grid beyond
// nodes, interlaced beyond nodes; without interlacing the node grid's two parts only.
template <bool BOTH> constexpr int dep_rec() { return BOTH ? kDN1 + kDN2 : kDN1; }
template <bool BOTH> constexpr int dep_beyond1() { return BOTH ? 2 * kDepOwn : kDepOwn; }
template <bool BOTH> constexpr int dep_beyond2() { return dep_beyond1<BOTH>() + kDepH1; }

// Beyond nodes in slabs, so that both directions are arithmetic.  Node grid (coordinates 0..8, at
// least one 8): z = 8 (81, x fastest), then y = 8 for z < 8 (72), then x = 8 for y, z < 8 (64).
// Interlaced (0..9, at least one 8 or 9): z in {8, 9} (2 x 100), then y in {8, 9} for z < 8
// (2 x 80), then x in {8, 9} for y, z < 8 (2 x 64).
constexpr __host__ __device__ inline int dep_beyond_slot1(int x, int y, int z) {
    return z == kTile ? y * kD1 + x : y == kTile ? 81 + z * kD1 + x : 153 + z * kTile + y;
}
constexpr __host__ __device__ inline void dep_beyond_node1(int s, int& x, int& y, int& z) {
    if (s < 81) {
        z = kTile; y = s / kD1; x = s % kD1;
    } else if (s < 153) {
        y = kTile; z = (s - 81) / kD1; x = (s - 81) % kD1;
    } else {
        x = kTile; z = (s - 153) / kTile; y = (s - 153) % kTile;
    }
}
constexpr __host__ __device__ inline int dep_beyond_slot2(int x, int y, int z) {
    return z >= kTile   ? (z - kTile) * 100 + y * kD2 + x
           : y >= kTile ? 200 + (y - kTile) * 80 + z * kD2 + x
                        : 360 + (x - kTile) * 64 + z * kTile + y;
}
constexpr __host__ __device__ inline void dep_beyond_node2(int s, int& x, int& y, int& z) {
    if (s < 200) {
        z = kTile + s / 100; y = s % 100 / kD2; x = s % kD2;
    } else if (s < 360) {
        const int q = s - 200;
        y = kTile + q / 80; z = q % 80 / kD2; x = q % kD2;
    } else {
        const int q = s - 360;
        x = kTile + q / 64; z = q % 64 / kTile; y = q % kTile;
    }
}
static_assert(81 + 72 + 64 == kDepH1 && 200 + 160 + 128 == kDepH2, "beyond slab sizes");

// LDS index of each entry of the interlaced tile record (the node-grid-only record is its entries
// 0..511 and 1024..1240).
struct DepRecMap {
    uint16_t lds[kDN1 + kDN2];
};
constexpr DepRecMap make_dep_rec_map() {
    DepRecMap m{};
    for (int j = 0; j < kDN1 + kDN2; ++j) {
        int x = 0, y = 0, z = 0, g2 = 0;
        if (j < 2 * kDepOwn) {
            const int k = j & (kDepOwn - 1);
            x = k & 7; y = (k >> 3) & 7; z = k >> 6; g2 = j >= kDepOwn;
        } else if (j < 2 * kDepOwn + kDepH1) {
            dep_beyond_node1(j - 2 * kDepOwn, x, y, z);
        } else {
            dep_beyond_node2(j - 2 * kDepOwn - kDepH1, x, y, z);
            g2 = 1;
        }
        m.lds[j] = (uint16_t)(g2 ? kDN1 + (z * kD2 + y) * kD2 + x : (z * kD1 + y) * kD1 + x);
    }
    return m;
}
__constant__ DepRecMap kDepRecMap = make_dep_rec_map();
constexpr int kDepRecIter = (kDN1 + kDN2 + 255) / 256;  // record entries per thread of a 256-thread block

// Compile-time checks: the record map is a bijection onto the LDS nodes, the beyond numbering inverts
// in both grids, and in k_dep_assemble's sums every beyond node of a tile is read exactly once, by the
// upper neighbour that owns the node (the one whose own coordinate is the node's minus 8 on the axes
// where it is 8 or more).
constexpr bool dep_maps_ok() {
    const DepRecMap m = make_dep_rec_map();
    bool hit[kDN1 + kDN2] = {};
    for (int j = 0; j < kDN1 + kDN2; ++j) {
        if (m.lds[j] >= kDN1 + kDN2 || hit[m.lds[j]]) return false;
        hit[m.lds[j]] = true;
    }
    for (int s = 0; s < kDepH1; ++s) {
        int x = 0, y = 0, z = 0;
        dep_beyond_node1(s, x, y, z);
        if (dep_beyond_slot1(x, y, z) != s || x > kTile || y > kTile || z > kTile) return false;
        if (x < kTile && y < kTile && z < kTile) return false;
    }
    for (int s = 0; s < kDepH2; ++s) {
        int x = 0, y = 0, z = 0;
        dep_beyond_node2(s, x, y, z);
        if (dep_beyond_slot2(x, y, z) != s || x > kTile + 1 || y > kTile + 1 || z > kTile + 1) return false;
        if (x < kTile && y < kTile && z < kTile) return false;
    }
    int r1[kDepH1] = {}, r2[kDepH2] = {};
    for (int g = 0; g < 2; ++g)
        for (int k = 0; k < kDepOwn; ++k) {
            const int x = k & 7, y = (k >> 3) & 7, z = k >> 6, lim = g ? 2 : 1;
            for (int c = 1; c < 8; ++c) {
                if (((c & 1) && x >= lim) || ((c & 2) && y >= lim) || ((c & 4) && z >= lim)) continue;
                const int ax = x + ((c & 1) ? kTile : 0), ay = y + ((c & 2) ? kTile : 0), az = z + ((c & 4) ? kTile : 0);
                if (g) ++r2[dep_beyond_slot2(ax, ay, az)];
                else ++r1[dep_beyond_slot1(ax, ay, az)];
            }
        }
    for (int s = 0; s < kDepH1; ++s)
        if (r1[s] != 1) return false;
    for (int s = 0; s < kDepH2; ++s)
        if (r2[s] != 1) return false;
    return true;
}
static_assert(dep_maps_ok(), "deposit record / beyond-node maps");

// tile t's origin cell, from its Morton index
__device__ inline void tile_origin(uint32_t t, int32_t o[3]) {
    const uint64_t base = (uint64_t)t << 9;
    o[0] = (int32_t)compact3_21(base);
    o[1] = (int32_t)compact3_21(base >> 1);
    o[2] = (int32_t)compact3_21(base >> 2);
}

__device__ inline float dep_value(unsigned long long v, double unit) { return (float)((double)(long long)v * unit); }

// Rounding by the last product's FMA: for |x| <= 2^22, x + 1.5 2^23 lies in [2^23, 2^24], where floats are
// the integers, so fma(a, b, kRoundMagic) is the nearest integer to a b (ties to even) and its bit
// pattern minus kRoundMagicBits is that integer.
constexpr float kRoundMagic = 12582912.0f;
constexpr uint32_t kRoundMagicBits = 0x4B400000u;
static_assert(kFixBits <= 22, "contributions must stay within the exact range of the rounding FMA");

// The CIC contributions of one particle as int32 fixed point plus kRoundMagicBits (the bits of the
// rounding FMA), v[0..7] for the node grid (corner (dx, dy, dz) = (c & 1, c >> 1 & 1, c >> 2) from its
// cell's lower corner) and v[8..15] for the interlaced grid (same corner order from the centre of cell
// (c - 1 + h) per axis, h = 1 if the particle is in the upper half of its cell); key = its (cell,
// half-cell octant), whose particles all reach the same 8 + 8 nodes; n1, n2 = the LDS index of the
// lower corner node in each grid.
template <bool BOTH>
__device__ inline void dep_contrib(const uint4 r, const int32_t o[3], int sh, int logG, float qp, float qm,
                                   uint32_t v[16], int& key, int& n1, int& n2) {
    const int cx = (int)(((r.x >> sh) - (uint32_t)o[0]) & (kTile - 1));
    const int cy = (int)(((r.y >> sh) - (uint32_t)o[1]) & (kTile - 1));
    const int cz = (int)(((r.z >> sh) - (uint32_t)o[2]) & (kTile - 1));
    const float fx = frac_of<float>(r.x, logG), fy = frac_of<float>(r.y, logG), fz = frac_of<float>(r.z, logG);
    const float q = (int32_t)r.w > 0 ? qp : qm;
    const bool hx = fx >= 0.5f, hy = fy >= 0.5f, hz = fz >= 0.5f;
    {
        const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {q * (1.0f - fz), q * fz};
#pragma unroll
        for (int c = 0; c < 8; ++c)
            v[c] = __float_as_uint(__builtin_fmaf(wz[c >> 2] * wy[(c >> 1) & 1], wx[c & 1], kRoundMagic));
    }
    if (BOTH) {
        const float vx[2] = {hx ? 1.5f - fx : 0.5f - fx, hx ? fx - 0.5f : 0.5f + fx};
        const float vy[2] = {hy ? 1.5f - fy : 0.5f - fy, hy ? fy - 0.5f : 0.5f + fy};
        const float vz[2] = {q * (hz ? 1.5f - fz : 0.5f - fz), q * (hz ? fz - 0.5f : 0.5f + fz)};
#pragma unroll
        for (int c = 0; c < 8; ++c)
            v[8 + c] = __float_as_uint(__builtin_fmaf(vz[c >> 2] * vy[(c >> 1) & 1], vx[c & 1], kRoundMagic));
    }
    key = ((cz * kTile + cy) * kTile + cx) << 3 | (hz ? 4 : 0) | (hy ? 2 : 0) | (hx ? 1 : 0);
    n1 = (cz * kD1 + cy) * kD1 + cx;
    // the centre of cell m has index m + 1; the lower node is the centre of cell c - 1 + h
    n2 = kDN1 + ((cz + (hz ? 1 : 0)) * kD2 + cy + (hy ? 1 : 0)) * kD2 + cx + (hx ? 1 : 0);
}

// Adds one run's summed contributions to the tile's LDS sums: acc holds the sums of the runs' FMA bit
// patterns, bias their count times kRoundMagicBits (both modulo 2^32; the difference, at most
// kDepPer 2^22 in magnitude, is exact as an int32).
template <bool BOTH>
__device__ inline void dep_flush(unsigned long long* a, const uint32_t acc[16], uint32_t bias, int n1, int n2) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int d = ((c >> 2) * kD1 + ((c >> 1) & 1)) * kD1 + (c & 1);
        atomicAdd(&a[n1 + d], (unsigned long long)(long long)(int32_t)(acc[c] - bias));
    }
    if (BOTH) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int d = ((c >> 2) * kD2 + ((c >> 1) & 1)) * kD2 + (c & 1);
            atomicAdd(&a[n2 + d], (unsigned long long)(long long)(int32_t)(acc[8 + c] - bias));
        }
    }
}

// A thread's contiguous share [p, pe) of a tile's particles (at most kDepPer): the contributions of
// consecutive particles of one (cell, octant) run are summed in registers (int32: at most
// kDepPer x 2^22) and added to LDS once per run, so a halo core costs one LDS atomic per node and
// kDepPer particles, and the lanes of a wavefront, kDepPer particles apart in Morton order, seldom
// meet at one LDS address.  Loads go four at a time; r holds the first four (loaded by the caller).
template <bool BOTH>
__device__ inline void dep_range(const uint4* __restrict__ parts, int32_t p, int32_t pe, uint4 (&r)[4],
                                 const int32_t o[3], int sh, int logG, float qp, float qm, unsigned long long* a) {
    uint32_t acc[16], bias = 0;
    int cur = -1, c1 = 0, c2 = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = 0;
    for (int32_t q0 = p; q0 < pe; q0 += 4) {
        if (q0 > p) {
#pragma unroll
            for (int u = 0; u < 4; ++u) r[u] = parts[q0 + u < pe ? q0 + u : p];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (q0 + u >= pe) break;
            uint32_t v[16];
            int key, n1, n2;
            dep_contrib<BOTH>(r[u], o, sh, logG, qp, qm, v, key, n1, n2);
            if (key != cur) {
                if (cur >= 0) dep_flush<BOTH>(a, acc, bias, c1, c2);
                cur = key;
                c1 = n1;
                c2 = n2;
                bias = kRoundMagicBits;
#pragma unroll
                for (int k = 0; k < 16; ++k) acc[k] = v[k];
            } else {
                bias += kRoundMagicBits;
#pragma unroll
                for (int k = 0; k < 16; ++k) acc[k] += v[k];
            }
        }
    }
    if (cur >= 0) dep_flush<BOTH>(a, acc, bias, c1, c2);
}

// Stores a tile's sums, converted to float, as its record (contiguous); rm = this thread's entries'
// LDS indices.
template <bool BOTH>
__device__ inline void dep_store_tile(const unsigned long long* a, const uint32_t (&rm)[kDepRecIter], double unit,
                                      float* __restrict__ tb) {
#pragma unroll
    for (int i = 0; i < kDepRecIter; ++i) {
        const int j = (int)threadIdx.x + 256 * i;
        if (j < dep_rec<BOTH>()) tb[j] = dep_value(a[rm[i]], unit);
    }
}

// LDS index of entry j of a tile record
template <bool BOTH>
__device__ inline uint32_t dep_rec_lds(int j) {
    return kDepRecMap.lds[(BOTH || j < kDepOwn) ? j : j + kDepOwn];
}

// Work items: block b < ntile takes tile (b & 7) ntile / 8 + b / 8 (contiguous Morton ranges per
// XCD) if it holds at most kInterpChunk particles; block ntile + e takes chunk e of the chunk list.
template <bool BOTH>
__global__ __launch_bounds__(256) void k_dep_tile(const uint4* __restrict__ parts, const int32_t* __restrict__ ts,
                                                  const int2* __restrict__ chunks, const int32_t* __restrict__ nchunk,
                                                  int32_t ntile, int logG, float qp, float qm, double unit,
                                                  float* __restrict__ tbuf, long long* __restrict__ cbuf) {
    constexpr int NN = kDN1 + (BOTH ? kDN2 : 0);
    constexpr int NR = dep_rec<BOTH>();
    __shared__ __align__(16) unsigned long long a[NN];
    const uint32_t b = blockIdx.x;
    int32_t t, pbeg, pend, e = -1;
    if (b < (uint32_t)ntile) {
        t = (int32_t)((b & 7u) * ((uint32_t)ntile >> 3) + (b >> 3));
        pbeg = ts[t];
        pend = ts[t + 1];
        if (pend - pbeg > kInterpChunk) return;  // chunked: the chunk blocks take it
    } else {
        e = (int32_t)(b - (uint32_t)ntile);
        if (e >= *nchunk) return;
        const int2 w = chunks[e];
        t = w.x;
        pbeg = w.y;
        pend = min(ts[t + 1], pbeg + kInterpChunk);
    }
    const int tid = threadIdx.x;
    // this thread's share of the run; its first four particles are loaded before the LDS is cleared
    // and its record entries' LDS indices before the second barrier, so their latency overlaps the
    // clearing and the barriers
    const int32_t per = (pend - pbeg + 255) >> 8;  // <= kInterpChunk / 256
    const int32_t p = pbeg + tid * per, pe = min(pend, p + per);
    uint4 r[4];
    if (p < pe) {
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] = parts[p + u < pe ? p + u : p];
    }
    uint4* a4 = reinterpret_cast<uint4*>(a);
    for (int i = tid; i < NN / 2; i += 256) a4[i] = make_uint4(0u, 0u, 0u, 0u);
    if ((NN & 1) && tid == 0) a[NN - 1] = 0ull;
    int32_t o[3];
    tile_origin((uint32_t)t, o);
    const int sh = 32 - logG;
    __syncthreads();
    dep_range<BOTH>(parts, p, pe, r, o, sh, logG, qp, qm, a);
    uint32_t rm[kDepRecIter];
#pragma unroll
    for (int i = 0; i < kDepRecIter; ++i) {
        const int j = tid + 256 * i;
        rm[i] = j < NR ? dep_rec_lds<BOTH>(j) : 0u;
    }
    __syncthreads();
    if (e >= 0) {  // a chunk: its whole grid, exact
        long long* out = cbuf + (int64_t)e * NN;
        for (int i = tid; i < NN; i += 256) out[i] = (long long)a[i];
        return;
    }
    dep_store_tile<BOTH>(a, rm, unit, tbuf + (int64_t)t * NR);
}

// The chunks of a chunked tile summed in chunk order (exact) into the tile's record: the tile's nc
// chunk blocks split the record into nc slices, block k summing slice k over all chunks.
template <bool BOTH>
__global__ __launch_bounds__(256) void k_dep_merge(const int32_t* __restrict__ ts, const int2* __restrict__ chunks,
                                                   const int32_t* __restrict__ nchunk, double unit,
                                                   const long long* __restrict__ cbuf, float* __restrict__ tbuf) {
    constexpr int NN = kDN1 + (BOTH ? kDN2 : 0);
    constexpr int NR = dep_rec<BOTH>();
    const int32_t e = (int32_t)blockIdx.x;
    if (e >= *nchunk) return;
    const int2 w = chunks[e];
    const int32_t t = w.x, beg = ts[t];
    const int32_t k = (w.y - beg) / kInterpChunk;  // this chunk's index in its tile
    const int32_t nc = (ts[t + 1] - beg + kInterpChunk - 1) / kInterpChunk;
    const int32_t first = e - k;  // the tile's chunks are consecutive entries
    const int32_t per = (NR + nc - 1) / nc;
    const int32_t j0 = k * per, j1 = min(NR, j0 + per);
    float* tb = tbuf + (int64_t)t * NR;
    for (int32_t j = j0 + (int32_t)threadIdx.x; j < j1; j += 256) {
        const uint32_t n = dep_rec_lds<BOTH>(j);
        long long acc = 0;
        for (int32_t q = 0; q < nc; ++q) acc += cbuf[(int64_t)(first + q) * NN + n];
        tb[j] = (float)((double)acc * unit);
    }
}

// Every tile's own blocks (nodes o .. o + 7 per axis of each grid), written as whole rows: own node
// (x, y, z) is the tile's own value plus, for every non-empty set c of its face axes (coordinate 0;
// interlaced 0 or 1), the beyond value of the lower neighbour across those axes (at coordinates
// + 8 there), summed in a fixed order (the eight terms pairwise, absent terms 0).
template <bool BOTH>
__global__ __launch_bounds__(256) void k_dep_assemble(const float* __restrict__ tbuf, int logG,
                                                      float* __restrict__ rho, float* __restrict__ rho2) {
    constexpr int NR = dep_rec<BOTH>();
    constexpr int NQ = BOTH ? 4 : 2;  // own nodes per thread: two per grid
    __shared__ uint32_t nbr[8];      // tile at offset -(c & 1, c >> 1 & 1, c >> 2)
    const uint32_t ntile = gridDim.x;
    const uint32_t b = blockIdx.x;
    const uint32_t t = (b & 7u) * (ntile >> 3) + (b >> 3);
    const int32_t tm = (1 << (logG - 3)) - 1;  // tiles per axis - 1
    const int tid = threadIdx.x;
    int32_t o[3];
    tile_origin(t, o);
    float v[NQ][8];
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q][0] = tbuf[(int64_t)t * NR + tid + 256 * q];  // own blocks: entries 0 .. NQ 256 - 1
    if (tid < 8)
        nbr[tid] = (uint32_t)(spread3_21((uint32_t)(((o[0] >> 3) - (tid & 1)) & tm)) |
                              (spread3_21((uint32_t)(((o[1] >> 3) - ((tid >> 1) & 1)) & tm)) << 1) |
                              (spread3_21((uint32_t)(((o[2] >> 3) - (tid >> 2)) & tm)) << 2));
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int g = q >> 1, k = tid + 256 * (q & 1);
        const int x = k & 7, y = (k >> 3) & 7, z = k >> 6, lim = g ? 2 : 1;
#pragma unroll
        for (int c = 1; c < 8; ++c) {
            v[q][c] = 0.0f;
            if (((c & 1) && x >= lim) || ((c & 2) && y >= lim) || ((c & 4) && z >= lim)) continue;
            const int ax = x + ((c & 1) ? kTile : 0), ay = y + ((c & 2) ? kTile : 0), az = z + ((c & 4) ? kTile : 0);
            const int sl = g ? dep_beyond2<BOTH>() + dep_beyond_slot2(ax, ay, az)
                             : dep_beyond1<BOTH>() + dep_beyond_slot1(ax, ay, az);
            v[q][c] = tbuf[(int64_t)nbr[c] * NR + sl];
        }
    }
    const int64_t G = 1ll << logG;
    const int64_t base = ((int64_t)o[2] * G + o[1]) * G + o[0];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int k = tid + 256 * (q & 1);
        const int32_t off = ((k >> 6) << (2 * logG)) + (((k >> 3) & 7) << logG) + (k & 7);
        const float sum = ((v[q][0] + v[q][1]) + (v[q][2] + v[q][3])) + ((v[q][4] + v[q][5]) + (v[q][6] + v[q][7]));
        ((q >> 1) ? rho2 : rho)[base + off] = sum;
    }
}

// Thread per grid node, nodes enumerated in Morton order so that the eight
// source cells of a wavefront's nodes are contiguous particle runs.
template <typename R>
__global__ void k_deposit(const uint4* __restrict__ parts, const int32_t* __restrict__ cs,
                          const int32_t* __restrict__ slot, const R* __restrict__ corner, int logG, R qp, R qm,
                          R* __restrict__ rho) {
    const int64_t G = 1ll << logG;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= G * G * G) return;
    const int32_t i = (int32_t)compact3_21((uint64_t)m);
    const int32_t j = (int32_t)compact3_21((uint64_t)m >> 1);
    const int32_t k = (int32_t)compact3_21((uint64_t)m >> 2);
    const int32_t msk = (int32_t)(G - 1);
    R acc = 0;
#pragma unroll
    for (int dz = -1; dz <= 0; ++dz) {
#pragma unroll
        for (int dy = -1; dy <= 0; ++dy) {
#pragma unroll
            for (int dx = -1; dx <= 0; ++dx) {
                const uint32_t ci = (uint32_t)((i + dx) & msk), cj = (uint32_t)((j + dy) & msk),
                               ck = (uint32_t)((k + dz) & msk);
                const uint64_t cm = spread3_21(ci) | (spread3_21(cj) << 1) | (spread3_21(ck) << 2);
                const int32_t s = cs[cm], e = cs[cm + 1];
                const int corner_idx = (-dx) + 2 * (-dy) + 4 * (-dz);  // node = cell corner (-dx, -dy, -dz)
                if (e - s > kDenseCell) {
                    acc += corner[8 * (int64_t)slot[cm] + corner_idx];
                    continue;
                }
                for (int32_t p = s; p < e; ++p) {
                    const uint4 r = parts[p];
                    const R fx = frac_of<R>(r.x, logG), fy = frac_of<R>(r.y, logG), fz = frac_of<R>(r.z, logG);
                    const R wx = dx == 0 ? (R)1 - fx : fx;
                    const R wy = dy == 0 ? (R)1 - fy : fy;
                    const R wz = dz == 0 ? (R)1 - fz : fz;
                    const R q = (int32_t)r.w > 0 ? qp : qm;
                    acc += q * (wx * wy * wz);
                }
            }
        }
    }
    rho[((int64_t)k * G + j) * G + i] = acc;
}

// ------------------------------------------------------------------ FFT core
// In-place mixed-radix Stockham over NL = 2^LP / G lines of length G = 2^LG held in LDS
// (element (line l, point p) at l * G + p, or at l + p * NL when LINE_FAST).  Every size is a
// compile-time constant (the kernels are instantiated per mesh size), so the index arithmetic is
// shifts and masks and the stage loop unrolls.  Stages: radix 8 while 8 Ns <= G, then one radix-4
// or radix-2 stage (3 LDS round trips at G = 256 and 512).  A stage of radix R with Ns = product
// of the earlier radices takes
// v_r = x[j + r G/R], multiplies by w^(r k) (k = j mod Ns, w = exp(-+2 pi i / (R Ns)), from the
// table tw[m] = e^{-2 pi i m / G}), does the R-point DFT and writes y_r to
// (j / Ns) R Ns + k + r Ns (autosorted order); reads go to registers, a barrier, then the writes.
template <int LG, int LP, bool LINE_FAST>
__device__ inline int fft_at(int l, int p) {
    constexpr int LNL = LP - LG;  // log2 of the line count (2^LP points per workgroup)
    return LINE_FAST ? l + (p << LNL) : (l << LG) + p;
}

template <typename R>
__device__ inline Cplx<R> cmul(Cplx<R> a, Cplx<R> w) {
#pragma clang fp contract(off)
    return {a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
}

// In-place DFT of RAD = 2, 4 or 8 points (forward e^{-2 pi i nk/RAD}; inverse the conjugate).  The
// 8-point form splits n = n1 + 4 n2, k = 2 k1 + k2: X[2 k1] = DFT4(v[n1] + v[n1 + 4]),
// X[2 k1 + 1] = DFT4((v[n1] - v[n1 + 4]) W8^n1).
template <typename R>
__device__ inline void dft4(const Cplx<R>* e, Cplx<R>* y, bool inv) {
#pragma clang fp contract(off)
    const Cplx<R> a0 = {e[0].x + e[2].x, e[0].y + e[2].y}, a1 = {e[0].x - e[2].x, e[0].y - e[2].y};
    const Cplx<R> a2 = {e[1].x + e[3].x, e[1].y + e[3].y}, a3 = {e[1].x - e[3].x, e[1].y - e[3].y};
    // forward: y1 = a1 - i a3, y3 = a1 + i a3 (inverse: the opposite signs)
    const Cplx<R> ia3 = inv ? Cplx<R>{-a3.y, a3.x} : Cplx<R>{a3.y, -a3.x};  // (-+i) a3
    y[0] = {a0.x + a2.x, a0.y + a2.y};
    y[2] = {a0.x - a2.x, a0.y - a2.y};
    y[1] = {a1.x + ia3.x, a1.y + ia3.y};
    y[3] = {a1.x - ia3.x, a1.y - ia3.y};
}

template <typename R, int RAD>
__device__ inline void dft_small(const Cplx<R>* v, Cplx<R>* y, bool inv) {
#pragma clang fp contract(off)
    if constexpr (RAD == 2) {
        y[0] = {v[0].x + v[1].x, v[0].y + v[1].y};
        y[1] = {v[0].x - v[1].x, v[0].y - v[1].y};
    } else if constexpr (RAD == 4) {
        dft4(v, y, inv);
    } else {
        const R s = (R)0.70710678118654752440;  // sqrt(1/2)
        Cplx<R> c[4], d[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            c[i] = {v[i].x + v[i + 4].x, v[i].y + v[i + 4].y};
            d[i] = {v[i].x - v[i + 4].x, v[i].y - v[i + 4].y};
        }
        const R x1 = d[1].x, y1 = d[1].y, x2 = d[2].x, y2 = d[2].y, x3 = d[3].x, y3 = d[3].y;
        if (inv) {  // W8^-1 = (s, s), W8^-2 = i, W8^-3 = (-s, s)
            d[1] = {s * (x1 - y1), s * (x1 + y1)};
            d[2] = {-y2, x2};
            d[3] = {-s * (x3 + y3), s * (x3 - y3)};
        } else {    // W8 = (s, -s), W8^2 = -i, W8^3 = (-s, -s)
            d[1] = {s * (x1 + y1), s * (y1 - x1)};
            d[2] = {y2, -x2};
            d[3] = {s * (y3 - x3), -s * (x3 + y3)};
        }
        Cplx<R> ce[4], de[4];
        dft4(c, ce, inv);
        dft4(d, de, inv);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            y[2 * i] = ce[i];
            y[2 * i + 1] = de[i];
        }
    }
}

// One Stockham stage of radix 2^LR after LS bits (Ns = 2^LS): butterfly j reads points
// j + r G / RAD, multiplies point r by W_{RAD Ns}^{r k} (k = j mod Ns, tw[m] = e^{-2 pi i m / G}),
// and writes output r to ((j / Ns) RAD Ns + k) + r Ns; all reads, a barrier, all writes.
template <typename R, int LG, int LP, bool LINE_FAST, int LR, int LS>
__device__ inline void fft_stage(Cplx<R>* buf, const Cplx<R>* tw, bool inv) {
#pragma clang fp contract(off)
    constexpr int RAD = 1 << LR, G = 1 << LG, LNL = LP - LG, Q = G / RAD, LQ = LG - LR, Ns = 1 << LS;
    constexpr int NB = (1 << LP) / RAD / 256;  // butterflies per thread
    static_assert(NB >= 1, "a workgroup holds at least 256 butterflies");
    const int tid = threadIdx.x;
    Cplx<R> y[NB][RAD];
    int ll[NB], oo[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const int q = tid + u * 256;
        const int l = LINE_FAST ? (q & ((1 << LNL) - 1)) : (q >> LQ);
        const int j = LINE_FAST ? (q >> LNL) : (q & (Q - 1));
        const int k = j & (Ns - 1);
        Cplx<R> v[RAD];
#pragma unroll
        for (int r = 0; r < RAD; ++r) v[r] = buf[fft_at<LG, LP, LINE_FAST>(l, j + r * Q)];
        if constexpr (LS > 0) {  // the first stage's twiddles are all 1
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                Cplx<R> w = tw[(r * k) << (LG - LS - LR)];
                if (inv) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small<R, RAD>(v, y[u], inv);
        ll[u] = l;
        oo[u] = ((j >> LS) << (LS + LR)) + k;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int r = 0; r < RAD; ++r) buf[fft_at<LG, LP, LINE_FAST>(ll[u], oo[u] + r * Ns)] = y[u][r];
    __syncthreads();
}

// Radix-8 stages while three or more bits remain, then one radix-4 or radix-2 stage
// (G = 512: 8 8 8; 256: 8 8 4; 128: 8 8 2; 64: 8 8; 32: 8 4; 16: 8 2).
template <typename R, int LG, int LP, bool LINE_FAST, int LS>
__device__ inline void fft_stages(Cplx<R>* buf, const Cplx<R>* tw, bool inv) {
    if constexpr (LG - LS >= 3) {
        fft_stage<R, LG, LP, LINE_FAST, 3, LS>(buf, tw, inv);
        fft_stages<R, LG, LP, LINE_FAST, LS + 3>(buf, tw, inv);
    } else if constexpr (LG - LS == 2) {
        fft_stage<R, LG, LP, LINE_FAST, 2, LS>(buf, tw, inv);
    } else if constexpr (LG - LS == 1) {
        fft_stage<R, LG, LP, LINE_FAST, 1, LS>(buf, tw, inv);
    }
}

// 2^LP / G lines of G points in LDS (LINE_FAST: point p of line l at l + p 2^(LP-LG), else l G + p),
// transformed in place; no FMA contraction, so every tile size rounds each butterfly identically.
template <typename R, int LG, int LP, bool LINE_FAST>
__device__ inline void lds_fft(Cplx<R>* buf, const Cplx<R>* __restrict__ tw, bool inv) {
    static_assert(LP >= 11 && LP >= LG, "one radix-8 butterfly per thread and one line per block");
    fft_stages<R, LG, LP, LINE_FAST, 0>(buf, tw, inv);
}

// X forward (R2C): pairs of real lines (y even, y odd) packed as one complex line.
template <typename R, int LG, int LP>
__global__ __launch_bounds__(256) void k_fft_x_r2c(const R* __restrict__ in, Cplx<R>* __restrict__ out,
                                                   const Cplx<R>* __restrict__ tw) {
#pragma clang fp contract(off)
    __shared__ Cplx<R> buf[1 << LP];
    __shared__ Cplx<R> stw[1 << LG];  // the twiddles, staged once (ordered by the barrier after the load)
    for (int i = threadIdx.x; i < (1 << LG); i += 256) stw[i] = tw[i];
    constexpr int G = 1 << LG, H = G / 2 + 1;
    constexpr int npair = (1 << LP) / G;                    // complex lines per block
    const int64_t line0 = (int64_t)blockIdx.x * npair * 2;  // first real line
    const R* src = in + line0 * G;
    // G = 16: the grid's G^2 = 256 real rows fill half of one block; the rest is zero and not stored
    constexpr int rows = 2 * npair < G * G ? 2 * npair : G * G;
    // 16-byte loads: VW consecutive x of one real line (G >= 16, so a line holds whole vectors)
    constexpr int VW = 16 / (int)sizeof(R);
    for (int e4 = threadIdx.x; e4 < (2 << LP) / VW; e4 += 256) {
        const int e = e4 * VW, line = e >> LG, x = e & (G - 1);
        R v[VW];
        if (line < rows) {
            const uint4 w = *(const uint4*)(src + e);
            __builtin_memcpy(v, &w, 16);
        } else {
#pragma unroll
            for (int i = 0; i < VW; ++i) v[i] = (R)0;
        }
        Cplx<R>* b = &buf[(line >> 1) * G + x];
#pragma unroll
        for (int i = 0; i < VW; ++i) {
            if (line & 1) b[i].y = v[i]; else b[i].x = v[i];
        }
    }
    __syncthreads();
    lds_fft<R, LG, LP, false>(buf, stw, false);
    Cplx<R>* dst = out + line0 * H;
    for (int e = threadIdx.x; e < rows * H; e += 256) {
        const int row = e / H, k = e - row * H;
        const int p = row >> 1;
        const Cplx<R> z = buf[p * G + k];
        const Cplx<R> zc = buf[p * G + ((G - k) & (G - 1))];
        Cplx<R> o;
        if ((row & 1) == 0) {  // A = (Z[k] + conj Z[G-k]) / 2
            o = {(R)0.5 * (z.x + zc.x), (R)0.5 * (z.y - zc.y)};
        } else {               // B = (Z[k] - conj Z[G-k]) / (2i)
            o = {(R)0.5 * (z.y + zc.y), (R)0.5 * (zc.x - z.x)};
        }
        dst[e] = o;
    }
}

// X inverse (C2R): two half-spectrum rows A, B -> Z = A + i B, inverse FFT, re / im = the two real rows.
template <typename R, int LG, int LP>
__global__ __launch_bounds__(256) void k_fft_x_c2r(const Cplx<R>* __restrict__ in, R* __restrict__ out,
                                                   const Cplx<R>* __restrict__ tw) {
    __shared__ Cplx<R> buf[1 << LP];
    __shared__ Cplx<R> stw[1 << LG];  // the twiddles, staged once (ordered by the barrier after the load)
    for (int i = threadIdx.x; i < (1 << LG); i += 256) stw[i] = tw[i];
    constexpr int G = 1 << LG, H = G / 2 + 1;
    constexpr int npair = (1 << LP) / G;
    const int64_t line0 = (int64_t)blockIdx.x * npair * 2;
    const Cplx<R>* src = in + line0 * H;
    constexpr int rows = 2 * npair < G * G ? 2 * npair : G * G;  // G = 16: half a block
    // each half-spectrum entry read once: it fills point k and, conjugated, point G - k
    for (int e = threadIdx.x; e < npair * H; e += 256) {
        const int p = e / H, k = e - p * H;
        Cplx<R> a = {0, 0}, b = {0, 0};
        if (2 * p < rows) {
            const Cplx<R>* ra = src + (int64_t)(2 * p) * H;
            a = ra[k];
            b = ra[H + k];
        }
        if (k == 0 || k == G / 2) {
            a.y = 0;
            b.y = 0;
        }
        buf[p * G + k] = {a.x - b.y, a.y + b.x};
        if (k > 0 && k < G / 2) buf[p * G + G - k] = {a.x + b.y, -a.y + b.x};  // conj A + i conj B
    }
    __syncthreads();
    lds_fft<R, LG, LP, false>(buf, stw, true);
    R* dst = out + line0 * G;
    constexpr int VW = 16 / (int)sizeof(R);  // 16-byte stores of VW consecutive x
    for (int e4 = threadIdx.x; e4 < rows * G / VW; e4 += 256) {
        const int e = e4 * VW, line = e >> LG, x = e & (G - 1);
        const Cplx<R>* z = &buf[(line >> 1) * G + x];
        R v[VW];
#pragma unroll
        for (int i = 0; i < VW; ++i) v[i] = (line & 1) ? z[i].y : z[i].x;
        uint4 w;
        __builtin_memcpy(&w, v, 16);
        *(uint4*)(dst + e) = w;
    }
}

// Block -> (plane, kx block) for the y and z passes.  A workgroup reads KB consecutive kx of every
// row, i.e. KB * sizeof(complex) bytes per row (64 B at G = 512 in FP32: half a 128-byte line),
// and the other half of each line belongs to the next kx block.  Blocks are dispatched round-robin
// over the 8 XCDs, so consecutive logical blocks are dealt to one XCD in sequence: the kx block
// that reads the other half of the lines runs next on the same XCD and finds them in its L2.
__device__ inline int fft_logical_block(int nblocks) {
    const int b = (int)blockIdx.x;
    return (nblocks & 7) ? b : (b & 7) * (nblocks >> 3) + (b >> 3);
}

// Y pass (complex, forward or inverse): block = (z, block of KB consecutive kx).
template <typename R, int LG, int LP>
__global__ __launch_bounds__(256) void k_fft_y(Cplx<R>* __restrict__ spec, const Cplx<R>* __restrict__ tw, int inv) {
    __shared__ Cplx<R> buf[1 << LP];
    __shared__ Cplx<R> stw[1 << LG];  // the twiddles, staged once (ordered by the barrier after the load)
    for (int i = threadIdx.x; i < (1 << LG); i += 256) stw[i] = tw[i];
    constexpr int G = 1 << LG, H = G / 2 + 1;
    constexpr int KB = (1 << LP) / G, LKB = LP - LG;
    constexpr int nkb = (H + KB - 1) / KB;
    const int lb = fft_logical_block(G * nkb);
    const int z = lb / nkb, kx0 = (lb % nkb) * KB;
    Cplx<R>* base = spec + (int64_t)z * G * H;
    for (int e = threadIdx.x; e < (1 << LP); e += 256) {
        const int y = e >> LKB, jb = e & (KB - 1);
        const int kx = kx0 + jb;
        buf[e] = (kx < H) ? base[(int64_t)y * H + kx] : Cplx<R>{0, 0};
    }
    __syncthreads();
    lds_fft<R, LG, LP, true>(buf, stw, inv != 0);
    for (int e = threadIdx.x; e < (1 << LP); e += 256) {
        const int y = e >> LKB, jb = e & (KB - 1);
        const int kx = kx0 + jb;
        if (kx < H) base[(int64_t)y * H + kx] = buf[e];
    }
}

// Z pass with the Green function: forward z FFT, multiply, inverse z FFT.
// green(kx, ky, kz) = norm * T[kx] T[ky] T[kz] / (K2[kx] + K2[ky] + K2[kz]); zero for DC and
// for max(|kx|,|ky|,|kz|) < k_min (apply_green_kernel, cufft_wrapper.cu:160-207).
template <typename R, int LG, int LP>
__global__ __launch_bounds__(256) void k_fft_z_green(Cplx<R>* __restrict__ spec, const Cplx<R>* __restrict__ tw,
                                                     const R* __restrict__ gtab, R norm, int k_min, int matched) {
#pragma clang fp contract(off)
    __shared__ Cplx<R> buf[1 << LP];
    __shared__ Cplx<R> stw[1 << LG];  // the twiddles, staged once (ordered by the barrier after the load)
    for (int i = threadIdx.x; i < (1 << LG); i += 256) stw[i] = tw[i];
    constexpr int G = 1 << LG, H = G / 2 + 1;
    constexpr int KB = (1 << LP) / G, LKB = LP - LG;
    constexpr int nkb = (H + KB - 1) / KB;
    const int lb = fft_logical_block(G * nkb);
    const int y = lb / nkb, kx0 = (lb % nkb) * KB;
    constexpr int64_t zs = (int64_t)G * H;
    Cplx<R>* base = spec + (int64_t)y * H;
    for (int e = threadIdx.x; e < (1 << LP); e += 256) {
        const int z = e >> LKB, jb = e & (KB - 1);
        const int kx = kx0 + jb;
        buf[e] = (kx < H) ? base[z * zs + kx] : Cplx<R>{0, 0};
    }
    __syncthreads();
    lds_fft<R, LG, LP, true>(buf, stw, false);
    const R* T = gtab;
    const R* K2 = gtab + G;
    const R* Ks = gtab + 2 * G;  // signed k_i
    const R* D = gtab + 3 * G;   // transfer function of the gradient stencil along one axis
    const int ky = y;
    const int aky = ky <= G / 2 ? ky : G - ky;
    const R Ty = T[ky], K2y = K2[ky], Ksy = Ks[ky], Dy = D[ky];  // the block's row: one ky
    // with KB <= 256 every element of a thread has the same kx (256 is a multiple of KB)
    const int kx_t = kx0 + (threadIdx.x & (KB - 1));
    const int kx_c = kx_t < H ? kx_t : 0;
    const R Tx_t = T[kx_c], K2x_t = K2[kx_c], Ksx_t = Ks[kx_c], Dx_t = D[kx_c];
    for (int e = threadIdx.x; e < (1 << LP); e += 256) {
        const int kz = e >> LKB, jb = e & (KB - 1);
        const int kx = kx0 + jb;
        const int akz = kz <= G / 2 ? kz : G - kz;
        const int akx = kx;
        const int kmax = max(akx, max(aky, akz));
        R f = 0;
        if (kx < H && !(akx == 0 && aky == 0 && akz == 0) && kmax >= k_min) {
            const bool same = KB <= 256;
            const R Tx = same ? Tx_t : T[kx], K2x = same ? K2x_t : K2[kx];
            const R k2 = K2x + K2y + K2[kz];
            f = norm * Tx * Ty * T[kz];
            if (matched) {
                // least-squares match of the stencil gradient D to i k: f *= (k . D) / |D|^2, with the
                // denominator floored at |k|^2 / 4 where D vanishes (Nyquist planes); one division
                const R Ksx = same ? Ksx_t : Ks[kx], Dx = same ? Dx_t : D[kx];
                const R dd = Dx * Dx + Dy * Dy + D[kz] * D[kz];
                const R fl = (R)0.25 * k2;
                f = f * (Ksx * Dx + Ksy * Dy + Ks[kz] * D[kz]) / (k2 * (dd > fl ? dd : fl));
            } else {
                f = f / k2;
            }
        }
        const Cplx<R> v = buf[e];
        buf[e] = {v.x * f, v.y * f};
    }
    __syncthreads();
    lds_fft<R, LG, LP, true>(buf, stw, true);
    for (int e = threadIdx.x; e < (1 << LP); e += 256) {
        const int z = e >> LKB, jb = e & (KB - 1);
        const int kx = kx0 + jb;
        if (kx < H) base[z * zs + kx] = buf[e];
    }
}

// Inverse z pass alone (the z pass of the solver without the forward transform and Green multiply).
template <typename R, int LG, int LP>
__global__ __launch_bounds__(256) void k_fft_z_inv(Cplx<R>* __restrict__ spec, const Cplx<R>* __restrict__ tw) {
    __shared__ Cplx<R> buf[1 << LP];
    __shared__ Cplx<R> stw[1 << LG];  // the twiddles, staged once (ordered by the barrier after the load)
    for (int i = threadIdx.x; i < (1 << LG); i += 256) stw[i] = tw[i];
    constexpr int G = 1 << LG, H = G / 2 + 1;
    constexpr int KB = (1 << LP) / G, LKB = LP - LG;
    constexpr int nkb = (H + KB - 1) / KB;
    const int lb = fft_logical_block(G * nkb);
    const int y = lb / nkb, kx0 = (lb % nkb) * KB;
    constexpr int64_t zs = (int64_t)G * H;
    Cplx<R>* base = spec + (int64_t)y * H;
    for (int e = threadIdx.x; e < (1 << LP); e += 256) {
        const int z = e >> LKB, jb = e & (KB - 1);
        const int kx = kx0 + jb;
        buf[e] = (kx < H) ? base[z * zs + kx] : Cplx<R>{0, 0};
    }
    __syncthreads();
    lds_fft<R, LG, LP, true>(buf, stw, true);
    for (int e = threadIdx.x; e < (1 << LP); e += 256) {
        const int z = e >> LKB, jb = e & (KB - 1);
        const int kx = kx0 + jb;
        if (kx < H) base[z * zs + kx] = buf[e];
    }
}

// One Poisson solve's device buffers and constants: the context's, or the scratch of the
// allocation-free entry point.
struct FftPlan {
    void* spec;           // half spectrum, complex [G][G][G/2+1]
    const void* twiddle;  // complex [G]
    const void* gtab;     // real [4 G]: T, k^2, k, stencil transfer D per axis index
    double green_norm;
    int logG, k_min, tile, precision;
};

static FftPlan plan_of(const Ctx& c) {
    return FftPlan{c.spec, c.twiddle, c.gtab, c.green_norm, c.logG, c.p.k_min, c.p.fft_tile, c.p.precision};
}

template <typename R, int LG, int LP>
static hipError_t inverse_fft_lg(const FftPlan& f, R* out, hipStream_t s) {
    constexpr int G = 1 << LG, H = G / 2 + 1, NPT = 1 << LP;
    constexpr int KB = NPT / G, nkb = (H + KB - 1) / KB;
    constexpr unsigned nx = (unsigned)(G * G > 2 * (NPT / G) ? (int64_t)G * G / (2 * (NPT / G)) : 1);
    const Cplx<R>* tw = (const Cplx<R>*)f.twiddle;
    Cplx<R>* spec = (Cplx<R>*)f.spec;
    k_fft_z_inv<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw);
    k_fft_y<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw, 1);
    k_fft_x_c2r<R, LG, LP><<<dim3(nx), dim3(256), 0, s>>>(spec, out, tw);
    return hipGetLastError();
}

template <typename R, int LG, int LP>
static hipError_t fft_solve_lg(const FftPlan& f, const R* rho, R* phi, hipStream_t s, int matched) {
    constexpr int G = 1 << LG, H = G / 2 + 1, NPT = 1 << LP;
    constexpr int KB = NPT / G, nkb = (H + KB - 1) / KB;
    constexpr unsigned nx = (unsigned)(G * G > 2 * (NPT / G) ? (int64_t)G * G / (2 * (NPT / G)) : 1);
    const Cplx<R>* tw = (const Cplx<R>*)f.twiddle;
    Cplx<R>* spec = (Cplx<R>*)f.spec;
    k_fft_x_r2c<R, LG, LP><<<dim3(nx), dim3(256), 0, s>>>(rho, spec, tw);
    k_fft_y<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw, 0);
    k_fft_z_green<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw, (const R*)f.gtab,
                                                                         (R)f.green_norm, f.k_min, matched);
    k_fft_y<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw, 1);
    k_fft_x_c2r<R, LG, LP><<<dim3(nx), dim3(256), 0, s>>>(spec, phi, tw);
    return hipGetLastError();
}

// the mesh sizes the context accepts (G = 16 .. 512) and FFT tiles (2^11, 2^12 points; 2^13 for FP32)
#define JTP_FFT_DISPATCH_LG(CALL, LP)          \
    switch (f.logG) {                          \
        case 4: return CALL(4, LP);            \
        case 5: return CALL(5, LP);            \
        case 6: return CALL(6, LP);            \
        case 7: return CALL(7, LP);            \
        case 8: return CALL(8, LP);            \
        case 9: return CALL(9, LP);            \
        default: return hipErrorInvalidValue;  \
    }

template <typename R>
static hipError_t inverse_fft_t(const FftPlan& f, R* out, hipStream_t s) {
#define JTP_INV(LG, LP) inverse_fft_lg<R, LG, LP>(f, out, s)
    if constexpr (sizeof(R) == 4) {
        if (f.tile == 13) JTP_FFT_DISPATCH_LG(JTP_INV, 13)
    }
    if (f.tile == 12) JTP_FFT_DISPATCH_LG(JTP_INV, 12)
    JTP_FFT_DISPATCH_LG(JTP_INV, 11)
#undef JTP_INV
}

hipError_t pm_inverse_fft(Ctx& c, void* out, hipStream_t s) {
    const FftPlan f = plan_of(c);
    if (c.p.precision == JTP_FP64) return inverse_fft_t<double>(f, (double*)out, s);
    return inverse_fft_t<float>(f, (float*)out, s);
}

template <typename R>
static hipError_t fft_solve_t(const FftPlan& f, const R* rho, R* phi, hipStream_t s, int matched) {
#define JTP_SOLVE(LG, LP) fft_solve_lg<R, LG, LP>(f, rho, phi, s, matched)
    if constexpr (sizeof(R) == 4) {
        if (f.tile == 13) JTP_FFT_DISPATCH_LG(JTP_SOLVE, 13)
    }
    if (f.tile == 12) JTP_FFT_DISPATCH_LG(JTP_SOLVE, 12)
    JTP_FFT_DISPATCH_LG(JTP_SOLVE, 11)
#undef JTP_SOLVE
}

static hipError_t plan_solve(const FftPlan& f, const void* rho, void* phi, hipStream_t s, int matched) {
    if (f.precision == JTP_FP64) return fft_solve_t<double>(f, (const double*)rho, (double*)phi, s, matched);
    return fft_solve_t<float>(f, (const float*)rho, (float*)phi, s, matched);
}

hipError_t pm_fft_solve(Ctx& c, const void* rho, void* phi, hipStream_t s, int matched) {
    if (matched < 0) matched = c.p.deconvolve;  // the force's Green function
    return plan_solve(plan_of(c), rho, phi, s, matched);
}

// deposit chunks of tiles holding more than kInterpChunk particles: sum over those tiles of
// ceil(np / kInterpChunk) <= 2 n / kInterpChunk
int64_t dep_chunk_cap(int64_t n) { return 2 * n / kInterpChunk + 2; }
int64_t dep_chunk_nodes() { return kDN1 + kDN2; }
int64_t dep_tile_record() { return dep_rec<true>(); }

static hipError_t deposit_fp32(Ctx& c, int64_t n, hipStream_t s) {
    const int64_t G3 = c.G3();
    const int32_t ntile = (int32_t)(G3 >> 9);  // 8^3-cell tiles (G >= 16: a multiple of 8)
    int32_t* count = c.dev_counters + kCtrInterp;  // light tiles, interpolation chunks, deposit chunks
    hipError_t e = launch_fill_i32(count, 3, 0, s);
    if (e != hipSuccess) return e;
    int32_t* light = (int32_t*)c.interp_list;
    int2* heavy = c.interp_list + (ntile + 1) / 2;
    k_tile_mark<<<dim3((unsigned)((ntile + 255) / 256)), dim3(256), 0, s>>>(c.tile_start, ntile, light, heavy,
                                                                           c.dep_chunks, count);
    // charges scaled by 2^k, |q| 2^k <= 2^kFixBits (exact: a power of two); the grids' unit is 2^-k
    int ex = 0;
    (void)std::frexp(c.m_plus > c.m_minus ? c.m_plus : c.m_minus, &ex);  // max |q| <= 2^ex
    const int k = kFixBits - ex;
    const float qp = (float)std::ldexp(c.m_plus, k), qm = (float)-std::ldexp(c.m_minus, k);
    const double unit = std::ldexp(1.0, -k);
    const unsigned nd = (unsigned)dep_chunk_cap(n);
    float* r1 = (float*)c.rho;
    float* r2 = (float*)c.rho2;
    if (c.p.interlace) {
        k_dep_tile<true><<<dim3((unsigned)ntile + nd), dim3(256), 0, s>>>(
            c.parts, c.tile_start, c.dep_chunks, count + 2, ntile, c.logG, qp, qm, unit, c.dep_tbuf, c.dep_cbuf);
        k_dep_merge<true><<<dim3(nd), dim3(256), 0, s>>>(c.tile_start, c.dep_chunks, count + 2, unit, c.dep_cbuf,
                                                         c.dep_tbuf);
        k_dep_assemble<true><<<dim3((unsigned)ntile), dim3(256), 0, s>>>(c.dep_tbuf, c.logG, r1, r2);
    } else {
        k_dep_tile<false><<<dim3((unsigned)ntile + nd), dim3(256), 0, s>>>(
            c.parts, c.tile_start, c.dep_chunks, count + 2, ntile, c.logG, qp, qm, unit, c.dep_tbuf, c.dep_cbuf);
        k_dep_merge<false><<<dim3(nd), dim3(256), 0, s>>>(c.tile_start, c.dep_chunks, count + 2, unit, c.dep_cbuf,
                                                          c.dep_tbuf);
        k_dep_assemble<false><<<dim3((unsigned)ntile), dim3(256), 0, s>>>(c.dep_tbuf, c.logG, r1, nullptr);
    }
    return hipGetLastError();
}

hipError_t pm_deposit(Ctx& c, int64_t n, hipStream_t s) {
    if (c.p.precision == JTP_FP32) return deposit_fp32(c, n, s);
    const int64_t G3 = c.G3();
    hipError_t e = launch_cell_start(c.keys0, n, 63 - 3 * c.logG, G3, c.cell_start, s);
    if (e != hipSuccess) return e;
    int32_t* count = c.dev_counters + kCtrDense;
    int32_t* list = c.dense_list;  // <= n / kDenseCell cells
    e = launch_fill_i32(count, 1, 0, s);
    if (e != hipSuccess) return e;
    const unsigned nb = (unsigned)((G3 + kBlock - 1) / kBlock);
    k_dense_mark<<<dim3(nb), dim3(kBlock), 0, s>>>(c.cell_start, G3, c.dense_slot, list, count);
    int64_t nd_max = n / kDenseCell + 1;
    unsigned nbd = (unsigned)((nd_max * 64 + kBlock - 1) / kBlock);
    if (nbd > 8192) nbd = 8192;
    const int64_t nd_cap = n / kDenseCell + 1;
    double* c8 = (double*)c.dense_corner;
    double* c27 = c8 + 8 * nd_cap;
    if (c.p.interlace) {
        k_dense_reduce_both<double><<<dim3(nbd), dim3(256), 0, s>>>(
            c.parts, c.cell_start, list, c.dense_slot, count, c.logG, c.m_plus, -c.m_minus, c8, c27);
        k_dense_combine<double><<<dim3(nbd), dim3(256), 0, s>>>(c.cell_start, list, c.dense_slot, count, 8, c8);
        k_dense_combine<double><<<dim3(nbd), dim3(256), 0, s>>>(c.cell_start, list, c.dense_slot, count, 27, c27);
        k_deposit_both<double><<<dim3(nb), dim3(kBlock), 0, s>>>(c.parts, c.cell_start, c.dense_slot, c8, c27,
                                                                 c.logG, c.m_plus, -c.m_minus, (double*)c.rho,
                                                                 (double*)c.rho2);
    } else {
        k_dense_reduce<double><<<dim3(nbd), dim3(256), 0, s>>>(c.parts, c.cell_start, list, c.dense_slot, count,
                                                               c.logG, c.m_plus, -c.m_minus, c8);
        k_dense_combine<double><<<dim3(nbd), dim3(256), 0, s>>>(c.cell_start, list, c.dense_slot, count, 8, c8);
        k_deposit<double><<<dim3(nb), dim3(kBlock), 0, s>>>(c.parts, c.cell_start, c.dense_slot, c8, c.logG,
                                                            c.m_plus, -c.m_minus, (double*)c.rho);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ interpolation
// a_i = -sigma_i sum_corners W * grad(phi)(corner), grad by the 4-point (or 2-point) stencil.
template <typename R, bool FOUR>
__device__ inline R dstencil(const R* __restrict__ phi, int64_t b, int64_t st, int c, int G, R inv) {
    // phi along one axis: b = index with that axis coordinate removed, st = stride, c = coordinate
    const int msk = G - 1;
    const R p1 = phi[b + (int64_t)((c + 1) & msk) * st], m1 = phi[b + (int64_t)((c - 1) & msk) * st];
    if (FOUR) {
        const R p2 = phi[b + (int64_t)((c + 2) & msk) * st], m2 = phi[b + (int64_t)((c - 2) & msk) * st];
        return ((R)8 * (p1 - m1) - (p2 - m2)) * inv;
    }
    return (p1 - m1) * inv;
}

// Gradient grid: grad[node] = (d/dx, d/dy, d/dz, 0) of phi by the stencil, one thread per node
// (x fastest, so the x-neighbours are coalesced and the y / z neighbours come from L2).
template <typename R, bool FOUR>
__global__ void k_grad(const R* __restrict__ phi, int logG, R inv, typename Vec4T<R>::T* __restrict__ grad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int G = 1 << logG;
    const int64_t G2 = (int64_t)G * G;
    if (i >= G2 * G) return;
    const int cx = (int)(i & (G - 1)), cy = (int)((i >> logG) & (G - 1)), cz = (int)(i >> (2 * logG));
    typename Vec4T<R>::T g;
    g.x = dstencil<R, FOUR>(phi, (int64_t)cz * G2 + (int64_t)cy * G, 1, cx, G, inv);
    g.y = dstencil<R, FOUR>(phi, (int64_t)cz * G2 + cx, G, cy, G, inv);
    g.z = dstencil<R, FOUR>(phi, (int64_t)cy * G + cx, G2, cz, G, inv);
    g.w = (R)0;
    grad[i] = g;
}

// CIC interpolation of the gradient grid at one particle (optionally shifted by half a cell).
template <typename R>
__device__ inline void interp_grid(const uint4 r, const typename Vec4T<R>::T* __restrict__ grad, int logG,
                                   uint32_t shift, R a[3]) {
    const int G = 1 << logG;
    const int64_t G2 = (int64_t)G * G;
    const uint32_t ux = r.x + shift, uy = r.y + shift, uz = r.z + shift;
    const int sh = 32 - logG;
    const int ix = (int)(ux >> sh), iy = (int)(uy >> sh), iz = (int)(uz >> sh);
    const R fx = frac_of<R>(ux, logG), fy = frac_of<R>(uy, logG), fz = frac_of<R>(uz, logG);
    const R wx[2] = {(R)1 - fx, fx}, wy[2] = {(R)1 - fy, fy}, wz[2] = {(R)1 - fz, fz};
    R ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz) {
        const int cz = (iz + dz) & (G - 1);
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            const int cy = (iy + dy) & (G - 1);
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int cx = (ix + dx) & (G - 1);
                const R w = wx[dx] * wy[dy] * wz[dz];
                const typename Vec4T<R>::T g = grad[(int64_t)cz * G2 + (int64_t)cy * G + cx];
                ax += w * g.x;
                ay += w * g.y;
                az += w * g.z;
            }
        }
    }
    a[0] = ax;
    a[1] = ay;
    a[2] = az;
}

// With interlacing, the shifted grid is sampled at x + h/2 (its nodes are the cell centres of
// the first grid) and the two accelerations are averaged.
template <typename R>
__global__ void k_interp(const uint4* __restrict__ parts, int64_t n, const typename Vec4T<R>::T* __restrict__ grad,
                         const typename Vec4T<R>::T* __restrict__ grad2, int logG, R* __restrict__ acc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 r = parts[i];
    R a[3];
    interp_grid<R>(r, grad, logG, 0u, a);
    R sg = (int32_t)r.w > 0 ? (R)-1 : (R)1;
    if (grad2) {
        R b[3];
        interp_grid<R>(r, grad2, logG, 1u << (31 - logG), b);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) a[ax] += b[ax];
        sg *= (R)0.5;
    }
    acc[3 * i] = sg * a[0];
    acc[3 * i + 1] = sg * a[1];
    acc[3 * i + 2] = sg * a[2];
}

// FP32: one workgroup per 8^3-cell tile, whose particles are one contiguous Morton run.  The
// potential around the tile is staged in LDS with the stencil's 2-cell halo, the gradient of every
// node the tile's particles reach is computed there once (as k_grad does per node), and each
// particle reads its 8 corner gradients per grid from LDS.  Grid 1: particles of cells 0..7 reach
// nodes 0..8, whose stencils read phi at -2..10; grid 2 (its node j at the centre of cell j - 1):
// nodes 0..9, phi2 at -2..11.  The arithmetic is k_grad's and interp_grid's, term for term.
constexpr int kIp1 = kTile + 5;  // phi nodes -2 .. 10 per axis
constexpr int kIp2 = kTile + 6;  // phi2 nodes -2 .. 11
constexpr int kIg1 = kTile + 1;  // gradient nodes 0 .. 8
constexpr int kIg2 = kTile + 2;  // gradient nodes 0 .. 9

template <bool FOUR>
__device__ inline float lds_stencil(const float* p, int b, int st, float inv) {
    const float p1 = p[b + st], m1 = p[b - st];
    if (FOUR) {
        const float p2 = p[b + 2 * st], m2 = p[b - 2 * st];
        return (8.0f * (p1 - m1) - (p2 - m2)) * inv;
    }
    return (p1 - m1) * inv;
}

// phi at nodes o - 2 .. o + NP - 3 per axis (periodic), staged row by row: each x-row is read as the
// four aligned float4 groups at x = o - 4 .. o + 11 (o is a multiple of 8 and G of 4, so no group
// straddles the periodic seam), four lanes per row, and the NP wanted values of a row are written to
// LDS (x stride 1, y stride NP, z stride NP^2).  Item j = threadIdx.x + NT k is group j & 3 of row
// j >> 2; every load is issued before the first is used, so a thread waits for one round trip.
template <int NP, int NT>
constexpr int stage_per_thread() { return (NP * NP * 4 + NT - 1) / NT; }

template <int NP, int NT>
__device__ inline void stage_load(const float* __restrict__ phi, int32_t ox, int32_t oy, int32_t oz, int logG,
                                  float4 (&v)[stage_per_thread<NP, NT>()]) {
    const int32_t msk = (1 << logG) - 1;
#pragma unroll
    for (int k = 0; k < stage_per_thread<NP, NT>(); ++k) {
        const int j = threadIdx.x + NT * k;
        if (j < NP * NP * 4) {
            const int row = j >> 2, ly = row % NP - 2, lz = row / NP - 2;
            const int64_t rb = (((int64_t)((oz + lz) & msk)) << logG | ((oy + ly) & msk)) << logG;
            v[k] = *(const float4*)(phi + rb + ((ox - 4 + 4 * (j & 3)) & msk));
        }
    }
}

template <int NP, int NT>
__device__ inline void stage_store(const float4 (&v)[stage_per_thread<NP, NT>()], float* p) {
#pragma unroll
    for (int k = 0; k < stage_per_thread<NP, NT>(); ++k) {
        const int j = threadIdx.x + NT * k;
        if (j < NP * NP * 4) {
            float* row = p + (j >> 2) * NP + 2;  // x = 0 of the row
            const int x0 = 4 * (j & 3) - 4;        // x of v[k].x
            const float e[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (x0 + q >= -2 && x0 + q <= NP - 3) row[x0 + q] = e[q];
        }
    }
}

// the stencil gradient at nodes o .. o + NG - 1 per axis from the staged potential p -> g
template <bool FOUR, int NP, int NG, int NT>
__device__ inline void stage_grad(const float* p, float inv, float4* g) {
    for (int i = threadIdx.x; i < NG * NG * NG; i += NT) {
        const int nx = i % NG, ny = (i / NG) % NG, nz = i / (NG * NG);
        const int b = ((nz + 2) * NP + ny + 2) * NP + nx + 2;
        float4 v;
        v.x = lds_stencil<FOUR>(p, b, 1, inv);
        v.y = lds_stencil<FOUR>(p, b, NP, inv);
        v.z = lds_stencil<FOUR>(p, b, NP * NP, inv);
        v.w = 0.0f;
        g[i] = v;
    }
}

// CIC sum of the LDS gradients around local node (ix, iy, iz), in interp_grid's order
template <int NG>
__device__ inline void tile_cic(const float4* g, int ix, int iy, int iz, float fx, float fy, float fz, float a[3]) {
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
    const float4* b = g + (iz * NG + iy) * NG + ix;
    float ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const float w = wx[dx] * wy[dy] * wz[dz];
                const float4 v = b[(dz * NG + dy) * NG + dx];
                ax += w * v.x;
                ay += w * v.y;
                az += w * v.z;
            }
    a[0] = ax;
    a[1] = ay;
    a[2] = az;
}

// Gradient along the axis of stride s summed over the two CIC corners of that axis (weights w0, w1)
// and bilinearly over the other two (strides s1, s2; weights u, v): the corners' stencils combine
// into one filter over the potential's lines k = -2 .. 3 (4-point) or -1 .. 2 (2-point) of the
// bilinear interpolation across the other axes.  Times the stencil's 1 / (12 h) (1 / (2 h)) later.
template <bool FOUR>
__device__ inline float tile_axis_sum(const float* p, int b, int s, int s1, int s2, float w0, float w1, float u0,
                                      float u1, float v0, float v1) {
    const float q00 = u0 * v0, q10 = u1 * v0, q01 = u0 * v1, q11 = u1 * v1;
    auto line = [&](int k) {
        const int o = b + k * s;
        return q00 * p[o] + q10 * p[o + s1] + q01 * p[o + s2] + q11 * p[o + s1 + s2];
    };
    if (FOUR) {
        const float lm2 = line(-2), lm1 = line(-1), l0 = line(0), l1 = line(1), l2 = line(2), l3 = line(3);
        return w0 * (8.0f * (l1 - lm1) - (l2 - lm2)) + w1 * (8.0f * (l2 - l0) - (l3 - lm1));
    }
    const float lm1 = line(-1), l0 = line(0), l1 = line(1), l2 = line(2);
    return w0 * (l1 - lm1) + w1 * (l2 - l0);
}

// CIC sum over the 8 corners around local node (ix, iy, iz) of each corner's stencil gradient of the
// staged potential p (NP nodes per axis, node 0 at local -2), one axis at a time (tile_axis_sum):
// light tiles, where computing every node's gradient first would cost more than it saves.
template <bool FOUR, int NP>
__device__ inline void tile_cic_phi(const float* p, int ix, int iy, int iz, float fx, float fy, float fz, float inv,
                                    float a[3]) {
    const float wx0 = 1.0f - fx, wy0 = 1.0f - fy, wz0 = 1.0f - fz;
    const int b0 = ((iz + 2) * NP + iy + 2) * NP + ix + 2;
    a[0] = inv * tile_axis_sum<FOUR>(p, b0, 1, NP, NP * NP, wx0, fx, wy0, fy, wz0, fz);
    a[1] = inv * tile_axis_sum<FOUR>(p, b0, NP, 1, NP * NP, wy0, fy, wx0, fx, wz0, fz);
    a[2] = inv * tile_axis_sum<FOUR>(p, b0, NP * NP, 1, NP, wz0, fz, wx0, fx, wy0, fy);
}


// List entry of block b, or -1: the entries are dealt to the 8 XCDs in contiguous ranges (block b
// runs on XCD b % 8), so neighbouring tiles share that XCD's L2.  Needs gridDim.x >= cnt + 7.
__device__ inline int32_t xcd_entry(uint32_t b, int32_t cnt) {
    const int32_t per = (cnt + 7) >> 3, k = (int32_t)(b >> 3);
    const int32_t e = (int32_t)(b & 7u) * per + k;
    return (k < per && e < cnt) ? e : -1;
}

// local cell of a particle of the tile with origin o, for grid 1 (0..7) and the half-cell-shifted
// grid 2 (0..8); the particle's cell is in the tile (its key is its position's), and the masks keep
// every LDS index in range whatever the input
__device__ inline int cell1(uint32_t u, int sh, int32_t o) { return (int)(((u >> sh) - (uint32_t)o) & (kTile - 1)); }
__device__ inline int cell2(uint32_t u, int sh, int32_t o, uint32_t msk) {
    return (int)min(((u >> sh) - (uint32_t)o) & msk, (uint32_t)kTile);
}

constexpr int kHeavyThreads = 512;  // 8 waves per workgroup: 24 waves per CU at 3 workgroups

template <bool FOUR, bool BOTH>
__global__ __launch_bounds__(kHeavyThreads) void k_interp_heavy(const uint4* __restrict__ parts, const int32_t* __restrict__ ts,
                                                      const int2* __restrict__ list, const int32_t* __restrict__ count,
                                                      const float* __restrict__ phi, const float* __restrict__ phi2,
                                                      int logG, float inv, float* __restrict__ acc) {
    __shared__ float p1[kIp1 * kIp1 * kIp1];
    __shared__ float p2[BOTH ? kIp2 * kIp2 * kIp2 : 1];
    __shared__ float4 g1[kIg1 * kIg1 * kIg1];
    __shared__ float4 g2[BOTH ? kIg2 * kIg2 * kIg2 : 1];
    const int32_t e = xcd_entry(blockIdx.x, *count);
    if (e < 0) return;  // the same for every thread of the block
    const int2 w = list[e];
    const uint64_t base = (uint64_t)w.x << 9;
    const int32_t pbeg = w.y, pend = min(ts[w.x + 1], pbeg + kInterpChunk);
    const int32_t ox = (int32_t)compact3_21(base), oy = (int32_t)compact3_21(base >> 1),
                  oz = (int32_t)compact3_21(base >> 2);
    {
        float4 v1[stage_per_thread<kIp1, kHeavyThreads>()], v2[stage_per_thread<kIp2, kHeavyThreads>()];
        stage_load<kIp1, kHeavyThreads>(phi, ox, oy, oz, logG, v1);
        if (BOTH) stage_load<kIp2, kHeavyThreads>(phi2, ox, oy, oz, logG, v2);
        stage_store<kIp1, kHeavyThreads>(v1, p1);
        if (BOTH) stage_store<kIp2, kHeavyThreads>(v2, p2);
    }
    __syncthreads();
    stage_grad<FOUR, kIp1, kIg1, kHeavyThreads>(p1, inv, g1);
    if (BOTH) stage_grad<FOUR, kIp2, kIg2, kHeavyThreads>(p2, inv, g2);
    __syncthreads();
    const int sh = 32 - logG;
    const uint32_t msk = (1u << logG) - 1u, half = 1u << (31 - logG);
    for (int32_t i = pbeg + (int32_t)threadIdx.x; i < pend; i += kHeavyThreads) {
        const uint4 r = parts[i];
        float a[3];
        tile_cic<kIg1>(g1, cell1(r.x, sh, ox), cell1(r.y, sh, oy), cell1(r.z, sh, oz), frac_of<float>(r.x, logG),
                       frac_of<float>(r.y, logG), frac_of<float>(r.z, logG), a);
        float sg = (int32_t)r.w > 0 ? -1.0f : 1.0f;
        if (BOTH) {
            const uint32_t ux = r.x + half, uy = r.y + half, uz = r.z + half;
            float b[3];
            tile_cic<kIg2>(g2, cell2(ux, sh, ox, msk), cell2(uy, sh, oy, msk), cell2(uz, sh, oz, msk),
                           frac_of<float>(ux, logG), frac_of<float>(uy, logG), frac_of<float>(uz, logG), b);
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) a[ax] += b[ax];
            sg *= 0.5f;
        }
        acc[3 * (int64_t)i] = sg * a[0];
        acc[3 * (int64_t)i + 1] = sg * a[1];
        acc[3 * (int64_t)i + 2] = sg * a[2];
    }
}

// Light tiles: only the potential is staged (~20 KB of LDS, so up to 8 workgroups per CU hide the
// loads' latency), the particle is loaded together with the halo, and each corner's gradient is
// taken from LDS directly.
template <bool FOUR, bool BOTH>
__global__ __launch_bounds__(256) void k_interp_light(const uint4* __restrict__ parts, const int32_t* __restrict__ ts,
                                                      const int32_t* __restrict__ list, const int32_t* __restrict__ count,
                                                      const float* __restrict__ phi, const float* __restrict__ phi2,
                                                      int logG, float inv, float* __restrict__ acc) {
    __shared__ float p1[kIp1 * kIp1 * kIp1];
    __shared__ float p2[BOTH ? kIp2 * kIp2 * kIp2 : 1];
    const int32_t e = xcd_entry(blockIdx.x, *count);
    if (e < 0) return;  // the same for every thread of the block
    const int32_t tl = list[e];
    const uint64_t base = (uint64_t)tl << 9;
    const int32_t pbeg = ts[tl], np = ts[tl + 1] - pbeg;  // 1 .. kLightMax
    const int32_t ox = (int32_t)compact3_21(base), oy = (int32_t)compact3_21(base >> 1),
                  oz = (int32_t)compact3_21(base >> 2);
    constexpr int kPer = kLightMax / 256;  // particles per thread
    static_assert(kLightMax % 256 == 0, "light tiles hold whole multiples of the block");
    uint4 rs[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u)
        if ((int32_t)threadIdx.x + 256 * u < np) rs[u] = parts[pbeg + (int32_t)threadIdx.x + 256 * u];
    {
        float4 v1[stage_per_thread<kIp1, 256>()], v2[stage_per_thread<kIp2, 256>()];
        stage_load<kIp1, 256>(phi, ox, oy, oz, logG, v1);
        if (BOTH) stage_load<kIp2, 256>(phi2, ox, oy, oz, logG, v2);
        stage_store<kIp1, 256>(v1, p1);
        if (BOTH) stage_store<kIp2, 256>(v2, p2);
    }
    __syncthreads();  // no barrier follows
    const int sh = 32 - logG;
    const uint32_t msk = (1u << logG) - 1u, half = 1u << (31 - logG);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        if ((int32_t)threadIdx.x + 256 * u >= np) continue;
        const int32_t i = pbeg + (int32_t)threadIdx.x + 256 * u;
        const uint4 r = rs[u];
        float a[3];
        tile_cic_phi<FOUR, kIp1>(p1, cell1(r.x, sh, ox), cell1(r.y, sh, oy), cell1(r.z, sh, oz),
                                 frac_of<float>(r.x, logG), frac_of<float>(r.y, logG), frac_of<float>(r.z, logG),
                                 inv, a);
        float sg = (int32_t)r.w > 0 ? -1.0f : 1.0f;
        if (BOTH) {
            const uint32_t ux = r.x + half, uy = r.y + half, uz = r.z + half;
            float b[3];
            tile_cic_phi<FOUR, kIp2>(p2, cell2(ux, sh, ox, msk), cell2(uy, sh, oy, msk), cell2(uz, sh, oz, msk),
                                     frac_of<float>(ux, logG), frac_of<float>(uy, logG), frac_of<float>(uz, logG),
                                     inv, b);
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) a[ax] += b[ax];
            sg *= 0.5f;
        }
        acc[3 * (int64_t)i] = sg * a[0];
        acc[3 * (int64_t)i + 1] = sg * a[1];
        acc[3 * (int64_t)i + 2] = sg * a[2];
    }
}

template <typename R>
static hipError_t interp_t(Ctx& c, int64_t n, hipStream_t s) {
    using V = typename Vec4T<R>::T;
    const double h = c.p.box_size / (double)c.G();
    const bool four = c.p.gradient == JTP_GRAD_4PT;
    const R inv = (R)(four ? 1.0 / (12.0 * h) : 1.0 / (2.0 * h));
    if constexpr (sizeof(R) == 4) {
        // the tile lists were built by the deposit of the same evaluation (k_tile_mark)
        const unsigned ntile = (unsigned)(c.G3() >> 9);  // 8^3-cell tiles (G >= 16)
        const int32_t* count = c.dev_counters + kCtrInterp;  // light tiles, heavy chunks
        const int32_t* light = (const int32_t*)c.interp_list;  // ntile entries, then the heavy chunks
        const int2* heavy = c.interp_list + (ntile + 1) / 2;
        const unsigned nheavy = (unsigned)(interp_list_size(c.G3(), n) - (ntile + 1) / 2);
        const float* p2 = c.p.interlace ? (const float*)c.phi2 : nullptr;
        auto kh = four ? (c.p.interlace ? k_interp_heavy<true, true> : k_interp_heavy<true, false>)
                       : (c.p.interlace ? k_interp_heavy<false, true> : k_interp_heavy<false, false>);
        auto kl = four ? (c.p.interlace ? k_interp_light<true, true> : k_interp_light<true, false>)
                       : (c.p.interlace ? k_interp_light<false, true> : k_interp_light<false, false>);
        kh<<<dim3(nheavy + 8), dim3(kHeavyThreads), 0, s>>>(c.parts, c.tile_start, heavy, count + 1, (const float*)c.phi, p2,
                                                  c.logG, (float)inv, (float*)c.pm_acc);
        kl<<<dim3(ntile + 8), dim3(256), 0, s>>>(c.parts, c.tile_start, light, count, (const float*)c.phi, p2,
                                                 c.logG, (float)inv, (float*)c.pm_acc);
    } else {  // FP64: gradient grids, then a gather per particle
        const unsigned ng = (unsigned)((c.G3() + kBlock - 1) / kBlock);
        auto kg = four ? k_grad<R, true> : k_grad<R, false>;
        kg<<<dim3(ng), dim3(kBlock), 0, s>>>((const R*)c.phi, c.logG, inv, (V*)c.grad);
        if (c.p.interlace) kg<<<dim3(ng), dim3(kBlock), 0, s>>>((const R*)c.phi2, c.logG, inv, (V*)c.grad2);
        const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
        k_interp<R><<<dim3(nb), dim3(kBlock), 0, s>>>(c.parts, n, (const V*)c.grad,
                                                      c.p.interlace ? (const V*)c.grad2 : nullptr, c.logG,
                                                      (R*)c.pm_acc);
    }
    return hipGetLastError();
}

hipError_t pm_interp(Ctx& c, int64_t n, hipStream_t s) {
    if (n == 0) return hipGetLastError();
    return c.p.precision == JTP_FP64 ? interp_t<double>(c, n, s) : interp_t<float>(c, n, s);
}

// ------------------------------------------------------------------ tables (host, FP64)
// Twiddles e^{-2 pi i m/G}; per-axis tables T[i] = exp(-k_i^2 r_s^2) / sinc(k_i h/2)^4 (the sinc^4
// only with deconvolution), K2[i] = k_i^2, Ks[i] = k_i and D[i] = the gradient stencil's transfer
// function (4-point: (8 sin kh - sin 2kh) / 6h; 2-point: sin kh / h), k_i = 2 pi i'/L with i' the
// signed frequency; norm = -4 pi G / (V_cell G^3) (density = mass/V_cell, unnormalised inverse FFT).
// Twiddle m (e^{-2 pi i m / G}) and the Green-function tables at axis index i (T, k^2, k, D), in FP64
// on the host (pm_tables) or the device (k_pm_tables).
__host__ __device__ inline void green_table_entry(int i, int G, double L, double r_s, int deconvolve, int four,
                                                  double tw[2], double tab[4]) {
    const double h = L / G;
    const double a = -2.0 * M_PI * (double)i / (double)G;
    tw[0] = cos(a);
    tw[1] = sin(a);
    const int is = i <= G / 2 ? i : i - G;
    const double k = (2.0 * M_PI / L) * is;
    double t = exp(-k * k * r_s * r_s);
    if (deconvolve && is != 0) {
        const double x = 0.5 * k * h;
        const double sc = sin(x) / x;
        t /= sc * sc * sc * sc;
    }
    tab[0] = t;
    tab[1] = k * k;
    tab[2] = k;
    tab[3] = four ? (8.0 * sin(k * h) - sin(2.0 * k * h)) / (6.0 * h) : sin(k * h) / h;
}

static double green_norm_of(const jtp_params& p) {
    const double h = p.box_size / p.n_cells;
    return -4.0 * M_PI * p.g_constant / (h * h * h * (double)p.n_cells * (double)p.n_cells * (double)p.n_cells);
}

hipError_t pm_tables(Ctx& c) {
    const int G = (int)c.G();
    std::vector<double> tw(2 * G), tab(4 * G);
    for (int i = 0; i < G; ++i) {
        double t2[2], t4[4];
        green_table_entry(i, G, c.p.box_size, c.p.r_s, c.p.deconvolve, c.p.gradient == JTP_GRAD_4PT, t2, t4);
        tw[2 * i] = t2[0];
        tw[2 * i + 1] = t2[1];
        for (int q = 0; q < 4; ++q) tab[q * G + i] = t4[q];
    }
    c.green_norm = green_norm_of(c.p);
    hipError_t e;
    if (c.p.precision == JTP_FP64) {
        e = hipMemcpy(c.twiddle, tw.data(), 2 * G * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c.gtab, tab.data(), 4 * G * sizeof(double), hipMemcpyHostToDevice);
    } else {
        std::vector<float> twf(2 * G), tabf(4 * G);
        for (int m = 0; m < 2 * G; ++m) twf[m] = (float)tw[m];
        for (int m = 0; m < 4 * G; ++m) tabf[m] = (float)tab[m];
        e = hipMemcpy(c.twiddle, twf.data(), 2 * G * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c.gtab, tabf.data(), 4 * G * sizeof(float), hipMemcpyHostToDevice);
    }
    return e;
}

// ------------------------------------------------------------------ allocation-free Poisson solve
// For callers that cannot allocate device memory (code under graph capture, serving loops): the
// half spectrum and the tables of one solve live in a file-scope device array, one per GPU, held
// until the process ends.  Its size is fixed by the largest mesh it serves: one FP64 half spectrum
// of kNoAllocMaxG^3 (17.0 MB at 128^3) plus the FP64 twiddle and Green tables.  Every call
// overwrites all of it before reading it (the tables kernel writes every table entry, the R2C pass
// every half-spectrum entry), so nothing needs clearing between calls; calls that overlap on one
// GPU share the array and race.
constexpr int kNoAllocMaxG = 128;
constexpr size_t kNoAllocSpecBytes = (size_t)kNoAllocMaxG * kNoAllocMaxG * (kNoAllocMaxG / 2 + 1) * 2 * sizeof(double);
constexpr size_t kNoAllocTabOff = kNoAllocSpecBytes;                                      // real [4 G]
constexpr size_t kNoAllocTwOff = kNoAllocTabOff + 4 * kNoAllocMaxG * sizeof(double);      // complex [G]
constexpr size_t kNoAllocBytes = kNoAllocTwOff + 2 * kNoAllocMaxG * sizeof(double);
static_assert(kNoAllocTabOff % 256 == 0 && kNoAllocTwOff % 256 == 0, "scratch sections stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on each GPU, looked up once per device with that device current.  In
// host code the array's name is only a stand-in for the symbol, never a device address.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static hipError_t scratch_ptr_current(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

template <typename R>
__global__ void k_pm_tables(R* __restrict__ tw, R* __restrict__ tab, int G, double L, double r_s, int deconvolve,
                            int four) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    double t2[2], t4[4];
    green_table_entry(i, G, L, r_s, deconvolve, four, t2, t4);
    tw[2 * i] = (R)t2[0];
    tw[2 * i + 1] = (R)t2[1];
#pragma unroll
    for (int q = 0; q < 4; ++q) tab[q * G + i] = (R)t4[q];
}

}  // namespace jtp

// Allocation-free form of jtp_pm_solve (no counterpart in the reference): see janus_treepm.h.
extern "C" int jtp_pm_solve_noalloc(const jtp_params* p, const void* d_rho, void* d_phi, void* stream) {
    using namespace jtp;
    const char* why = nullptr;
    if (!p) why = "params is null";
    else if (!d_rho || !d_phi) why = "null grid pointer";
    else if (p->n_cells < 16 || (p->n_cells & (p->n_cells - 1))) why = "n_cells must be a power of two >= 16";
    else if (p->n_cells > kNoAllocMaxG) why = "n_cells exceeds 128, the largest mesh the allocation-free scratch holds";
    else if (p->precision != JTP_FP32 && p->precision != JTP_FP64) why = "precision must be JTP_FP32 or JTP_FP64";
    else if (!(p->box_size > 0) || !std::isfinite(p->box_size)) why = "box_size must be > 0";
    else if (!(p->r_s >= 0) || !std::isfinite(p->r_s)) why = "r_s must be >= 0";
    else if (!std::isfinite(p->g_constant)) why = "g_constant must be finite";
    else if (p->k_min < 0) why = "k_min must be >= 0";
    else if (p->gradient != JTP_GRAD_4PT && p->gradient != JTP_GRAD_2PT) why = "bad gradient";
    else if (p->deconvolve != 0 && p->deconvolve != 1) why = "deconvolve must be 0 or 1";
    else if (p->fft_tile != 0 && p->fft_tile != 11 && p->fft_tile != 12 &&
             !(p->fft_tile == 13 && p->precision == JTP_FP32))
        why = "fft_tile must be 0, 11, 12 or (FP32) 13";
    if (why) {
        std::fprintf(stderr, "jtp_pm_solve_noalloc: %s\n", why);
        return JTP_ERR_ARG;
    }
    void* base = nullptr;
    if (scratch_ptr_current(&base) != hipSuccess) return JTP_ERR_HIP;
    unsigned char* b = (unsigned char*)base;
    hipStream_t s = (hipStream_t)stream;
    const int G = p->n_cells;
    int logG = 0;
    while ((1 << logG) < G) ++logG;
    const FftPlan f{b, b + kNoAllocTwOff, b + kNoAllocTabOff, green_norm_of(*p), logG, p->k_min, p->fft_tile,
                    p->precision};
    const int four = p->gradient == JTP_GRAD_4PT;
    if (p->precision == JTP_FP64)
        k_pm_tables<double><<<dim3((G + 127) / 128), dim3(128), 0, s>>>((double*)f.twiddle, (double*)f.gtab, G,
                                                                         p->box_size, p->r_s, p->deconvolve, four);
    else
        k_pm_tables<float><<<dim3((G + 127) / 128), dim3(128), 0, s>>>((float*)f.twiddle, (float*)f.gtab, G,
                                                                       p->box_size, p->r_s, p->deconvolve, four);
    if (hipGetLastError() != hipSuccess) return JTP_ERR_HIP;
    if (plan_solve(f, d_rho, d_phi, s, p->deconvolve) != hipSuccess) return JTP_ERR_HIP;
    return JTP_OK;
}
