// Long-range (PM) part: deterministic CIC deposit of rho+ - rho-, 3-D real FFT
// Poisson solve with the split Green function, gradient + CIC interpolation.
//
// Reference behaviour replaced:
//   cic_scatter (FP64 CAS atomics into two grids), nbody_gpu_twopass.rs:85-158,
//     and PmGrid::assign_mass, src/treepm/pm_grid.rs:63-103 -> k_deposit: one net grid
//     rho+ - rho- gathered per node from Morton-sorted particles; no float atomics.
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
#include <vector>

namespace jtp {

template <typename R>
__device__ inline R frac_of(uint32_t u, int logG) {
    const uint32_t fb = 32 - logG;
    const uint32_t m = (fb >= 32) ? 0xffffffffu : ((1u << fb) - 1u);
    return (R)(u & m) * (R)(1.0 / (double)(1ull << fb));
}

// ------------------------------------------------------------------ deposit
// Deterministic CIC deposit ("sort + segmented reduction").  The particles are
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

// FP32 form: one workgroup per tile of 8^3 cells (a Morton-aligned block), owning the 8^3 nodes of
// each grid that take mass from its cells: node-grid nodes 0..7 per axis (cell corners) and the
// interlaced nodes at the centres of cells 0..7.  Those nodes also take mass from the one-cell
// shell around the tile (cells -1 .. 8 per axis), so the workgroup reads the particles of all
// 10^3 cells and keeps only the contributions to its own nodes: every node is written by one
// workgroup with a plain store, no global atomics and no zeroing pass.  Within the workgroup the
// sparse cells' contributions are summed in LDS as 32-bit fixed point: integer addition is
// associative, so the sums do not depend on the order in which the LDS atomics land and the grid
// is bit-reproducible.  Charges arrive pre-scaled by 2^k with |q| 2^k <= 2^22, and a node takes at
// most 2 x 2 x 2 sparse cells' worth of unit weight per axis (node grid: cells n-1, n with weights
// <= 1; interlaced grid: cells m-1, m, m+1 with weights <= 1/2, 1, 1/2) of <= kDenseCell = 32
// particles each, so |sum| <= 8 * 32 * 2^22 = 2^30 fits an int32; the unit 2^-k is 2^-22 of the
// largest |q| or finer, and rounding each contribution to it adds a relative error of ~2^-23 per
// particle, as FP32 itself does.  Dense cells (> kDenseCell particles) enter through their corner
// sums from k_dense_reduce*, computed with the same pre-scaled charges and added per owned node in
// a fixed cell order when the node is stored.  Each thread takes a contiguous run of the tile's
// sparse particles, so the lanes of a wavefront sit in different cells and seldom hit one LDS
// address together.  Tiles are dealt to the 8 XCDs in contiguous Morton ranges (blocks are
// dispatched round-robin over the XCDs), so the shell cells a tile reads were mostly read by its
// predecessor on the same XCD and sit in that XCD's L2.  LDS: 20 KB, so 7 workgroups per CU hide
// the latency of the cell-table and particle reads.
constexpr int kTile = 8;
constexpr int kShell = kTile + 2;                      // cells -1 .. 8 per axis
constexpr int kShellCells = kShell * kShell * kShell;  // 1000
constexpr int kN1 = kTile + 3;                         // node-grid nodes -1 .. 9 per axis
constexpr int kN2 = kTile + 4;                         // centres of cells -2 .. 9 per axis
constexpr int kFixBits = 22;                           // |q| 2^k <= 2^kFixBits
static_assert(kFixBits + 8 <= 30, "a node sum must fit an int32");
static_assert(kDenseCell <= 32, "the int32 bound assumes <= 32 particles per sparse cell");
// the largest LDS index a shell particle reaches: cell 8 (index 9) plus corner offset 1 per axis in
// a1; cell 8 above its centre (index 10) plus offset 1 per axis in a2
static_assert(((kTile + 2) * kN1 + kTile + 2) * kN1 + kTile + 2 < kN1 * kN1 * kN1, "a1 too small");
static_assert(((kTile + 3) * kN2 + kTile + 3) * kN2 + kTile + 3 < kN2 * kN2 * kN2, "a2 too small");

template <bool BOTH>
__global__ __launch_bounds__(256) void k_deposit_tile(const uint4* __restrict__ parts, const int32_t* __restrict__ cs,
                                                      const int32_t* __restrict__ slot,
                                                      const float* __restrict__ corner8,
                                                      const float* __restrict__ corner27, int logG, float qp,
                                                      float qm, double unit, float* __restrict__ rho,
                                                      float* __restrict__ rho2) {
    // every node the shell's particles reach, so that the corner offsets are compile-time constants
    // (LDS instruction offsets) and no contribution needs a bounds test; only the tile's own nodes
    // are stored at the end
    __shared__ int32_t a1[kN1 * kN1 * kN1];           // node-grid nodes -1 .. 9, index l + 1
    __shared__ int32_t a2[BOTH ? kN2 * kN2 * kN2 : 1];  // centres of cells -2 .. 9, index l + 2
    __shared__ int32_t cst[kShellCells];      // particle start of a sparse shell cell; -1 - slot of a dense one
    __shared__ int32_t pre[kShellCells + 1];  // prefix of the sparse cells' counts
    __shared__ int32_t wsum[4];
    __shared__ int32_t any_dense;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t G = 1ll << logG;
    const int32_t msk = (int32_t)G - 1;
    const unsigned ntile = gridDim.x;
    const unsigned tile = (blockIdx.x & 7u) * (ntile >> 3) + (blockIdx.x >> 3);
    const uint64_t base = (uint64_t)tile << 9;  // Morton index of the tile's first cell
    const int32_t ox = (int32_t)compact3_21(base), oy = (int32_t)compact3_21(base >> 1),
                  oz = (int32_t)compact3_21(base >> 2);
    for (int i = tid; i < kN1 * kN1 * kN1; i += 256) a1[i] = 0;
    if (BOTH)
        for (int i = tid; i < kN2 * kN2 * kN2; i += 256) a2[i] = 0;
    if (tid == 0) any_dense = 0;
    // shell cells 4 tid .. 4 tid + 3 (s = (cz + 1) 100 + (cy + 1) 10 + cx + 1)
    int cnt[4] = {0, 0, 0, 0};
    bool dense = false;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int sc = 4 * tid + u;
        if (sc >= kShellCells) continue;
        const int cx = sc % kShell - 1, cy = (sc / kShell) % kShell - 1, cz = sc / (kShell * kShell) - 1;
        const uint64_t cm = spread3_21((uint32_t)((ox + cx) & msk)) | (spread3_21((uint32_t)((oy + cy) & msk)) << 1) |
                            (spread3_21((uint32_t)((oz + cz) & msk)) << 2);
        const int32_t st = cs[cm], k = cs[cm + 1] - st;
        if (k <= kDenseCell) {
            cst[sc] = st;
            cnt[u] = k;
        } else {
            cst[sc] = -1 - slot[cm];
            dense = true;
        }
    }
    // block exclusive scan of the sparse counts in shell-cell order
    const int mine = cnt[0] + cnt[1] + cnt[2] + cnt[3];
    int incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();  // also orders the zeroing before the atomics below
    if (dense) any_dense = 1;
    int woff = 0;
    for (int w = 0; w < wv; ++w) woff += wsum[w];
    int run = woff + incl - mine;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int sc = 4 * tid + u;
        if (sc < kShellCells) pre[sc] = run;
        run += cnt[u];
    }
    if (tid == 0) pre[kShellCells] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    const int32_t total = pre[kShellCells];
    const int32_t per = (total + 255) >> 8;
    int32_t t = tid * per;
    const int32_t t_end = min(total, t + per);
    if (t < t_end) {
        int lo = 0, hi = kShellCells - 1;  // last shell cell with pre <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        int32_t cell_end = pre[lo + 1];
        for (; t < t_end; ++t) {
            while (t >= cell_end) cell_end = pre[++lo + 1];
            const int cx = lo % kShell - 1, cy = (lo / kShell) % kShell - 1, cz = lo / (kShell * kShell) - 1;
            const uint4 r = parts[cst[lo] + (t - pre[lo])];
            const float fx = frac_of<float>(r.x, logG), fy = frac_of<float>(r.y, logG), fz = frac_of<float>(r.z, logG);
            const float q = (int32_t)r.w > 0 ? qp : qm;
            {
                const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {q * (1.0f - fz), q * fz};
                int32_t* a = &a1[((cz + 1) * kN1 + cy + 1) * kN1 + cx + 1];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                    atomicAdd(a + (dz * kN1 + dy) * kN1 + dx, __float2int_rn(wz[dz] * wy[dy] * wx[dx]));
                }
            }
            if (BOTH) {
                // the two interlaced nodes around the particle per axis: centres of cells (c-1, c) if
                // f < 1/2, else (c, c+1); the centre of cell m has index m + 2
                const bool lox = fx < 0.5f, loy = fy < 0.5f, loz = fz < 0.5f;
                const int bx = cx + (lox ? 1 : 2), by = cy + (loy ? 1 : 2), bz = cz + (loz ? 1 : 2);
                const float vx[2] = {lox ? 0.5f - fx : 1.5f - fx, lox ? 0.5f + fx : fx - 0.5f};
                const float vy[2] = {loy ? 0.5f - fy : 1.5f - fy, loy ? 0.5f + fy : fy - 0.5f};
                const float vz[2] = {q * (loz ? 0.5f - fz : 1.5f - fz), q * (loz ? 0.5f + fz : fz - 0.5f)};
                int32_t* a = &a2[(bz * kN2 + by) * kN2 + bx];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                    atomicAdd(a + (dz * kN2 + dy) * kN2 + dx, __float2int_rn(vz[dz] * vy[dy] * vx[dx]));
                }
            }
        }
    }
    __syncthreads();
    // node (x, y, z) of the tile: the node grid's node at cell corner o + l, the interlaced grid's
    // node at the centre of cell o + l; the dense cells around it add their corner sums in a fixed
    // order (node grid: cells l - e, e in {0,1}^3; interlaced: cells l - e, e in {-1,0,1}^3)
    const bool dn = any_dense != 0;
    for (int i = tid; i < kTile * kTile * kTile; i += 256) {
        const int lx = i % kTile, ly = (i / kTile) % kTile, lz = i / (kTile * kTile);
        double v1 = (double)a1[((lz + 1) * kN1 + ly + 1) * kN1 + lx + 1];
        if (dn) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {  // cell l - e (shell index l - e + 1), its corner e
                const int ex = q & 1, ey = (q >> 1) & 1, ez = q >> 2;
                const int32_t c = cst[((lz - ez + 1) * kShell + ly - ey + 1) * kShell + lx - ex + 1];
                if (c < 0) v1 += (double)corner8[8 * (int64_t)(-1 - c) + q];
            }
        }
        rho[((int64_t)(oz + lz) * G + (oy + ly)) * G + (ox + lx)] = (float)(v1 * unit);
        if (BOTH) {
            double v2 = (double)a2[((lz + 2) * kN2 + ly + 2) * kN2 + lx + 2];
            if (dn) {
                for (int q = 0; q < 27; ++q) {  // cell l - e, e = (q % 3, q / 3 % 3, q / 9) - 1: its entry q
                    const int ex = q % 3 - 1, ey = (q / 3) % 3 - 1, ez = q / 9 - 1;
                    const int32_t c = cst[((lz - ez + 1) * kShell + ly - ey + 1) * kShell + lx - ex + 1];
                    if (c < 0) v2 += (double)corner27[27 * (int64_t)(-1 - c) + q];
                }
            }
            rho2[((int64_t)((oz + lz + 1) & msk) * G + ((oy + ly + 1) & msk)) * G + ((ox + lx + 1) & msk)] =
                (float)(v2 * unit);
        }
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

template <typename R, int LG, int LP>
static hipError_t inverse_fft_lg(Ctx& c, R* out, hipStream_t s) {
    constexpr int G = 1 << LG, H = G / 2 + 1, NPT = 1 << LP;
    constexpr int KB = NPT / G, nkb = (H + KB - 1) / KB;
    constexpr unsigned nx = (unsigned)(G * G > 2 * (NPT / G) ? (int64_t)G * G / (2 * (NPT / G)) : 1);
    const Cplx<R>* tw = (const Cplx<R>*)c.twiddle;
    Cplx<R>* spec = (Cplx<R>*)c.spec;
    k_fft_z_inv<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw);
    k_fft_y<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw, 1);
    k_fft_x_c2r<R, LG, LP><<<dim3(nx), dim3(256), 0, s>>>(spec, out, tw);
    return hipGetLastError();
}

template <typename R, int LG, int LP>
static hipError_t fft_solve_lg(Ctx& c, const R* rho, R* phi, hipStream_t s, int matched) {
    constexpr int G = 1 << LG, H = G / 2 + 1, NPT = 1 << LP;
    constexpr int KB = NPT / G, nkb = (H + KB - 1) / KB;
    constexpr unsigned nx = (unsigned)(G * G > 2 * (NPT / G) ? (int64_t)G * G / (2 * (NPT / G)) : 1);
    const Cplx<R>* tw = (const Cplx<R>*)c.twiddle;
    Cplx<R>* spec = (Cplx<R>*)c.spec;
    k_fft_x_r2c<R, LG, LP><<<dim3(nx), dim3(256), 0, s>>>(rho, spec, tw);
    k_fft_y<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw, 0);
    k_fft_z_green<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw, (const R*)c.gtab,
                                                                         (R)c.green_norm, c.p.k_min, matched);
    k_fft_y<R, LG, LP><<<dim3((unsigned)(G * nkb)), dim3(256), 0, s>>>(spec, tw, 1);
    k_fft_x_c2r<R, LG, LP><<<dim3(nx), dim3(256), 0, s>>>(spec, phi, tw);
    return hipGetLastError();
}

// the mesh sizes the context accepts (G = 16 .. 512) and FFT tiles (2^11, 2^12 points; 2^13 for FP32)
#define JTP_FFT_DISPATCH_LG(CALL, LP)          \
    switch (c.logG) {                          \
        case 4: return CALL(4, LP);            \
        case 5: return CALL(5, LP);            \
        case 6: return CALL(6, LP);            \
        case 7: return CALL(7, LP);            \
        case 8: return CALL(8, LP);            \
        case 9: return CALL(9, LP);            \
        default: return hipErrorInvalidValue;  \
    }

template <typename R>
static hipError_t inverse_fft_t(Ctx& c, R* out, hipStream_t s) {
#define JTP_INV(LG, LP) inverse_fft_lg<R, LG, LP>(c, out, s)
    if constexpr (sizeof(R) == 4) {
        if (c.p.fft_tile == 13) JTP_FFT_DISPATCH_LG(JTP_INV, 13)
    }
    if (c.p.fft_tile == 12) JTP_FFT_DISPATCH_LG(JTP_INV, 12)
    JTP_FFT_DISPATCH_LG(JTP_INV, 11)
#undef JTP_INV
}

hipError_t pm_inverse_fft(Ctx& c, void* out, hipStream_t s) {
    if (c.p.precision == JTP_FP64) return inverse_fft_t<double>(c, (double*)out, s);
    return inverse_fft_t<float>(c, (float*)out, s);
}

template <typename R>
static hipError_t fft_solve_t(Ctx& c, const R* rho, R* phi, hipStream_t s, int matched) {
#define JTP_SOLVE(LG, LP) fft_solve_lg<R, LG, LP>(c, rho, phi, s, matched)
    if constexpr (sizeof(R) == 4) {
        if (c.p.fft_tile == 13) JTP_FFT_DISPATCH_LG(JTP_SOLVE, 13)
    }
    if (c.p.fft_tile == 12) JTP_FFT_DISPATCH_LG(JTP_SOLVE, 12)
    JTP_FFT_DISPATCH_LG(JTP_SOLVE, 11)
#undef JTP_SOLVE
}

hipError_t pm_fft_solve(Ctx& c, const void* rho, void* phi, hipStream_t s, int matched) {
    if (matched < 0) matched = c.p.deconvolve;  // the force's Green function
    if (c.p.precision == JTP_FP64) return fft_solve_t<double>(c, (const double*)rho, (double*)phi, s, matched);
    return fft_solve_t<float>(c, (const float*)rho, (float*)phi, s, matched);
}

hipError_t pm_deposit(Ctx& c, int64_t n, hipStream_t s) {
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
    if (c.p.precision == JTP_FP64) {
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
            k_dense_reduce<double><<<dim3(nbd), dim3(256), 0, s>>>(c.parts, c.cell_start, list, c.dense_slot,
                                                                   count, c.logG, c.m_plus, -c.m_minus, c8);
            k_dense_combine<double><<<dim3(nbd), dim3(256), 0, s>>>(c.cell_start, list, c.dense_slot, count, 8, c8);
            k_deposit<double><<<dim3(nb), dim3(kBlock), 0, s>>>(c.parts, c.cell_start, c.dense_slot, c8, c.logG,
                                                                c.m_plus, -c.m_minus, (double*)c.rho);
        }
    } else {
        float* c8 = (float*)c.dense_corner;
        float* c27 = c8 + 8 * nd_cap;
        // charges scaled by 2^k, |q| 2^k <= 2^kFixBits (exact: a power of two); the grids' unit is 2^-k
        int ex = 0;
        (void)std::frexp(c.m_plus > c.m_minus ? c.m_plus : c.m_minus, &ex);  // max |q| <= 2^ex
        const int k = kFixBits - ex;
        const float qp = (float)std::ldexp(c.m_plus, k), qm = (float)-std::ldexp(c.m_minus, k);
        const double unit = std::ldexp(1.0, -k);
        const unsigned ntile = (unsigned)(G3 >> 9);  // 8^3-cell tiles (G >= 16: a multiple of 8)
        if (c.p.interlace) {
            k_dense_reduce_both<float><<<dim3(nbd), dim3(256), 0, s>>>(c.parts, c.cell_start, list, c.dense_slot,
                                                                       count, c.logG, qp, qm, c8, c27);
            k_dense_combine<float><<<dim3(nbd), dim3(256), 0, s>>>(c.cell_start, list, c.dense_slot, count, 8, c8);
            k_dense_combine<float><<<dim3(nbd), dim3(256), 0, s>>>(c.cell_start, list, c.dense_slot, count, 27, c27);
            k_deposit_tile<true><<<dim3(ntile), dim3(256), 0, s>>>(c.parts, c.cell_start, c.dense_slot, c8, c27,
                                                                   c.logG, qp, qm, unit, (float*)c.rho,
                                                                   (float*)c.rho2);
        } else {
            k_dense_reduce<float><<<dim3(nbd), dim3(256), 0, s>>>(c.parts, c.cell_start, list, c.dense_slot,
                                                                  count, c.logG, qp, qm, c8);
            k_dense_combine<float><<<dim3(nbd), dim3(256), 0, s>>>(c.cell_start, list, c.dense_slot, count, 8, c8);
            k_deposit_tile<false><<<dim3(ntile), dim3(256), 0, s>>>(c.parts, c.cell_start, c.dense_slot, c8, nullptr,
                                                                    c.logG, qp, qm, unit, (float*)c.rho, nullptr);
        }
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

// CIC sum over the 8 corners around local node (ix, iy, iz) with each corner's gradient taken from
// the staged potential p (NP nodes per axis, node 0 at local -2) by the stencil: light tiles, where
// computing every node's gradient first would cost more than it saves.
template <bool FOUR, int NP>
__device__ inline void tile_cic_phi(const float* p, int ix, int iy, int iz, float fx, float fy, float fz, float inv,
                                    float a[3]) {
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
    const int b0 = ((iz + 2) * NP + iy + 2) * NP + ix + 2;
    float ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const float w = wx[dx] * wy[dy] * wz[dz];
                const int b = b0 + (dz * NP + dy) * NP + dx;
                ax += w * lds_stencil<FOUR>(p, b, 1, inv);
                ay += w * lds_stencil<FOUR>(p, b, NP, inv);
                az += w * lds_stencil<FOUR>(p, b, NP * NP, inv);
            }
    a[0] = ax;
    a[1] = ay;
    a[2] = az;
}

// Tiles are classed by particle count: light tiles (1 .. kLightMax particles, kLightMax / 256 per thread) are
// listed for k_interp_light; every chunk of kInterpChunk particles of a heavier tile (a halo core
// holds 1e4-1e5) is listed, as (tile, first particle), for a k_interp_heavy workgroup.  The list
// order varies between runs; each particle is still computed by one thread with the same
// arithmetic, so the results do not.
__global__ __launch_bounds__(256) void k_interp_mark(const int32_t* __restrict__ cs, int32_t ntile,
                                                     int32_t* __restrict__ light, int2* __restrict__ heavy,
                                                     int32_t* __restrict__ count) {
    const int32_t t = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    int32_t beg = 0, np = 0;
    if (t < ntile) {
        beg = cs[(int64_t)t << 9];
        np = cs[((int64_t)t + 1) << 9] - beg;
    }
    const bool lt = np > 0 && np <= kLightMax;
    const int32_t nh = np > kLightMax ? (np + kInterpChunk - 1) / kInterpChunk : 0;
    const int at_l = wave_reserve(lt ? 1 : 0, count);
    if (lt) light[at_l] = t;
    const int at_h = wave_reserve(nh, count + 1);
    for (int32_t k = 0; k < nh; ++k) heavy[at_h + k] = make_int2(t, beg + k * kInterpChunk);
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
__global__ __launch_bounds__(kHeavyThreads) void k_interp_heavy(const uint4* __restrict__ parts, const int32_t* __restrict__ cs,
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
    const int32_t pbeg = w.y, pend = min(cs[base + 512], pbeg + kInterpChunk);
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
__global__ __launch_bounds__(256) void k_interp_light(const uint4* __restrict__ parts, const int32_t* __restrict__ cs,
                                                      const int32_t* __restrict__ list, const int32_t* __restrict__ count,
                                                      const float* __restrict__ phi, const float* __restrict__ phi2,
                                                      int logG, float inv, float* __restrict__ acc) {
    __shared__ float p1[kIp1 * kIp1 * kIp1];
    __shared__ float p2[BOTH ? kIp2 * kIp2 * kIp2 : 1];
    const int32_t e = xcd_entry(blockIdx.x, *count);
    if (e < 0) return;  // the same for every thread of the block
    const uint64_t base = (uint64_t)list[e] << 9;
    const int32_t pbeg = cs[base], np = cs[base + 512] - pbeg;  // 1 .. kLightMax
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
        const unsigned ntile = (unsigned)(c.G3() >> 9);  // 8^3-cell tiles (G >= 16)
        int32_t* count = c.dev_counters + kCtrInterp;    // light tiles, heavy chunks
        hipError_t e = launch_fill_i32(count, 2, 0, s);
        if (e != hipSuccess) return e;
        int32_t* light = (int32_t*)c.interp_list;  // ntile entries, then the heavy chunks
        int2* heavy = c.interp_list + (ntile + 1) / 2;
        k_interp_mark<<<dim3((ntile + 255) / 256), dim3(256), 0, s>>>(c.cell_start, (int32_t)ntile, light, heavy,
                                                                       count);
        const unsigned nheavy = (unsigned)(interp_list_size(c.G3(), n) - (ntile + 1) / 2);
        const float* p2 = c.p.interlace ? (const float*)c.phi2 : nullptr;
        auto kh = four ? (c.p.interlace ? k_interp_heavy<true, true> : k_interp_heavy<true, false>)
                       : (c.p.interlace ? k_interp_heavy<false, true> : k_interp_heavy<false, false>);
        auto kl = four ? (c.p.interlace ? k_interp_light<true, true> : k_interp_light<true, false>)
                       : (c.p.interlace ? k_interp_light<false, true> : k_interp_light<false, false>);
        kh<<<dim3(nheavy + 8), dim3(kHeavyThreads), 0, s>>>(c.parts, c.cell_start, heavy, count + 1, (const float*)c.phi, p2,
                                                  c.logG, (float)inv, (float*)c.pm_acc);
        kl<<<dim3(ntile + 8), dim3(256), 0, s>>>(c.parts, c.cell_start, light, count, (const float*)c.phi, p2,
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
hipError_t pm_tables(Ctx& c) {
    const int G = (int)c.G();
    const double L = c.p.box_size, h = L / G;
    std::vector<double> tw(2 * G), tab(4 * G);
    for (int m = 0; m < G; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)G;
        tw[2 * m] = std::cos(a);
        tw[2 * m + 1] = std::sin(a);
    }
    const double dk = 2.0 * M_PI / L;
    const bool four = c.p.gradient == JTP_GRAD_4PT;
    for (int i = 0; i < G; ++i) {
        const int is = i <= G / 2 ? i : i - G;
        const double k = dk * is;
        double t = std::exp(-k * k * c.p.r_s * c.p.r_s);
        if (c.p.deconvolve && is != 0) {
            const double x = 0.5 * k * h;
            const double sc = std::sin(x) / x;
            t /= sc * sc * sc * sc;
        }
        tab[i] = t;
        tab[G + i] = k * k;
        tab[2 * G + i] = k;
        tab[3 * G + i] = four ? (8.0 * std::sin(k * h) - std::sin(2.0 * k * h)) / (6.0 * h) : std::sin(k * h) / h;
    }
    c.green_norm = -4.0 * M_PI * c.p.g_constant / (h * h * h * (double)G * (double)G * (double)G);
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

}  // namespace jtp
