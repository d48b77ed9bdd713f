// Kick and drift kernels of the integrator step around the force.
//
// Reference behaviour replaced: kick_f32 / drift_f32 (src/nbody_gpu_twopass.rs:33-64), launched
// by step_treepm_gpu (:4486-4649).
//   kick:  v += (a - H v dtau_dt) dt, the same explicit Hubble friction, evaluated in FP64 with
//          H chosen per species (the reference passes one H for both);
//   drift: the displacement v dt is formed in FP64 and added to the 32-bit fixed-point position
//          in integer arithmetic, so the periodic wrap is exact (drift_f32 adds in f32 and wraps
//          once per axis).
// The step scalars are read from device memory (StepParams), so a captured graph of the step
// stays valid when dt or H change between steps.
#include "jtp_internal.h"

namespace jtp {

template <typename V>
__global__ void k_kick(V* __restrict__ vel, const V* __restrict__ acc, const int8_t* __restrict__ species,
                       int64_t n, const StepParams* __restrict__ prm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dt = prm->kick_dt;
    const double h = (species[i] > 0 ? prm->hubble_plus : prm->hubble_minus) * prm->dtau_dt;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = (double)vel[3 * i + d];
        vel[3 * i + d] = (V)(v + ((double)acc[3 * i + d] - h * v) * dt);
    }
}

template <typename V>
__global__ void k_drift(uint32_t* __restrict__ pos, const V* __restrict__ vel, int64_t n3,
                        const StepParams* __restrict__ prm) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n3) return;
    // displacement in fixed-point units, reduced to one period before the integer conversion
    double d = (double)vel[j] * prm->drift_units;
    d -= 4294967296.0 * rint(d * (1.0 / 4294967296.0));
    const int64_t du = (int64_t)rint(d);
    pos[j] = (uint32_t)((uint64_t)pos[j] + (uint64_t)du);
}

hipError_t launch_kick(Ctx& c, void* vel, const void* acc, const int8_t* species, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    if (c.p.precision == JTP_FP64)
        k_kick<double><<<dim3(nb), dim3(kBlock), 0, s>>>((double*)vel, (const double*)acc, species, n, c.dev_step);
    else
        k_kick<float><<<dim3(nb), dim3(kBlock), 0, s>>>((float*)vel, (const float*)acc, species, n, c.dev_step);
    return hipGetLastError();
}

hipError_t launch_drift(Ctx& c, uint32_t* pos, const void* vel, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t n3 = 3 * n;
    const unsigned nb = (unsigned)((n3 + kBlock - 1) / kBlock);
    if (c.p.precision == JTP_FP64)
        k_drift<double><<<dim3(nb), dim3(kBlock), 0, s>>>(pos, (const double*)vel, n3, c.dev_step);
    else
        k_drift<float><<<dim3(nb), dim3(kBlock), 0, s>>>(pos, (const float*)vel, n3, c.dev_step);
    return hipGetLastError();
}

}  // namespace jtp

// ------------------------------------------------------------------ FP64 diagnostics
// Deterministic two-level reductions (fixed grid, fixed in-block order) of
//   velocities: per species kinetic energy sum |m| v^2 / 2 and momentum sum |m| v;
//   grid: sum rho * phi over the nodes (twice the long-range potential energy of the grid).
// The reference copies velocities and grids to the host for these sums
// (nbody_gpu_twopass.rs:5071-5077, 5293-5307).
namespace jtp {

constexpr int kDiagBlocks = 1024;
constexpr int kDiagVals = 8;  // ke+, ke-, p+ xyz, p- xyz

template <int NV>
__device__ inline void block_reduce_store(double (&v)[NV], double* __restrict__ out) {
    __shared__ double lds[NV][256];
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int k = 0; k < NV; ++k) lds[k][threadIdx.x] += lds[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < NV) out[threadIdx.x] = lds[threadIdx.x][0];
}

template <typename V>
__global__ __launch_bounds__(256) void k_diag_vel(const V* __restrict__ vel, const int8_t* __restrict__ species,
                                                  int64_t n, double mp, double mm, double* __restrict__ partial) {
    double v[kDiagVals] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool plus = species[i] > 0;
        const double m = plus ? mp : mm;
        const double vx = (double)vel[3 * i], vy = (double)vel[3 * i + 1], vz = (double)vel[3 * i + 2];
        const double ke = 0.5 * m * (vx * vx + vy * vy + vz * vz);
        const int o = plus ? 0 : 1;
        v[o] += ke;
        v[2 + 3 * o] += m * vx;
        v[3 + 3 * o] += m * vy;
        v[4 + 3 * o] += m * vz;
    }
    block_reduce_store<kDiagVals>(v, partial + (int64_t)blockIdx.x * kDiagVals);
}

template <typename R>
__global__ __launch_bounds__(256) void k_diag_grid(const R* __restrict__ rho, const R* __restrict__ phi, int64_t m,
                                                   double* __restrict__ partial) {
    double v[1] = {0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        v[0] += (double)rho[i] * (double)phi[i];
    block_reduce_store<1>(v, partial + blockIdx.x);
}

// Sums `nb` partial vectors of `nv` values in block order (one workgroup, fixed order).
__global__ __launch_bounds__(256) void k_diag_final(const double* __restrict__ partial, int nb, int nv,
                                                    double* __restrict__ out) {
    __shared__ double lds[256];
    for (int k = 0; k < nv; ++k) {
        double s = 0;
        for (int b = threadIdx.x; b < nb; b += 256) s += partial[(int64_t)b * nv + k];
        lds[threadIdx.x] = s;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = lds[0];
        __syncthreads();
    }
}

// out[0..7] = ke+, ke-, p+ xyz, p- xyz; out[8] = sum rho phi of grid 1, out[9] of grid 2 (0 if absent).
hipError_t launch_diagnostics(Ctx& c, int64_t n, const void* vel, const int8_t* species, double* scratch,
                              double* out, hipStream_t s) {
    double* part = scratch;  // kDiagBlocks * kDiagVals doubles
    if (n > 0) {
        if (c.p.precision == JTP_FP64)
            k_diag_vel<double><<<dim3(kDiagBlocks), dim3(256), 0, s>>>((const double*)vel, species, n, c.m_plus,
                                                                       c.m_minus, part);
        else
            k_diag_vel<float><<<dim3(kDiagBlocks), dim3(256), 0, s>>>((const float*)vel, species, n, c.m_plus,
                                                                      c.m_minus, part);
        k_diag_final<<<dim3(1), dim3(256), 0, s>>>(part, kDiagBlocks, kDiagVals, out);
    } else {
        const hipError_t e = launch_fill_i32((int32_t*)out, 2 * kDiagVals, 0, s);
        if (e != hipSuccess) return e;
    }
    const void* grids[2][2] = {{c.rho, c.phi}, {c.rho2, c.phi2}};
    for (int g = 0; g < 2; ++g) {
        if (!c.p.pm_enabled || (g == 1 && !c.p.interlace)) {
            const hipError_t e = launch_fill_i32((int32_t*)(out + kDiagVals + g), 2, 0, s);
            if (e != hipSuccess) return e;
            continue;
        }
        if (c.p.precision == JTP_FP64)
            k_diag_grid<double><<<dim3(kDiagBlocks), dim3(256), 0, s>>>((const double*)grids[g][0],
                                                                        (const double*)grids[g][1], c.G3(), part);
        else
            k_diag_grid<float><<<dim3(kDiagBlocks), dim3(256), 0, s>>>((const float*)grids[g][0],
                                                                       (const float*)grids[g][1], c.G3(), part);
        k_diag_final<<<dim3(1), dim3(256), 0, s>>>(part, kDiagBlocks, 1, out + kDiagVals + g);
    }
    return hipGetLastError();
}

}  // namespace jtp

// ------------------------------------------------------------------ potential energy (FP64)
// out = {sum |m| sigma Phi_long, sum |m| sigma Phi_short, sum m^2 S, Q = sum |m| sigma}, from which
// w_i = |m_i| sigma_i (Phi_long(x_i) + Phi_short_i) - m_i^2 S_i, W = 1/2 sum_i w_i, where Phi_long is
// the CIC interpolation of the mesh potential (mean of the two interlaced grids), Phi_short the
// walk's short-range potential, and S_i = sum_{a,b} w_a w_b K(n_a - n_b) the mesh potential a unit
// mass deposited with particle i's CIC weights produces at the same weights (K: the potential of a
// unit mass at one node, at offsets in {-1, 0, 1}^3), i.e. the self-energy the mesh includes.
namespace jtp {

__device__ inline void cic_w(uint32_t u, int logG, int& i0, double w[2]) {
    const int sh = 32 - logG;
    i0 = (int)(u >> sh);
    const double f = (double)(u & ((1u << sh) - 1u)) * (1.0 / (double)(1ull << sh));
    w[0] = 1.0 - f;
    w[1] = f;
}

__device__ inline void grid_terms(const uint4 r, uint32_t shift, const double* __restrict__ phi, int logG,
                                  const double* __restrict__ kself, double& pot, double& self) {
    const int G = 1 << logG;
    int ix, iy, iz;
    double wx[2], wy[2], wz[2];
    cic_w(r.x + shift, logG, ix, wx);
    cic_w(r.y + shift, logG, iy, wy);
    cic_w(r.z + shift, logG, iz, wz);
    double p = 0, sf = 0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int ax = a & 1, ay = (a >> 1) & 1, az = a >> 2;
        const double wa = wx[ax] * wy[ay] * wz[az];
        const int64_t node = ((int64_t)((iz + az) & (G - 1)) * G + ((iy + ay) & (G - 1))) * G + ((ix + ax) & (G - 1));
        p += wa * phi[node];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int bx = b & 1, by = (b >> 1) & 1, bz = b >> 2;
            const double wb = wx[bx] * wy[by] * wz[bz];
            sf += wa * wb * kself[(ax - bx + 1) + 3 * (ay - by + 1) + 9 * (az - bz + 1)];
        }
    }
    pot = p;
    self = sf;
}

__global__ __launch_bounds__(256) void k_energy_terms(const uint4* __restrict__ parts, const int32_t* __restrict__ perm,
                                                      int64_t n, const double* __restrict__ phi,
                                                      const double* __restrict__ phi2, int logG,
                                                      const double* __restrict__ kself,
                                                      const double* __restrict__ pot_short, double mp, double mm,
                                                      double* __restrict__ partial) {
    double v[5] = {0, 0, 0, 0, 0};  // |m| sigma Phi_long, |m| sigma Phi_short, m^2 S, |m| sigma, m^2
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint4 r = parts[i];
        const bool plus = (int32_t)r.w > 0;
        const double m = plus ? mp : mm, sg = plus ? 1.0 : -1.0;
        double p1, s1;
        grid_terms(r, 0u, phi, logG, kself, p1, s1);
        if (phi2) {
            double p2, s2;
            grid_terms(r, 1u << (31 - logG), phi2, logG, kself, p2, s2);
            p1 = 0.5 * (p1 + p2);
            s1 = 0.5 * (s1 + s2);
        }
        v[0] += m * sg * p1;
        v[1] += m * sg * pot_short[perm[i]];
        v[2] += m * m * s1;
        v[3] += m * sg;
        v[4] += m * m;
    }
    block_reduce_store<5>(v, partial + (int64_t)blockIdx.x * 5);
}

hipError_t launch_energy_terms(Ctx& c, int64_t n, const double* kself, const double* pot_short, double* scratch,
                               double* out, hipStream_t s) {
    k_energy_terms<<<dim3(kDiagBlocks), dim3(256), 0, s>>>(c.parts, c.vals0, n, (const double*)c.phi,
                                                           c.p.interlace ? (const double*)c.phi2 : nullptr, c.logG,
                                                           kself, pot_short, c.m_plus, c.m_minus, scratch);
    k_diag_final<<<dim3(1), dim3(256), 0, s>>>(scratch, kDiagBlocks, 5, out);
    return hipGetLastError();
}

}  // namespace jtp

// ------------------------------------------------------------------ segregation metrics
// Counts per nearest-grid-point cell and species (integer atomics: exact, order-free), totals per
// species; then, per cell, with mean = total / cells: (d - (mean+ - mean-))^2 for d = n+ - n-,
// delta+, delta-, delta+^2, delta-^2, delta+ delta-, (n- - mean-)^2, reduced in FP64 in fixed order.
namespace jtp {

__device__ inline int64_t seg_cell(uint32_t u, int32_t nc) { return (int64_t)(((uint64_t)u * (uint32_t)nc) >> 32); }

__global__ __launch_bounds__(256) void k_seg_count(const uint32_t* __restrict__ pos,
                                                   const int8_t* __restrict__ species, int64_t n, int32_t nc,
                                                   int32_t* __restrict__ counts, int32_t* __restrict__ totals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const bool plus = live && species[i] > 0;
    if (live) {
        const int64_t cell =
            (seg_cell(pos[3 * i + 2], nc) * nc + seg_cell(pos[3 * i + 1], nc)) * nc + seg_cell(pos[3 * i], nc);
        atomicAdd(&counts[2 * cell + (plus ? 0 : 1)], 1);
    }
    const uint64_t mp = __ballot(plus), mm = __ballot(live && !plus);
    if ((threadIdx.x & 63) == 0) {
        if (mp) atomicAdd(&totals[0], (int32_t)__popcll(mp));
        if (mm) atomicAdd(&totals[1], (int32_t)__popcll(mm));
    }
}

__global__ __launch_bounds__(256) void k_seg_moments(const int32_t* __restrict__ counts,
                                                     const int32_t* __restrict__ totals, int64_t cells,
                                                     double* __restrict__ partial) {
    const double mp = (double)totals[0] / (double)cells, mm = (double)totals[1] / (double)cells;
    const double ed = mp - mm;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (int64_t)gridDim.x * blockDim.x) {
        const double np = (double)counts[2 * c], nm = (double)counts[2 * c + 1];
        const double dd = np - nm - ed;
        const double dp = mp > 0 ? (np - mp) / mp : 0.0, dm = mm > 0 ? (nm - mm) / mm : 0.0;
        v[0] += dd * dd;
        v[1] += dp;
        v[2] += dm;
        v[3] += dp * dp;
        v[4] += dm * dm;
        v[5] += dp * dm;
        v[6] += (nm - mm) * (nm - mm);
    }
    block_reduce_store<7>(v, partial + (int64_t)blockIdx.x * 7);
}

// out[0..6] = the sums above, out[7], out[8] = the species totals (as doubles)
__global__ void k_seg_totals(const int32_t* __restrict__ totals, double* __restrict__ out) {
    if (threadIdx.x < 2) out[7 + threadIdx.x] = (double)totals[threadIdx.x];
}

hipError_t launch_segregation(int64_t n, const uint32_t* pos, const int8_t* species, int32_t n_cells,
                              int32_t* counts, double* scratch, double* out, hipStream_t s) {
    const int64_t cells = (int64_t)n_cells * n_cells * n_cells;
    hipError_t e = launch_fill_i32(counts, 2 * cells + 2, 0, s);
    if (e != hipSuccess) return e;
    int32_t* totals = counts + 2 * cells;
    if (n > 0)
        k_seg_count<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(pos, species, n, n_cells, counts, totals);
    k_seg_moments<<<dim3(kDiagBlocks), dim3(256), 0, s>>>(counts, totals, cells, scratch);
    k_diag_final<<<dim3(1), dim3(256), 0, s>>>(scratch, kDiagBlocks, 7, out);
    k_seg_totals<<<dim3(1), dim3(64), 0, s>>>(totals, out);
    return hipGetLastError();
}

}  // namespace jtp
