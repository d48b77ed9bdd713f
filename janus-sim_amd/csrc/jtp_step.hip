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
