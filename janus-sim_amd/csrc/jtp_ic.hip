// Zel'dovich initial conditions on the GPU (one particle per cell of an n^3 lattice, two species).
//
// Reference behaviour replaced: generate_zeldovich_ics / generate_displacement_field,
// src/ic_gen.rs:51-169, 179-309 (host loops, rustfft, StdRng):
//   * delta(k) = (re + i im) e^{i phase}, re, im ~ N(0, 1) sqrt(P(k) / 2) D, phase ~ U[0, 2 pi),
//     P(k) = k^(n_s - 4) delta_rms^2, delta(0) = 0 (ic_gen.rs:196-221);
//   * psi_i = Re IFFT[-i k_i delta(k) / k^2] / n^3 with the reference's signed k per index
//     (ic_gen.rs:224-296);
//   * displacements scaled so that max |psi| = 0.3 lattice spacing, velocities psi x H0 sqrt(1 + z)
//     (ic_gen.rs:82-95); lattice at cell centres, periodic wrap (ic_gen.rs:117-149);
//   * a uniformly random set of round(n^3 / (1 + mu)) cells holds m+, the rest m- (ic_gen.rs:58-60,
//     99-110).
// The real part of a complex inverse FFT equals the inverse real FFT of the input's Hermitian part,
// H(k) = (F(k) + conj(F(k')))/2 with k' the index of -k, so each component is produced as a half
// spectrum and run through the Poisson solver's inverse z, y and C2R x passes.
// Random numbers: SplitMix64 (output j of the generator seeded with s is mix(s + (j + 1) gamma)),
// three per mode at j = 3 m .. 3 m + 2 for mode index m = (iz G + iy) G + ix: a uniform phase and a
// Box-Muller pair.  The sign keys are output m of the generator seeded with seed + 12345; the n+
// cells with the smallest (key, index) are m+.  These streams are not the reference's StdRng
// (ChaCha12) streams, so a given seed gives other fields with the same statistics.
#include "jtp_internal.h"

#include <cmath>

namespace jtp {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

__host__ __device__ inline uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// output j (0-based) of SplitMix64 seeded with s
__host__ __device__ inline uint64_t splitmix_at(uint64_t s, uint64_t j) { return splitmix_mix(s + (j + 1) * kGamma); }

__device__ inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

__device__ inline double signed_k(int i, int G, double dk) { return (i <= G / 2 ? (double)i : (double)(i - G)) * dk; }

// F(k) M_dim(k) for the full-grid mode (ix, iy, iz), M = -i k_dim / k^2 (zero at k = 0)
__device__ inline Cplx<double> mode_term(uint64_t seed, int ix, int iy, int iz, int G, double dk, double n_s,
                                         double drms2, double growth, int dim) {
    const double kx = signed_k(ix, G, dk), ky = signed_k(iy, G, dk), kz = signed_k(iz, G, dk);
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (!(k2 > 0)) return {0.0, 0.0};
    const uint64_t m = ((uint64_t)iz * G + iy) * G + ix;
    const double phase = u01(splitmix_at(seed, 3 * m)) * 2.0 * M_PI;
    const double u1 = u01(splitmix_at(seed, 3 * m + 1)), u2 = u01(splitmix_at(seed, 3 * m + 2));
    const double rad = sqrt(-2.0 * log(1.0 - u1));
    const double amp = sqrt(pow(sqrt(k2), n_s - 4.0) * drms2 / 2.0) * growth;
    const double re = rad * cos(2.0 * M_PI * u2) * amp, im = rad * sin(2.0 * M_PI * u2) * amp;
    const double c = cos(phase), s = sin(phase);
    const double fr = re * c - im * s, fi = re * s + im * c;
    const double kd = dim == 0 ? kx : (dim == 1 ? ky : kz);
    const double f = -kd / k2;  // F * (i f): (fr + i fi)(i f) = -fi f + i fr f
    return {-fi * f, fr * f};
}

// Half spectrum [z][y][kx], kx in [0, G/2], of component dim, times 1 / G^3.
template <typename R>
__global__ __launch_bounds__(256) void k_ic_spectrum(Cplx<R>* __restrict__ spec, int logG, uint64_t seed, double dk,
                                                     double n_s, double drms2, double growth, int dim) {
    const int G = 1 << logG, H = G / 2 + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G * G * H) return;
    const int kx = (int)(t % H), ky = (int)((t / H) % G), kz = (int)(t / ((int64_t)H * G));
    const Cplx<double> a = mode_term(seed, kx, ky, kz, G, dk, n_s, drms2, growth, dim);
    const Cplx<double> b = mode_term(seed, (G - kx) & (G - 1), (G - ky) & (G - 1), (G - kz) & (G - 1), G, dk, n_s,
                                     drms2, growth, dim);
    const double inv = 1.0 / ((double)G * G * G);
    spec[t] = {(R)(0.5 * (a.x + b.x) * inv), (R)(0.5 * (a.y - b.y) * inv)};
}

// max over cells of |psi|^2, per block, then one block over the partials
template <typename R>
__global__ __launch_bounds__(256) void k_ic_maxdisp(const R* __restrict__ px, const R* __restrict__ py,
                                                    const R* __restrict__ pz, int64_t m, double* __restrict__ part) {
    __shared__ double lds[256];
    double v = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const double d = (double)px[i] * px[i] + (double)py[i] * py[i] + (double)pz[i] * pz[i];
        v = d > v ? d : v;
    }
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}

__global__ __launch_bounds__(256) void k_ic_maxfinal(const double* __restrict__ part, int nb, double* __restrict__ out) {
    __shared__ double lds[256];
    double v = 0;
    for (int b = threadIdx.x; b < nb; b += 256) v = fmax(v, part[b]);
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(lds[0]);
}

__device__ inline uint32_t fixed_of(double x, double L) {
    double t = x / L + 0.5;
    t -= floor(t);
    uint64_t w = (uint64_t)(t * 4294967296.0);
    if (w >= 4294967296ull) w -= 4294967296ull;
    return (uint32_t)w;
}

// Particle m = cell m: lattice centre + psi scale, wrapped; velocity psi scale vel_scale.
template <typename R>
__global__ __launch_bounds__(256) void k_ic_place(const R* __restrict__ px, const R* __restrict__ py,
                                                  const R* __restrict__ pz, const double* __restrict__ maxd, int logG,
                                                  double L, double vel_scale, uint32_t* __restrict__ pos,
                                                  void* __restrict__ vel, int vel_f64) {
    const int G = 1 << logG;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (int64_t)G * G * G) return;
    const double spacing = L / G, half = L / 2;
    const double md = maxd[0];
    const double scale = md > 1e-10 ? spacing * 0.3 / md : 1.0;
    const int ix = (int)(m % G), iy = (int)((m / G) % G), iz = (int)(m / ((int64_t)G * G));
    const double ps[3] = {(double)px[m] * scale, (double)py[m] * scale, (double)pz[m] * scale};
    const int ii[3] = {ix, iy, iz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double x = ((double)ii[a] + 0.5) * spacing - half + ps[a];
        if (x > half) x -= L;
        if (x < -half) x += L;
        pos[3 * m + a] = fixed_of(x, L);
        if (vel_f64) ((double*)vel)[3 * m + a] = ps[a] * vel_scale;
        else ((float*)vel)[3 * m + a] = (float)(ps[a] * vel_scale);
    }
}

__global__ __launch_bounds__(256) void k_ic_sign_keys(uint64_t* __restrict__ keys, int32_t* __restrict__ vals,
                                                      int64_t m, uint64_t seed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    keys[i] = splitmix_at(seed, (uint64_t)i);
    vals[i] = (int32_t)i;
}

__global__ __launch_bounds__(256) void k_ic_signs(const int32_t* __restrict__ order, int64_t m, int64_t n_plus,
                                                  int8_t* __restrict__ species) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    species[order[r]] = r < n_plus ? 1 : -1;
}

template <typename R>
static hipError_t ic_component(Ctx& c, const jtp_ic_params& p, int dim, R* out, hipStream_t s) {
    const int G = (int)c.G(), H = G / 2 + 1;
    const int64_t nh = (int64_t)G * G * H;
    const double dk = 2.0 * M_PI / p.box_size;
    const double growth = 1.0 / (1.0 + p.z_init);  // D = a_init (ic_gen.rs:75-76)
    k_ic_spectrum<R><<<dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s>>>(
        (Cplx<R>*)c.spec, c.logG, p.seed, dk, p.n_s, p.delta_rms * p.delta_rms, growth, dim);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return pm_inverse_fft(c, out, s);
}

template <typename R>
static hipError_t generate_t(Ctx& c, const jtp_ic_params& p, uint32_t* pos, void* vel, int vel_f64, int8_t* species,
                             int64_t n_plus, hipStream_t s) {
    const int64_t m = c.G3();
    R* px = (R*)c.rho;
    R* py = (R*)c.phi;
    R* pz = (R*)c.grad;
    hipError_t e = ic_component<R>(c, p, 0, px, s);
    if (e == hipSuccess) e = ic_component<R>(c, p, 1, py, s);
    if (e == hipSuccess) e = ic_component<R>(c, p, 2, pz, s);
    if (e != hipSuccess) return e;
    double* part = c.diag_scratch;
    double* maxd = c.diag_scratch + 1024 * 8;
    k_ic_maxdisp<R><<<dim3(1024), dim3(256), 0, s>>>(px, py, pz, m, part);
    k_ic_maxfinal<<<dim3(1), dim3(256), 0, s>>>(part, 1024, maxd);
    const double vel_scale = 69.9 / 977.8 * std::sqrt(1.0 + p.z_init);  // ic_gen.rs:88-91
    const unsigned nb = (unsigned)((m + 255) / 256);
    k_ic_place<R><<<dim3(nb), dim3(256), 0, s>>>(px, py, pz, maxd, c.logG, p.box_size, vel_scale, pos, vel, vel_f64);
    k_ic_sign_keys<<<dim3(nb), dim3(256), 0, s>>>(c.keys0, c.vals0, m, p.seed + 12345);
    e = hipGetLastError();
    if (e == hipSuccess) e = radix_sort_pairs(c, m, s);
    if (e != hipSuccess) return e;
    k_ic_signs<<<dim3(nb), dim3(256), 0, s>>>(c.vals0, m, n_plus, species);
    return hipGetLastError();
}

hipError_t generate_ics(Ctx& c, const jtp_ic_params& p, uint32_t* pos, void* vel, int vel_f64, int8_t* species,
                        int64_t n_plus, hipStream_t s) {
    if (c.p.precision == JTP_FP64) return generate_t<double>(c, p, pos, vel, vel_f64, species, n_plus, s);
    return generate_t<float>(c, p, pos, vel, vel_f64, species, n_plus, s);
}

}  // namespace jtp
