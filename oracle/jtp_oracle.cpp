// ORACLE — test infrastructure only.  CPU (FP64 / integer) restatements of the
// reference TreePM path, used by tests/ and by bench.py's cpu_baseline leg as
// the checker.  Nothing in the product library links or calls this code.
//
// Contents and the reference code each part restates:
//   orc_fixed_from_f64     box convention of pm_grid.rs:64-70 as 32-bit fixed point
//   orc_morton_fixed       21-bit/axis interleave of compute_morton_all (nbody_gpu_twopass.rs:1466-1484)
//   orc_morton_f32_ref     compute_morton_all on f32 inputs, same FP32 arithmetic (:1458-1462)
//   orc_radix_sort_ref     stable LSD radix sort, 8 passes x 8 bits (radix_histogram/prefix/scatter :1521-1658)
//   orc_karras_ref         build_bvh_tp / delta_tp / range_tp / split_tp (:356-447)
//   orc_node_sums          per-node (n+, sum x+, n-, sum x-) and AABB by brute force over the node range
//   orc_deposit            PmGrid::assign_mass (pm_grid.rs:63-103) on one net grid, fixed-point fractions
//   orc_fft3d              complex 3-D FFT by 1-D passes along x, y, z (pm_grid.rs:146-181,230-263)
//   orc_poisson            solve_poisson_filtered Green function (pm_grid.rs:183-228, cufft_wrapper.cu:160-207)
//                          plus optional CIC deconvolution; density = mass / V_cell
//   orc_interp             interpolate_force (pm_grid.rs:275-359) with 2- or 4-point differences, net potential
//   orc_short_direct       theta = 0 limit of the short-range walk: minimum-image direct sum, same pair law
//   orc_ewald              FP64 Ewald direct summation with Plummer correction of the nearest image
//   orc_refcpu_*           TreePMForce (treepm_force.rs:35-131) + TreePMTree (tree_short.rs:49-247) +
//                          PmGrid two-grid solve, restated for the CPU baseline and compat parity
//   orc_splitmix64         SplitMix64 (Steele, Lea, Flood 2014; Vigna's reference C), the IC random stream
//   orc_ic                 generate_zeldovich_ics / generate_displacement_field (src/ic_gen.rs:51-169,
//                          179-309) step by step with the SplitMix64 streams in place of StdRng: full
//                          complex spectrum, complex inverse FFT per component, real part / n^3
//   orc_segregation*       compute_segregation_metrics (src/bin/janus_mu19_production.rs:219-273): the
//                          _f32_ref form bins f32 positions exactly as the reference does, the fixed form
//                          bins 32-bit fixed-point coordinates by floor(u n / 2^32)
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <complex>
#include <omp.h>

typedef std::complex<double> cd;

static inline double fixed_scale(double L) { return L / 4294967296.0; }
static inline double fixed_disp(uint32_t uj, uint32_t ui, double L) { return (double)(int32_t)(uj - ui) * fixed_scale(L); }

static inline uint64_t spread3(uint64_t v) {
    // bit-by-bit interleave (independent of the magic-mask form used by the kernels)
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (3 * b);
    return r;
}

extern "C" {

// ------------------------------------------------------------------ fixed point, keys, sort
void orc_fixed_from_f64(int64_t n3, const double* x, double L, uint32_t* out) {
    for (int64_t i = 0; i < n3; ++i) {
        double t = x[i] / L + 0.5;
        t -= std::floor(t);
        uint64_t w = (uint64_t)(t * 4294967296.0);
        if (w >= 4294967296ull) w -= 4294967296ull;
        out[i] = (uint32_t)w;
    }
}

void orc_morton_fixed(int64_t n, const uint32_t* pos, uint64_t* keys) {
    for (int64_t i = 0; i < n; ++i)
        keys[i] = spread3(pos[3 * i] >> 11) | (spread3(pos[3 * i + 1] >> 11) << 1) | (spread3(pos[3 * i + 2] >> 11) << 2);
}

void orc_morton_f32_ref(int64_t n, const float* pos, float box_half, float inv_cell, uint64_t* keys) {
    for (int64_t i = 0; i < n; ++i) {
        uint64_t c[3];
        for (int a = 0; a < 3; ++a) {
            volatile float s = pos[3 * i + a] + box_half;   // no contraction: add, then multiply
            volatile float g = s * inv_cell;
            float cl = std::fmin(std::fmax((float)g, 0.0f), 2097151.0f);
            c[a] = (uint64_t)(uint32_t)cl;
        }
        keys[i] = spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
    }
}

// compute_morton_f32 (nbody_gpu_twopass.rs:302-326): scale = 1024 / (2 box_half) in f32, per axis
// (unsigned)((x + box_half) * scale) (CUDA's conversion saturates: negative and NaN -> 0), min 1023,
// the bit loop expand10_tp, key = mc << 32 | i.
void orc_morton30_f32_ref(int64_t n, const float* pos, float box_half, uint64_t* keys, int32_t* idx) {
    volatile float two_bh = 2.0f * box_half;
    volatile float scale = 1024.0f / two_bh;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t mc = 0;
        for (int a = 0; a < 3; ++a) {
            volatile float s = pos[3 * i + a] + box_half;
            volatile float g = s * scale;
            const float x = g;
            uint32_t c;
            if (!(x > 0.0f)) c = 0;
            else if (x >= 4294967296.0f) c = 0xffffffffu;
            else c = (uint32_t)x;
            if (c > 0x3ffu) c = 0x3ffu;
            uint64_t r = 0;
            for (int b = 0; b < 10; ++b) r |= (uint64_t)((c >> b) & 1u) << (3 * b);
            mc |= r << a;
        }
        keys[i] = (mc << 32) | (uint64_t)(uint32_t)i;
        idx[i] = (int32_t)i;
    }
}

void orc_radix_sort_ref(int64_t n, const uint64_t* keys_in, int32_t* perm) {
    std::vector<uint64_t> ka(keys_in, keys_in + n), kb(n);
    std::vector<int32_t> va(n), vb(n);
    for (int64_t i = 0; i < n; ++i) va[i] = (int32_t)i;
    for (int pass = 0; pass < 8; ++pass) {
        const int sh = 8 * pass;
        uint64_t cnt[257] = {0};
        for (int64_t i = 0; i < n; ++i) cnt[((ka[i] >> sh) & 255) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (int64_t i = 0; i < n; ++i) {
            const uint64_t d = (ka[i] >> sh) & 255;
            const uint64_t o = cnt[d]++;
            kb[o] = ka[i];
            vb[o] = va[i];
        }
        ka.swap(kb);
        va.swap(vb);
    }
    std::memcpy(perm, va.data(), n * sizeof(int32_t));
}

static inline int clz64(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }
static inline int kdelta(const uint64_t* k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    if (k[i] == k[j]) return 64 + clz64((uint64_t)(i ^ j));
    return clz64(k[i] ^ k[j]);
}

// nodes4[i] = {first, last, left, right}; a leaf child is encoded ~sorted_index
// (the reference stores it as n - 1 + index).
void orc_karras_ref(int64_t n, const uint64_t* k, int32_t* nodes4) {
    for (int64_t i = 0; i + 1 < n; ++i) {
        int64_t first, last;
        if (i == 0) {
            first = 0;
            last = n - 1;
        } else {
            const int dl = kdelta(k, n, i, i - 1), dr = kdelta(k, n, i, i + 1);
            const int64_t d = (dr > dl) ? 1 : -1;
            const int dmin = (d > 0) ? dl : dr;
            int64_t lmax = 2;
            while (kdelta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
            int64_t l = 0;
            for (int64_t t = lmax / 2; t >= 1; t /= 2)
                if (kdelta(k, n, i, i + (l + t) * d) > dmin) l += t;
            const int64_t j = i + l * d;
            first = d > 0 ? i : j;
            last = d > 0 ? j : i;
        }
        const int dn = kdelta(k, n, first, last);
        int64_t s = 0, t = last - first;
        while (t > 1) {
            t = (t + 1) / 2;
            if (kdelta(k, n, first, first + s + t) > dn) s += t;
        }
        const int64_t g = first + s;
        const int64_t ref_left = (first == g) ? (n - 1 + first) : g;
        const int64_t ref_right = (g + 1 == last) ? (n - 1 + last) : g + 1;
        nodes4[4 * i] = (int32_t)first;
        nodes4[4 * i + 1] = (int32_t)last;
        nodes4[4 * i + 2] = ref_left >= n - 1 ? ~(int32_t)(ref_left - (n - 1)) : (int32_t)ref_left;
        nodes4[4 * i + 3] = ref_right >= n - 1 ? ~(int32_t)(ref_right - (n - 1)) : (int32_t)ref_right;
    }
}

// sums8[i] = {n+, sx+, sy+, sz+, n-, sx-, sy-, sz-}; boxes6[i] = {min xyz, max xyz}
void orc_node_sums(int64_t n, const uint32_t* spos, const int8_t* ssp, const int32_t* nodes4, uint64_t* sums8,
                   uint32_t* boxes6) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n - 1; ++i) {
        uint64_t s[8] = {0};
        uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0, 0, 0};
        for (int64_t p = nodes4[4 * i]; p <= nodes4[4 * i + 1]; ++p) {
            const int o = ssp[p] > 0 ? 0 : 4;
            s[o] += 1;
            for (int a = 0; a < 3; ++a) {
                const uint32_t u = spos[3 * p + a];
                s[o + 1 + a] += u;
                mn[a] = std::min(mn[a], u);
                mx[a] = std::max(mx[a], u);
            }
        }
        std::memcpy(sums8 + 8 * i, s, sizeof(s));
        for (int a = 0; a < 3; ++a) {
            boxes6[6 * i + a] = mn[a];
            boxes6[6 * i + 3 + a] = mx[a];
        }
    }
}

// ------------------------------------------------------------------ PM
// rho[i + G(j + G k)] += q W, q = +mp for species +1, -mm for species -1.
void orc_deposit(int64_t n, const uint32_t* pos, const int8_t* sp, double mp, double mm, int G, double* rho) {
    int logG = 0;
    while ((1 << logG) < G) ++logG;
    const int fb = 32 - logG;
    std::fill(rho, rho + (int64_t)G * G * G, 0.0);
    for (int64_t p = 0; p < n; ++p) {
        int ic[3];
        double f[3];
        for (int a = 0; a < 3; ++a) {
            const uint32_t u = pos[3 * p + a];
            ic[a] = (int)(u >> fb);
            f[a] = (double)(u & ((1u << fb) - 1u)) / (double)(1ull << fb);
        }
        const double q = sp[p] > 0 ? mp : -mm;
        for (int dk = 0; dk < 2; ++dk)
            for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di) {
                    const int64_t ci = (ic[0] + di) % G, cj = (ic[1] + dj) % G, ck = (ic[2] + dk) % G;
                    const double w = (di ? f[0] : 1.0 - f[0]) * (dj ? f[1] : 1.0 - f[1]) * (dk ? f[2] : 1.0 - f[2]);
                    rho[ci + G * (cj + G * ck)] += q * w;
                }
    }
}

static void fft1d(cd* a, int n, bool inv) {
    // iterative radix-2 Cooley-Tukey, unnormalised; forward uses e^{-2 pi i jk/n}
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = 2.0 * M_PI / len * (inv ? 1 : -1);
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                const cd w(std::cos(ang * k), std::sin(ang * k));
                const cd u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

// data[i + G(j + G k)], passes along x, then y, then z (inverse: z, y, x)
static void fft3d(std::vector<cd>& d, int G, bool inv) {
    const int64_t G2 = (int64_t)G * G;
    auto pass = [&](int axis) {
#pragma omp parallel
        {
            std::vector<cd> line(G);
#pragma omp for
            for (int64_t q = 0; q < G2; ++q) {
                int64_t base, st;
                if (axis == 0) { base = q * G; st = 1; }
                else if (axis == 1) { base = (q % G) + G2 * (q / G); st = G; }
                else { base = q; st = G2; }
                for (int t = 0; t < G; ++t) line[t] = d[base + t * st];
                fft1d(line.data(), G, inv);
                for (int t = 0; t < G; ++t) d[base + t * st] = line[t];
            }
        }
    };
    if (!inv) { pass(0); pass(1); pass(2); }
    else { pass(2); pass(1); pass(0); }
}

void orc_fft3d(int G, double* re, double* im, int inverse) {
    const int64_t G3 = (int64_t)G * G * G;
    std::vector<cd> d(G3);
    for (int64_t i = 0; i < G3; ++i) d[i] = cd(re[i], im[i]);
    fft3d(d, G, inverse != 0);
    for (int64_t i = 0; i < G3; ++i) { re[i] = d[i].real(); im[i] = d[i].imag(); }
}

// phi = IFFT[ -4 pi g / k^2 * exp(-k^2 r_s^2) / W(k)^2 * FFT[rho / V_cell] ], W = prod sinc^2(k_i h/2)
// (W only when deconv != 0); DC -> 0; max|k_i| < k_min -> 0.  match >= 0 additionally multiplies by
// (k . D) / max(|D|^2, |k|^2 / 4) with D the transfer function of the gradient stencil (0: 4-point,
// 1: 2-point); the floor keeps the factor bounded where D vanishes (the Nyquist planes).
void orc_poisson(int G, double L, double g, double r_s, int k_min, int deconv, int match, const double* rho,
                 double* phi) {
    const int64_t G3 = (int64_t)G * G * G;
    const double h = L / G, vcell = h * h * h;
    std::vector<cd> d(G3);
    for (int64_t i = 0; i < G3; ++i) d[i] = cd(rho[i] / vcell, 0.0);
    fft3d(d, G, false);
    const double dk = 2.0 * M_PI / L;
#pragma omp parallel for
    for (int64_t k = 0; k < G; ++k)
        for (int64_t j = 0; j < G; ++j)
            for (int64_t i = 0; i < G; ++i) {
                const int64_t idx = i + G * (j + G * k);
                const double ki = (double)(i <= G / 2 ? i : i - G), kj = (double)(j <= G / 2 ? j : j - G),
                             kk = (double)(k <= G / 2 ? k : k - G);
                const int kint = (int)std::max(std::fabs(ki), std::max(std::fabs(kj), std::fabs(kk)));
                const double k2 = (ki * ki + kj * kj + kk * kk) * dk * dk;
                if (k2 < 1e-300 || kint < k_min) { d[idx] = 0.0; continue; }
                double gr = -4.0 * M_PI * g / k2;
                if (r_s > 0) gr *= std::exp(-k2 * r_s * r_s);
                if (deconv) {
                    double w = 1.0;
                    for (double kc : {ki, kj, kk}) {
                        if (kc != 0) {
                            const double x = 0.5 * kc * dk * h;
                            const double s = std::sin(x) / x;
                            w *= s * s;
                        }
                    }
                    gr /= w * w;
                }
                if (match >= 0) {
                    double kd = 0, dd = 0, kk2 = 0;
                    for (double kc : {ki, kj, kk}) {
                        const double kv = kc * dk;
                        const double D = match == 0 ? (8.0 * std::sin(kv * h) - std::sin(2.0 * kv * h)) / (6.0 * h)
                                                    : std::sin(kv * h) / h;
                        kd += kv * D;
                        dd += D * D;
                        kk2 += kv * kv;
                    }
                    gr *= kd / std::max(dd, 0.25 * kk2);
                }
                d[idx] *= gr;
            }
    fft3d(d, G, true);
    for (int64_t i = 0; i < G3; ++i) phi[i] = d[i].real() / (double)G3;
}

// a_i = -sigma_i sum_corners W grad(phi)
void orc_interp(int64_t n, const uint32_t* pos, const int8_t* sp, int G, double L, const double* phi, int four,
                double* acc) {
    int logG = 0;
    while ((1 << logG) < G) ++logG;
    const int fb = 32 - logG;
    const double h = L / G;
    auto at = [&](int64_t i, int64_t j, int64_t k) {
        i = ((i % G) + G) % G; j = ((j % G) + G) % G; k = ((k % G) + G) % G;
        return phi[i + G * (j + G * k)];
    };
    auto der = [&](int64_t i, int64_t j, int64_t k, int ax) {
        int64_t e[3] = {0, 0, 0};
        e[ax] = 1;
        const double p1 = at(i + e[0], j + e[1], k + e[2]), m1 = at(i - e[0], j - e[1], k - e[2]);
        if (!four) return (p1 - m1) / (2.0 * h);
        const double p2 = at(i + 2 * e[0], j + 2 * e[1], k + 2 * e[2]), m2 = at(i - 2 * e[0], j - 2 * e[1], k - 2 * e[2]);
        return (8.0 * (p1 - m1) - (p2 - m2)) / (12.0 * h);
    };
#pragma omp parallel for
    for (int64_t p = 0; p < n; ++p) {
        int64_t ic[3];
        double f[3];
        for (int a = 0; a < 3; ++a) {
            const uint32_t u = pos[3 * p + a];
            ic[a] = u >> fb;
            f[a] = (double)(u & ((1u << fb) - 1u)) / (double)(1ull << fb);
        }
        double a3[3] = {0, 0, 0};
        for (int dk = 0; dk < 2; ++dk)
            for (int dj = 0; dj < 2; ++dj)
                for (int di = 0; di < 2; ++di) {
                    const double w = (di ? f[0] : 1 - f[0]) * (dj ? f[1] : 1 - f[1]) * (dk ? f[2] : 1 - f[2]);
                    for (int ax = 0; ax < 3; ++ax) a3[ax] += w * der(ic[0] + di, ic[1] + dj, ic[2] + dk, ax);
                }
        const double sg = sp[p] > 0 ? -1.0 : 1.0;
        for (int ax = 0; ax < 3; ++ax) acc[3 * p + ax] = sg * a3[ax];
    }
}

// ------------------------------------------------------------------ short range, theta = 0
static inline double split_factor(int mode, double r2, double r2s, double rcut2, double rs) {
    if (mode == 2) {
        const double u = r2 / rcut2;
        return 1.0 - u * u;
    }
    const double x = std::sqrt(r2s) / (2.0 * rs);
    double S = std::erfc(x);
    if (mode == 0) S += 2.0 / std::sqrt(M_PI) * x * std::exp(-x * x);
    return S;
}

}  // extern "C"

struct CellList {
    int nc = 1;
    std::vector<int64_t> start, idx;
    int reach = 1;  // neighbour cells searched on each side
    void build(int64_t n, const uint32_t* pos, double L, double rcut, int sub = 1) {
        // cells of edge >= rcut / sub, searched +-sub cells around the target's cell
        reach = sub;
        nc = (int)std::floor(sub * L / rcut);
        if (nc < 2 * sub + 1) { nc = 1; reach = 0; }
        if (nc > 256) nc = 256;
        const int64_t nc3 = (int64_t)nc * nc * nc;
        start.assign(nc3 + 1, 0);
        idx.resize(n);
        std::vector<int64_t> cell(n);
        for (int64_t p = 0; p < n; ++p) {
            int64_t c[3];
            for (int a = 0; a < 3; ++a) c[a] = (int64_t)(((uint64_t)pos[3 * p + a] * (uint64_t)nc) >> 32);
            cell[p] = c[0] + nc * (c[1] + nc * c[2]);
            start[cell[p] + 1]++;
        }
        for (int64_t c = 0; c < nc3; ++c) start[c + 1] += start[c];
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t p = 0; p < n; ++p) idx[fill[cell[p]]++] = p;
    }
    template <class F>
    void for_neighbours(const uint32_t* u, F&& f) const {
        const int64_t c = cell_of(u);
        const int64_t cx = c % nc, cy = (c / nc) % nc, cz = c / ((int64_t)nc * nc);
        for (int dz = -reach; dz <= reach; ++dz)
            for (int dy = -reach; dy <= reach; ++dy)
                for (int dx = -reach; dx <= reach; ++dx) {
                    const int64_t q = ((cx + dx + nc) % nc) + nc * (((cy + dy + nc) % nc) + nc * ((cz + dz + nc) % nc));
                    for (int64_t s = start[q]; s < start[q + 1]; ++s) f(idx[s]);
                }
    }
    int64_t cell_of(const uint32_t* u) const {
        int64_t c[3];
        for (int a = 0; a < 3; ++a) c[a] = (int64_t)(((uint64_t)u[a] * (uint64_t)nc) >> 32);
        return c[0] + nc * (c[1] + nc * c[2]);
    }
};

extern "C" {

void orc_short_direct(int64_t n, const uint32_t* pos, const int8_t* sp, double qp, double qm, double L, double eps,
                      double rcut, double rs, int mode, int64_t nt, const int64_t* tidx, double* acc) {
    CellList cl;
    cl.build(n, pos, L, rcut, 2);
    const int nc = cl.nc;
    const double rcut2 = rcut * rcut, eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t i = tidx ? tidx[t] : t;
        const uint32_t* ui = pos + 3 * i;
        double a[3] = {0, 0, 0};
        auto visit = [&](int64_t j) {
            const double d[3] = {fixed_disp(pos[3 * j], ui[0], L), fixed_disp(pos[3 * j + 1], ui[1], L),
                                 fixed_disp(pos[3 * j + 2], ui[2], L)};
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(r2 < rcut2)) return;
            const double r2s = r2 + eps2;
            const double f = (sp[j] > 0 ? qp : -qm) * split_factor(mode, r2, r2s, rcut2, rs) / (r2s * std::sqrt(r2s));
            for (int ax = 0; ax < 3; ++ax) a[ax] += f * d[ax];
        };
        if (nc == 1) {
            for (int64_t j = 0; j < n; ++j) visit(j);
        } else {
            cl.for_neighbours(ui, visit);
        }
        const double sg = sp[i] > 0 ? 1.0 : -1.0;
        for (int ax = 0; ax < 3; ++ax) acc[3 * t + ax] = sg * a[ax];
    }
}

// ------------------------------------------------------------------ Ewald
// Periodic force on target i from every source j (signed charge q_j = sigma_j G |m_j|):
//   a_i = sigma_i sum_j q_j [ sum_n B(|d+nL|) (d+nL)/|d+nL|^3                          (real space, |d+nL| < rc)
//                            + (Plummer - Newton)(d_min)                                 (softened nearest image)
//                            - (4 pi / L^3) sum_{k != 0} k/k^2 e^{-k^2/4 a^2} sin(k.(x_i - x_j)) ]
// with d = x_j - x_i and B(r) = erfc(a r) + 2 a r / sqrt(pi) e^{-a^2 r^2}; the k = 0 term is dropped
// (the Plummer correction is applied to nearest images inside rc; beyond rc it is below 1.5 eps^2/rc^2
// of the pair force)
// (uniform neutralising background per species).  Reciprocal space uses structure factors
// C(k) = sum q cos(k.x), S(k) = sum q sin(k.x) over half the k-sphere |n| <= nk.
void orc_ewald(int64_t n, const uint32_t* pos, const int8_t* sp, double qp, double qm, double L, double eps,
               double alpha, double rc, int nk, int64_t nt, const int64_t* tidx, double* acc) {
    const double sc = fixed_scale(L);
    // reciprocal vectors (half space: first nonzero component positive)
    std::vector<int> kv;
    for (int a = -nk; a <= nk; ++a)
        for (int b = -nk; b <= nk; ++b)
            for (int c = -nk; c <= nk; ++c) {
                if (a * a + b * b + c * c == 0 || a * a + b * b + c * c > nk * nk) continue;
                if (a < 0 || (a == 0 && b < 0) || (a == 0 && b == 0 && c < 0)) continue;
                kv.push_back(a); kv.push_back(b); kv.push_back(c);
            }
    const int64_t nkv = (int64_t)kv.size() / 3;
    const double tpl = 2.0 * M_PI / L;
    std::vector<double> Ck(nkv, 0.0), Sk(nkv, 0.0);
    const int nth = omp_get_max_threads();
    std::vector<std::vector<double>> Cl(nth, std::vector<double>(nkv, 0.0)), Sl(nth, std::vector<double>(nkv, 0.0));
#pragma omp parallel
    {
        const int th = omp_get_thread_num();
        std::vector<cd> ex(2 * nk + 1), ey(2 * nk + 1), ez(2 * nk + 1);
        double* C = Cl[th].data();
        double* S = Sl[th].data();
#pragma omp for schedule(static)
        for (int64_t j = 0; j < n; ++j) {
            const double q = sp[j] > 0 ? qp : -qm;
            const double x[3] = {pos[3 * j] * sc, pos[3 * j + 1] * sc, pos[3 * j + 2] * sc};
            cd* e[3] = {ex.data(), ey.data(), ez.data()};
            for (int a = 0; a < 3; ++a) {
                const cd w(std::cos(tpl * x[a]), std::sin(tpl * x[a]));
                e[a][nk] = 1.0;
                for (int m = 1; m <= nk; ++m) {
                    e[a][nk + m] = e[a][nk + m - 1] * w;
                    e[a][nk - m] = std::conj(e[a][nk + m]);
                }
            }
            for (int64_t v = 0; v < nkv; ++v) {
                const cd z = ex[nk + kv[3 * v]] * ey[nk + kv[3 * v + 1]] * ez[nk + kv[3 * v + 2]];
                C[v] += q * z.real();
                S[v] += q * z.imag();
            }
        }
    }
    for (int th = 0; th < nth; ++th)
        for (int64_t v = 0; v < nkv; ++v) { Ck[v] += Cl[th][v]; Sk[v] += Sl[th][v]; }
    std::vector<double> kfac(nkv);
    for (int64_t v = 0; v < nkv; ++v) {
        const double k2 = tpl * tpl * (double)(kv[3 * v] * kv[3 * v] + kv[3 * v + 1] * kv[3 * v + 1] + kv[3 * v + 2] * kv[3 * v + 2]);
        // half-space sum: each +-k pair counted once, so factor 2
        kfac[v] = 2.0 * 4.0 * M_PI / (L * L * L) * std::exp(-k2 / (4.0 * alpha * alpha)) / k2;
    }
    // real space via cell list with cell >= rc
    CellList cl;
    cl.build(n, pos, L, rc, 2);
    const int nc = cl.nc;
    // image shells beyond the minimum image (none needed when rc <= L/2)
    const int nimg = rc <= 0.5 * L ? 0 : (int)std::ceil(rc / L);
    const double rc2 = rc * rc, eps2 = eps * eps;
    const double two_over_sqrtpi = 2.0 / std::sqrt(M_PI);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t i = tidx ? tidx[t] : t;
        const uint32_t* ui = pos + 3 * i;
        double a[3] = {0, 0, 0};
        auto visit = [&](int64_t j) {
            const double q = sp[j] > 0 ? qp : -qm;
            const double d0[3] = {fixed_disp(pos[3 * j], ui[0], L), fixed_disp(pos[3 * j + 1], ui[1], L),
                                  fixed_disp(pos[3 * j + 2], ui[2], L)};
            for (int ix = -nimg; ix <= nimg; ++ix)
                for (int iy = -nimg; iy <= nimg; ++iy)
                    for (int iz = -nimg; iz <= nimg; ++iz) {
                        const double d[3] = {d0[0] + ix * L, d0[1] + iy * L, d0[2] + iz * L};
                        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        if (r2 >= rc2 || r2 == 0.0) continue;
                        const double r = std::sqrt(r2);
                        double f = (std::erfc(alpha * r) + two_over_sqrtpi * alpha * r * std::exp(-alpha * alpha * r2)) / (r2 * r);
                        if (ix == 0 && iy == 0 && iz == 0) {
                            const double r2s = r2 + eps2;
                            f += 1.0 / (r2s * std::sqrt(r2s)) - 1.0 / (r2 * r);
                        }
                        f *= q;
                        for (int ax = 0; ax < 3; ++ax) a[ax] += f * d[ax];
                    }
        };
        if (nc == 1 || nimg > 0) {
            for (int64_t j = 0; j < n; ++j) visit(j);
        } else {
            cl.for_neighbours(ui, visit);
        }
        // reciprocal: -(4pi/L^3) sum k/k^2 e^{..} [sin(k.xi) C(k) - cos(k.xi) S(k)]
        const double xi[3] = {ui[0] * sc, ui[1] * sc, ui[2] * sc};
        for (int64_t v = 0; v < nkv; ++v) {
            const double kx = tpl * kv[3 * v], ky = tpl * kv[3 * v + 1], kz = tpl * kv[3 * v + 2];
            const double ph = kx * xi[0] + ky * xi[1] + kz * xi[2];
            const double s = std::sin(ph), c = std::cos(ph);
            const double w = -kfac[v] * (s * Ck[v] - c * Sk[v]);
            a[0] += w * kx;
            a[1] += w * ky;
            a[2] += w * kz;
        }
        const double sg = sp[i] > 0 ? 1.0 : -1.0;
        for (int ax = 0; ax < 3; ++ax) acc[3 * t + ax] = sg * a[ax];
    }
}

// Ewald potential (electrostatic sign convention, sum q/r) at a site of a simple cubic lattice of
// period L with one unit charge per cell, from all OTHER sites plus the neutralising background:
// the Madelung value -2.837297479 / L.
double orc_ewald_sc_potential(double L, double alpha, int nreal, int nk) {
    double phi = 0.0;
    for (int a = -nreal; a <= nreal; ++a)
        for (int b = -nreal; b <= nreal; ++b)
            for (int c = -nreal; c <= nreal; ++c) {
                if (!a && !b && !c) continue;
                const double r = L * std::sqrt((double)(a * a + b * b + c * c));
                phi += std::erfc(alpha * r) / r;
            }
    for (int a = -nk; a <= nk; ++a)
        for (int b = -nk; b <= nk; ++b)
            for (int c = -nk; c <= nk; ++c) {
                if (!a && !b && !c) continue;
                const double k2 = (2 * M_PI / L) * (2 * M_PI / L) * (double)(a * a + b * b + c * c);
                phi += 4.0 * M_PI / (L * L * L) * std::exp(-k2 / (4 * alpha * alpha)) / k2;
            }
    phi += -M_PI / (alpha * alpha * L * L * L) - 2.0 * alpha / std::sqrt(M_PI);
    return phi;
}

// ------------------------------------------------------------------ reference CPU TreePM
struct V3 { double x, y, z; };
struct Box { V3 mn, mx; };
static inline V3 center(const Box& b) { return {(b.mn.x + b.mx.x) * 0.5, (b.mn.y + b.mx.y) * 0.5, (b.mn.z + b.mx.z) * 0.5}; }
static inline double bsize(const Box& b) { return std::max(b.mx.x - b.mn.x, std::max(b.mx.y - b.mn.y, b.mx.z - b.mn.z)); }
static inline Box child_box(const Box& b, int o) {
    const V3 c = center(b);
    Box r;
    r.mn = {(o & 1) ? c.x : b.mn.x, (o & 2) ? c.y : b.mn.y, (o & 4) ? c.z : b.mn.z};
    r.mx = {(o & 1) ? b.mx.x : c.x, (o & 2) ? b.mx.y : c.y, (o & 4) ? b.mx.z : c.z};
    return r;
}

struct RNode {
    int kind = 0;         // 0 empty, 1 leaf, 2 internal
    int64_t leaf = -1;    // leaf particle (kind 1)
    int32_t child[8];
    V3 cp{0, 0, 0}, cm{0, 0, 0};
    double mp = 0, mm = 0;
};

struct RefCpu {
    double r_cut, L, theta, eps, r_s, g_pm, g_tree;
    int G;
    std::vector<double> rho_p, rho_m, phi_p, phi_m;
    std::vector<RNode> nodes;
    Box root_box;
    const double* pos = nullptr;
    const double* mass = nullptr;
    const int8_t* sign = nullptr;

    int32_t build(std::vector<int64_t>& idx, const Box& b, int depth) {
        if (idx.empty()) { nodes.push_back(RNode()); return (int32_t)nodes.size() - 1; }
        if (idx.size() == 1 || depth > 60) {
            // depth guard: coincident particles would recurse forever in the reference
            RNode nd; nd.kind = 1; nd.leaf = idx[0];
            nodes.push_back(nd);
            return (int32_t)nodes.size() - 1;
        }
        std::vector<int64_t> ch[8];
        const V3 c = center(b);
        for (int64_t p : idx) {
            const int o = (pos[3 * p] >= c.x) | ((pos[3 * p + 1] >= c.y) << 1) | ((pos[3 * p + 2] >= c.z) << 2);
            ch[o].push_back(p);
        }
        RNode nd; nd.kind = 2;
        for (int p = 0; p < 8; ++p) nd.child[p] = -1;
        int32_t kids[8];
        for (int o = 0; o < 8; ++o) kids[o] = build(ch[o], child_box(b, o), depth + 1);
        for (int o = 0; o < 8; ++o) nd.child[o] = kids[o];
        for (int64_t p : idx) {
            const double m = mass[p];
            if (sign[p] > 0) { nd.cp.x += pos[3 * p] * m; nd.cp.y += pos[3 * p + 1] * m; nd.cp.z += pos[3 * p + 2] * m; nd.mp += m; }
            else { nd.cm.x += pos[3 * p] * m; nd.cm.y += pos[3 * p + 1] * m; nd.cm.z += pos[3 * p + 2] * m; nd.mm += m; }
        }
        if (nd.mp > 0) { nd.cp.x /= nd.mp; nd.cp.y /= nd.mp; nd.cp.z /= nd.mp; }
        if (nd.mm > 0) { nd.cm.x /= nd.mm; nd.cm.y /= nd.mm; nd.cm.z /= nd.mm; }
        nodes.push_back(nd);
        return (int32_t)nodes.size() - 1;
    }

    V3 pair(const V3& pi, int si, const V3& pj, double mj, int sj) const {
        if (mj == 0.0) return {0, 0, 0};
        const V3 r = {pj.x - pi.x, pj.y - pi.y, pj.z - pi.z};
        const double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
        const double rr = std::sqrt(r2);
        if (rr >= r_cut) return {0, 0, 0};
        const double r2s = r2 + eps * eps;
        if (r2s < 1e-20) return {0, 0, 0};
        const double inter = (si == sj ? 1.0 : -1.0) * g_tree;
        const double inv3 = 1.0 / (r2s * std::sqrt(r2s));
        const double x = rr / r_cut, x2 = x * x;
        const double w = 1.0 - x2 * x2;
        const double f = inter * mj * inv3 * w;
        return {r.x * f, r.y * f, r.z * f};
    }

    V3 walk(int32_t ni, const Box& b, const V3& p, int s, int64_t excl) const {
        const RNode& nd = nodes[ni];
        if (nd.kind == 0) return {0, 0, 0};
        if (nd.kind == 1) {
            if (nd.leaf == excl) return {0, 0, 0};
            const int64_t j = nd.leaf;
            return pair(p, s, {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]}, mass[j], sign[j]);
        }
        const V3 c = center(b);
        const double dr = std::sqrt((p.x - c.x) * (p.x - c.x) + (p.y - c.y) * (p.y - c.y) + (p.z - c.z) * (p.z - c.z));
        const double cs = bsize(b);
        const double mind = std::max(dr - cs * 0.866, 0.0);
        if (mind >= r_cut) return {0, 0, 0};
        if (cs / dr < theta && dr < r_cut) {
            const V3 a = pair(p, s, nd.cp, nd.mp, 1);
            const V3 bb = pair(p, s, nd.cm, nd.mm, -1);
            return {a.x + bb.x, a.y + bb.y, a.z + bb.z};
        }
        V3 acc = {0, 0, 0};
        for (int o = 0; o < 8; ++o) {
            const V3 a = walk(nd.child[o], child_box(b, o), p, s, excl);
            acc.x += a.x; acc.y += a.y; acc.z += a.z;
        }
        return acc;
    }

    void assign(double x, double y, double z, double m, int s) {
        const double half = L / 2, gs = (double)G;
        auto remE = [](double a, double b) { double r = std::fmod(a, b); return r < 0 ? r + b : r; };
        const double gx = remE((x + half) / L * gs, gs), gy = remE((y + half) / L * gs, gs), gz = remE((z + half) / L * gs, gs);
        const int64_t ix = (int64_t)std::floor(gx), iy = (int64_t)std::floor(gy), iz = (int64_t)std::floor(gz);
        const double fx = gx - ix, fy = gy - iy, fz = gz - iz;
        const double wx[2] = {1 - fx, fx}, wy[2] = {1 - fy, fy}, wz[2] = {1 - fz, fz};
        std::vector<double>& gr = s > 0 ? rho_p : rho_m;
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
                for (int dk = 0; dk < 2; ++dk) {
                    const int64_t ci = (ix + di) % G, cj = (iy + dj) % G, ck = (iz + dk) % G;
                    gr[ci + G * (cj + G * ck)] += std::fabs(m) * wx[di] * wy[dj] * wz[dk];
                }
    }

    void solve(const std::vector<double>& rho, std::vector<double>& phi) {
        const int64_t G3 = (int64_t)G * G * G;
        std::vector<cd> d(G3);
        for (int64_t i = 0; i < G3; ++i) d[i] = cd(rho[i], 0);
        fft3d(d, G, false);
        const double dk = 2 * M_PI / L;
#pragma omp parallel for
        for (int64_t i = 0; i < G; ++i)
            for (int64_t j = 0; j < G; ++j)
                for (int64_t k = 0; k < G; ++k) {
                    const int64_t idx = i + G * (j + G * k);
                    const double ki = i <= G / 2 ? (double)i : (double)i - G, kj = j <= G / 2 ? (double)j : (double)j - G,
                                 kk = k <= G / 2 ? (double)k : (double)k - G;
                    const double k2 = (ki * ki + kj * kj + kk * kk) * dk * dk;
                    if (k2 > 1e-10) {
                        double gr = -4.0 * M_PI * g_pm / k2;
                        if (r_s > 0) gr *= std::exp(-k2 * r_s * r_s);
                        d[idx] *= gr;
                    } else {
                        d[idx] = 0;
                    }
                }
        fft3d(d, G, true);
        for (int64_t i = 0; i < G3; ++i) phi[i] = d[i].real() / (double)G3;
    }

    V3 pm_force(double x, double y, double z, int s) const {
        const double half = L / 2, gs = (double)G;
        auto remE = [](double a, double b) { double r = std::fmod(a, b); return r < 0 ? r + b : r; };
        const double gx = remE((x + half) / L * gs, gs), gy = remE((y + half) / L * gs, gs), gz = remE((z + half) / L * gs, gs);
        const int64_t ix = (int64_t)std::floor(gx), iy = (int64_t)std::floor(gy), iz = (int64_t)std::floor(gz);
        const double fx = gx - ix, fy = gy - iy, fz = gz - iz;
        const std::vector<double>& pa = s > 0 ? phi_p : phi_m;
        const std::vector<double>& pr = s > 0 ? phi_m : phi_p;
        const double h = L / G;
        const double wx[2] = {1 - fx, fx}, wy[2] = {1 - fy, fy}, wz[2] = {1 - fz, fz};
        V3 f = {0, 0, 0};
        auto I = [&](int64_t a, int64_t b, int64_t c) { return a + G * (b + G * c); };
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
                for (int dk = 0; dk < 2; ++dk) {
                    const int64_t ci = (ix + di) % G, cj = (iy + dj) % G, ck = (iz + dk) % G;
                    const double w = wx[di] * wy[dj] * wz[dk];
                    const int64_t cip = (ci + 1) % G, cim = (ci + G - 1) % G, cjp = (cj + 1) % G, cjm = (cj + G - 1) % G,
                                  ckp = (ck + 1) % G, ckm = (ck + G - 1) % G;
                    const double dax = (pa[I(cip, cj, ck)] - pa[I(cim, cj, ck)]) / (2 * h);
                    const double day = (pa[I(ci, cjp, ck)] - pa[I(ci, cjm, ck)]) / (2 * h);
                    const double daz = (pa[I(ci, cj, ckp)] - pa[I(ci, cj, ckm)]) / (2 * h);
                    const double drx = (pr[I(cip, cj, ck)] - pr[I(cim, cj, ck)]) / (2 * h);
                    const double dry = (pr[I(ci, cjp, ck)] - pr[I(ci, cjm, ck)]) / (2 * h);
                    const double drz = (pr[I(ci, cj, ckp)] - pr[I(ci, cj, ckm)]) / (2 * h);
                    f.x += w * (-dax + drx);
                    f.y += w * (-day + dry);
                    f.z += w * (-daz + drz);
                }
        return f;
    }
};

void* orc_refcpu_create(double r_cut, int G, double L, double theta, double eps, double r_s, double g_pm,
                        double g_tree) {
    RefCpu* r = new RefCpu();
    r->r_cut = r_cut; r->G = G; r->L = L; r->theta = theta; r->eps = eps; r->r_s = r_s;
    r->g_pm = g_pm; r->g_tree = g_tree;
    const int64_t G3 = (int64_t)G * G * G;
    r->rho_p.assign(G3, 0); r->rho_m.assign(G3, 0); r->phi_p.assign(G3, 0); r->phi_m.assign(G3, 0);
    return r;
}
void orc_refcpu_destroy(void* h) { delete (RefCpu*)h; }

// TreePMForce::update: deposit both grids, solve both, rebuild the tree (pm_enabled / tree_enabled flags)
void orc_refcpu_update(void* h, int64_t n, const double* pos, const double* mass, const int8_t* sign, int pm,
                       int tree) {
    RefCpu& r = *(RefCpu*)h;
    r.pos = pos; r.mass = mass; r.sign = sign;
    std::fill(r.rho_p.begin(), r.rho_p.end(), 0.0);
    std::fill(r.rho_m.begin(), r.rho_m.end(), 0.0);
    std::fill(r.phi_p.begin(), r.phi_p.end(), 0.0);
    std::fill(r.phi_m.begin(), r.phi_m.end(), 0.0);
    if (pm) {
        for (int64_t p = 0; p < n; ++p) r.assign(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], mass[p], sign[p]);
        r.solve(r.rho_p, r.phi_p);
        r.solve(r.rho_m, r.phi_m);
    }
    r.nodes.clear();
    if (tree && n > 0) {
        V3 mn = {1e300, 1e300, 1e300}, mx = {-1e300, -1e300, -1e300};
        for (int64_t p = 0; p < n; ++p) {
            mn.x = std::min(mn.x, pos[3 * p]); mn.y = std::min(mn.y, pos[3 * p + 1]); mn.z = std::min(mn.z, pos[3 * p + 2]);
            mx.x = std::max(mx.x, pos[3 * p]); mx.y = std::max(mx.y, pos[3 * p + 1]); mx.z = std::max(mx.z, pos[3 * p + 2]);
        }
        mn.x -= 1; mn.y -= 1; mn.z -= 1; mx.x += 1; mx.y += 1; mx.z += 1;
        const double size = std::max(mx.x - mn.x, std::max(mx.y - mn.y, mx.z - mn.z));
        const V3 c = {(mn.x + mx.x) / 2, (mn.y + mx.y) / 2, (mn.z + mx.z) / 2};
        r.root_box.mn = {c.x - size / 2, c.y - size / 2, c.z - size / 2};
        r.root_box.mx = {c.x + size / 2, c.y + size / 2, c.z + size / 2};
        std::vector<int64_t> idx(n);
        for (int64_t p = 0; p < n; ++p) idx[p] = p;
        r.nodes.reserve(2 * n + 8);
        r.build(idx, r.root_box, 0);
    }
}

// compute_all_forces restricted to targets tidx (nullptr: all), with self-exclusion
void orc_refcpu_forces(void* h, int64_t nt, const int64_t* tidx, int pm, int tree, double* acc) {
    RefCpu& r = *(RefCpu*)h;
    const int32_t root = (int32_t)r.nodes.size() - 1;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t i = tidx ? tidx[t] : t;
        const V3 p = {r.pos[3 * i], r.pos[3 * i + 1], r.pos[3 * i + 2]};
        V3 f = {0, 0, 0};
        if (pm) f = r.pm_force(p.x, p.y, p.z, r.sign[i]);
        if (tree && root >= 0) {
            const V3 a = r.walk(root, r.root_box, p, r.sign[i], i);
            f.x += a.x; f.y += a.y; f.z += a.z;
        }
        acc[3 * t] = f.x; acc[3 * t + 1] = f.y; acc[3 * t + 2] = f.z;
    }
}

int orc_num_threads(void) { return omp_get_max_threads(); }

// ------------------------------------------------------------------ initial conditions
static uint64_t sm_next(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// output j (0-based) of the generator seeded with s: j + 1 advances of the state
static uint64_t sm_at(uint64_t s, uint64_t j) {
    uint64_t st = s + j * 0x9E3779B97F4A7C15ull;
    return sm_next(st);
}

void orc_splitmix64(uint64_t seed, int64_t count, uint64_t* out) {
    uint64_t st = seed;
    for (int64_t i = 0; i < count; ++i) out[i] = sm_next(st);
}

// pos [n^3][3] in [-L/2, L/2], vel [n^3][3], sign [n^3]; returns n+ (ic_gen.rs:51-169)
int64_t orc_ic(int n, double L, double z_init, uint64_t seed, double mu, double n_s, double delta_rms, double* pos,
               double* vel, int8_t* sign) {
    const int64_t n3 = (int64_t)n * n * n;
    const int64_t n_plus = (int64_t)std::llround((double)n3 / (1.0 + mu));
    const double dk = 2.0 * M_PI / L, spacing = L / n, half = L / 2.0, d_growth = 1.0 / (1.0 + z_init);
    const int half_n = n / 2;
    auto kof = [&](int i) { return (i <= half_n ? (double)i : (double)(i - n)) * dk; };
    auto u01 = [](uint64_t r) { return (double)(r >> 11) / 9007199254740992.0; };
    // delta(k) over the full grid (ic_gen.rs:196-221)
    std::vector<cd> dk_field(n3, cd(0, 0));
    for (int iz = 0; iz < n; ++iz)
        for (int iy = 0; iy < n; ++iy)
            for (int ix = 0; ix < n; ++ix) {
                const double kx = kof(ix), ky = kof(iy), kz = kof(iz);
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (!(k2 > 0)) continue;
                const int64_t idx = ((int64_t)iz * n + iy) * n + ix;
                const double pk = std::pow(std::sqrt(k2), n_s - 4.0) * delta_rms * delta_rms;
                const double amp = std::sqrt(pk / 2.0);
                const double phase = u01(sm_at(seed, 3 * idx)) * 2.0 * M_PI;
                const double u1 = u01(sm_at(seed, 3 * idx + 1)), u2 = u01(sm_at(seed, 3 * idx + 2));
                const double r = std::sqrt(-2.0 * std::log(1.0 - u1));
                const double re = r * std::cos(2.0 * M_PI * u2) * amp * d_growth;
                const double im = r * std::sin(2.0 * M_PI * u2) * amp * d_growth;
                dk_field[idx] = cd(re * std::cos(phase) - im * std::sin(phase), re * std::sin(phase) + im * std::cos(phase));
            }
    // psi_i = Re IFFT[-i k_i delta / k^2] / n^3 (ic_gen.rs:224-296)
    std::vector<double> psi[3];
    for (int dim = 0; dim < 3; ++dim) {
        std::vector<cd> f(n3);
        for (int iz = 0; iz < n; ++iz)
            for (int iy = 0; iy < n; ++iy)
                for (int ix = 0; ix < n; ++ix) {
                    const int64_t idx = ((int64_t)iz * n + iy) * n + ix;
                    const double kx = kof(ix), ky = kof(iy), kz = kof(iz);
                    const double k2 = kx * kx + ky * ky + kz * kz;
                    const double ki = dim == 0 ? kx : (dim == 1 ? ky : kz);
                    f[idx] = k2 > 0 ? dk_field[idx] * cd(0.0, -ki / k2) : cd(0, 0);
                }
        fft3d(f, n, true);
        psi[dim].resize(n3);
        for (int64_t i = 0; i < n3; ++i) psi[dim][i] = f[i].real() / (double)n3;
    }
    // scale to max |psi| = 0.3 spacing; velocity scale H0 sqrt(1 + z) (ic_gen.rs:82-95)
    double maxd = 0;
    for (int64_t i = 0; i < n3; ++i)
        maxd = std::max(maxd, std::sqrt(psi[0][i] * psi[0][i] + psi[1][i] * psi[1][i] + psi[2][i] * psi[2][i]));
    const double scale = maxd > 1e-10 ? spacing * 0.3 / maxd : 1.0;
    const double vel_scale = 69.9 / 977.8 * std::sqrt(1.0 + z_init);
    // signs: the n+ cells of smallest (key, index) (ic_gen.rs:99-110 draws a random subset by shuffling)
    std::vector<std::pair<uint64_t, int64_t>> keys(n3);
    for (int64_t i = 0; i < n3; ++i) keys[i] = {sm_at(seed + 12345, (uint64_t)i), i};
    std::sort(keys.begin(), keys.end());
    for (int64_t r = 0; r < n3; ++r) sign[keys[r].second] = r < n_plus ? 1 : -1;
    // lattice + displacement, periodic wrap (ic_gen.rs:117-149)
    for (int iz = 0; iz < n; ++iz)
        for (int iy = 0; iy < n; ++iy)
            for (int ix = 0; ix < n; ++ix) {
                const int64_t g = ((int64_t)iz * n + iy) * n + ix;
                const int ii[3] = {ix, iy, iz};
                for (int a = 0; a < 3; ++a) {
                    double x = (ii[a] + 0.5) * spacing - half + psi[a][g] * scale;
                    while (x > half) x -= L;
                    while (x < -half) x += L;
                    pos[3 * g + a] = x;
                    vel[3 * g + a] = psi[a][g] * scale * vel_scale;
                }
            }
    return n_plus;
}

// ------------------------------------------------------------------ segregation metrics
// janus_mu19_production.rs:238-272, from per-cell counts
static void seg_metrics(const std::vector<int32_t>& np, const std::vector<int32_t>& nm, double* out) {
    const size_t nc = np.size();
    double tp = 0, tm = 0;
    for (size_t i = 0; i < nc; ++i) {
        tp += np[i];
        tm += nm[i];
    }
    const double mean_p = tp / (double)nc, mean_m = tm / (double)nc;
    double var_diff = 0;
    for (size_t i = 0; i < nc; ++i) {
        const double d = (double)np[i] - (double)nm[i] - (mean_p - mean_m);
        var_diff += d * d;
    }
    var_diff /= (double)nc;
    const double pois = mean_p + mean_m;
    out[0] = pois > 0 ? var_diff / pois : 0.0;
    double sp = 0, sm = 0, spp = 0, smm = 0, spm = 0;
    for (size_t i = 0; i < nc; ++i) {
        const double dp = mean_p > 0 ? ((double)np[i] - mean_p) / mean_p : 0.0;
        const double dm = mean_m > 0 ? ((double)nm[i] - mean_m) / mean_m : 0.0;
        sp += dp;
        sm += dm;
        spp += dp * dp;
        smm += dm * dm;
        spm += dp * dm;
    }
    const double n = (double)nc;
    const double vp = spp / n - (sp / n) * (sp / n), vm = smm / n - (sm / n) * (sm / n);
    const double cov = spm / n - (sp / n) * (sm / n);
    out[1] = (vp > 0 && vm > 0) ? cov / (std::sqrt(vp) * std::sqrt(vm)) : 0.0;
    double vmc = 0;
    for (size_t i = 0; i < nc; ++i) vmc += ((double)nm[i] - mean_m) * ((double)nm[i] - mean_m);
    vmc /= (double)nc;
    out[2] = mean_m > 0 ? (vmc - mean_m) / mean_m : 0.0;
}

// the reference's binning (janus_mu19_production.rs:220-236) of f32 positions in [-L/2, L/2)
void orc_segregation_f32_ref(int64_t n, const float* pos, const int8_t* sp, double L, int nc, double* out) {
    const double cell = L / nc, half = L / 2.0;
    std::vector<int32_t> np((size_t)nc * nc * nc, 0), nm((size_t)nc * nc * nc, 0);
    auto bin = [&](float v) {
        double w = std::fmod((double)v + half, L);
        if (w < 0) w += L;  // rem_euclid
        return (size_t)(w / cell) % (size_t)nc;
    };
    for (int64_t i = 0; i < n; ++i) {
        const size_t idx = (bin(pos[3 * i]) * nc + bin(pos[3 * i + 1])) * nc + bin(pos[3 * i + 2]);
        if (sp[i] > 0) ++np[idx];
        else ++nm[idx];
    }
    seg_metrics(np, nm, out);
}

// fixed-point binning: cell = floor(u nc / 2^32)
void orc_segregation(int64_t n, const uint32_t* pos, const int8_t* sp, int nc, double* out) {
    std::vector<int32_t> np((size_t)nc * nc * nc, 0), nm((size_t)nc * nc * nc, 0);
    auto bin = [&](uint32_t u) { return (size_t)(((uint64_t)u * (uint64_t)nc) >> 32); };
    for (int64_t i = 0; i < n; ++i) {
        const size_t idx = (bin(pos[3 * i]) * nc + bin(pos[3 * i + 1])) * nc + bin(pos[3 * i + 2]);
        if (sp[i] > 0) ++np[idx];
        else ++nm[idx];
    }
    seg_metrics(np, nm, out);
}

}  // extern "C"
