// C ABI of the TreePM force step: context lifecycle, buffer management and
// the phase sequence of one force evaluation.
//
// Phase order on the caller's stream (each phase bracketed by a HIP event):
//   keys + radix sort + particle gather | Karras tree + monopoles + groups |
//   CIC deposit | FFT Poisson solve | interpolation | group walk (+ PM add, scatter to caller order)
// The reference runs the same pieces with a device synchronisation after
// every launch (src/nbody_gpu_twopass.rs:4486-4649); here only the final
// status read synchronises.
#include "jtp_internal.h"

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

namespace jtp {

static int fail(Ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
static int hip_fail(Ctx* c, hipError_t e, const char* where) {
    return fail(c, JTP_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}
#define JTP_HIP(c, call)                                    \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail((c), e_, #call); \
    } while (0)

template <typename T>
static hipError_t dalloc(T*& p, size_t count) {
    if (p) {
        (void)hipFree((void*)p);
        p = nullptr;
    }
    if (count == 0) count = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e == hipSuccess) p = (T*)q;
    return e;
}
static hipError_t dalloc_bytes(void*& p, size_t bytes) {
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
    if (bytes == 0) bytes = 1;
    return hipMalloc(&p, bytes);
}

// Chebyshev interpolants of the split factor for the FP32 walk (forms and degrees in
// jtp_internal.h), with maximum errors at r_cut = 4.5 r_s, theta = 0.5 (FP32 Horner included,
// relative to the Newtonian pair force):
//   particles, erfc + Gaussian: 1 / rho^3 + H(rho^2), degree 8: 6e-6;
//   particles, erfc only: S(x), degree 10: 6e-6;
//   monopoles: exp(-x^2) P(x), P = erfcx(x) [+ 2x/sqrt(pi)] on [0, Xm], Xm = Xp (1 + sqrt(3) theta),
//              degree 8: 1.5e-4 (well below the monopole truncation error at theta = 0.5).
// Degree-deg Chebyshev interpolant of f on [0, X], returned as monomial coefficients of
// t = 2 y / X - 1 (zeros above deg, up to kPolyDeg).
template <typename F>
static void cheb_fit(double X, int deg, float* out, F f) {
    const int N = deg + 1;
    std::vector<double> fv(N), cheb(N, 0.0);
    for (int k = 0; k < N; ++k) {
        const double t = std::cos(M_PI * (k + 0.5) / N);
        fv[k] = f(0.5 * X * (t + 1.0));
    }
    for (int j = 0; j < N; ++j) {
        double s = 0;
        for (int k = 0; k < N; ++k) s += fv[k] * std::cos(M_PI * j * (k + 0.5) / N);
        cheb[j] = 2.0 * s / N;
    }
    cheb[0] *= 0.5;
    std::vector<std::vector<double>> T(N, std::vector<double>(N, 0.0));  // T_j(t) as monomials
    T[0][0] = 1.0;
    if (N > 1) T[1][1] = 1.0;
    for (int j = 1; j + 1 < N; ++j)
        for (int d = 0; d < N; ++d) T[j + 1][d] = -T[j - 1][d] + (d > 0 ? 2.0 * T[j][d - 1] : 0.0);
    for (int d = 0; d <= kPolyDeg; ++d) {
        double s = 0;
        if (d < N)
            for (int j = 0; j < N; ++j) s += cheb[j] * T[j][d];
        out[d] = (float)s;
    }
}
// S(x) = erfc(x) [+ 2x/sqrt(pi) exp(-x^2)]: the short-range share of the pair force.
static double split_direct(double x, bool gauss) {
    return std::erfc(x) + (gauss ? 2.0 * x / std::sqrt(M_PI) * std::exp(-x * x) : 0.0);
}
// exp(x^2) S(x) = erfcx(x) [+ 2x/sqrt(pi)].
static double split_scaled(double x, bool gauss) {
    return std::erfc(x) * std::exp(x * x) + (gauss ? 2.0 * x / std::sqrt(M_PI) : 0.0);
}
// (S(x) - 1) / x^3 for the erfc + Gaussian law.  Its Taylor series follows from those of erfc and
// x exp(-x^2): S(x) - 1 = (2/sqrt(pi)) sum_{n>=1} (-1)^n x^{2n+1} / n! * 2n / (2n+1); the series is
// used below x = 0.5, where the direct form cancels.
static double split_excess_x3(double x) {
    if (x < 0.5) {
        const double x2 = x * x;
        double term = 1.0, sum = 0.0;  // term = x^{2n-2} / n!
        for (int nn = 1; nn < 30; ++nn) {
            term = (nn == 1) ? 1.0 : term * x2 / nn;
            sum += ((nn & 1) ? -1.0 : 1.0) * term * (2.0 * nn) / (2.0 * nn + 1.0);
        }
        return 2.0 / std::sqrt(M_PI) * sum;
    }
    return (split_direct(x, true) - 1.0) / (x * x * x);
}

void make_split_poly(Ctx& c) {
    if (!(c.p.r_s > 0)) return;
    const double rho_max = std::sqrt(c.p.r_cut * c.p.r_cut + c.p.softening * c.p.softening);
    const double Xp = rho_max / (2.0 * c.p.r_s) * 1.0001 + 1e-12;
    const double Xm = Xp * (1.0 + std::sqrt(3.0) * c.p.theta);
    const bool gauss = c.p.split_mode == JTP_SPLIT_ERFC_GAUSS;
    if (gauss) {
        // H(v) = (S(x) - 1) / rho^3 with rho = sqrt(v), x = rho / (2 r_s)
        const double Vp = rho_max * rho_max * 1.0002 + 1e-30;
        const double cinv = 1.0 / (2.0 * c.p.r_s);
        cheb_fit(Vp, kPolyDegPair, c.poly.cp,
                 [&](double v) { return cinv * cinv * cinv * split_excess_x3(std::sqrt(v) * cinv); });
        c.poly_tscale = (float)(2.0 / Vp);
        // monopoles: the same H out to kTailExt r_cut
        const double Vm = (kTailExt * kTailExt * c.p.r_cut * c.p.r_cut + c.p.softening * c.p.softening) * 1.0002 + 1e-30;
        cheb_fit(Vm, kPolyDegTail, c.poly.cm,
                 [&](double v) { return cinv * cinv * cinv * split_excess_x3(std::sqrt(v) * cinv); });
        c.poly_tscale_m = (float)(2.0 / Vm);
        c.poly_rho2_tail = (float)(Vm / 1.0002);
    } else {
        cheb_fit(Xp, kPolyDeg, c.poly.cp, [&](double x) { return split_direct(x, false); });
        c.poly_tscale = (float)(1.0 / (c.p.r_s * Xp));
        cheb_fit(Xm, kPolyDeg, c.poly.cm, [&](double x) { return split_scaled(x, false); });
        c.poly_tscale_m = (float)(1.0 / (c.p.r_s * Xm));
    }
}

static void free_particle_buffers(Ctx& c) {
    void* ptrs[] = {c.keys0, c.keys1, c.vals0, c.vals1, c.hist, c.hist_blk, c.parts, c.node_geo, c.node_kids,
                    c.up, c.flags, c.node_sum, c.node_mono, c.groups, c.pm_acc,
                    c.frontier0, c.frontier1, c.dense_list, c.dense_corner, c.interp_list, c.dep_chunks,
                    c.dep_cbuf};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    c.keys0 = c.keys1 = nullptr;
    c.vals0 = c.vals1 = nullptr;
    c.hist = c.hist_blk = nullptr;
    c.parts = nullptr;
    c.node_geo = nullptr;
    c.node_kids = nullptr;
    c.up = nullptr;
    c.flags = nullptr;
    c.frontier0 = c.frontier1 = c.dense_list = nullptr;
    c.dense_corner = nullptr;
    c.interp_list = nullptr;
    c.dep_chunks = nullptr;
    c.dep_cbuf = nullptr;
    c.node_sum = c.node_mono = nullptr;
    c.groups = nullptr;
    c.pm_acc = nullptr;
    c.cap_n = 0;
}

static int ensure_capacity(Ctx& c, int64_t n) {
    if (n <= c.cap_n) return JTP_OK;
    free_particle_buffers(c);
    const int64_t cap = n < 1024 ? 1024 : n + n / 8;
    const int64_t nblk = (cap + kSortTile - 1) / kSortTile;
    const int64_t m = nblk * 256;
    const size_t rs = c.real_size();
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) {
        if (e == hipSuccess && x != hipSuccess) e = x;
    };
    chk(dalloc(c.keys0, cap));
    chk(dalloc(c.keys1, cap));
    chk(dalloc(c.vals0, cap));
    chk(dalloc(c.vals1, cap));
    chk(dalloc(c.hist, m));
    chk(dalloc(c.hist_blk, (m + 2047) / 2048 + 1));
    chk(dalloc(c.parts, cap));
    chk(dalloc(c.node_geo, 2 * cap));
    chk(dalloc(c.node_kids, cap));
    chk(dalloc(c.up, cap));
    chk(dalloc(c.flags, cap));
    chk(dalloc(c.frontier0, cap));
    chk(dalloc(c.frontier1, cap));
    chk(dalloc(c.interp_list, (size_t)interp_list_size(c.G3(), cap)));
    if (c.p.precision == JTP_FP64) {  // dense-cell pre-reduction of the gather deposit
        chk(dalloc(c.dense_list, cap / kDenseCell + 1));
        chk(dalloc_bytes(c.dense_corner, (size_t)(cap / kDenseCell + 1) * (8 + 27) * rs));
    } else {  // chunks of the heavy tiles of the read-once deposit
        chk(dalloc(c.dep_chunks, (size_t)dep_chunk_cap(cap)));
        chk(dalloc(c.dep_cbuf, (size_t)dep_chunk_cap(cap) * dep_chunk_nodes()));
    }
    chk(dalloc_bytes(c.node_sum, (size_t)cap * 64));
    chk(dalloc_bytes(c.node_mono, (size_t)cap * (rs == 8 ? 64 : 32)));
    chk(dalloc(c.groups, cap + 4096 + 64));
    chk(dalloc_bytes(c.pm_acc, (size_t)cap * 3 * rs));
    if (e != hipSuccess) {
        free_particle_buffers(c);
        (void)hipGetLastError();
        return fail(&c, JTP_ERR_ALLOC, std::string("device allocation failed: ") + hipGetErrorString(e));
    }
    c.cap_n = cap;
    return JTP_OK;
}

// The three segregation numbers from the device sums of launch_segregation (janus_mu19_production.rs:240-272).
void segregation_metrics(const double v[9], int64_t cells, double out[3]) {
    const double C = (double)cells, mp = v[7] / C, mm = v[8] / C;
    const double var_diff = v[0] / C, pois = mp + mm;
    out[0] = pois > 0 ? var_diff / pois : 0.0;
    const double var_p = v[3] / C - (v[1] / C) * (v[1] / C), var_m = v[4] / C - (v[2] / C) * (v[2] / C);
    const double cov = v[5] / C - (v[1] / C) * (v[2] / C);
    out[1] = (var_p > 0 && var_m > 0) ? cov / (std::sqrt(var_p) * std::sqrt(var_m)) : 0.0;
    out[2] = mm > 0 ? (v[6] / C - mm) / mm : 0.0;
}

static int check_params(const jtp_params* p, std::string& msg) {
    if (!p) { msg = "params is null"; return JTP_ERR_ARG; }
    if (!(p->box_size > 0) || !std::isfinite(p->box_size)) { msg = "box_size must be > 0"; return JTP_ERR_ARG; }
    const int G = p->n_cells;
    if (G < 16 || G > 512 || (G & (G - 1))) { msg = "n_cells must be a power of two in [16, 512]"; return JTP_ERR_ARG; }
    if (!(p->r_s >= 0) || !std::isfinite(p->r_s)) { msg = "r_s must be >= 0"; return JTP_ERR_ARG; }
    if (!(p->r_cut > 0) || p->r_cut > p->box_size / 5.0) { msg = "r_cut must be in (0, L/5]"; return JTP_ERR_ARG; }
    if (!(p->theta >= 0) || p->theta > 1.0) { msg = "theta must be in [0, 1]"; return JTP_ERR_ARG; }
    if (!(p->softening > 0) || !std::isfinite(p->softening)) { msg = "softening must be > 0"; return JTP_ERR_ARG; }
    if (!std::isfinite(p->g_constant)) { msg = "g_constant must be finite"; return JTP_ERR_ARG; }
    if (p->leaf_size < 1 || p->leaf_size > 16) { msg = "leaf_size must be in [1, 16]"; return JTP_ERR_ARG; }
    if (p->group_size < 1 || p->group_size > 64) { msg = "group_size must be in [1, 64]"; return JTP_ERR_ARG; }
    if (p->precision != JTP_FP32 && p->precision != JTP_FP64) { msg = "precision must be JTP_FP32 or JTP_FP64"; return JTP_ERR_ARG; }
    if (p->split_mode < JTP_SPLIT_ERFC_GAUSS || p->split_mode > JTP_SPLIT_POLY4) { msg = "bad split_mode"; return JTP_ERR_ARG; }
    if (p->split_mode != JTP_SPLIT_POLY4 && !(p->r_s > 0)) { msg = "erfc split modes need r_s > 0"; return JTP_ERR_ARG; }
    if (p->gradient != JTP_GRAD_4PT && p->gradient != JTP_GRAD_2PT) { msg = "bad gradient"; return JTP_ERR_ARG; }
    if (p->k_min < 0) { msg = "k_min must be >= 0"; return JTP_ERR_ARG; }
    if (p->interlace != 0 && p->interlace != 1) { msg = "interlace must be 0 or 1"; return JTP_ERR_ARG; }
    if (!(p->theta_cut >= 0) || !std::isfinite(p->theta_cut)) { msg = "theta_cut must be >= 0"; return JTP_ERR_ARG; }
    if (p->fft_tile != 0 && p->fft_tile != 11 && p->fft_tile != 12 && !(p->fft_tile == 13 && p->precision == JTP_FP32)) {
        msg = "fft_tile must be 0, 11, 12 or (FP32) 13";
        return JTP_ERR_ARG;
    }
    return JTP_OK;
}

}  // namespace jtp

using namespace jtp;

struct jtp_ctx {
    Ctx c;
};

extern "C" {

const char* jtp_version(void) { return "janus_treepm 0.1 (gfx950)"; }

void jtp_default_params(jtp_params* p, double box_size, int32_t n_cells) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->box_size = box_size;
    p->n_cells = n_cells;
    p->k_min = 0;
    const double h = n_cells > 0 ? box_size / n_cells : 1.0;
    p->r_s = 1.25 * h;
    p->r_cut = 4.5 * p->r_s;
    p->theta = 0.5;
    p->softening = h / 30.0;
    p->g_constant = 1.0;
    p->leaf_size = 4;
    p->group_size = 32;
    p->precision = JTP_FP32;
    p->split_mode = JTP_SPLIT_ERFC_GAUSS;
    p->gradient = JTP_GRAD_4PT;
    p->deconvolve = 1;
    p->pm_enabled = 1;
    p->tree_enabled = 1;
    p->interlace = 1;
    p->theta_cut = 0.66 * p->theta;
    p->fft_tile = 0;
}

int jtp_params_from_config(const jtp_config_fields* f, jtp_params* p) {
    if (!f || !p) return JTP_ERR_ARG;
    if (!(f->box_size_mpc > 0) || !std::isfinite(f->box_size_mpc) || f->n_cells <= 0) return JTP_ERR_ARG;
    jtp_default_params(p, f->box_size_mpc, f->n_cells);
    const double h = f->box_size_mpc / f->n_cells;
    // r_cut = 2 L / n_cells (janus_optim.rs:74) and r_s = r_cut / 3 (nbody_gpu_twopass.rs:1694) unless the
    // optional split fields are present
    const double ratio = f->r_cut_over_r_s > 0 ? f->r_cut_over_r_s : 3.0;
    const double rs_cells = f->split_scale_cells > 0 ? f->split_scale_cells : 2.0 / ratio;
    p->r_s = rs_cells * h;
    p->r_cut = ratio * p->r_s;
    p->theta = f->theta < 0 ? 0.7 : f->theta;  // default_theta (config.rs:42)
    p->theta_cut = 0.66 * p->theta;
    if (f->softening_mpc > 0) {
        p->softening = f->softening_mpc;
    } else {  // config.rs:191-196
        if (f->n_particles <= 0) return JTP_ERR_ARG;
        p->softening = f->box_size_mpc / std::pow((double)f->n_particles, 1.0 / 3.0) / 30.0;
    }
    p->k_min = f->k_min;
    std::string msg;
    return check_params(p, msg);
}

int jtp_create(const jtp_params* p, jtp_ctx** out) {
    if (!out) return JTP_ERR_ARG;
    *out = nullptr;
    std::string msg;
    int rc = check_params(p, msg);
    if (rc != JTP_OK) {
        std::fprintf(stderr, "jtp_create: %s\n", msg.c_str());
        return rc;
    }
    jtp_ctx* h = new (std::nothrow) jtp_ctx();
    if (!h) return JTP_ERR_ALLOC;
    Ctx& c = h->c;
    c.p = *p;
    int G = p->n_cells;
    while ((1 << c.logG) < G) ++c.logG;
    hipError_t e = hipGetDevice(&c.device);
    if (e != hipSuccess) {
        std::fprintf(stderr, "jtp_create: hipGetDevice: %s\n", hipGetErrorString(e));
        delete h;
        return JTP_ERR_HIP;
    }
    const size_t rs = c.real_size();
    const int64_t G3 = c.G3();
    e = dalloc_bytes(c.rho, G3 * rs);
    if (e == hipSuccess) e = dalloc_bytes(c.phi, G3 * rs);
    if (e == hipSuccess && p->interlace) e = dalloc_bytes(c.rho2, G3 * rs);
    if (e == hipSuccess && p->interlace) e = dalloc_bytes(c.phi2, G3 * rs);
    if (p->precision == JTP_FP64) {  // gradient grids (the FP32 interpolation keeps them in LDS tiles)
        if (e == hipSuccess) e = dalloc_bytes(c.grad, G3 * 4 * rs);
        if (e == hipSuccess && p->interlace) e = dalloc_bytes(c.grad2, G3 * 4 * rs);
    }
    if (e == hipSuccess) e = dalloc_bytes(c.spec, (size_t)G * G * c.Hc() * 2 * rs);
    if (e == hipSuccess) e = dalloc_bytes(c.twiddle, (size_t)2 * G * rs);
    if (e == hipSuccess) e = dalloc_bytes(c.gtab, (size_t)4 * G * rs);
    if (p->precision == JTP_FP64) {
        if (e == hipSuccess) e = dalloc(c.cell_start, (size_t)G3 + 1);
        if (e == hipSuccess) e = dalloc(c.dense_slot, (size_t)G3 + 1);
    } else {
        if (e == hipSuccess) e = dalloc(c.tile_start, (size_t)(G3 >> 9) + 1);
        if (e == hipSuccess) e = dalloc(c.dep_tbuf, (size_t)(G3 >> 9) * dep_tile_record());
    }
    if (e == hipSuccess) e = dalloc(c.cell4, (size_t)2 * 4097 + 64);
    if (e == hipSuccess) e = dalloc(c.cellnode, (size_t)4096);
    c.start_level = 2;  // L/4 > L/5 >= r_cut
    while (c.start_level < kGroupLevelBits && p->box_size / (double)(1 << (c.start_level + 1)) >= p->r_cut)
        ++c.start_level;
    if (e == hipSuccess) e = dalloc(c.dev_counters, kCtrTotal);
    if (e == hipSuccess) e = dalloc(c.dev_inter, 2);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c.host_pinned, 4 * sizeof(unsigned long long), 0);
    if (e == hipSuccess) e = dalloc(c.dev_step, 1);
    if (e == hipSuccess) e = dalloc(c.diag_scratch, kDiagScratch);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c.host_step, sizeof(StepParams), 0);
    if (e == hipSuccess) e = pm_tables(c);
    if (e == hipSuccess) {
        for (int i = 0; i < 10 && e == hipSuccess; ++i) e = hipEventCreate(&c.ev[i]);
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreate(&c.ev_build[i]);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c.fork_ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c.join_ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking);  // default priority
        c.events_ok = (e == hipSuccess);
    }
    if (e != hipSuccess) {
        std::fprintf(stderr, "jtp_create: %s\n", hipGetErrorString(e));
        (void)hipGetLastError();
        jtp_destroy(h);
        return JTP_ERR_ALLOC;
    }
    make_split_poly(c);
    *out = h;
    return JTP_OK;
}

int jtp_destroy(jtp_ctx* h) {
    if (!h) return JTP_OK;
    Ctx& c = h->c;
    (void)hipDeviceSynchronize();
    free_particle_buffers(c);
    if (c.graph.exec) (void)hipGraphExecDestroy(c.graph.exec);
    if (c.graph.capture_stream) (void)hipStreamDestroy(c.graph.capture_stream);
    void* ptrs[] = {c.rho, c.phi, c.rho2, c.phi2, c.grad, c.grad2, c.spec, c.twiddle, c.gtab, c.cell_start,
                    c.tile_start, c.dep_tbuf,
                    c.cell4, c.cellnode, c.dev_counters, c.dev_inter, c.dense_slot, c.dev_step, c.diag_scratch,
                    c.pot_short, c.acc_scratch, c.kself, c.seg};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c.host_pinned) (void)hipHostFree(c.host_pinned);
    if (c.host_step) (void)hipHostFree(c.host_step);
    for (int i = 0; i < 10; ++i)
        if (c.ev[i]) (void)hipEventDestroy(c.ev[i]);
    for (int i = 0; i < 2; ++i)
        if (c.ev_build[i]) (void)hipEventDestroy(c.ev_build[i]);
    if (c.fork_ev) (void)hipEventDestroy(c.fork_ev);
    if (c.join_ev) (void)hipEventDestroy(c.join_ev);
    if (c.side) (void)hipStreamDestroy(c.side);
    delete h;
    return JTP_OK;
}

const char* jtp_last_error(const jtp_ctx* h) { return h ? h->c.err.c_str() : "null context"; }

int jtp_set_species_mass(jtp_ctx* h, double m_plus, double m_minus_abs) {
    if (!h) return JTP_ERR_ARG;
    if (!(m_plus >= 0) || !(m_minus_abs >= 0) || !std::isfinite(m_plus) || !std::isfinite(m_minus_abs))
        return fail(&h->c, JTP_ERR_ARG, "species masses must be finite and >= 0");
    h->c.m_plus = m_plus;
    h->c.m_minus = m_minus_abs;
    h->c.graph_dirty = true;  // the masses are kernel arguments of a captured step
    return JTP_OK;
}

// Enqueues one force evaluation and the copies of its status words to pinned host memory; no
// synchronisation, so it can be captured into a graph.  `timed` brackets the phases with events.
// Everything before the walk: keys, sort, gather, tree build (side stream) and PM phases, joined
// back into `s`.
static int enqueue_prep(Ctx& c, int64_t n, const uint32_t* d_pos, const int8_t* d_species, hipStream_t s,
                        bool timed) {
    if (timed) JTP_HIP(&c, hipEventRecord(c.ev[0], s));
    JTP_HIP(&c, launch_fill_i32((int32_t*)c.dev_inter, 4, 0, s));
    // the packed records live in the node-sum buffer, which the tree build (after the gather) overwrites
    uint4* rec = (uint4*)c.node_sum;
    JTP_HIP(&c, launch_morton_keys(d_pos, d_species, c.keys0, c.vals0, rec, n, s));
    JTP_HIP(&c, radix_sort_pairs(c, n, s, kKeyLowBits));
    const bool tiles = c.p.pm_enabled && c.p.precision == JTP_FP32;  // tile starts for the FP32 PM
    JTP_HIP(&c, launch_gather_parts(rec, c.vals0, c.parts, n, tiles ? c.tile_start : nullptr, c.logG, s));
    if (timed) JTP_HIP(&c, hipEventRecord(c.ev[1], s));
    // The tree build and the PM phases only share read-only inputs (sorted keys and records), so the
    // build runs on the side stream while the PM phases run on `s`; the walk waits for both.
    const bool fork = c.p.tree_enabled && c.p.pm_enabled;
    if (fork) {
        JTP_HIP(&c, hipEventRecord(c.fork_ev, s));
        JTP_HIP(&c, hipStreamWaitEvent(c.side, c.fork_ev, 0));
        if (timed) JTP_HIP(&c, hipEventRecord(c.ev_build[0], c.side));
        JTP_HIP(&c, build_tree(c, n, c.side));
        if (timed) JTP_HIP(&c, hipEventRecord(c.ev_build[1], c.side));
        JTP_HIP(&c, hipEventRecord(c.join_ev, c.side));
    } else if (c.p.tree_enabled) {
        if (timed) JTP_HIP(&c, hipEventRecord(c.ev_build[0], s));
        JTP_HIP(&c, build_tree(c, n, s));
        if (timed) JTP_HIP(&c, hipEventRecord(c.ev_build[1], s));
    }
    if (timed) JTP_HIP(&c, hipEventRecord(c.ev[2], s));
    if (c.p.pm_enabled) JTP_HIP(&c, pm_deposit(c, n, s));
    if (timed) JTP_HIP(&c, hipEventRecord(c.ev[3], s));
    if (c.p.pm_enabled) JTP_HIP(&c, pm_fft_solve(c, c.rho, c.phi, s));
    if (c.p.pm_enabled && c.p.interlace) JTP_HIP(&c, pm_fft_solve(c, c.rho2, c.phi2, s));
    if (timed) JTP_HIP(&c, hipEventRecord(c.ev[4], s));
    if (c.p.pm_enabled) JTP_HIP(&c, pm_interp(c, n, s));
    if (timed) JTP_HIP(&c, hipEventRecord(c.ev[5], s));
    if (fork) JTP_HIP(&c, hipStreamWaitEvent(s, c.join_ev, 0));
    return JTP_OK;
}

// Status words of an evaluation: interaction counts, group count, walk error flag, root flag.
static int enqueue_status(Ctx& c, int64_t n, hipStream_t s) {
    JTP_HIP(&c, hipMemcpyAsync(c.host_pinned, c.dev_inter, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    JTP_HIP(&c, hipMemcpyAsync(c.host_pinned + 3, c.dev_inter + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    JTP_HIP(&c, hipMemcpyAsync(c.host_pinned + 1, c.dev_counters, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (c.p.tree_enabled && n >= 2)
        JTP_HIP(&c, hipMemcpyAsync((int32_t*)(c.host_pinned + 2), c.dev_counters + kCtrRoot, sizeof(int32_t),
                                   hipMemcpyDeviceToHost, s));
    return JTP_OK;
}

static int enqueue_forces(Ctx& c, int64_t n, const uint32_t* d_pos, const int8_t* d_species, void* d_acc,
                          hipStream_t s, bool timed, double* pot = nullptr) {
    int rc = enqueue_prep(c, n, d_pos, d_species, s, timed);
    if (rc != JTP_OK) return rc;
    if (timed) JTP_HIP(&c, hipEventRecord(c.ev[7], s));
    if (c.p.tree_enabled) JTP_HIP(&c, walk(c, n, d_acc, s, pot));
    else JTP_HIP(&c, write_pm_only(c, n, d_acc, s));
    if (timed) JTP_HIP(&c, hipEventRecord(c.ev[6], s));
    return enqueue_status(c, n, s);
}

// Interprets the status words of the last force evaluation (after the stream synchronised).
static int check_status(Ctx& c, int64_t n) {
    c.have_result = true;
    if (c.p.tree_enabled) {
        const int32_t* cnt = (const int32_t*)(c.host_pinned + 1);
        if (n >= 2 && ((const int32_t*)(c.host_pinned + 2))[0] != 1)
            return fail(&c, JTP_ERR_STATE, "bottom-up pass did not reach the root (malformed hierarchy)");
        if (cnt[1] == 1) return fail(&c, JTP_ERR_OVERFLOW, "walk stack overflow: a group's traversal exceeded the stack bound");
        if (cnt[1] != 0) return fail(&c, JTP_ERR_STATE, "walk found a malformed hierarchy (code " + std::to_string(cnt[1]) + ")");
    }
    return JTP_OK;
}

static int prepare(Ctx& c, int64_t n) {
    const int64_t old_cap = c.cap_n;
    int rc = ensure_capacity(c, n);
    if (rc != JTP_OK) return rc;
    if (c.cap_n != old_cap) c.graph_dirty = true;  // buffers moved
    c.n_groups_max = (n + c.p.group_size - 1) / c.p.group_size + 4096;
    return JTP_OK;
}

int jtp_compute_forces(jtp_ctx* h, int64_t n, const uint32_t* d_pos, const int8_t* d_species, void* d_acc,
                       void* stream) {
    if (!h) return JTP_ERR_ARG;
    Ctx& c = h->c;
    if (n < 0 || n > (int64_t)0x7fffffff - 8192) return fail(&c, JTP_ERR_ARG, "n out of range");
    if (n > 0 && (!d_pos || !d_species || !d_acc)) return fail(&c, JTP_ERR_ARG, "null device pointer");
    hipStream_t s = (hipStream_t)stream;
    c.have_result = false;
    c.last_n = n;
    c.last_stream = s;
    c.last_kind = 0;
    if (n == 0) {
        c.have_result = true;
        return JTP_OK;
    }
    int rc = prepare(c, n);
    if (rc != JTP_OK) return rc;
    rc = enqueue_forces(c, n, d_pos, d_species, d_acc, s, true);
    if (rc != JTP_OK) return rc;
    JTP_HIP(&c, hipStreamSynchronize(s));
    return check_status(c, n);
}

// One step: kicks and drifts around an enqueued force evaluation (see jtp_step in the header).
static int enqueue_step(Ctx& c, int64_t n, uint32_t* d_pos, void* d_vel, const int8_t* d_species, void* d_acc,
                        int scheme, hipStream_t s, bool timed) {
    if (scheme == JTP_STEP_DKD) {
        JTP_HIP(&c, launch_drift(c, d_pos, d_vel, n, s));
        int rc = enqueue_forces(c, n, d_pos, d_species, d_acc, s, timed);
        if (rc != JTP_OK) return rc;
        JTP_HIP(&c, launch_kick(c, d_vel, d_acc, d_species, n, s));
        JTP_HIP(&c, launch_drift(c, d_pos, d_vel, n, s));
    } else {
        JTP_HIP(&c, launch_kick(c, d_vel, d_acc, d_species, n, s));
        JTP_HIP(&c, launch_drift(c, d_pos, d_vel, n, s));
        int rc = enqueue_forces(c, n, d_pos, d_species, d_acc, s, timed);
        if (rc != JTP_OK) return rc;
        JTP_HIP(&c, launch_kick(c, d_vel, d_acc, d_species, n, s));
    }
    return JTP_OK;
}

static void drop_graph(Ctx& c) {
    if (c.graph.exec) (void)hipGraphExecDestroy(c.graph.exec);
    c.graph.exec = nullptr;
    c.graph.n = -1;
}

// Captures the step into c.graph (on the context's own capture stream: the null stream cannot be
// captured) unless the cached graph was captured for the same arguments.
static int ensure_graph(Ctx& c, int64_t n, uint32_t* d_pos, void* d_vel, const int8_t* d_species, void* d_acc,
                        int scheme, hipStream_t s) {
    StepGraph& g = c.graph;
    if (g.exec && !c.graph_dirty && g.n == n && g.pos == d_pos && g.vel == d_vel && g.species == d_species &&
        g.acc == d_acc && g.scheme == scheme && g.stream == s)
        return JTP_OK;
    drop_graph(c);
    if (!g.capture_stream) JTP_HIP(&c, hipStreamCreateWithFlags(&g.capture_stream, hipStreamNonBlocking));
    JTP_HIP(&c, hipStreamBeginCapture(g.capture_stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(c, n, d_pos, d_vel, d_species, d_acc, scheme, g.capture_stream, false);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g.capture_stream, &graph);
    if (rc != JTP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(&c, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g.exec = nullptr;
        return hip_fail(&c, e, "hipGraphInstantiate");
    }
    g.n = n;
    g.pos = d_pos;
    g.vel = d_vel;
    g.species = d_species;
    g.acc = d_acc;
    g.scheme = scheme;
    g.stream = s;
    c.graph_dirty = false;
    return JTP_OK;
}

int jtp_step(jtp_ctx* h, int64_t n, uint32_t* d_pos, void* d_vel, const int8_t* d_species, void* d_acc,
             int32_t flags, double dt, double hubble_plus, double hubble_minus, double dtau_dt, void* stream) {
    if (!h) return JTP_ERR_ARG;
    Ctx& c = h->c;
    const int scheme = flags & 0xff;
    if ((flags & ~(0xff | JTP_STEP_GRAPH)) || (scheme != JTP_STEP_DKD && scheme != JTP_STEP_KDK))
        return fail(&c, JTP_ERR_ARG, "unknown step flags");
    if (n < 0 || n > (int64_t)0x7fffffff - 8192) return fail(&c, JTP_ERR_ARG, "n out of range");
    if (n > 0 && (!d_pos || !d_vel || !d_species || !d_acc)) return fail(&c, JTP_ERR_ARG, "null device pointer");
    if (!std::isfinite(dt) || !std::isfinite(hubble_plus) || !std::isfinite(hubble_minus) || !std::isfinite(dtau_dt))
        return fail(&c, JTP_ERR_ARG, "step scalars must be finite");
    hipStream_t s = (hipStream_t)stream;
    const bool use_graph = (flags & JTP_STEP_GRAPH) != 0;
    c.have_result = false;
    c.last_n = n;
    c.last_stream = s;
    c.last_kind = use_graph ? 2 : 1;
    if (n == 0) {
        c.have_result = true;
        return JTP_OK;
    }
    int rc = prepare(c, n);
    if (rc != JTP_OK) return rc;

    StepParams& sp = *c.host_step;  // the previous step synchronised, so the staging buffer is free
    std::memset(&sp, 0, sizeof(sp));
    const double units = 4294967296.0 / c.p.box_size;
    sp.kick_dt = scheme == JTP_STEP_DKD ? dt : 0.5 * dt;
    sp.drift_units = (scheme == JTP_STEP_DKD ? 0.5 * dt : dt) * units;
    sp.hubble_plus = hubble_plus;
    sp.hubble_minus = hubble_minus;
    sp.dtau_dt = dtau_dt;
    if (use_graph) {
        rc = ensure_graph(c, n, d_pos, d_vel, d_species, d_acc, scheme, s);
        if (rc != JTP_OK) return rc;
    }
    JTP_HIP(&c, hipEventRecord(c.ev[8], s));
    JTP_HIP(&c, hipMemcpyAsync(c.dev_step, c.host_step, sizeof(StepParams), hipMemcpyHostToDevice, s));
    if (use_graph) {
        JTP_HIP(&c, hipGraphLaunch(c.graph.exec, s));
    } else {
        rc = enqueue_step(c, n, d_pos, d_vel, d_species, d_acc, scheme, s, true);
        if (rc != JTP_OK) return rc;
    }
    JTP_HIP(&c, hipEventRecord(c.ev[9], s));
    JTP_HIP(&c, hipStreamSynchronize(s));
    return check_status(c, n);
}

int jtp_diagnostics(jtp_ctx* h, int64_t n, const void* d_vel, const int8_t* d_species, jtp_diag* out,
                    void* stream) {
    if (!h || !out) return JTP_ERR_ARG;
    Ctx& c = h->c;
    if (n < 0 || (n > 0 && (!d_vel || !d_species))) return fail(&c, JTP_ERR_ARG, "bad diagnostics arguments");
    hipStream_t s = (hipStream_t)stream;
    double* res = c.diag_scratch + 1024 * 8;
    JTP_HIP(&c, launch_diagnostics(c, n, d_vel, d_species, c.diag_scratch, res, s));
    double host[10];
    JTP_HIP(&c, hipMemcpyAsync(host, res, sizeof(host), hipMemcpyDeviceToHost, s));
    JTP_HIP(&c, hipStreamSynchronize(s));
    out->kinetic[0] = host[0];
    out->kinetic[1] = host[1];
    for (int k = 0; k < 3; ++k) {
        out->momentum[0][k] = host[2 + k];
        out->momentum[1][k] = host[5 + k];
    }
    out->pm_energy = c.p.interlace ? 0.25 * (host[8] + host[9]) : 0.5 * host[8];
    return JTP_OK;
}

void jtp_ic_default_params(jtp_ic_params* p) {
    if (!p) return;
    p->box_size = 50.0;
    p->n_grid = 128;
    p->z_init = 10.0;
    p->seed = 42;
    p->mu = 19.0;
    p->n_s = 0.965;
    p->delta_rms = 0.10;
}

int jtp_generate_ics(jtp_ctx* h, const jtp_ic_params* p, uint32_t* d_pos, void* d_vel, int32_t vel_f64,
                     int8_t* d_species, int64_t* n_plus_out, void* stream) {
    if (!h || !p) return JTP_ERR_ARG;
    Ctx& c = h->c;
    if (p->n_grid != c.p.n_cells) return fail(&c, JTP_ERR_ARG, "n_grid must equal the context's n_cells");
    if (p->box_size != c.p.box_size) return fail(&c, JTP_ERR_ARG, "box_size must equal the context's box_size");
    if (!(p->mu >= 0) || !std::isfinite(p->n_s) || !(p->delta_rms > 0) || !(p->z_init > -1))
        return fail(&c, JTP_ERR_ARG, "bad IC parameters");
    if (!d_pos || !d_vel || !d_species) return fail(&c, JTP_ERR_ARG, "null device pointer");
    const int64_t n = c.G3();
    int rc = prepare(c, n);  // sort buffers for the sign keys
    if (rc != JTP_OK) return rc;
    if (!c.grad) JTP_HIP(&c, dalloc_bytes(c.grad, (size_t)n * c.real_size()));  // third displacement grid
    const int64_t n_plus = (int64_t)std::llround((double)n / (1.0 + p->mu));  // ic_gen.rs:58-60
    hipStream_t s = (hipStream_t)stream;
    c.have_result = false;
    JTP_HIP(&c, generate_ics(c, *p, d_pos, d_vel, vel_f64, d_species, n_plus, s));
    JTP_HIP(&c, hipStreamSynchronize(s));
    if (n_plus_out) *n_plus_out = n_plus;
    return JTP_OK;
}

int jtp_segregation(jtp_ctx* h, int64_t n, const uint32_t* d_pos, const int8_t* d_species, int32_t n_cells,
                    double out[3], void* stream) {
    if (!h || !out) return JTP_ERR_ARG;
    Ctx& c = h->c;
    if (n_cells < 1 || n_cells > 1024) return fail(&c, JTP_ERR_ARG, "n_cells must be in [1, 1024]");
    if (n < 0 || n > (int64_t)0x7fffffff) return fail(&c, JTP_ERR_ARG, "n out of range");
    if (n > 0 && (!d_pos || !d_species)) return fail(&c, JTP_ERR_ARG, "null device pointer");
    hipStream_t s = (hipStream_t)stream;
    const int64_t cells = (int64_t)n_cells * n_cells * n_cells;
    if (c.seg_cap < cells) {
        JTP_HIP(&c, hipStreamSynchronize(s));
        if (c.seg) (void)hipFree(c.seg);
        c.seg = nullptr;
        c.seg_cap = 0;
        JTP_HIP(&c, dalloc(c.seg, (size_t)(2 * cells + 2)));
        c.seg_cap = cells;
    }
    double* res = c.diag_scratch + 1024 * 8;
    JTP_HIP(&c, launch_segregation(n, d_pos, d_species, n_cells, c.seg, c.diag_scratch, res, s));
    double v[9];
    JTP_HIP(&c, hipMemcpyAsync(v, res, sizeof(v), hipMemcpyDeviceToHost, s));
    JTP_HIP(&c, hipStreamSynchronize(s));
    segregation_metrics(v, cells, out);
    return JTP_OK;
}

int jtp_potential_energy(jtp_ctx* h, int64_t n, const uint32_t* d_pos, const int8_t* d_species, double* w_out,
                         void* stream) {
    if (!h || !w_out) return JTP_ERR_ARG;
    Ctx& c = h->c;
    if (c.p.precision != JTP_FP64 || c.p.split_mode != JTP_SPLIT_ERFC_GAUSS || !c.p.tree_enabled || !c.p.pm_enabled)
        return fail(&c, JTP_ERR_ARG, "potential energy needs an FP64 context with the erfc + Gaussian split, tree and PM");
    if (n < 0 || n > (int64_t)0x7fffffff - 8192) return fail(&c, JTP_ERR_ARG, "n out of range");
    if (n > 0 && (!d_pos || !d_species)) return fail(&c, JTP_ERR_ARG, "null device pointer");
    *w_out = 0;
    hipStream_t s = (hipStream_t)stream;
    c.have_result = false;
    c.last_n = n;
    c.last_kind = 0;
    if (n == 0) {
        c.have_result = true;
        return JTP_OK;
    }
    int rc = prepare(c, n);
    if (rc != JTP_OK) return rc;
    if (c.pot_cap < c.cap_n) {
        JTP_HIP(&c, dalloc(c.pot_short, (size_t)c.cap_n));
        JTP_HIP(&c, dalloc(c.acc_scratch, (size_t)3 * c.cap_n));
        c.pot_cap = c.cap_n;
    }
    if (!c.kself_ready) {
        // the mesh potential of a unit mass at node 0, at the 27 nodes around it
        JTP_HIP(&c, dalloc(c.kself, 27));
        JTP_HIP(&c, launch_fill_i32((int32_t*)c.rho, 2 * c.G3(), 0, s));
        const double one = 1.0;
        JTP_HIP(&c, hipStreamSynchronize(s));
        JTP_HIP(&c, hipMemcpy(c.rho, &one, sizeof(double), hipMemcpyHostToDevice));
        JTP_HIP(&c, pm_fft_solve(c, c.rho, c.phi, s, 0));
        JTP_HIP(&c, hipStreamSynchronize(s));
        double k27[27];
        const int64_t G = c.G();
        for (int d = 0; d < 27; ++d) {
            const int64_t dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
            const int64_t node = (((dz + G) % G) * G + (dy + G) % G) * G + (dx + G) % G;
            JTP_HIP(&c, hipMemcpy(&k27[d], (const double*)c.phi + node, sizeof(double), hipMemcpyDeviceToHost));
        }
        JTP_HIP(&c, hipMemcpy(c.kself, k27, sizeof(k27), hipMemcpyHostToDevice));
        c.kself_ready = true;
    }
    rc = enqueue_forces(c, n, d_pos, d_species, c.acc_scratch, s, true, c.pot_short);
    if (rc != JTP_OK) return rc;
    // The force's Green function is matched to the gradient stencil, which distorts the potential at
    // a few cells; the energy uses the potential of the plain deconvolved Gaussian-split Green function
    // (the same one the self-energy kernel was computed with).
    JTP_HIP(&c, pm_fft_solve(c, c.rho, c.phi, s, 0));
    if (c.p.interlace) JTP_HIP(&c, pm_fft_solve(c, c.rho2, c.phi2, s, 0));
    double* res = c.diag_scratch + 1024 * 8;
    JTP_HIP(&c, launch_energy_terms(c, n, c.kself, c.pot_short, c.diag_scratch, res, s));
    double host[5];
    JTP_HIP(&c, hipMemcpyAsync(host, res, sizeof(host), hipMemcpyDeviceToHost, s));
    JTP_HIP(&c, hipStreamSynchronize(s));
    rc = check_status(c, n);
    if (rc != JTP_OK) return rc;
    // The mesh drops the k = 0 mode (a uniform background of net density -Q/V); its long-range part
    // is thereby accounted for, but the short-range sum runs over particles only, so the background's
    // short-range potential G (Q/V) 4 pi r_s^2 (the integral of erfc(r / 2 r_s) / r) is added:
    // W_bg = 1/2 Q x 4 pi G r_s^2 Q / V (the Ewald background term pi Q^2 / (V alpha^2), alpha = 1/2 r_s).
    const double L = c.p.box_size, Q = host[3], V = L * L * L;
    const double w_bg = 2.0 * M_PI * c.p.g_constant * c.p.r_s * c.p.r_s * Q * Q / V;
    // Subtracting the mesh's self-potential S_i removes each particle's long-range interaction with
    // its own periodic images too, which belongs to the periodic energy.  A lone mass in the box has
    // W = 1/2 G m^2 xi / L (xi = 2.837297479, the simple cubic Madelung constant; the gravitational
    // sign is +); of that, the background's short-range share is 4 pi r_s^2 / V (in w_bg) and the
    // images' short-range share, -sum_n erfc(|n| L / 2 r_s) / |n| L, lies beyond r_cut <= L / 5 and
    // is not part of the cut force law, so the long-range share is xi / L - 4 pi r_s^2 / V plus the
    // image sum (below 1e-12 of it for r_s <= L / 10).
    double img = 2.837297479480619 / L - 4.0 * M_PI * c.p.r_s * c.p.r_s / V;
    for (int i = -2; i <= 2; ++i)
        for (int j = -2; j <= 2; ++j)
            for (int k = -2; k <= 2; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                const double r = L * std::sqrt((double)(i * i + j * j + k * k));
                img += std::erfc(r * 0.5 / c.p.r_s) / r;
            }
    const double w_img = 0.5 * c.p.g_constant * host[4] * img;
    *w_out = 0.5 * (host[0] + host[1] - host[2]) + w_bg + w_img;
    return JTP_OK;
}

int jtp_pos_to_fixed_f64(jtp_ctx* h, int64_t n, const double* d_pos, uint32_t* d_fixed, void* stream) {
    if (!h || n < 0 || (n > 0 && (!d_pos || !d_fixed))) return JTP_ERR_ARG;
    JTP_HIP(&h->c, launch_fixed_from_f64(d_pos, d_fixed, 3 * n, h->c.p.box_size, (hipStream_t)stream));
    return JTP_OK;
}
int jtp_pos_to_fixed_f32(jtp_ctx* h, int64_t n, const float* d_pos, uint32_t* d_fixed, void* stream) {
    if (!h || n < 0 || (n > 0 && (!d_pos || !d_fixed))) return JTP_ERR_ARG;
    JTP_HIP(&h->c, launch_fixed_from_f32(d_pos, d_fixed, 3 * n, h->c.p.box_size, (hipStream_t)stream));
    return JTP_OK;
}

int jtp_pm_solve(jtp_ctx* h, const void* d_rho, void* d_phi, void* stream) {
    if (!h) return JTP_ERR_ARG;
    if (!d_rho || !d_phi) return fail(&h->c, JTP_ERR_ARG, "null grid pointer");
    hipStream_t s = (hipStream_t)stream;
    JTP_HIP(&h->c, pm_fft_solve(h->c, d_rho, d_phi, s));
    JTP_HIP(&h->c, hipStreamSynchronize(s));
    return JTP_OK;
}

int jtp_morton_keys_f32(int64_t n, const float* d_pos, float box_half, float inv_cell_size, uint64_t* d_keys,
                        void* stream) {
    if (n < 0 || (n > 0 && (!d_pos || !d_keys))) return JTP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (launch_morton_keys_f32(d_pos, d_keys, n, box_half, inv_cell_size, s) != hipSuccess) return JTP_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return JTP_ERR_HIP;
    return JTP_OK;
}

int jtp_morton30_keys_f32(int64_t n, const float* d_pos, float box_half, float /*inv_cell_size*/, uint64_t* d_keys,
                          int32_t* d_indices, void* stream) {
    if (n < 0 || n > (int64_t)0x7fffffff || (n > 0 && (!d_pos || !d_keys || !d_indices))) return JTP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (launch_morton30_keys_f32(d_pos, d_keys, d_indices, n, box_half, s) != hipSuccess) return JTP_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return JTP_ERR_HIP;
    return JTP_OK;
}

int jtp_get_stats(jtp_ctx* h, jtp_stats* out) {
    if (!h || !out) return JTP_ERR_ARG;
    Ctx& c = h->c;
    std::memset(out, 0, sizeof(*out));
    if (!c.have_result) return fail(&c, JTP_ERR_STATE, "no completed force evaluation");
    out->n_particles = c.last_n;
    if (c.last_n <= 0) return JTP_OK;
    // force phases from ev[0..6] (not recorded inside a captured step: reported as -1);
    // the total covers the force call, or the whole step with its kicks and drifts
    float ms[6] = {-1, -1, -1, -1, -1, -1};
    if (c.last_kind < 2) {
        // sort ev0-ev1; build on its own events (side stream); deposit ev2-ev3, FFT ev3-ev4,
        // interpolation ev4-ev5; walk ev7-ev6 (ev7 follows the wait for the build)
        JTP_HIP(&c, hipEventSynchronize(c.ev[6]));
        JTP_HIP(&c, hipEventElapsedTime(&ms[0], c.ev[0], c.ev[1]));
        if (c.p.tree_enabled) JTP_HIP(&c, hipEventElapsedTime(&ms[1], c.ev_build[0], c.ev_build[1]));
        else ms[1] = 0;
        for (int i = 2; i < 5; ++i) JTP_HIP(&c, hipEventElapsedTime(&ms[i], c.ev[i], c.ev[i + 1]));
        JTP_HIP(&c, hipEventElapsedTime(&ms[5], c.ev[7], c.ev[6]));
    }
    out->ms_sort = ms[0];
    out->ms_build = ms[1];
    out->ms_deposit = ms[2];
    out->ms_fft = ms[3];
    out->ms_interp = ms[4];
    out->ms_walk = ms[5];
    float tot = -1;
    if (c.last_kind == 0) {
        JTP_HIP(&c, hipEventElapsedTime(&tot, c.ev[0], c.ev[6]));
    } else if (c.last_kind != 3) {
        JTP_HIP(&c, hipEventSynchronize(c.ev[9]));
        JTP_HIP(&c, hipEventElapsedTime(&tot, c.ev[8], c.ev[9]));
    }
    out->ms_total = tot;
    out->interactions = c.host_pinned[0];
    out->interactions_mono = c.host_pinned[3];
    out->n_groups = ((const int32_t*)(c.host_pinned + 1))[0];
    return JTP_OK;
}

int jtp_stage_size(jtp_ctx* h, int32_t stage, size_t* bytes) {
    if (!h || !bytes) return JTP_ERR_ARG;
    Ctx& c = h->c;
    const int64_t n = c.last_n < 0 ? 0 : c.last_n;
    const int64_t nI = n > 1 ? n - 1 : 0;
    const size_t rs = c.real_size();
    switch (stage) {
        case JTP_STAGE_KEYS: *bytes = n * 8; break;
        case JTP_STAGE_PERM: *bytes = n * 4; break;
        case JTP_STAGE_NODES: *bytes = nI * 16; break;
        case JTP_STAGE_MONO: *bytes = nI * 64; break;
        case JTP_STAGE_SUMS: *bytes = nI * 64; break;
        case JTP_STAGE_RHO:
        case JTP_STAGE_PHI: *bytes = c.G3() * rs; break;
        case JTP_STAGE_RHO2:
        case JTP_STAGE_PHI2: *bytes = c.p.interlace ? c.G3() * rs : 0; break;
        case JTP_STAGE_ACC_PM: *bytes = n * 3 * rs; break;
        case JTP_STAGE_BOXES: *bytes = nI * 24; break;
        default: return fail(&c, JTP_ERR_ARG, "unknown stage");
    }
    return JTP_OK;
}

int jtp_stage_copy(jtp_ctx* h, int32_t stage, void* host_dst, size_t bytes) {
    if (!h || !host_dst) return JTP_ERR_ARG;
    Ctx& c = h->c;
    size_t need = 0;
    int rc = jtp_stage_size(h, stage, &need);
    if (rc != JTP_OK) return rc;
    if (bytes < need) return fail(&c, JTP_ERR_ARG, "destination too small");
    if (!c.have_result) return fail(&c, JTP_ERR_STATE, "no completed force evaluation");
    if (need == 0) return JTP_OK;
    JTP_HIP(&c, hipDeviceSynchronize());
    const int64_t n = c.last_n;
    const int64_t nI = n - 1;
    switch (stage) {
        case JTP_STAGE_KEYS: JTP_HIP(&c, hipMemcpy(host_dst, c.keys0, need, hipMemcpyDeviceToHost)); break;
        case JTP_STAGE_PERM: JTP_HIP(&c, hipMemcpy(host_dst, c.vals0, need, hipMemcpyDeviceToHost)); break;
        case JTP_STAGE_RHO: JTP_HIP(&c, hipMemcpy(host_dst, c.rho, need, hipMemcpyDeviceToHost)); break;
        case JTP_STAGE_PHI: JTP_HIP(&c, hipMemcpy(host_dst, c.phi, need, hipMemcpyDeviceToHost)); break;
        case JTP_STAGE_RHO2: JTP_HIP(&c, hipMemcpy(host_dst, c.rho2, need, hipMemcpyDeviceToHost)); break;
        case JTP_STAGE_PHI2: JTP_HIP(&c, hipMemcpy(host_dst, c.phi2, need, hipMemcpyDeviceToHost)); break;
        case JTP_STAGE_NODES:
        case JTP_STAGE_BOXES: {
            if (!c.p.tree_enabled) return fail(&c, JTP_ERR_STATE, "tree disabled");
            std::vector<uint4> geo(2 * nI);
            std::vector<int2> kids(nI);
            JTP_HIP(&c, hipMemcpy(geo.data(), c.node_geo, geo.size() * sizeof(uint4), hipMemcpyDeviceToHost));
            JTP_HIP(&c, hipMemcpy(kids.data(), c.node_kids, kids.size() * sizeof(int2), hipMemcpyDeviceToHost));
            if (stage == JTP_STAGE_NODES) {
                int32_t* o = (int32_t*)host_dst;
                for (int64_t i = 0; i < nI; ++i) {
                    o[4 * i] = (int32_t)geo[2 * i].w;
                    o[4 * i + 1] = (int32_t)geo[2 * i + 1].w;
                    o[4 * i + 2] = kids[i].x;
                    o[4 * i + 3] = kids[i].y;
                }
            } else {
                uint32_t* o = (uint32_t*)host_dst;
                for (int64_t i = 0; i < nI; ++i) {  // stored as (centre, extent): min = c - e / 2, max = min + e
                    const uint32_t c3[3] = {geo[2 * i].x, geo[2 * i].y, geo[2 * i].z};
                    const uint32_t e3[3] = {geo[2 * i + 1].x, geo[2 * i + 1].y, geo[2 * i + 1].z};
                    for (int a = 0; a < 3; ++a) {
                        o[6 * i + a] = c3[a] - e3[a] / 2;
                        o[6 * i + 3 + a] = c3[a] - e3[a] / 2 + e3[a];
                    }
                }
            }
            break;
        }
        case JTP_STAGE_SUMS: {
            if (!c.p.tree_enabled) return fail(&c, JTP_ERR_STATE, "tree disabled");
            std::vector<uint64_t> raw(nI * 8);
            JTP_HIP(&c, hipMemcpy(raw.data(), c.node_sum, nI * 64, hipMemcpyDeviceToHost));
            uint64_t* o = (uint64_t*)host_dst;
            for (int64_t i = 0; i < nI; ++i) {
                const uint64_t* r = &raw[8 * i];
                const uint32_t* cnt = (const uint32_t*)(r + 6);
                o[8 * i] = cnt[0];
                o[8 * i + 1] = r[0]; o[8 * i + 2] = r[1]; o[8 * i + 3] = r[2];
                o[8 * i + 4] = cnt[1];
                o[8 * i + 5] = r[3]; o[8 * i + 6] = r[4]; o[8 * i + 7] = r[5];
            }
            break;
        }
        case JTP_STAGE_MONO: {
            if (!c.p.tree_enabled) return fail(&c, JTP_ERR_STATE, "tree disabled");
            double* o = (double*)host_dst;
            if (c.p.precision == JTP_FP64) {
                std::vector<double> raw(nI * 8);
                JTP_HIP(&c, hipMemcpy(raw.data(), c.node_mono, nI * 64, hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < nI; ++i) {
                    o[8 * i] = raw[8 * i + 3];
                    o[8 * i + 1] = raw[8 * i]; o[8 * i + 2] = raw[8 * i + 1]; o[8 * i + 3] = raw[8 * i + 2];
                    o[8 * i + 4] = raw[8 * i + 7];
                    o[8 * i + 5] = raw[8 * i + 4]; o[8 * i + 6] = raw[8 * i + 5]; o[8 * i + 7] = raw[8 * i + 6];
                }
            } else {
                std::vector<uint32_t> raw(nI * 8);
                JTP_HIP(&c, hipMemcpy(raw.data(), c.node_mono, nI * 32, hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < nI; ++i) {
                    const uint32_t* r = &raw[8 * i];
                    float qp, qm;
                    std::memcpy(&qp, &r[3], 4);
                    std::memcpy(&qm, &r[7], 4);
                    o[8 * i] = qp;
                    o[8 * i + 1] = r[0]; o[8 * i + 2] = r[1]; o[8 * i + 3] = r[2];
                    o[8 * i + 4] = qm;
                    o[8 * i + 5] = r[4]; o[8 * i + 6] = r[5]; o[8 * i + 7] = r[6];
                }
            }
            break;
        }
        case JTP_STAGE_ACC_PM: {
            if (!c.p.pm_enabled) return fail(&c, JTP_ERR_STATE, "PM disabled");
            const size_t rs = c.real_size();
            std::vector<char> raw(n * 3 * rs);
            std::vector<int32_t> perm(n);
            JTP_HIP(&c, hipMemcpy(raw.data(), c.pm_acc, raw.size(), hipMemcpyDeviceToHost));
            JTP_HIP(&c, hipMemcpy(perm.data(), c.vals0, n * 4, hipMemcpyDeviceToHost));
            char* o = (char*)host_dst;
            for (int64_t i = 0; i < n; ++i) std::memcpy(o + perm[i] * 3 * rs, raw.data() + i * 3 * rs, 3 * rs);
            break;
        }
        default: return fail(&c, JTP_ERR_ARG, "unknown stage");
    }
    return JTP_OK;
}

void* jtp_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}
void jtp_free(void* p) {
    if (p) (void)hipFree(p);
}
int jtp_copy_h2d(void* d, const void* h, size_t b) { return hipMemcpy(d, h, b, hipMemcpyHostToDevice) == hipSuccess ? JTP_OK : JTP_ERR_HIP; }
int jtp_copy_d2h(void* h, const void* d, size_t b) { return hipMemcpy(h, d, b, hipMemcpyDeviceToHost) == hipSuccess ? JTP_OK : JTP_ERR_HIP; }
int jtp_copy_d2d(void* d, const void* s, size_t b) {
    if (b && (!d || !s)) return JTP_ERR_ARG;
    if (device_copy(d, s, b, nullptr) != hipSuccess) return JTP_ERR_HIP;
    return hipStreamSynchronize(nullptr) == hipSuccess ? JTP_OK : JTP_ERR_HIP;
}
int jtp_device_memory(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return JTP_ERR_ARG;
    return hipMemGetInfo(free_bytes, total_bytes) == hipSuccess ? JTP_OK : JTP_ERR_HIP;
}
int jtp_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? JTP_OK : JTP_ERR_HIP; }
int jtp_device_count(int* out) {
    if (!out) return JTP_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        *out = 0;
        return JTP_ERR_HIP;
    }
    *out = n;
    return JTP_OK;
}

// ---------------------------------------------------------------- ensembles of boxes
constexpr int kMaxEnsStreams = 8;

struct jtp_ensemble {
    std::vector<jtp_ctx*> boxes;
    std::vector<hipStream_t> streams;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    hipEvent_t join[kMaxEnsStreams]{};
    std::string err;
    // captured batch
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
    std::vector<int64_t> g_n;
    std::vector<const void*> g_ptrs;
    hipStream_t g_stream = nullptr;
    int last_mode = 0;  // 0 concurrent, 1 sequential, 2 graph
    bool have_result = false;
    // one walk launch over all boxes: per-box walk arguments (pinned host copy, device copy) and the
    // launch's group offsets and queue heads
    unsigned char* h_args = nullptr;
    void* d_args = nullptr;
    int32_t* d_ctrl = nullptr;
};

static int ens_fail(jtp_ensemble* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    return code;
}
#define ENS_HIP(e, call)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return ens_fail((e), JTP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int jtp_ensemble_destroy(jtp_ensemble* e) {
    if (!e) return JTP_OK;
    (void)hipDeviceSynchronize();
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->capture_stream) (void)hipStreamDestroy(e->capture_stream);
    for (hipStream_t s : e->streams)
        if (s) (void)hipStreamDestroy(s);
    hipEvent_t evs[] = {e->fork, e->t0, e->t1};
    for (hipEvent_t v : evs)
        if (v) (void)hipEventDestroy(v);
    for (int k = 0; k < kMaxEnsStreams; ++k)
        if (e->join[k]) (void)hipEventDestroy(e->join[k]);
    for (jtp_ctx* b : e->boxes) jtp_destroy(b);
    if (e->h_args) (void)hipHostFree(e->h_args);
    if (e->d_args) (void)hipFree(e->d_args);
    if (e->d_ctrl) (void)hipFree(e->d_ctrl);
    delete e;
    return JTP_OK;
}

int jtp_ensemble_create(int32_t n_boxes, const jtp_params* params, const double* m_plus, const double* m_minus_abs,
                        int32_t n_streams, jtp_ensemble** out) {
    if (!out) return JTP_ERR_ARG;
    *out = nullptr;
    if (n_boxes < 1 || !params || !m_plus || !m_minus_abs) return JTP_ERR_ARG;
    if (n_streams == 0) n_streams = 4;
    if (n_streams < 1 || n_streams > kMaxEnsStreams) return JTP_ERR_ARG;
    jtp_ensemble* e = new (std::nothrow) jtp_ensemble();
    if (!e) return JTP_ERR_ALLOC;
    for (int b = 0; b < n_boxes; ++b) {
        jtp_ctx* h = nullptr;
        int rc = jtp_create(&params[b], &h);
        if (rc == JTP_OK) rc = jtp_set_species_mass(h, m_plus[b], m_minus_abs[b]);
        if (rc != JTP_OK) {
            if (h) jtp_destroy(h);
            jtp_ensemble_destroy(e);
            return rc;
        }
        e->boxes.push_back(h);
    }
    hipError_t er = hipSuccess;
    for (int k = 0; k < n_streams && er == hipSuccess; ++k) {
        hipStream_t s = nullptr;
        er = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);  // default priority
        if (er == hipSuccess) e->streams.push_back(s);
        if (er == hipSuccess) er = hipEventCreateWithFlags(&e->join[k], hipEventDisableTiming);
    }
    if (er == hipSuccess) er = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
    if (er == hipSuccess) er = hipEventCreate(&e->t0);
    if (er == hipSuccess) er = hipEventCreate(&e->t1);
    const size_t ab = walk_args_bytes() * (size_t)n_boxes;
    if (er == hipSuccess) er = hipHostMalloc((void**)&e->h_args, ab, hipHostMallocDefault);
    if (er == hipSuccess) er = hipMalloc(&e->d_args, ab);
    if (er == hipSuccess) er = hipMalloc((void**)&e->d_ctrl, sizeof(int32_t) * (size_t)n_boxes * (2 + kXcds));
    if (er != hipSuccess) {
        (void)hipGetLastError();
        jtp_ensemble_destroy(e);
        return JTP_ERR_HIP;
    }
    *out = e;
    return JTP_OK;
}

const char* jtp_ensemble_last_error(const jtp_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

// Enqueues every box.  concurrent: fork from `s`, box b on streams[b % K], join back into `s`.
// one_walk: the boxes' preparation (keys, sort, build, PM) as above, then one walk launch per
// group of boxes with equal walk_batch_key over all their groups, on `s`.
static int ens_enqueue(jtp_ensemble* e, const int64_t* n, const uint32_t* const* pos, const int8_t* const* sp,
                       void* const* acc, bool concurrent, bool one_walk, hipStream_t s, bool timed) {
    const int K = (int)e->streams.size();
    if (concurrent) {
        ENS_HIP(e, hipEventRecord(e->fork, s));
        for (int k = 0; k < K; ++k) ENS_HIP(e, hipStreamWaitEvent(e->streams[k], e->fork, 0));
    }
    std::vector<int> batched;  // boxes whose walk joins the shared launch
    for (size_t b = 0; b < e->boxes.size(); ++b) {
        if (n[b] == 0) continue;
        Ctx& c = e->boxes[b]->c;
        hipStream_t bs = concurrent ? e->streams[b % K] : s;
        int rc;
        if (one_walk && c.p.tree_enabled) {
            rc = enqueue_prep(c, n[b], pos[b], sp[b], bs, false);
            batched.push_back((int)b);
        } else {
            rc = enqueue_forces(c, n[b], pos[b], sp[b], acc[b], bs, timed);
        }
        if (rc != JTP_OK) return ens_fail(e, rc, "box " + std::to_string(b) + ": " + c.err);
    }
    if (concurrent) {
        for (int k = 0; k < K; ++k) {
            ENS_HIP(e, hipEventRecord(e->join[k], e->streams[k]));
            ENS_HIP(e, hipStreamWaitEvent(s, e->join[k], 0));
        }
    }
    if (batched.empty()) return JTP_OK;
    // the walk arguments of the batched boxes, grouped by key (each group contiguous)
    const size_t ab = walk_args_bytes();
    std::vector<int> keys;
    for (int b : batched) {
        const int k = walk_batch_key(e->boxes[b]->c);
        if (std::find(keys.begin(), keys.end(), k) == keys.end()) keys.push_back(k);
    }
    size_t slot = 0, ctrl = 0;
    std::vector<std::pair<size_t, int>> launches;  // first slot, box count
    std::vector<int64_t> items;
    for (int k : keys) {
        const size_t first = slot;
        int64_t gmax = 0;
        for (int b : batched) {
            Ctx& c = e->boxes[b]->c;
            if (walk_batch_key(c) != k) continue;
            walk_fill_args(c, n[b], acc[b], e->h_args + ab * slot++);
            gmax += c.n_groups_max;
        }
        launches.emplace_back(first, (int)(slot - first));
        items.push_back(gmax);
    }
    ENS_HIP(e, hipMemcpyAsync(e->d_args, e->h_args, ab * slot, hipMemcpyHostToDevice, s));
    for (size_t i = 0; i < launches.size(); ++i) {
        const int m = launches[i].second;
        ENS_HIP(e, walk_boxes((const unsigned char*)e->d_args + ab * launches[i].first, m, e->d_ctrl + ctrl,
                              items[i], keys[i], s));
        ctrl += (size_t)m + 1 + kXcds;
    }
    for (int b : batched) {
        int rc = enqueue_status(e->boxes[b]->c, n[b], s);
        if (rc != JTP_OK) return ens_fail(e, rc, "box " + std::to_string(b) + ": " + e->boxes[b]->c.err);
    }
    return JTP_OK;
}

int jtp_ensemble_compute_forces(jtp_ensemble* e, const int64_t* n, const uint32_t* const* pos,
                                const int8_t* const* sp, void* const* acc, int32_t flags, void* stream) {
    if (!e) return JTP_ERR_ARG;
    if (!n || !pos || !sp || !acc) return ens_fail(e, JTP_ERR_ARG, "null array");
    if (flags & ~(JTP_ENS_SEQUENTIAL | JTP_ENS_WALK_PER_BOX | JTP_ENS_GRAPH)) return ens_fail(e, JTP_ERR_ARG, "unknown flags");
    const bool graph = (flags & JTP_ENS_GRAPH) != 0;
    const bool concurrent = !(flags & JTP_ENS_SEQUENTIAL);
    const bool one_walk = !(flags & JTP_ENS_WALK_PER_BOX);
    hipStream_t s = (hipStream_t)stream;
    const size_t B = e->boxes.size();
    e->have_result = false;
    bool dirty = false;
    for (size_t b = 0; b < B; ++b) {
        if (n[b] < 0 || n[b] > (int64_t)0x7fffffff - 8192) return ens_fail(e, JTP_ERR_ARG, "n out of range");
        if (n[b] > 0 && (!pos[b] || !sp[b] || !acc[b])) return ens_fail(e, JTP_ERR_ARG, "null device pointer");
        Ctx& c = e->boxes[b]->c;
        c.have_result = false;
        c.last_n = n[b];
        c.last_kind = (graph || (one_walk && c.p.tree_enabled)) ? 3 : 0;  // 3: phases not timed per box
        c.last_stream = s;
        if (n[b] > 0) {
            int rc = prepare(c, n[b]);  // allocations happen here, outside any capture
            if (rc != JTP_OK) return ens_fail(e, rc, "box " + std::to_string(b) + ": " + c.err);
            dirty = dirty || c.graph_dirty;
        }
    }
    e->last_mode = graph ? 2 : (concurrent ? 0 : 1);
    if (graph) {
        std::vector<const void*> ptrs;
        for (size_t b = 0; b < B; ++b) {
            ptrs.push_back(pos[b]);
            ptrs.push_back(sp[b]);
            ptrs.push_back(acc[b]);
        }
        std::vector<int64_t> nv(n, n + B);
        ptrs.push_back((const void*)(intptr_t)((concurrent ? 1 : 0) | (one_walk ? 2 : 0)));
        if (!e->exec || dirty || nv != e->g_n || ptrs != e->g_ptrs || s != e->g_stream) {
            if (e->exec) (void)hipGraphExecDestroy(e->exec);
            e->exec = nullptr;
            if (!e->capture_stream) ENS_HIP(e, hipStreamCreateWithFlags(&e->capture_stream, hipStreamNonBlocking));
            ENS_HIP(e, hipStreamBeginCapture(e->capture_stream, hipStreamCaptureModeThreadLocal));
            int rc = ens_enqueue(e, n, pos, sp, acc, concurrent, one_walk, e->capture_stream, false);
            hipGraph_t g = nullptr;
            hipError_t er = hipStreamEndCapture(e->capture_stream, &g);
            if (rc != JTP_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (er != hipSuccess) return ens_fail(e, JTP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(er));
            er = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (er != hipSuccess) {
                e->exec = nullptr;
                return ens_fail(e, JTP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(er));
            }
            e->g_n = nv;
            e->g_ptrs = ptrs;
            e->g_stream = s;
            // the captured kernels bake in every box's current masses and buffers
            for (size_t b = 0; b < B; ++b) e->boxes[b]->c.graph_dirty = false;
        }
        ENS_HIP(e, hipEventRecord(e->t0, s));
        ENS_HIP(e, hipGraphLaunch(e->exec, s));
    } else {
        ENS_HIP(e, hipEventRecord(e->t0, s));
        int rc = ens_enqueue(e, n, pos, sp, acc, concurrent, one_walk, s, true);
        if (rc != JTP_OK) return rc;
    }
    ENS_HIP(e, hipEventRecord(e->t1, s));
    ENS_HIP(e, hipStreamSynchronize(s));
    e->have_result = true;
    for (size_t b = 0; b < B; ++b) {
        Ctx& c = e->boxes[b]->c;
        if (n[b] == 0) {
            c.have_result = true;
            continue;
        }
        int rc = check_status(c, n[b]);
        if (rc != JTP_OK) return ens_fail(e, rc, "box " + std::to_string(b) + ": " + c.err);
    }
    return JTP_OK;
}

int jtp_ensemble_get_stats(jtp_ensemble* e, int32_t box, jtp_stats* out) {
    if (!e || !out) return JTP_ERR_ARG;
    if (box >= (int32_t)e->boxes.size() || box < -1) return ens_fail(e, JTP_ERR_ARG, "box out of range");
    if (!e->have_result) return ens_fail(e, JTP_ERR_STATE, "no completed ensemble evaluation");
    if (box >= 0) {
        int rc = jtp_get_stats(e->boxes[box], out);
        if (rc != JTP_OK) return ens_fail(e, rc, e->boxes[box]->c.err);
        return JTP_OK;  // inside the batch graph nothing is timed per box (all -1)
    }
    std::memset(out, 0, sizeof(*out));
    ENS_HIP(e, hipEventSynchronize(e->t1));
    float ms = 0;
    ENS_HIP(e, hipEventElapsedTime(&ms, e->t0, e->t1));
    out->ms_total = ms;
    out->ms_sort = out->ms_build = out->ms_deposit = out->ms_fft = out->ms_interp = out->ms_walk = -1;
    for (jtp_ctx* h : e->boxes) {
        const Ctx& c = h->c;
        if (c.last_n <= 0) continue;
        out->interactions += c.host_pinned[0];
        out->interactions_mono += c.host_pinned[3];
        out->n_groups += ((const int32_t*)(c.host_pinned + 1))[0];
        out->n_particles += c.last_n;
    }
    return JTP_OK;
}

}  // extern "C"
