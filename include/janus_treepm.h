/*
 * janus_treepm.h — C ABI of the MI355X (gfx950) bimetric TreePM force step.
 *
 * One context owns every plan, table and scratch buffer of one force
 * evaluation on one GPU; nothing is process-global, so distinct contexts may
 * be driven from distinct host threads.  Every function returns an int:
 * JTP_OK (0) on success or a negative JTP_ERR_* code, with the message
 * available from jtp_last_error().  No function aborts the process.
 *
 * Particle arrays are caller-owned DEVICE pointers.  Positions are 32-bit
 * fixed point per axis, u = floor((x/L + 1/2) * 2^32) mod 2^32 for x in
 * [-L/2, L/2) (the reference's box convention, src/treepm/pm_grid.rs:64-70),
 * stored xyz-interleaved (u[3*i+0..2]) like the reference's f32 particle
 * arrays (src/nbody_gpu_twopass.rs:1817-1823).  Species are int8 +1 / -1
 * (MassSign, src/lib.rs:82-86).  Same species attract, opposite species repel.
 *
 * Streams are passed as `void*` holding a hipStream_t (NULL = null stream).
 */
#ifndef JANUS_TREEPM_H
#define JANUS_TREEPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (reference: cufft_wrapper.cu:13-27 returns -1 on any error) */
#define JTP_OK            0
#define JTP_ERR_ARG      -1   /* invalid argument (null pointer, bad size, bad parameter) */
#define JTP_ERR_HIP      -2   /* a HIP runtime call failed                               */
#define JTP_ERR_ALLOC    -3   /* device allocation failed                                */
#define JTP_ERR_OVERFLOW -4   /* a bounded on-device structure overflowed (walk stack)   */
#define JTP_ERR_STATE    -5   /* call out of order (e.g. stage read before a compute)    */

/* ---- enumerations */
#define JTP_FP32 0            /* fixed-point positions, FP32 walk/PM, Kahan accumulation */
#define JTP_FP64 1            /* plain FP64 form of the same kernels                     */

#define JTP_SPLIT_ERFC_GAUSS 0 /* S = erfc(x) + 2x/sqrt(pi) e^{-x^2}, x = rho/(2 r_s): force-consistent with PM e^{-k^2 r_s^2} */
#define JTP_SPLIT_ERFC       1 /* S = erfc(x) only: reference GPU walk (nbody_gpu_twopass.rs:1776-1786) */
#define JTP_SPLIT_POLY4      2 /* S = 1-(r/r_cut)^4: reference CPU tree (tree_short.rs:17-31, splitting.rs:20-38) */

#define JTP_GRAD_4PT 0        /* (8(f[+1]-f[-1]) - (f[+2]-f[-2])) / 12h */
#define JTP_GRAD_2PT 1        /* (f[+1]-f[-1]) / 2h: reference pm_grid.rs:329-348, nbody_gpu_twopass.rs:233-244 */

/* ---- parameters (YAML drop-in: simulation.box_size_mpc config.rs:23, pm_grid.n_cells :98,
 *      pm_grid.k_min :101, simulation.theta :36, simulation.softening_mpc :39) */
typedef struct jtp_params {
    double  box_size;    /* L                                                        */
    int32_t n_cells;     /* G, PM grid per axis; power of two in [16, 512]           */
    int32_t k_min;       /* zero modes with max|k_i| < k_min (0 = only DC); cufft_wrapper.cu:170-181 */
    double  r_s;         /* Gaussian split scale (length units)                      */
    double  r_cut;       /* short-range cutoff (length units), 0 < r_cut <= L/5      */
    double  theta;       /* opening angle, 0 <= theta <= 1 (0 = direct sum)          */
    double  softening;   /* Plummer epsilon > 0                                      */
    double  g_constant;  /* G                                                        */
    int32_t leaf_size;   /* particles per leaf bucket, 1..16                         */
    int32_t group_size;  /* targets per walk group, 1..64                            */
    int32_t precision;   /* JTP_FP32 / JTP_FP64                                      */
    int32_t split_mode;  /* JTP_SPLIT_*                                              */
    int32_t gradient;    /* JTP_GRAD_*                                               */
    int32_t deconvolve;  /* 1: divide the Green function by the CIC window W(k)^2 and match it to
                            the gradient stencil D: G *= (k . D) / max(|D|^2, |k|^2/4) (least
                            squares; the floor only acts near the Nyquist planes)          */
    int32_t pm_enabled;  /* 0: skip the long-range part                              */
    int32_t tree_enabled;/* 0: skip the short-range part                             */
    int32_t interlace;   /* 1: average the PM force of two CIC grids offset by half a cell
                            (cancels the leading aliasing error; one extra deposit/solve/interp) */
    double  theta_cut;   /* opening angle for nodes that straddle r_cut, >= 0 (0 = always open
                            them); the walk uses min(theta, theta_cut)                  */
    int32_t fft_tile;    /* log2 of the complex points one FFT workgroup holds in LDS: 11 (16 KB
                            FP32 / 32 KB FP64: 8 workgroups per CU), 12 (32 / 64 KB) or 13 (FP32
                            only, 64 KB: 128-byte rows in the y / z passes at G = 512); 0 = 11.
                            Every tile gives bit-identical results.                   */
} jtp_params;
/* Walk acceptance: a node's two monopoles are used when extent < theta * gap, gap being the
 * periodic group-box to node-box distance; if the node also reaches beyond r_cut from the
 * group box, it needs extent < theta_cut * gap. */

/* Per-phase timings of the last jtp_compute_forces (HIP events on the call's stream). */
typedef struct jtp_stats {
    double  ms_total;     /* force call, or the whole step (kicks, drifts) after jtp_step */
    double  ms_sort;      /* Morton keys + radix sort + particle gather        */
    double  ms_build;     /* Karras hierarchy + bottom-up monopoles + groups   */
    double  ms_deposit;   /* deterministic CIC deposit (FP32: tile lists, per-tile
                             integer sums, face merge; FP64: cell offsets, gather) */
    double  ms_fft;       /* R2C, Green, C2R                                   */
    double  ms_interp;    /* gradient + CIC interpolation                      */
    double  ms_walk;      /* group walk (one kernel launch)                    */
    uint64_t interactions;/* pair + monopole interactions evaluated by the walk */
    uint64_t interactions_mono; /* of which monopole interactions                */
    int64_t n_groups;
    int64_t n_particles;
} jtp_stats;

/* Stages readable with jtp_stage_copy after a jtp_compute_forces (parity tests). */
#define JTP_STAGE_KEYS      1 /* uint64[n] sorted Morton keys, low 16 bits cleared              */
#define JTP_STAGE_PERM      2 /* int32[n]  sorted slot -> original particle index               */
#define JTP_STAGE_NODES     3 /* int32[n-1][4] internal nodes: first, last, left, right
                                 (child >= 0: internal node; child < 0: leaf ~sorted_index)     */
#define JTP_STAGE_MONO      4 /* double[n-1][8]: q+, com+ xyz, q-, com- xyz as used by the walk
                                 (q = sigma G |m| count; com in fixed-point units)              */
#define JTP_STAGE_RHO       5 /* real[G^3]  deposited mass per cell sum_p q_p W, x fastest      */
#define JTP_STAGE_PHI       6 /* real[G^3]  long-range potential of rho+ - rho-                 */
#define JTP_STAGE_ACC_PM    7 /* real[n][3] long-range acceleration, ORIGINAL particle order    */
#define JTP_STAGE_SUMS      8 /* uint64[n-1][8]: n+, sum x+ y+ z+, n-, sum x- y- z- (exact, fixed units) */
#define JTP_STAGE_BOXES     9 /* uint32[n-1][6] node bounding boxes bmin xyz, bmax xyz (fixed)  */
#define JTP_STAGE_RHO2     10 /* real[G^3]  interlaced grid: mass deposited at positions + h/2    */
#define JTP_STAGE_PHI2     11 /* real[G^3]  potential of the interlaced grid                      */

/* ---- lifecycle (reference: CuFFTPoisson::new / drop, cufft_ffi.rs:85-118,339-353;
 *      TreePMForce::new, treepm_force.rs:35-51) */
typedef struct jtp_ctx jtp_ctx;
/* Defaults: r_s = 1.25 cells, r_cut = 4.5 r_s, theta = 0.5, softening = L/G/30,
 * G = 1, leaf 4, group 32, FP32, erfc+Gaussian split, 4-point gradient, deconvolution on,
 * interlacing on, theta_cut = 0.66 theta. */
void jtp_default_params(jtp_params* p, double box_size, int32_t n_cells);
/* ---- YAML drop-in: the reference's JanusConfig fields that set the force (config.rs:21-40,94-102)
 * plus two optional split fields the reference lacks.  jtp_params_from_config fills every jtp_params
 * field: the ones below from the config, the rest as jtp_default_params.  Without the split fields it
 * reproduces the reference's derivation: r_cut = 2 L / n_cells (janus_optim.rs:74) and
 * r_s = r_cut / 3 (nbody_gpu_twopass.rs:1694, treepm_force.rs:36), i.e. split_scale_cells = 2/3 and
 * r_cut_over_r_s = 3.  Returns JTP_ERR_ARG (and leaves *p unspecified) if the result is invalid. */
typedef struct jtp_config_fields {
    double  box_size_mpc;      /* simulation.box_size_mpc (config.rs:23)                            */
    int64_t n_particles;       /* simulation.n_particles (config.rs:25); only used for the default
                                  softening                                                         */
    double  theta;             /* simulation.theta (config.rs:36); < 0: the YAML default 0.7 (:42)   */
    double  softening_mpc;     /* simulation.softening_mpc (config.rs:39); <= 0 (absent):
                                  box_size_mpc / n_particles^(1/3) / 30 (config.rs:191-196)          */
    int32_t n_cells;           /* pm_grid.n_cells (config.rs:98)                                    */
    int32_t k_min;             /* pm_grid.k_min (config.rs:101)                                     */
    double  split_scale_cells; /* optional pm_grid.split_scale_cells: r_s in PM cells;
                                  <= 0 (absent): r_cut_cells / r_cut_over_r_s with r_cut_cells = 2   */
    double  r_cut_over_r_s;    /* optional pm_grid.r_cut_over_r_s; <= 0 (absent): 3                 */
} jtp_config_fields;
int  jtp_params_from_config(const jtp_config_fields* f, jtp_params* p);

int  jtp_create(const jtp_params* p, jtp_ctx** out);
int  jtp_destroy(jtp_ctx* ctx);
const char* jtp_last_error(const jtp_ctx* ctx);
const char* jtp_version(void);

/* Species masses: m_plus > 0 for species +1, m_minus_abs = |m-| > 0 for species -1.
 * Replaces the reference's per-particle mass (Particle.mass, nbody.rs:79-91) and
 * mass_factor (nbody_gpu_twopass.rs:2319-2322). */
int  jtp_set_species_mass(jtp_ctx* ctx, double m_plus, double m_minus_abs);

/* ---- the integrator's force call.
 * Replaces TreePMForce::update + compute_all_forces (treepm_force.rs:66-87,124-131) and the
 * force part of GpuNBodyTwoPass::step_treepm_gpu (nbody_gpu_twopass.rs:4529-4616, 3961-4138).
 * d_acc receives float[n*3] (JTP_FP32) or double[n*3] (JTP_FP64) in the caller's order:
 * a_i = sigma_i * sum_j sigma_j G |m_j| (x_j - x_i)/|.|^3, periodic, mean density removed. */
int  jtp_compute_forces(jtp_ctx* ctx, int64_t n, const uint32_t* d_pos_fixed,
                        const int8_t* d_species, void* d_acc, void* stream);

/* ---- one integrator step around the force.
 * Replaces GpuNBodyTwoPass::step_treepm_gpu (nbody_gpu_twopass.rs:4486-4649) with its kernels
 * drift_f32 / kick_f32 / add_pm_forces (:33-78).
 *   JTP_STEP_DKD: drift dt/2, force, kick dt, drift dt/2 (the reference's order; d_acc is output);
 *   JTP_STEP_KDK: kick dt/2 with d_acc = a(t) (input), drift dt, force -> d_acc, kick dt/2.
 * Kick, per species s: v += (a - H_s v dtau_dt) dt_kick (kick_f32's explicit Hubble friction;
 * the reference passes one H for both species, JanusExpansion::at_time h_plus).
 * Drift: u += round(v dt_drift 2^32 / L) in 64-bit integer arithmetic from an FP64 displacement,
 * so the periodic wrap is exact (drift_f32 wraps f32 positions once per axis).
 * d_vel and d_acc: float[n*3] (JTP_FP32) or double[n*3] (JTP_FP64), caller order.
 * JTP_STEP_GRAPH: the step's launches are captured once into a hipGraph and replayed; the
 * graph is re-captured when n, a pointer, the scheme, the stream, the masses or the buffer
 * capacity change (dt and H live in device memory and never force a re-capture).
 * Synchronises the stream before returning (status check), like jtp_compute_forces. */
#define JTP_STEP_DKD   0
#define JTP_STEP_KDK   1
#define JTP_STEP_GRAPH 0x100
int  jtp_step(jtp_ctx* ctx, int64_t n, uint32_t* d_pos_fixed, void* d_vel, const int8_t* d_species,
              void* d_acc, int32_t flags, double dt, double hubble_plus, double hubble_minus,
              double dtau_dt, void* stream);

/* ---- FP64 diagnostics on the device (reference: kinetic_energy and the grid potential-energy sum,
 * nbody_gpu_twopass.rs:5071-5077, 5293-5307, which copy to the host and sum there).
 * Deterministic reductions (fixed order) in FP64 of d_vel (float or double per the precision):
 * kinetic[s] = sum |m| v^2 / 2 and momentum[s] = sum |m| v per species (s = 0: +1, 1: -1;
 * total sum_s momentum[s] is conserved by the Janus pair law); pm_energy = 1/2 sum rho phi over the
 * mesh of the last force evaluation (mean of the two interlaced grids), the long-range potential
 * energy including its self-energy part. */
typedef struct jtp_diag {
    double kinetic[2];
    double momentum[2][3];
    double pm_energy;
} jtp_diag;
int  jtp_diagnostics(jtp_ctx* ctx, int64_t n, const void* d_vel, const int8_t* d_species, jtp_diag* out,
                     void* stream);

/* Total potential energy of the two-species system, FP64 (for energy-drift diagnostics; needs a
 * JTP_FP64 context with JTP_SPLIT_ERFC_GAUSS, tree and PM enabled).  Runs one force evaluation whose
 * walk also sums the short-range potential -sum_j q_j erfc(rho/2r_s)/rho exactly (every node opened),
 * interpolates the mesh potential at the particles and removes the mesh self-energy of each particle:
 * W = 1/2 sum_i |m_i| sigma_i (Phi_mesh(x_i) + Phi_short_i) - 1/2 sum_i m_i^2 S_i
 *     + 2 pi G r_s^2 Q^2 / V (background) + 1/2 G sum_i m_i^2 C (own periodic images),
 * Q = sum_i |m_i| sigma_i, C = 2.837297479 / L - 4 pi r_s^2 / V + sum_n erfc(|n| L / 2 r_s) / |n| L.
 * (The reference estimates PE from the grid only: potential_energy_binding_sampled,
 * nbody_gpu_twopass.rs:5280-5307.) */
int  jtp_potential_energy(jtp_ctx* ctx, int64_t n, const uint32_t* d_pos_fixed, const int8_t* d_species,
                          double* energy, void* stream);

/* Segregation of the two species on an n_cells^3 nearest-grid-point count grid (1 <= n_cells <= 1024;
 * cell of a fixed-point coordinate u: floor(u n_cells / 2^32)):
 *   out[0] = Var(n+ - n-) / (mean+ + mean-)          ("Diff/Pois")
 *   out[1] = corr(delta+, delta-), delta = n / mean - 1
 *   out[2] = (Var(n-) - mean-) / mean-               (excess variance of the m- counts)
 * Replaces compute_segregation_metrics, src/bin/janus_mu19_production.rs:219-273 (host loops over
 * f32 copies of the positions): counts with integer atomics, FP64 moments reduced in fixed order. */
int  jtp_segregation(jtp_ctx* ctx, int64_t n, const uint32_t* d_pos_fixed, const int8_t* d_species,
                     int32_t n_cells, double out[3], void* stream);

/* ---- ensembles of independent boxes (no counterpart in the reference, which runs one box per
 * process; the batch replaces a loop of step_treepm_gpu / TreePMForce calls over boxes).
 * Each box has its own parameters (L, G, r_s, ...) and species masses and its own buffers.
 * jtp_ensemble_compute_forces evaluates the force of every box: each box's preparation (keys,
 * sort, tree build, PM) is spread over n_streams internal streams that fork from and join back into
 * the caller's stream, so small boxes run concurrently, and then the short-range walk of all boxes
 * runs as one persistent launch pulling (box, particle-group) work items (one launch per set of
 * boxes sharing precision, split law and group size).  JTP_ENS_WALK_PER_BOX launches each box's walk
 * on its own stream after its preparation instead; JTP_ENS_SEQUENTIAL runs the boxes one after
 * another on the caller's stream; JTP_ENS_GRAPH captures the whole batch (all streams) into one
 * hipGraph and replays it (re-captured when a count, a pointer or the stream changes).  Results per
 * box are identical in every mode and to jtp_compute_forces on that box.  Arrays are indexed by box. */
typedef struct jtp_ensemble jtp_ensemble;
#define JTP_ENS_SEQUENTIAL   0x1
#define JTP_ENS_WALK_PER_BOX 0x2
#define JTP_ENS_GRAPH        0x100
int  jtp_ensemble_create(int32_t n_boxes, const jtp_params* params, const double* m_plus,
                         const double* m_minus_abs, int32_t n_streams, jtp_ensemble** out);
int  jtp_ensemble_destroy(jtp_ensemble* ens);
const char* jtp_ensemble_last_error(const jtp_ensemble* ens);
int  jtp_ensemble_compute_forces(jtp_ensemble* ens, const int64_t* n, const uint32_t* const* d_pos_fixed,
                                 const int8_t* const* d_species, void* const* d_acc, int32_t flags,
                                 void* stream);
/* box >= 0: that box's phases (-1 when graph-replayed); box = -1: ms_total of the whole batch,
 * summed interactions, groups and particles. */
int  jtp_ensemble_get_stats(jtp_ensemble* ens, int32_t box, jtp_stats* out);

/* Position conversion to fixed point on the device (x in any period; wrapped). */
int  jtp_pos_to_fixed_f64(jtp_ctx* ctx, int64_t n, const double* d_pos, uint32_t* d_fixed, void* stream);
int  jtp_pos_to_fixed_f32(jtp_ctx* ctx, int64_t n, const float* d_pos, uint32_t* d_fixed, void* stream);

/* ---- PM-only Poisson solve on caller grids (reference: cufft_solve_device[_filtered],
 * cufft_wrapper.cu:336-398, bound at cufft_ffi.rs:291-337).  d_rho: real[G^3] mass per cell
 * (x fastest); d_phi: real[G^3] potential of that mass with the context's Green function
 * (split r_s, k_min, deconvolution).  real = float or double per the context precision. */
int  jtp_pm_solve(jtp_ctx* ctx, const void* d_rho, void* d_phi, void* stream);

/* ---- allocation-free PM-only Poisson solve.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same result as jtp_pm_solve on a context created with *p (the Green function
 * from p's box_size, n_cells, r_s, g_constant, k_min, gradient, deconvolve; precision and fft_tile
 * as there; the other fields are ignored), without a context: the half spectrum and the tables live
 * in a static device array of 17 MB per GPU (one FP64 half spectrum of 128^3), so n_cells <= 128
 * (larger meshes return JTP_ERR_ARG).  Calls no allocator.  Asynchronous on the stream: launch
 * errors are returned, nothing is synchronised.  Calls that overlap on one GPU share the array and
 * race; calls on one stream are ordered. */
int  jtp_pm_solve_noalloc(const jtp_params* p, const void* d_rho, void* d_phi, void* stream);

/* ---- reference-compatible Morton keys of f32 positions (compute_morton_all,
 * nbody_gpu_twopass.rs:1447-1487): 21 bits/axis of (x + box_half) * inv_cell_size. */
int  jtp_morton_keys_f32(int64_t n, const float* d_pos, float box_half, float inv_cell_size,
                         uint64_t* d_keys, void* stream);

/* ---- reference-compatible 30-bit keys with the index tie-break (compute_morton_f32,
 * nbody_gpu_twopass.rs:302-326, the key of the reference's per-species tree build
 * build_single_sign_tree :2338-2545): per axis c = min(sat_u32((x + box_half) * (1024 / (2 box_half))), 1023)
 * in f32 arithmetic (sat_u32: truncation toward zero, negative and NaN -> 0), the 10-bit interleave
 * mc (x in bit 0), key[i] = mc << 32 | i and indices[i] = i.  inv_cell_size is ignored, as in the
 * reference kernel.  0 <= n <= 2^31 - 1 (the index is a 32-bit field). */
int  jtp_morton30_keys_f32(int64_t n, const float* d_pos, float box_half, float inv_cell_size,
                           uint64_t* d_keys, int32_t* d_indices, void* stream);

/* ---- diagnostics */
int  jtp_get_stats(jtp_ctx* ctx, jtp_stats* out);            /* synchronises the last call */
int  jtp_stage_size(jtp_ctx* ctx, int32_t stage, size_t* bytes);
int  jtp_stage_copy(jtp_ctx* ctx, int32_t stage, void* host_dst, size_t bytes);

/* ---- device memory helpers (reference: cufft_alloc/free/copy_*, cufft_wrapper.cu:251-288) */
void* jtp_alloc(size_t bytes);
void  jtp_free(void* p);
int   jtp_copy_h2d(void* d_dst, const void* h_src, size_t bytes);
int   jtp_copy_d2h(void* h_dst, const void* d_src, size_t bytes);
int   jtp_copy_d2d(void* d_dst, const void* d_src, size_t bytes);
int   jtp_device_synchronize(void);
int   jtp_device_count(int* out);
int   jtp_device_memory(size_t* free_bytes, size_t* total_bytes);  /* hipMemGetInfo of the current device */

/* ---- initial conditions
 * Zel'dovich lattice ICs on the GPU: n_grid^3 particles, one per cell centre, displaced by
 * psi = Re IFFT[-i k delta(k) / k^2] / n^3 with P(k) = k^(n_s - 4) delta_rms^2 and D = 1/(1 + z_init),
 * rescaled so that max |psi| = 0.3 lattice spacing; velocity = psi x H0 sqrt(1 + z_init) in Mpc/Gyr
 * (H0 = 69.9 km/s/Mpc); round(n^3 / (1 + mu)) randomly chosen cells hold m+, the rest m-.
 * Replaces generate_zeldovich_ics (src/ic_gen.rs:51-169, 179-309).  The context must have
 * n_cells == n_grid and box_size == p->box_size; its FFT precision is the computation's.  Random
 * numbers come from SplitMix64 streams keyed by the seed (not the reference's StdRng streams: the
 * same seed gives a different realisation with the same statistics).  Particle i is cell i,
 * i = (iz n + iy) n + ix as in the reference.  *n_plus receives the m+ count. */
typedef struct jtp_ic_params {
    double box_size;   /* Mpc */
    int32_t n_grid;    /* particles per dimension (a power of two in [16, 512]) */
    double z_init;
    uint64_t seed;
    double mu;         /* |m-| / m+: sets the m+ fraction 1 / (1 + mu) */
    double n_s;        /* spectral index */
    double delta_rms;  /* overdensity amplitude */
} jtp_ic_params;
/* IcParams::default (ic_gen.rs:24-36): L = 50, n = 128, z = 10, seed 42, mu 19, n_s 0.965, 0.10 */
void  jtp_ic_default_params(jtp_ic_params* p);
int   jtp_generate_ics(jtp_ctx* ctx, const jtp_ic_params* p, uint32_t* d_pos_fixed, void* d_vel, int32_t vel_f64,
                       int8_t* d_species, int64_t* n_plus, void* stream);

/* ---- snapshots
 * The reference's two on-disk formats (little-endian):
 *   JTP_SNAP_PRODUCTION  save_snapshot, src/bin/janus_mu19_production.rs:199-216: u32 N, f32 L, u32 step,
 *                        f32 z, then per particle 3 f32 position, 3 f32 velocity, i8 sign (25 bytes, packed).
 *   JTP_SNAP_V3          src/snapshot_v3.rs:14-16,108-157,313-335: a 408-byte header (magic "JANUSV3\n",
 *                        version 3, header_size, n_total, a, t_gyr, l_box, h0, mu, omega_b, base masses and
 *                        softenings, n_split_max, seed_ic, z_init, n_stars, z_start_run, sfr, rho_max,
 *                        256-byte run label), then per particle 3 f32 position, 3 f32 velocity, f32 mass,
 *                        f32 epsilon, u8 sign (1 = m+, 255 = m-), u8 split_level, u8 is_star, u8 flags (36 B).
 * Positions on disk are f32 in [-L/2, L/2) (this library's box convention), converted from / to the
 * fixed-point device arrays on the GPU; velocities are stored as f32 of (device value x velocity_scale)
 * and read back divided by it (device f32 or f64, vel_f64 = 0 / 1).  Records are packed and unpacked
 * on the GPU in chunks of 2^20 and the file moves through pinned host buffers.  Writing v3 fills
 * mass and epsilon per species from the metadata and split_level = is_star = flags = 0 (as the
 * reference's writers do); reading v3 loads position, velocity and sign only (one mass per species
 * drives the force path).  A v3 header_size above 408 is skipped over (snapshot_v3.rs:425-431).
 * Errors of these calls (no context) are reported by jtp_io_last_error(). */
#define JTP_SNAP_PRODUCTION 1
#define JTP_SNAP_V3         3
typedef struct jtp_snap_meta {
    int32_t format;            /* JTP_SNAP_* */
    int64_t n;                 /* particle count (written: the arrays' length; read: the file's) */
    double box_size;           /* L (v3: l_box) */
    uint32_t step;             /* production format */
    double z;                  /* production: redshift; v3 read: 1/a - 1 */
    double velocity_scale;     /* file velocity = device velocity x velocity_scale (default 1) */
    double a, t_gyr, h0, mu, omega_b;                                      /* v3 header */
    double m_part_plus_base, m_part_minus_base, eps_plus_base, eps_minus_base;
    uint32_t n_split_max, seed_ic;
    double z_init;
    uint64_t n_stars;
    double z_start_run, sfr, rho_max;
    char run_label[256];
    float mass_plus, mass_minus, eps_plus, eps_minus;                      /* v3 per-particle fields */
} jtp_snap_meta;

/* Defaults of SnapshotHeaderV3::default (snapshot_v3.rs:48-76), velocity_scale 1, the given format and L. */
void  jtp_snap_meta_default(jtp_snap_meta* meta, int32_t format, double box_size);
/* Writes meta->n particles (meta->format) to path.  Replaces save_snapshot / write_snapshot_v3. */
int   jtp_snapshot_write(const char* path, const jtp_snap_meta* meta, const uint32_t* d_pos_fixed,
                         const void* d_vel, int32_t vel_f64, const int8_t* d_species, void* stream);
/* Header only (read_header_only, snapshot_v3.rs:442-463; magic and version checked). */
int   jtp_snapshot_read_meta(const char* path, int32_t format, jtp_snap_meta* meta);
/* Loads n particles (n must equal the file's count) into device arrays.  Replaces read_snapshot_v3. */
int   jtp_snapshot_read(const char* path, int32_t format, int64_t n, uint32_t* d_pos_fixed, void* d_vel,
                        int32_t vel_f64, int8_t* d_species, double velocity_scale, void* stream);
const char* jtp_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* JANUS_TREEPM_H */
