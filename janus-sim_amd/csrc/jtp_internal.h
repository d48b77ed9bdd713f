// Internal declarations shared by the HIP translation units of libjanus_treepm.
// Device helpers, launch wrappers and the context structure.  Nothing here is
// part of the C ABI (include/janus_treepm.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <type_traits>

#include "../../include/janus_treepm.h"

namespace jtp {

constexpr int kWave = 64;            // CDNA wavefront width
constexpr int kBlock = 256;          // threads per workgroup for streaming kernels
constexpr int kSortIPT = 24;         // keys per thread per radix-sort tile (6144-key tiles: 2.35 ms for
                                      // the 2.2e7-key sort against 2.87 / 2.50 / 2.73 ms with 2048 / 4096 / 8192)
constexpr int kSortTile = kBlock * kSortIPT;
constexpr int kGroupLevelBits = 4;   // walk groups never straddle a level-4 Morton cell (L/16)
constexpr int kPolyDeg = 10;         // FP32 split-function polynomial degree
constexpr int kDenseCell = 32;       // PM cells with more particles are pre-reduced by wavefronts,
constexpr int kDenseChunk = 1024;    // one per chunk of this many particles
constexpr int kInterpChunk = 4096;   // FP32 interpolation and deposit: particles per workgroup of a heavy tile,
constexpr int kLightMax = 1024;      // tiles with at most this many are light (four particles per thread)
// int2 entries of Ctx::interp_list for G^3 cells and up to n particles: the light-tile list, then
// sum over tiles of ceil(np / kInterpChunk) <= n / kInterpChunk + n / (kLightMax + 1) + 1 chunks
inline int64_t interp_list_size(int64_t G3, int64_t n) {
    return (G3 / 512 + 1) / 2 + n / kInterpChunk + n / (kLightMax + 1) + 1;
}
constexpr int kMaxRounds = 128;      // size of the bottom-up control block (frontier count ... root flag)
constexpr int kXcds = 8;             // MI355X accelerator complex dies; workgroup i runs on XCD i % 8
// device counter block (int32): offsets
constexpr int kCtrNGroups = 0, kCtrErr = 1, kCtrDense = 3, kCtrInterp = 4;  // kCtrInterp: 3 words
constexpr int kCtrRounds = 8;                       // [0] global frontier size, [kMaxRounds+1] root flag
constexpr int kCtrRoot = kCtrRounds + kMaxRounds + 1;
constexpr int kCtrXcd = kCtrRounds + kMaxRounds + 4;  // kXcds walk-queue heads
constexpr int kCtrTotal = kCtrXcd + kXcds + 4;

// ---------------------------------------------------------------- small types
template <typename R> struct Cplx { R x, y; };

// Morton interleave of a 21-bit integer into every third bit (x in bit 0).
__host__ __device__ inline uint64_t spread3_21(uint64_t v) {
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8))  & 0x100f00f00f00f00fULL;
    v = (v | (v << 4))  & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2))  & 0x1249249249249249ULL;
    return v;
}
__host__ __device__ inline uint32_t compact3_21(uint64_t v) {
    v &= 0x1249249249249249ULL;
    v = (v ^ (v >> 2))  & 0x10c30c30c30c30c3ULL;
    v = (v ^ (v >> 4))  & 0x100f00f00f00f00fULL;
    v = (v ^ (v >> 8))  & 0x1f0000ff0000ffULL;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffULL;
    v = (v ^ (v >> 32)) & 0x1fffffULL;
    return (uint32_t)v;
}
__host__ __device__ inline uint64_t morton_fixed(uint32_t ux, uint32_t uy, uint32_t uz) {
    return spread3_21(ux >> 11) | (spread3_21(uy >> 11) << 1) | (spread3_21(uz >> 11) << 2);
}
// The particle sort orders by the Morton key with its low kKeyLowBits bits cleared (15-16 bits per
// axis, L / 2^16 and finer left in input order): every user of the order reads key prefixes (PM
// cells and tiles, walk seed cells, the radix tree, whose boxes are exact whatever the order inside a
// key), so the two radix passes over those bits are skipped.
constexpr int kKeyLowBits = 16;
constexpr uint64_t kKeySortMask = ~((1ull << kKeyLowBits) - 1);
static_assert(kKeyLowBits % 8 == 0 && (64 - kKeyLowBits) / 8 % 2 == 0,
              "whole 8-bit radix passes, an even number of them (the result lands in the first buffer)");

// Split-function coefficients for the FP32 walk, rho^2 = r^2 + eps^2, x = rho / (2 r_s):
//  * particle pairs under erfc + Gaussian (always cut at r_cut): S(x) - 1 = x^3 h(x^2) with h entire,
//    so S / rho^3 = 1 / rho^3 + H(rho^2), H(v) = (S(x) - 1) / rho^3 interpolated by the degree
//    kPolyDegPair Chebyshev interpolant Pp(t), t = 2 rho^2 / Vp - 1 on [0, Vp = rho_max^2]:
//    no rho, no separate S, and a faster-converging fit than S(x) itself;
//  * particle pairs under erfc only: S(x) ~= Pp(t), t = 2x/Xp - 1, degree kPolyDeg;
//  * monopoles under erfc + Gaussian (the node straddles r_cut; no cut at r_cut): the same H form
//    fitted on [0, Vm = (kTailExt r_cut)^2 + eps^2] with degree kPolyDegTail, the force dropped beyond
//    Vm where S(kTailExt r_cut) < 2e-4;
//  * monopoles under erfc only (x up to Xm ~ r_cut (1 + sqrt(3) theta) / 2 r_s):
//    S(x) ~= exp(-x^2) Pm(t), t = min(2x/Xm - 1, 1), Pm interpolating erfcx(x) (degree kPolyDeg),
//    which stays bounded past the fitted range.
constexpr int kPolyDegPair = 8;
constexpr int kPolyDegTail = 10;
constexpr double kTailExt = 1.4;
struct SplitPoly {
    float cp[kPolyDeg + 1];
    float cm[kPolyDeg + 1];
};

// Lanes of a wavefront that share one walk target: groups of <= 16 targets use 4 lanes per
// target, <= 32 use 2, larger groups 1.
inline int lanes_per_target(int group_size) { return group_size <= 16 ? 4 : (group_size <= 32 ? 2 : 1); }

// Scalars of one integrator step, read by the kick / drift kernels from device memory.
struct StepParams {
    double kick_dt;      // time step of each kick
    double hubble_plus;  // H for species +1
    double hubble_minus; // H for species -1
    double dtau_dt;
    double drift_units;  // drift time step * 2^32 / L (velocity -> fixed-point displacement)
    double pad[3];
};

// Captured step graph and what it was captured for.
struct StepGraph {
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
    int64_t n = -1;
    const void *pos = nullptr, *vel = nullptr, *species = nullptr, *acc = nullptr;
    int scheme = -1;
    hipStream_t stream = nullptr;
};

// ---------------------------------------------------------------- context
struct Ctx {
    jtp_params p{};
    double m_plus = 1.0, m_minus = 1.0;   // |m| per species
    int logG = 0;
    std::string err;
    int device = 0;

    // capacities (elements)
    int64_t cap_n = 0;
    // particle-side buffers
    uint64_t *keys0 = nullptr, *keys1 = nullptr;  // keys1: the sort's second buffer, then the build's int2 leaf ranges
    int32_t *vals0 = nullptr, *vals1 = nullptr;
    uint32_t *hist = nullptr;        // 256 * nblocks
    uint32_t *hist_blk = nullptr;    // scan partials
    uint4 *parts = nullptr;          // sorted records: fixed xyz + species
    // tree
    uint4 *node_geo = nullptr;       // 2 per internal node: (centre xyz, first leaf), (extent xyz, last leaf)
    int2 *node_kids = nullptr;
    int2 *up = nullptr;              // internal node: (parent, sibling); the root's parent is -1
    int32_t *flags = nullptr;        // bottom-up arrival counters
    int32_t *frontier0 = nullptr;    // bottom-up: nodes whose range crosses a chunk edge, ready to climb
    int32_t *frontier1 = nullptr;    // bottom-up: nodes whose range crosses a superchunk edge, ready to climb
    int32_t *dense_list = nullptr;   // deposit: the PM cells holding more than kDenseCell particles
    int2 *interp_list = nullptr;     // FP32 interpolation: the light tiles (G^3 / 512 int32), then the
                                     // (tile, first particle) of the heavy tiles' chunks; interp_list_size
    int2 *dep_chunks = nullptr;      // FP32 deposit: (tile, first particle) of the chunks of tiles holding
                                     // more than kInterpChunk particles; dep_chunk_cap entries
    long long *dep_cbuf = nullptr;   // FP32 deposit: int64 LDS grid of each such chunk (1729 per entry)
    void *node_sum = nullptr;        // NodeSum per internal node
    void *node_mono = nullptr;       // 2 x (uint4 | double4) per internal node
    int2 *groups = nullptr;
    int32_t *cell4 = nullptr;        // level-4 cell offsets (4097)
    int32_t *cellnode = nullptr;     // node holding each Morton cell of level start_level (<= 4096)
    int start_level = 4;             // finest level whose cell edge L / 2^level is >= r_cut (2..4)
    int32_t *dev_counters = nullptr; // kCtr* layout
    unsigned long long *dev_inter = nullptr;  // interaction counter
    // PM
    int32_t *cell_start = nullptr;   // FP64: G^3 + 1 first sorted particle of each cell
    int32_t *tile_start = nullptr;   // FP32: G^3 / 512 + 1 first sorted particle of each 8^3-cell tile
    float *dep_tbuf = nullptr;       // FP32 deposit: each tile's record, its sums as floats (1729 per tile)
    int32_t *dense_slot = nullptr;   // G^3: slot of a dense cell's corner sums (written for dense cells only)
    void *dense_corner = nullptr;    // real [n / kDenseCell + 1][8 + 27] corner sums of dense-cell chunks
    void *rho = nullptr;             // real G^3 deposited mass per cell
    void *phi = nullptr;             // real G^3 potential
    void *rho2 = nullptr, *phi2 = nullptr;  // interlaced grid (nodes at the cell centres of the first)
    void *grad = nullptr, *grad2 = nullptr;  // FP64: real4 [G^3] stencil gradient of phi / phi2 (w unused); FP32: grad
                                             // is allocated only by jtp_generate_ics (one real grid)
    void *spec = nullptr;            // complex G^2 (G/2+1)
    void *twiddle = nullptr;         // complex G  e^{-2 pi i m/G}
    void *gtab = nullptr;            // real 2G: per-axis window/Gaussian factor, k_i^2
    void *pm_acc = nullptr;          // real 3n (sorted order)
    SplitPoly poly{};
    float poly_tscale = 0.0f;    // particles: t = rho^2 * poly_tscale - 1 (erfc + Gaussian), rho * poly_tscale - 1
    float poly_tscale_m = 0.0f;  // monopoles
    float poly_rho2_tail = 0.0f; // erfc + Gaussian monopoles: rho^2 beyond which the force is dropped
    double green_norm = 0.0;
    unsigned long long* host_pinned = nullptr;  // [0] interactions, [1] (n_groups | err << 32)

    double* diag_scratch = nullptr;   // FP64 reduction partials (kDiagBlocks x 8) + 10 results
    double* pot_short = nullptr;      // energy diagnostic: short-range potential per particle (cap_n)
    double* acc_scratch = nullptr;    // energy diagnostic: forces of that evaluation (3 cap_n)
    int64_t pot_cap = 0;
    double* kself = nullptr;          // 27 values of the mesh potential of a unit mass, offsets {-1,0,1}^3
    int32_t* seg = nullptr;           // segregation counts: int32 [cells][2] + 2 totals
    int64_t seg_cap = 0;              // cells the buffer holds
    bool kself_ready = false;
    StepParams* dev_step = nullptr;   // device copy of the step scalars
    StepParams* host_step = nullptr;  // pinned staging of the step scalars
    StepGraph graph;
    bool graph_dirty = true;          // masses or buffers changed since the capture

    int64_t last_n = -1;
    bool have_result = false;
    int last_kind = 0;                // 0 force call, 1 eager step, 2 graph step, 3 ensemble graph
    // ev[0..6]: force phase boundaries, ev[7]: walk start; ev[8], ev[9]: step begin / end;
    // ev_build[2]: the tree build on the side stream
    hipEvent_t ev[10]{};
    hipEvent_t ev_build[2]{};
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    hipStream_t side = nullptr;       // the tree build runs here, concurrently with the PM phases
    bool events_ok = false;
    hipStream_t last_stream = nullptr;
    int64_t n_groups_max = 0;

    size_t real_size() const { return p.precision == JTP_FP64 ? 8 : 4; }
    int64_t G() const { return p.n_cells; }
    int64_t G3() const { return (int64_t)p.n_cells * p.n_cells * p.n_cells; }
    int64_t Hc() const { return p.n_cells / 2 + 1; }
};

// ---------------------------------------------------------------- launchers
// All return hipError_t from hipGetLastError after the launch(es).
hipError_t launch_fixed_from_f64(const double* pos, uint32_t* out, int64_t n3, double L, hipStream_t s);
hipError_t launch_fixed_from_f32(const float* pos, uint32_t* out, int64_t n3, double L, hipStream_t s);
// keys, identity values and the packed (x, y, z, species) records in input order
hipError_t launch_morton_keys(const uint32_t* pos, const int8_t* species, uint64_t* keys, int32_t* vals,
                              uint4* rec, int64_t n, hipStream_t s);
hipError_t launch_morton_keys_f32(const float* pos, uint64_t* keys, int64_t n, float box_half,
                                  float inv_cell, hipStream_t s);
hipError_t launch_morton30_keys_f32(const float* pos, uint64_t* keys, int32_t* idx, int64_t n, float box_half,
                                    hipStream_t s);
// keys0/vals0 -> keys0/vals0 sorted by the key bits from first_bit (a multiple of 16) up, stable
hipError_t radix_sort_pairs(Ctx& c, int64_t n, hipStream_t s, int first_bit = 0);
hipError_t device_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
// parts[i] = rec[perm[i]]; with tile_start (FP32 PM), also the first sorted particle of every
// Morton-aligned 8^3-cell tile of the 2^logG mesh, tile_start[ntile] = n
hipError_t launch_gather_parts(const uint4* rec, const int32_t* perm, uint4* parts, int64_t n, int32_t* tile_start,
                               int logG, hipStream_t s);

hipError_t build_tree(Ctx& c, int64_t n, hipStream_t s);            // Karras + bottom-up + mono + groups
hipError_t launch_fill_i32(int32_t* out, int64_t m, int32_t v, hipStream_t s);
hipError_t launch_cell_start(const uint64_t* keys, int64_t n, int shift, int64_t ncells, int32_t* out,
                             hipStream_t s);

hipError_t pm_deposit(Ctx& c, int64_t n, hipStream_t s);           // grid <- sum q W
int64_t dep_chunk_cap(int64_t n);  // FP32 deposit: chunk entries for up to n particles
int64_t dep_chunk_nodes();         // FP32 deposit: int64 node sums per chunk entry
int64_t dep_tile_record();         // FP32 deposit: floats per tile record (own blocks, beyond nodes)
// phi <- Green * rho; matched = 1 multiplies by the gradient-stencil match (the force's Green
// function when deconvolving; -1 = as the context's force), 0 leaves the potential unmatched
hipError_t pm_fft_solve(Ctx& c, const void* rho, void* phi, hipStream_t s, int matched = -1);
// out <- unnormalised inverse real 3-D FFT of the Hermitian half spectrum in c.spec (overwritten)
hipError_t pm_inverse_fft(Ctx& c, void* out, hipStream_t s);
// Zel'dovich lattice initial conditions of G^3 particles (G = the context's mesh)
hipError_t generate_ics(Ctx& c, const jtp_ic_params& p, uint32_t* pos, void* vel, int vel_f64, int8_t* species,
                        int64_t n_plus, hipStream_t s);
hipError_t pm_interp(Ctx& c, int64_t n, hipStream_t s);            // pm_acc <- -sigma grad phi
hipError_t pm_tables(Ctx& c);                                      // twiddles + Green tables (sync)

// pot != nullptr (FP64 contexts, erfc + Gaussian law): also writes the short-range potential
hipError_t walk(Ctx& c, int64_t n, void* acc_out, hipStream_t s, double* pot = nullptr);
// One walk launch over several boxes: each box's arguments (walk_args_bytes() each, filled on the
// host by walk_fill_args after its build is enqueued) in device memory; boxes with equal
// walk_batch_key share a launch.  d_ctrl holds nbox + 1 + kXcds int32 (group offsets, queue heads).
size_t walk_args_bytes();
void walk_fill_args(Ctx& c, int64_t n, void* acc_out, void* dst);
int walk_batch_key(const Ctx& c);
hipError_t walk_boxes(const void* d_args, int32_t nbox, int32_t* d_ctrl, int64_t items_max, int key, hipStream_t s);
hipError_t launch_kick(Ctx& c, void* vel, const void* acc, const int8_t* species, int64_t n, hipStream_t s);
hipError_t launch_drift(Ctx& c, uint32_t* pos, const void* vel, int64_t n, hipStream_t s);
hipError_t launch_diagnostics(Ctx& c, int64_t n, const void* vel, const int8_t* species, double* scratch,
                              double* out, hipStream_t s);
constexpr int kDiagScratch = 1024 * 8 + 16;  // doubles
hipError_t launch_segregation(int64_t n, const uint32_t* pos, const int8_t* species, int32_t n_cells,
                              int32_t* counts, double* scratch, double* out, hipStream_t s);
// sums out[0..8] of launch_segregation over `cells` cells -> the three metrics of jtp_segregation
void segregation_metrics(const double v[9], int64_t cells, double out[3]);
hipError_t launch_energy_terms(Ctx& c, int64_t n, const double* kself, const double* pot_short, double* scratch,
                               double* out, hipStream_t s);
hipError_t write_pm_only(Ctx& c, int64_t n, void* acc_out, hipStream_t s);

// host-side: FP32 split polynomial for the context's split mode and x range
void make_split_poly(Ctx& c);

}  // namespace jtp
