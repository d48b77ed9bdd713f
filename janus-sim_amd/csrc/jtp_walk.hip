// Short-range group walk.
//
// Reference behaviour replaced: forces_treepm_short_range (src/nbody_gpu_twopass.rs:1670-1811),
// a per-warp shared-stack walk run once per species tree, and the CPU recursion
// TreePMTree::acc_recursive + pairwise_acc_with_split (src/treepm/tree_short.rs:172-247).
//
// One wavefront owns one group of <= 64 Morton-consecutive targets.  The
// traversal is lane-parallel: each step pops up to 64 nodes from a per-wave LDS
// stack, every lane classifies one node against the group's bounding box
// (prune beyond r_cut / accept the two monopoles / take the bucket's particles /
// open), and a wavefront prefix sum gives every source its slot in an LDS
// interaction list (bucket particles are then listed one lane per particle)
// while a ballot pushes the opened nodes' children.  When the list is long
// enough every lane evaluates its target against all entries (broadcast LDS
// reads).  Sources are stored relative to the group centre, so the pair
// displacement is a plain subtraction; periodic minimum image comes from
// wrapping 32-bit fixed-point differences when the entry is built.  The groups
// of one box (k_walk) or of several boxes (k_walk_boxes) are pulled from queues
// by persistent workgroups.
//
// Pair law (per entry j with signed charge q_j = sigma_j G |m_j|):
//   a_i += sigma_i q_j S(.) (x_j - x_i) / (r^2 + eps^2)^{3/2},  zero unless r^2 < r_cut^2
// Same species attract, opposite species repel (tree_short.rs:238, nbody_gpu_twopass.rs:1711).
#include "jtp_internal.h"
#include "jtp_device.h"

namespace jtp {

// LDS per wave (FP32): 448 x 16 B list + 576 x 4 B stack = 9.4 KB, so four 4-wave workgroups fit
// a CU (150 of its 160 KB, 16 waves); a flush every 224 entries.  Measured at N = 2.2e7: a 256 /
// 128 list (six workgroups per CU) and a 512 / 256 list (three) are both ~10 % slower; 384 / 192
// is 3 % slower.  Pops shrink when the stack nears kStackCap, so a smaller stack costs
// parallelism in deep traversals, never correctness.
constexpr int kStackCap = 448;          // logical stack bound per wave
constexpr int kStackPhys = kStackCap + 128;
constexpr int kListCap = 448;
constexpr int kFlushAt = 224;
constexpr int kPad = 4;  // list slots kept free for the lane-sharing padding
constexpr int32_t kNoNodeW = (int32_t)0x80000000;  // empty start cell

struct WalkArgs {
    const uint4* parts;
    const uint4* geo;
    const int2* kids;
    const void* mono;
    const int2* groups;
    int32_t* counters;  // kCtr* layout: group count, error flag, per-XCD queue heads
    unsigned long long* inter;
    const void* pm_acc;  // sorted order, nullptr when the PM part is disabled
    double* pot;         // FP64 energy diagnostic: short-range potential per particle (caller order)
    const int32_t* perm;
    void* out;
    int64_t n;
    int32_t root;
    int32_t leaf;
    const int32_t* cellnode;  // node of each Morton cell of level start_level (nullptr: start at the root)
    int32_t start_level;
    double theta_cut2;  // theta_cut^2 for nodes whose box reaches beyond r_cut
    int64_t max_steps;
    double scale;   // length per fixed-point unit, L / 2^32
    double eps2;
    double rcut2;
    double theta2;
    double inv2rs;  // 1 / (2 r_s)
    double qp, qm;  // signed particle charges sigma G |m| for species +1 / -1
    SplitPoly poly;
    float tscale;    // particles: t = rho^2 * tscale - 1 (erfc + Gaussian) or rho * tscale - 1
    float tscale_m;  // monopoles: t = rho^2 * tscale_m - 1 (erfc + Gaussian) or rho * tscale_m - 1
    float rho2_tail; // erfc + Gaussian monopoles: rho^2 beyond which the force is dropped
    float escale;    // exp(-x^2) = exp2(r2s * escale)
    float theta2_f, theta_cut2_f;
};

__device__ inline void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

template <typename R> struct EntryT;
template <> struct EntryT<float> { using T = float4; };
template <> struct EntryT<double> { using T = double4; };

// Pair evaluation: returns the force factor f such that a += f * d for the displacement d.
// CUT applies the hard cutoff r < r_cut (particle entries, and monopoles under the x^4 law whose
// weight turns negative beyond r_cut); monopoles under the erfc laws keep the smooth tail of S.
// rcut2s = r_cut^2 + eps^2: the cut tested on rho^2 = r^2 + eps^2.
template <typename R, int MODE, bool CUT>
struct Pair;

template <int DEG>
__device__ inline float horner(const float* c, float t) {
    float P = c[DEG];
#pragma unroll
    for (int d = DEG - 1; d >= 0; --d) P = __builtin_fmaf(P, t, c[d]);
    return P;
}

template <int MODE, bool CUT>
struct Pair<float, MODE, CUT> {
    __device__ static inline float f(float q, float dx, float dy, float dz, const WalkArgs& a, float eps2,
                                     float rcut2, float rcut2s) {
        if (MODE == JTP_SPLIT_POLY4) {
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float inv = __builtin_amdgcn_rsqf(r2 + eps2);
            const float u = r2 * (1.0f / rcut2);
            const float fv = q * (1.0f - u * u) * (inv * inv * inv);
            return (!CUT || r2 < rcut2) ? fv : 0.0f;
        }
        const float r2s = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2s);
        const float inv3 = inv * inv * inv;
        float fv;
        if (CUT && MODE == JTP_SPLIT_ERFC_GAUSS) {
            // particle pair: S / rho^3 = 1 / rho^3 + H(rho^2)
            const float t = __builtin_fmaf(r2s, a.tscale, -1.0f);
            fv = q * (inv3 + horner<kPolyDegPair>(a.poly.cp, t));
        } else if (MODE == JTP_SPLIT_ERFC_GAUSS) {
            // monopole: the same form with H fitted out to kTailExt r_cut, where S is below 2e-4 and
            // the monopole's force is dropped
            const float t = __builtin_fmaf(r2s, a.tscale_m, -1.0f);
            fv = q * (inv3 + horner<kPolyDegTail>(a.poly.cm, t));
            return r2s < a.rho2_tail ? fv : 0.0f;
        } else if (CUT) {
            // particle pair, erfc law: direct interpolant of S(x)
            const float t = r2s * inv * a.tscale - 1.0f;
            fv = q * horner<kPolyDeg>(a.poly.cp, t) * inv3;
        } else {
            // monopole, erfc law: exp(-x^2) times the interpolant of the scaled factor (bounded
            // beyond its range)
            const float t = fminf(r2s * inv * a.tscale_m - 1.0f, 1.0f);
            fv = q * (__builtin_amdgcn_exp2f(r2s * a.escale) * horner<kPolyDeg>(a.poly.cm, t)) * inv3;
        }
        return (!CUT || r2s < rcut2s) ? fv : 0.0f;
    }
};

template <int MODE, bool CUT>
struct Pair<double, MODE, CUT> {
    __device__ static inline double f(double q, double dx, double dy, double dz, const WalkArgs& a, double eps2,
                                      double rcut2, double) {
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r2s = r2 + eps2;
        const double inv = rsqrt(r2s);
        const double inv3 = inv * inv * inv;
        double S;
        if (MODE == JTP_SPLIT_POLY4) {
            const double u = r2 / rcut2;
            S = 1.0 - u * u;
        } else {
            const double x = r2s * inv * a.inv2rs;
            S = erfc(x);
            if (MODE == JTP_SPLIT_ERFC_GAUSS) S += 1.1283791670955126 * x * exp(-x * x);
        }
        const double fv = q * S * inv3;
        return (!CUT || r2 < rcut2) ? fv : 0.0;
    }
};

// Short-range potential of a pair under the erfc + Gaussian law (FP64 energy diagnostic):
// q erfc(rho / 2 r_s) / rho, whose gradient is the pair force; zero for the target itself (r = 0)
// and beyond the cut where the force is cut.
template <bool CUT>
__device__ inline double pair_potential(double q, double dx, double dy, double dz, const WalkArgs& a, double eps2,
                                        double rcut2) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r2s = r2 + eps2;
    const double inv = rsqrt(r2s);
    const double x = r2s * inv * a.inv2rs;
    return (r2 > 0.0 && (!CUT || r2 < rcut2)) ? q * erfc(x) * inv : 0.0;
}

template <typename R>
__device__ inline R fixed_rel(uint32_t u, uint32_t c, R scale) {
    return (R)(int32_t)(u - c) * scale;
}

// Particle entry (relative to the group centre gc, minimum image from the 32-bit wrap).
template <typename R, typename E>
__device__ inline E part_entry(const uint4& pr, const uint32_t gc[3], R scale, const R qpart[2]) {
    E e;
    e.x = fixed_rel<R>(pr.x, gc[0], scale);
    e.y = fixed_rel<R>(pr.y, gc[1], scale);
    e.z = fixed_rel<R>(pr.z, gc[2], scale);
    e.w = qpart[(int32_t)pr.w > 0 ? 1 : 0];
    return e;
}

// Monopole entry (relative to the group centre gc) from its record.
__device__ inline float4 mono_entry(const uint4& m, const uint32_t gc[3], float scale) {
    return make_float4(fixed_rel<float>(m.x, gc[0], scale), fixed_rel<float>(m.y, gc[1], scale),
                       fixed_rel<float>(m.z, gc[2], scale), __uint_as_float(m.w));
}
__device__ inline double wrap_fixed(double d) {
    // d in fixed units, bring into [-2^31, 2^31)
    return d - 4294967296.0 * rint(d * (1.0 / 4294967296.0));
}
__device__ inline double4 mono_entry(const double4& m, const uint32_t gc[3], double scale) {
    return make_double4(wrap_fixed(m.x - (double)gc[0]) * scale, wrap_fixed(m.y - (double)gc[1]) * scale,
                        wrap_fixed(m.z - (double)gc[2]) * scale, m.w);
}

// LPT lanes share one target: a group holds up to 64 / LPT targets and each target's
// interaction list is split over its LPT lanes (entries j = sub, sub + LPT, ...).
template <typename R, int LPT>
struct WalkLds {
    using E = typename EntryT<R>::T;
    E list[kListCap];
    int32_t stack[kStackPhys];
    uint32_t owner[64];  // bucket-particle owner lanes, one round at a time
};

// One group g of one box: traversal, list evaluation and the output write.  Adds the group's
// interaction counts to my_inter / my_mono (lane 0's copies are the wavefront's).
template <typename R, int MODE, int LPT, bool POT>
__device__ __forceinline__ void walk_group(const WalkArgs& a, int g, WalkLds<R, LPT>& lds, unsigned long long& my_inter,
                                           unsigned long long& my_mono) {
    constexpr int GL = 64 / LPT;  // targets per wavefront
    using E = typename EntryT<R>::T;
    using M = typename std::conditional<sizeof(R) == 4, uint4, double4>::type;
    const int lane = threadIdx.x & 63;
    E* L = lds.list;
    int32_t* S = lds.stack;
    uint32_t* O = lds.owner;
    const M* mono = (const M*)a.mono;
    const R scale = (R)a.scale;
    const R eps2 = (R)a.eps2, rcut2 = (R)a.rcut2;
    // the opening angles as arguments of the arithmetic type (no per-step conversion)
    const R theta2 = sizeof(R) == 4 ? (R)a.theta2_f : (R)a.theta2;
    const R thetac2 = sizeof(R) == 4 ? (R)a.theta_cut2_f : (R)a.theta_cut2;
    const R rcut2s = (R)(a.rcut2 + a.eps2);
    const R rcut2u = (R)(a.rcut2 / (a.scale * a.scale));  // r_cut^2 in fixed-point units
    const R qpart[2] = {(R)a.qm, (R)a.qp};
    // entries one popped node can append, rounded up to a power of two so the pop bound is a shift
    int app_shift = 1;
    while ((1 << app_shift) < a.leaf) ++app_shift;
    const uint64_t lt = lanemask_lt();

    const int2 gr = a.groups[g];
    const int tgt = lane % GL, sub = lane / GL;
    const bool valid = tgt < gr.y;
    const int64_t ti = (int64_t)gr.x + (valid ? tgt : 0);
    const uint4 rec = a.parts[ti];
    const int nvalid = gr.y < GL ? gr.y : GL;

    uint32_t gmn[3], gmx[3], gc[3], gh[3];
    {
        const uint32_t p[3] = {rec.x, rec.y, rec.z};
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            gmn[ax] = wave_min_u32(valid ? p[ax] : 0xffffffffu);
            gmx[ax] = wave_max_u32(valid ? p[ax] : 0u);
            gc[ax] = gmn[ax] + (gmx[ax] - gmn[ax]) / 2;
            gh[ax] = (gmx[ax] - gmn[ax]) / 2 + ((gmx[ax] - gmn[ax]) & 1u);
        }
    }
    const R tx = fixed_rel<R>(rec.x, gc[0], scale), ty = fixed_rel<R>(rec.y, gc[1], scale),
            tz = fixed_rel<R>(rec.z, gc[2], scale);
    const R ghu[3] = {(R)gh[0] + (R)0.5, (R)gh[1] + (R)0.5, (R)gh[2] + (R)0.5};
    R sx = 0, sy = 0, sz = 0, cx = 0, cy = 0, cz = 0;
    double spot = 0;  // POT: short-range potential sum

    int sp = 1, lp = 0, lm = 0;
    int pidx = -1, pdst = 0;  // a bucket particle to list in the next step (index, list slot)
    int mono_inside = 0;  // monopole entries placed on the particle side since the last flush
    int64_t steps = 0;
    if (a.cellnode) {
        // seed with the 27 cells around the group's cell: a cell's edge is >= r_cut, so every
        // source within r_cut of the group lies in them (the group lies inside one cell)
        const int lv = a.start_level, sh = 32 - lv, msk = (1 << lv) - 1;
        int32_t nd = kNoNodeW;
        if (lane < 27) {
            const int cx = (int)((gc[0] >> sh) + (lane % 3) - 1) & msk;
            const int cy = (int)((gc[1] >> sh) + ((lane / 3) % 3) - 1) & msk;
            const int cz = (int)((gc[2] >> sh) + (lane / 9) - 1) & msk;
            const uint64_t cm = spread3_21((uint32_t)cx) | (spread3_21((uint32_t)cy) << 1) |
                                (spread3_21((uint32_t)cz) << 2);
            nd = a.cellnode[cm];
        }
        const uint64_t have = __ballot(nd != kNoNodeW);
        if (nd != kNoNodeW) S[__popcll(have & lt)] = nd;
        sp = __popcll(have);
    } else if (lane == 0) {
        S[0] = a.root;
    }
    wave_lds_fence();

    // particles and cut-law monopoles grow up from L[0], tail-law monopoles down from L[kListCap - 1]
    auto drain = [&]() {  // list the pending bucket particles before the list is read
        if (pidx >= 0) L[pdst] = part_entry<R, E>(a.parts[pidx], gc, scale, qpart);
        pidx = -1;
        wave_lds_fence();
    };
    auto flush = [&](int np_in, int nm_in) {
        // the counts are wave-uniform: keep the loop control in scalar registers
        const int np_ = __builtin_amdgcn_readfirstlane(np_in);
        const int nm_ = __builtin_amdgcn_readfirstlane(nm_in);
        const int itp = (np_ + LPT - 1) / LPT, itm = (nm_ + LPT - 1) / LPT;
        if (LPT > 1) {  // zero-charge padding so every lane reads a finite entry
            if (lane < LPT - 1) {
                E z;
                z.x = z.y = z.z = (R)0;
                z.w = (R)0;
                L[np_ + lane] = z;
                L[kListCap - nm_ - 1 - lane] = z;
            }
            wave_lds_fence();
        }
        R px = 0, py = 0, pz = 0;
        double ppot = 0;
#pragma unroll 4
        for (int jj = 0; jj < itp; ++jj) {
            const E e = L[jj * LPT + sub];
            const R dx = e.x - tx, dy = e.y - ty, dz = e.z - tz;
            const R f = Pair<R, MODE, true>::f(e.w, dx, dy, dz, a, eps2, rcut2, rcut2s);
            px += f * dx;
            py += f * dy;
            pz += f * dz;
            if (POT) ppot += pair_potential<true>((double)e.w, dx, dy, dz, a, (double)eps2, (double)rcut2);
        }
        const int m0 = kListCap - itm * LPT + sub;
#pragma unroll 4
        for (int jj = 0; jj < itm; ++jj) {
            const E e = L[m0 + jj * LPT];
            const R dx = e.x - tx, dy = e.y - ty, dz = e.z - tz;
            const R f = Pair<R, MODE, MODE == JTP_SPLIT_POLY4>::f(e.w, dx, dy, dz, a, eps2, rcut2, rcut2s);
            px += f * dx;
            py += f * dy;
            pz += f * dz;
            if (POT) ppot += pair_potential<false>((double)e.w, dx, dy, dz, a, (double)eps2, (double)rcut2);
        }
#pragma unroll
        for (int off = GL; off < 64; off <<= 1) {  // combine the LPT partial sums of a target
            px += __shfl_xor(px, off, 64);
            py += __shfl_xor(py, off, 64);
            pz += __shfl_xor(pz, off, 64);
            if (POT) ppot += __shfl_xor(ppot, off, 64);
        }
        if (POT) spot += ppot;
        if (sizeof(R) == 4) {  // compensated accumulation of the chunk partials
            R y = px - cx, t = sx + y;
            cx = (t - sx) - y;
            sx = t;
            y = py - cy; t = sy + y; cy = (t - sy) - y; sy = t;
            y = pz - cz; t = sz + y; cz = (t - sz) - y; sz = t;
        } else {
            sx += px;
            sy += py;
            sz += pz;
        }
        my_inter += (unsigned long long)(np_ + nm_) * (unsigned long long)nvalid;
        my_mono += (unsigned long long)(nm_ + mono_inside) * (unsigned long long)nvalid;
        mono_inside = 0;
    };

    while (sp > 0) {
        if (++steps > a.max_steps) {  // a finite hierarchy never gets here; report instead of spinning
            if (lane == 0) atomicExch(&a.counters[kCtrErr], 2);
            break;
        }
        int k = sp < 64 ? sp : 64;
        const int kl = (kListCap - 2 * kPad - lp - lm) >> app_shift;
        k = k < kl ? k : kl;
        const int ks = kStackCap - sp > 1 ? kStackCap - sp : 1;
        k = k < ks ? k : ks;
        sp -= k;
        const bool has = lane < k;
        const int32_t popped = has ? S[sp + lane] : a.root;
        wave_lds_fence();
        const bool bad = has && (popped >= (int32_t)(a.n - 1) || ~popped >= (int32_t)a.n);
        if (bad) atomicExch(&a.counters[kCtrErr], 3);  // malformed child index: skipped
        const bool live = has && !bad;
        const int32_t node = live ? popped : a.root;

        // Classification, free of branches except the memory traffic: an internal node's record is
        // (box centre, first) and (box extent, last); a particle child's is its own record and a
        // zero extent.
        const bool isnode = node >= 0;
        const uint4* rec0 = isnode ? a.geo + 2 * node : a.parts + ~node;
        const uint4 g0 = rec0[0];
        const uint4 g1r = (isnode ? rec0 + 1 : rec0)[0];
        // an internal node's monopoles and children are read with its box (one memory round
        // trip per step for all three; particle children read node 0's, a shared cache line)
        const int32_t snode = isnode ? node : 0;
        const M mr0 = mono[2 * snode], mr1 = mono[2 * snode + 1];
        const int2 kd = a.kids[snode];
        if (pidx >= 0) L[pdst] = part_entry<R, E>(a.parts[pidx], gc, scale, qpart);  // previous step's
        const uint4 g1 = isnode ? g1r : make_uint4(0u, 0u, 0u, 0u);
        const int32_t first = isnode ? (int32_t)g0.w : ~node;
        const int32_t cnt = isnode ? (int32_t)g1.w - first + 1 : 1;
        // box-to-box gap and farthest-corner distance in fixed-point units: the node spans
        // centre -+ (extent / 2 + 1/2), the group centre -+ (gh + 1/2) (ghu); the periodic minimum
        // image comes from the 32-bit wrap of the centre offset
        R gap2 = 0, far2 = 0;
        {
            const uint32_t nc[3] = {g0.x, g0.y, g0.z}, ne[3] = {g1.x, g1.y, g1.z};
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const R adc = fabs((R)(int32_t)(nc[ax] - gc[ax]));
                const R t = (R)ne[ax] * (R)0.5 + ghu[ax];
                const R gp = fmax(adc - t, (R)0);
                gap2 += gp * gp;
                const R fr = adc + t;
                far2 += fr * fr;
            }
        }
        const R ext = (R)max(max(g1.x, g1.y), g1.z);
        const bool reach = live && gap2 < rcut2u;
        const bool inside = far2 < rcut2u;
        // accept the two monopoles / take the bucket's particles / open
        const bool acc = reach && isnode && ext * ext < (inside ? theta2 : thetac2) * gap2;
        const bool bucket = reach && !acc && cnt <= a.leaf;
        const bool open = reach && !acc && !bucket;
        const E m0 = mono_entry(mr0, gc, scale), m1 = mono_entry(mr1, gc, scale);
        const bool h0 = acc && m0.w != (R)0, h1 = acc && m1.w != (R)0;
        const int nq = (h0 ? 1 : 0) + (h1 ? 1 : 0);
        // FP32: monopoles of a node entirely within r_cut take the (exact there) cut pair law and sit
        // on the particle side; the others take the tail law from the top of the list
        const bool cutlaw = sizeof(R) == 4 && inside && MODE != JTP_SPLIT_POLY4;
        const int nb = bucket ? cnt : 0;
        const int n4 = cutlaw ? nq : 0;
        const int nt = cutlaw ? 0 : nq;
        // one scan over (bucket particles | cut-law monopoles << 11 | tail monopoles << 19):
        // per wave at most 64 x 16, 128 and 128 entries
        const int packed = nb | (n4 << 11) | (nt << 19);
        const int incl = wave_incl_scan_dpp(packed);
        const int totl = __builtin_amdgcn_readlane(incl, 63);
        const int btot = totl & 0x7ff, ptot = btot + ((totl >> 11) & 0xff), ttot = totl >> 19;
        const int bexcl = (incl & 0x7ff) - nb;
        if (nq) {  // particle side after the bucket particles, or top down
            const bool top = !cutlaw;
            int o = top ? kListCap - 1 - (lm + (incl >> 19) - nt) : lp + btot + ((incl >> 11) & 0xff) - n4;
            const int d = top ? -1 : 1;
            if (h0) {
                L[o] = m0;
                o += d;
            }
            if (h1) L[o] = m1;
        }
        // bucket particles, one lane per particle: item w of the step belongs to the bucket lane
        // whose run [bexcl, bexcl + nb) holds it, found by a running maximum over lane-indexed
        // run starts (64 items per round)
        // The first round's particles are read and listed during the next step (their loads
        // overlap that step's node reads); later rounds are listed now.
        int nidx = -1, ndst = 0;
        for (int r0 = 0; r0 < btot; r0 += 64) {
            O[lane] = 0u;
            wave_lds_fence();
            if (nb && bexcl < r0 + 64 && bexcl + nb > r0) O[bexcl > r0 ? bexcl - r0 : 0] = (uint32_t)lane + 1u;
            wave_lds_fence();
            // every lane takes part in the lane reads (a read from an inactive lane returns 0)
            const int ow = max((int)wave_incl_max_dpp(O[lane]) - 1, 0);
            const int fb = __shfl(first, ow, 64), eb = __shfl(bexcl, ow, 64);
            const int w = r0 + lane;
            if (w < btot) {
                if (r0 == 0) {
                    nidx = fb + (w - eb);
                    ndst = lp + w;
                } else {
                    L[lp + w] = part_entry<R, E>(a.parts[fb + (w - eb)], gc, scale, qpart);
                }
            }
        }
        // pushes
        const uint64_t om = __ballot(open);
        const int nopen = __popcll(om);
        if (sp + 2 * nopen > kStackPhys) {
            if (lane == 0) atomicExch(&a.counters[kCtrErr], 1);
        } else if (open) {
            const int rank = __popcll(om & lt);
            S[sp + 2 * rank] = kd.y;
            S[sp + 2 * rank + 1] = kd.x;
        }
        if (sp + 2 * nopen <= kStackPhys) sp += 2 * nopen;
        // the counts are wave-uniform: keep them in scalar registers
        pidx = nidx;
        pdst = ndst;
        lp = __builtin_amdgcn_readfirstlane(lp + ptot);
        lm = __builtin_amdgcn_readfirstlane(lm + ttot);
        sp = __builtin_amdgcn_readfirstlane(sp);
        mono_inside += ptot - btot;
        wave_lds_fence();
        if (lp + lm >= kFlushAt) {
            drain();
            flush(lp, lm);
            lp = 0;
            lm = 0;
            wave_lds_fence();
        }
    }
    drain();
    if (lp + lm > 0) flush(lp, lm);
    wave_lds_fence();

    if (valid && sub == 0) {
        const R sg = (int32_t)rec.w > 0 ? (R)1 : (R)-1;
        R ox = sg * sx, oy = sg * sy, oz = sg * sz;
        if (a.pm_acc) {
            const R* pm = (const R*)a.pm_acc;
            ox += pm[3 * ti];
            oy += pm[3 * ti + 1];
            oz += pm[3 * ti + 2];
        }
        const int64_t o = (int64_t)a.perm[ti];
        R* out = (R*)a.out;
        out[3 * o] = ox;
        out[3 * o + 1] = oy;
        out[3 * o + 2] = oz;
        if (POT) a.pot[o] = -spot;  // Phi_short = -sum_j q_j erfc(rho / 2 r_s) / rho
    }
}

// XCD-local queues: the work items (in Morton order) are cut into kXcds contiguous ranges and a
// workgroup first drains the range of the XCD it runs on (workgroup i -> XCD i % 8), so
// concurrent groups on one XCD share tree nodes in that XCD's L2; then it steals from the
// following ranges.  Returns the next item or -1.  Each item's result does not depend on who
// computes it.
__device__ inline int next_item(int32_t n_items, int32_t* heads, int xcd, int& steal) {
    const int lane = threadIdx.x & 63;
    while (steal < kXcds) {
        const int q = (xcd + steal) % kXcds;
        const int32_t start = (int32_t)((int64_t)n_items * q / kXcds);
        const int32_t end = (int32_t)((int64_t)n_items * (q + 1) / kXcds);
        int t = 0;
        if (lane == 0) t = atomicAdd(&heads[q], 1);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
        if (start + t < end) return start + t;
        ++steal;
    }
    return -1;
}

// One wavefront per group of one box.
template <typename R, int MODE, int WAVES, int LPT, bool POT = false>
__global__ __launch_bounds__(64 * WAVES) void k_walk(WalkArgs a) {
    __shared__ WalkLds<R, LPT> lds_s[WAVES];
    WalkLds<R, LPT>& lds = lds_s[threadIdx.x >> 6];
    unsigned long long my_inter = 0, my_mono = 0;
    const int32_t n_groups = a.counters[kCtrNGroups];
    const int xcd = blockIdx.x % kXcds;
    int steal = 0;
    for (int g; (g = next_item(n_groups, a.counters + kCtrXcd, xcd, steal)) >= 0;)
        walk_group<R, MODE, LPT, POT>(a, g, lds, my_inter, my_mono);
    if ((threadIdx.x & 63) == 0 && my_inter) {
        atomicAdd(a.inter, my_inter);
        atomicAdd(a.inter + 1, my_mono);
    }
}

// One persistent launch over the groups of several boxes: work item t is group t - goff[b] of box
// b (goff[b] <= t < goff[b + 1], boxes in order); each box's arguments, counters and outputs are
// its own, so every box's result is the one its own launch gives.
template <typename R, int MODE, int WAVES, int LPT>
__global__ __launch_bounds__(64 * WAVES) void k_walk_boxes(const WalkArgs* __restrict__ boxes,
                                                         const int32_t* __restrict__ goff, int32_t nbox,
                                                         int32_t* heads) {
    __shared__ WalkLds<R, LPT> lds_s[WAVES];
    WalkLds<R, LPT>& lds = lds_s[threadIdx.x >> 6];
    const int32_t n_items = goff[nbox];
    const int xcd = blockIdx.x % kXcds;
    int steal = 0;
    for (int t; (t = next_item(n_items, heads, xcd, steal)) >= 0;) {
        int lo = 0, hi = nbox;  // largest b with goff[b] <= t
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (goff[mid] <= t) lo = mid; else hi = mid;
        }
        const int b = __builtin_amdgcn_readfirstlane(lo);
        const WalkArgs& a = boxes[b];
        unsigned long long my_inter = 0, my_mono = 0;
        walk_group<R, MODE, LPT, false>(a, t - goff[b], lds, my_inter, my_mono);
        if ((threadIdx.x & 63) == 0 && my_inter) {
            atomicAdd(a.inter, my_inter);
            atomicAdd(a.inter + 1, my_mono);
        }
    }
}

template <typename R>
__global__ void k_pm_only(const R* __restrict__ pm, const int32_t* __restrict__ perm, int64_t n, R* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = perm[i];
    R v0 = 0, v1 = 0, v2 = 0;
    if (pm) {
        v0 = pm[3 * i];
        v1 = pm[3 * i + 1];
        v2 = pm[3 * i + 2];
    }
    out[3 * o] = v0;
    out[3 * o + 1] = v1;
    out[3 * o + 2] = v2;
}

hipError_t write_pm_only(Ctx& c, int64_t n, void* acc_out, hipStream_t s) {
    if (n == 0) return hipGetLastError();
    const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    if (c.p.precision == JTP_FP64)
        k_pm_only<double><<<dim3(nb), dim3(kBlock), 0, s>>>(c.p.pm_enabled ? (const double*)c.pm_acc : nullptr,
                                                            c.vals0, n, (double*)acc_out);
    else
        k_pm_only<float><<<dim3(nb), dim3(kBlock), 0, s>>>(c.p.pm_enabled ? (const float*)c.pm_acc : nullptr,
                                                           c.vals0, n, (float*)acc_out);
    return hipGetLastError();
}

// Persistent grid: enough workgroups for every work item, at most 16 per CU (what fits).
template <int WAVES>
static unsigned walk_blocks(int64_t items) {
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int64_t blocks = (items + WAVES - 1) / WAVES;
    const int64_t cap = (int64_t)cus * 16;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

template <typename R, int MODE, int WAVES, int LPT, bool POT = false>
static hipError_t launch_walk(const WalkArgs& a, int64_t gmax, hipStream_t s) {
    k_walk<R, MODE, WAVES, LPT, POT><<<dim3(walk_blocks<WAVES>(gmax)), dim3(64 * WAVES), 0, s>>>(a);
    return hipGetLastError();
}

template <typename R, int WAVES, int LPT>
static hipError_t launch_walk_mode(int mode, const WalkArgs& a, int64_t gmax, hipStream_t s) {
    if (mode == JTP_SPLIT_ERFC_GAUSS) return launch_walk<R, JTP_SPLIT_ERFC_GAUSS, WAVES, LPT>(a, gmax, s);
    if (mode == JTP_SPLIT_ERFC) return launch_walk<R, JTP_SPLIT_ERFC, WAVES, LPT>(a, gmax, s);
    return launch_walk<R, JTP_SPLIT_POLY4, WAVES, LPT>(a, gmax, s);
}

static void fill_args(Ctx& c, int64_t n, void* acc_out, WalkArgs& a) {
    a = WalkArgs{};
    a.parts = c.parts;
    a.geo = c.node_geo;
    a.kids = c.node_kids;
    a.mono = c.node_mono;
    a.groups = c.groups;
    a.counters = c.dev_counters;
    a.inter = c.dev_inter;
    a.pm_acc = c.p.pm_enabled ? c.pm_acc : nullptr;
    a.perm = c.vals0;
    a.out = acc_out;
    a.n = n;
    a.root = n >= 2 ? 0 : ~0;
    a.leaf = c.p.leaf_size;
    a.cellnode = n >= 2 ? c.cellnode : nullptr;
    a.start_level = c.start_level;
    a.max_steps = 4 * n + 1024;  // each step pops >= 1 of the <= 2n-1 nodes
    const double tc = c.p.theta_cut < c.p.theta ? c.p.theta_cut : c.p.theta;
    a.theta_cut2 = tc * tc;
    a.scale = c.p.box_size / 4294967296.0;
    a.eps2 = c.p.softening * c.p.softening;
    a.rcut2 = c.p.r_cut * c.p.r_cut;
    a.theta2 = c.p.theta * c.p.theta;
    a.theta2_f = (float)a.theta2;
    a.theta_cut2_f = (float)a.theta_cut2;
    a.inv2rs = c.p.r_s > 0 ? 0.5 / c.p.r_s : 0.0;
    a.qp = c.p.g_constant * c.m_plus;
    a.qm = -c.p.g_constant * c.m_minus;
    a.poly = c.poly;
    // t = x * 2/X - 1 with x = rho/(2 r_s) and X the fit range (make_split_poly)
    a.tscale = c.poly_tscale;
    a.tscale_m = c.poly_tscale_m;
    a.rho2_tail = c.poly_rho2_tail;
    a.escale = (float)(c.p.r_s > 0 ? -1.4426950408889634 / (4.0 * c.p.r_s * c.p.r_s) : 0.0);
}

hipError_t walk(Ctx& c, int64_t n, void* acc_out, hipStream_t s, double* pot) {
    WalkArgs a;
    fill_args(c, n, acc_out, a);
    const int64_t gmax = c.n_groups_max;
    const int mode = c.p.split_mode;
    const int lpt = lanes_per_target(c.p.group_size);
    if (c.p.precision == JTP_FP64 && pot) {  // energy diagnostic: erfc + Gaussian law only
        // The short-range kernel is not harmonic (its Laplacian is 4 pi G times the Gaussian density
        // of the split, > 0 off the origin), so a monopole evaluated at a node's centre of mass falls
        // short of the node's mean potential by ~ <d^2>/6 times that Laplacian: a bias of one sign that
        // the forces do not show but the potential sums up (about 2 % of the short-range potential
        // at theta = 0.5 on a uniform lattice).  The potential therefore opens every node (theta = 0).
        a.theta2 = a.theta_cut2 = 0.0;
        a.theta2_f = a.theta_cut2_f = 0.0f;
        a.pot = pot;
        if (lpt == 4) return launch_walk<double, JTP_SPLIT_ERFC_GAUSS, 2, 4, true>(a, gmax, s);
        if (lpt == 2) return launch_walk<double, JTP_SPLIT_ERFC_GAUSS, 2, 2, true>(a, gmax, s);
        return launch_walk<double, JTP_SPLIT_ERFC_GAUSS, 2, 1, true>(a, gmax, s);
    }
    if (c.p.precision == JTP_FP64) {
        if (lpt == 4) return launch_walk_mode<double, 2, 4>(mode, a, gmax, s);
        if (lpt == 2) return launch_walk_mode<double, 2, 2>(mode, a, gmax, s);
        return launch_walk_mode<double, 2, 1>(mode, a, gmax, s);
    }
    if (lpt == 4) return launch_walk_mode<float, 4, 4>(mode, a, gmax, s);
    if (lpt == 2) return launch_walk_mode<float, 4, 2>(mode, a, gmax, s);
    return launch_walk_mode<float, 4, 1>(mode, a, gmax, s);
}

// ---- one walk launch over several boxes ----

size_t walk_args_bytes() { return sizeof(WalkArgs); }

void walk_fill_args(Ctx& c, int64_t n, void* acc_out, void* dst) { fill_args(c, n, acc_out, *(WalkArgs*)dst); }

int walk_batch_key(const Ctx& c) {
    return (c.p.precision * 4 + c.p.split_mode) * 8 + lanes_per_target(c.p.group_size);
}

// goff[b] = groups of the boxes before b (each box's count as its own build wrote it), goff[nbox] the
// total; the kXcds queue heads after it start at 0.  One workgroup.
__global__ __launch_bounds__(256) void k_walk_boxes_offsets(const WalkArgs* __restrict__ boxes, int32_t nbox,
                                                            int32_t* __restrict__ ctrl) {
    __shared__ int32_t cnt_s[256];
    int32_t acc = 0;  // thread 0's running total
    for (int b0 = 0; b0 < nbox; b0 += 256) {
        const int b = b0 + (int)threadIdx.x;
        cnt_s[threadIdx.x] = b < nbox ? boxes[b].counters[kCtrNGroups] : 0;
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < 256 && b0 + i < nbox; ++i) {
                ctrl[b0 + i] = acc;
                acc += cnt_s[i];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) ctrl[nbox] = acc;
    if (threadIdx.x < kXcds) ctrl[nbox + 1 + threadIdx.x] = 0;
}

template <typename R, int MODE, int WAVES, int LPT>
static hipError_t launch_walk_boxes(const WalkArgs* d_args, int32_t nbox, int32_t* d_ctrl, int64_t items_max,
                                    hipStream_t s) {
    k_walk_boxes<R, MODE, WAVES, LPT><<<dim3(walk_blocks<WAVES>(items_max)), dim3(64 * WAVES), 0, s>>>(
        d_args, d_ctrl, nbox, d_ctrl + nbox + 1);
    return hipGetLastError();
}

template <typename R, int WAVES, int LPT>
static hipError_t launch_walk_boxes_mode(int mode, const WalkArgs* d, int32_t nbox, int32_t* ctrl, int64_t items,
                                         hipStream_t s) {
    if (mode == JTP_SPLIT_ERFC_GAUSS) return launch_walk_boxes<R, JTP_SPLIT_ERFC_GAUSS, WAVES, LPT>(d, nbox, ctrl, items, s);
    if (mode == JTP_SPLIT_ERFC) return launch_walk_boxes<R, JTP_SPLIT_ERFC, WAVES, LPT>(d, nbox, ctrl, items, s);
    return launch_walk_boxes<R, JTP_SPLIT_POLY4, WAVES, LPT>(d, nbox, ctrl, items, s);
}

hipError_t walk_boxes(const void* d_args, int32_t nbox, int32_t* d_ctrl, int64_t items_max, int key, hipStream_t s) {
    if (nbox < 1) return hipSuccess;
    const WalkArgs* d = (const WalkArgs*)d_args;
    k_walk_boxes_offsets<<<dim3(1), dim3(256), 0, s>>>(d, nbox, d_ctrl);
    const int lpt = key % 8, mode = (key / 8) % 4, prec = key / 32;
    if (prec == JTP_FP64) {
        if (lpt == 4) return launch_walk_boxes_mode<double, 2, 4>(mode, d, nbox, d_ctrl, items_max, s);
        if (lpt == 2) return launch_walk_boxes_mode<double, 2, 2>(mode, d, nbox, d_ctrl, items_max, s);
        return launch_walk_boxes_mode<double, 2, 1>(mode, d, nbox, d_ctrl, items_max, s);
    }
    if (lpt == 4) return launch_walk_boxes_mode<float, 4, 4>(mode, d, nbox, d_ctrl, items_max, s);
    if (lpt == 2) return launch_walk_boxes_mode<float, 4, 2>(mode, d, nbox, d_ctrl, items_max, s);
    return launch_walk_boxes_mode<float, 4, 1>(mode, d, nbox, d_ctrl, items_max, s);
}

}  // namespace jtp
