// Karras LBVH over the sorted Morton keys, integer-exact bottom-up dual
// (m+, m-) monopoles, walk groups and cell offset tables.
//
// Reference behaviour replaced:
//   build_bvh_tp + delta_tp / range_tp / split_tp / clz64_tp (src/nbody_gpu_twopass.rs:356-447):
//     same radix-tree topology (same delta with the index tie-break, same range and split search);
//     leaves are encoded here as ~sorted_index instead of n-1+index.
//   init_leaves_tp + reduce_tp (nbody_gpu_twopass.rs:450-561) build ONE species per tree in FP32
//     with atomics; here one tree holds both species and every node keeps separate (n+, sum x+) and
//     (n-, sum x-) as exact 64-bit integer sums of fixed-point coordinates, so the monopoles are
//     independent of summation order and bit-reproducible.  The dual-monopole node is the design of
//     src/nbody_gpu.rs:603-874 (init_leaves/reduce_com); no periodic unfolding is needed because a
//     radix-tree node is a Morton-prefix range and never straddles the periodic boundary.
#include "jtp_internal.h"
#include "jtp_device.h"

namespace jtp {

struct alignas(16) NodeSum {  // 64 bytes: the stage copy reads this layout
    uint64_t sp[3];
    uint64_t sm[3];
    uint32_t np, nm;
    uint32_t pad[2];
};
static_assert(sizeof(NodeSum) == 64, "NodeSum layout");

__device__ inline int clz64_dev(uint64_t x) { return x == 0 ? 64 : __clzll((long long)x); }

__device__ inline int kdelta(const uint64_t* __restrict__ k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + clz64_dev((uint64_t)(i ^ j));
    return clz64_dev(a ^ b);
}

__global__ void k_karras(const uint64_t* __restrict__ keys, int64_t n, int2* __restrict__ rng,
                         int2* __restrict__ kids, int2* __restrict__ up, int32_t* __restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int64_t first, last;
    if (i == 0) {
        first = 0;
        last = n - 1;
    } else {
        const int dl = kdelta(keys, n, i, i - 1);
        const int dr = kdelta(keys, n, i, i + 1);
        const int64_t d = (dr > dl) ? 1 : -1;
        const int dmin = (d > 0) ? dl : dr;
        int64_t lmax = 2;
        while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        const int64_t j = i + l * d;
        first = d > 0 ? i : j;
        last = d > 0 ? j : i;
    }
    // split search (split_tp)
    const int dn = kdelta(keys, n, first, last);
    int64_t s = 0, t = last - first;
    while (t > 1) {
        t = (t + 1) / 2;
        if (kdelta(keys, n, first, first + s + t) > dn) s += t;
    }
    const int64_t gamma = first + s;
    const int32_t lc = (first == gamma) ? ~(int32_t)first : (int32_t)gamma;
    const int32_t rc = (gamma + 1 == last) ? ~(int32_t)last : (int32_t)(gamma + 1);
    kids[i] = make_int2(lc, rc);
    rng[i] = make_int2((int32_t)first, (int32_t)last);  // the bottom-up pass writes the full records
    // an internal child's parent and sibling, so that the bottom-up climb reads one record per level
    if (lc >= 0) up[lc] = make_int2((int32_t)i, rc);
    if (rc >= 0) up[rc] = make_int2((int32_t)i, lc);
    if (i == 0) up[0] = make_int2(-1, 0);  // the root
    flags[i] = (lc < 0 ? 1 : 0) + (rc < 0 ? 1 : 0);  // arrivals still missing: 2 - this
}

struct ChildData {
    uint32_t mn[3], mx[3];
    uint64_t sp[3], sm[3];
    uint32_t np, nm;
    uint32_t first, last;  // sorted leaf range
};

__device__ inline void load_child(int32_t c, const uint4* parts, uint4* geo, NodeSum* sums, ChildData& o) {
    if (c < 0) {
        const uint4 r = parts[~c];
        const uint32_t p[3] = {r.x, r.y, r.z};
        const bool plus = (int32_t)r.w > 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o.mn[a] = p[a];
            o.mx[a] = p[a];
            o.sp[a] = plus ? p[a] : 0;
            o.sm[a] = plus ? 0 : p[a];
        }
        o.np = plus ? 1u : 0u;
        o.nm = plus ? 0u : 1u;
        o.first = o.last = (uint32_t)~c;
    } else {
        const uint4 g0 = geo[2 * c], g1 = geo[2 * c + 1];
        o.first = g0.w;
        o.last = g1.w;
        const uint32_t nc[3] = {g0.x, g0.y, g0.z}, e[3] = {g1.x, g1.y, g1.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o.mn[a] = nc[a] - e[a] / 2;
            o.mx[a] = o.mn[a] + e[a];
        }
        const NodeSum s = sums[c];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o.sp[a] = s.sp[a];
            o.sm[a] = s.sm[a];
        }
        o.np = s.np;
        o.nm = s.nm;
    }
}

// Rounded integer mean of a fixed-point coordinate sum.
__device__ inline uint32_t mean_fixed(uint64_t s, uint32_t cnt) {
    return (uint32_t)((s + (uint64_t)(cnt >> 1)) / (uint64_t)cnt);
}
__device__ inline double mean_fixed_f64(uint64_t s, uint32_t cnt) {
    const uint64_t q = s / cnt, r = s % cnt;
    return (double)q + (double)r / (double)cnt;
}

// Walk records of one node, written as soon as its sums and box are final.
//   FP32: {com+ xyz (fixed), q+ (float bits)}, {com- xyz, q-}
//   FP64: {com+ xyz (fixed units, as double), q+}, {com- xyz, q-}
// q = sigma G |m| n; an empty species gets the box centre and q = 0.
template <int PREC>
__device__ inline void write_mono(int32_t p, const NodeSum& s, const uint32_t mn[3], const uint32_t mx[3],
                                  void* mono, double qp, double qm) {
    if (PREC == JTP_FP64) {
        const double c[3] = {(double)mn[0] + 0.5 * (double)(mx[0] - mn[0]),
                             (double)mn[1] + 0.5 * (double)(mx[1] - mn[1]),
                             (double)mn[2] + 0.5 * (double)(mx[2] - mn[2])};
        double4 a, b;
        if (s.np)
            a = make_double4(mean_fixed_f64(s.sp[0], s.np), mean_fixed_f64(s.sp[1], s.np),
                             mean_fixed_f64(s.sp[2], s.np), qp * (double)s.np);
        else
            a = make_double4(c[0], c[1], c[2], 0.0);
        if (s.nm)
            b = make_double4(mean_fixed_f64(s.sm[0], s.nm), mean_fixed_f64(s.sm[1], s.nm),
                             mean_fixed_f64(s.sm[2], s.nm), -qm * (double)s.nm);
        else
            b = make_double4(c[0], c[1], c[2], 0.0);
        ((double4*)mono)[2 * p] = a;
        ((double4*)mono)[2 * p + 1] = b;
    } else {
        const uint32_t c[3] = {mn[0] + (mx[0] - mn[0]) / 2, mn[1] + (mx[1] - mn[1]) / 2, mn[2] + (mx[2] - mn[2]) / 2};
        const float qpf = (float)qp, qmf = (float)qm;
        uint4 a, b;
        if (s.np)
            a = make_uint4(mean_fixed(s.sp[0], s.np), mean_fixed(s.sp[1], s.np), mean_fixed(s.sp[2], s.np),
                           __float_as_uint(qpf * (float)s.np));
        else
            a = make_uint4(c[0], c[1], c[2], __float_as_uint(0.0f));
        if (s.nm)
            b = make_uint4(mean_fixed(s.sm[0], s.nm), mean_fixed(s.sm[1], s.nm), mean_fixed(s.sm[2], s.nm),
                           __float_as_uint(-qmf * (float)s.nm));
        else
            b = make_uint4(c[0], c[1], c[2], __float_as_uint(0.0f));
        ((uint4*)mono)[2 * p] = a;
        ((uint4*)mono)[2 * p + 1] = b;
    }
}

// Node p from its two children in the fixed order (left, right): exact integer sums, bounding
// box, leaf range and walk records, written out; returns the node's data for its parent.  The
// result does not depend on which thread or pass completes the node.
template <int PREC>
__device__ inline ChildData combine_write(int32_t p, const ChildData& a, const ChildData& b, uint4* __restrict__ geo,
                                          NodeSum* __restrict__ sums, void* mono, double qp, double qm) {
    ChildData o;
    NodeSum s;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        s.sp[ax] = o.sp[ax] = a.sp[ax] + b.sp[ax];
        s.sm[ax] = o.sm[ax] = a.sm[ax] + b.sm[ax];
        o.mn[ax] = min(a.mn[ax], b.mn[ax]);
        o.mx[ax] = max(a.mx[ax], b.mx[ax]);
    }
    s.np = o.np = a.np + b.np;
    s.nm = o.nm = a.nm + b.nm;
    s.pad[0] = s.pad[1] = 0;
    o.first = a.first;
    o.last = b.last;
    sums[p] = s;
    // box as (centre, first) and (extent, last): the walk's box test needs no further arithmetic,
    // and min = centre - extent / 2, max = min + extent recover the corners exactly
    uint32_t e[3], nc[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        e[ax] = o.mx[ax] - o.mn[ax];
        nc[ax] = o.mn[ax] + e[ax] / 2;
    }
    geo[2 * p] = make_uint4(nc[0], nc[1], nc[2], o.first);
    geo[2 * p + 1] = make_uint4(e[0], e[1], e[2], o.last);
    write_mono<PREC>(p, s, o.mn, o.mx, mono, qp, qm);
    return o;
}

// A node whose two children are leaves: both loaded.
template <int PREC>
__device__ inline ChildData finish_start(int32_t p, const uint4* __restrict__ parts, uint4* __restrict__ geo,
                                         const int2* __restrict__ kids, NodeSum* __restrict__ sums, void* mono,
                                         double qp, double qm) {
    const int2 k = kids[p];
    ChildData a, b;
    load_child(k.x, parts, geo, sums, a);
    load_child(k.y, parts, geo, sums, b);
    return combine_write<PREC>(p, a, b, geo, sums, mono, qp, qm);
}

// The second arrival at node `par` climbs with the data of its child `p` in registers and loads
// only the sibling `sib` (up[p].y).  The left child's index is the split gamma and the right one's
// gamma + 1, leaves and internal nodes alike, so p is the left child iff p < the sibling's index.
template <int PREC>
__device__ inline ChildData finish_up(int32_t par, int32_t p, int32_t sib, const ChildData& cur,
                                      const uint4* __restrict__ parts, uint4* __restrict__ geo,
                                      NodeSum* __restrict__ sums, void* mono, double qp, double qm) {
    const bool left = p < (sib < 0 ? ~sib : sib);
    ChildData o;
    load_child(sib, parts, geo, sums, o);
    return left ? combine_write<PREC>(par, cur, o, geo, sums, mono, qp, qm)
                : combine_write<PREC>(par, o, cur, geo, sums, mono, qp, qm);
}

// Bottom-up pass in two launches, with no device-scope fence (the reference climbs with a
// device-scope fence per level, nbody_gpu_twopass.rs:485-560; on MI355X a device-scope release
// writes back the per-XCD L2, which made that pattern ~100x slower than this one).
//
// 1. k_bu_local: one workgroup per chunk of kBuChunk consecutive sorted leaves.  A Karras node
//    whose leaf range lies inside the chunk ("local") has only local descendants and its index
//    lies in the chunk.  Threads climb from the local nodes with two leaf children: finish the
//    node, then bump its parent's arrival count in LDS; the arrival that completes the parent
//    continues with it, the other stops.  Only waves of one workgroup meet at a local node, so
//    workgroup-scope release / acquire fences order the node data.  A local node whose parent is
//    not local bumps the parent's global arrival count (flags[], preset by k_karras to the number
//    of leaf children); the arrival that completes that parent appends it to a global list.
// 2. k_bu_mid: the same one level up, one workgroup per superchunk of 64 chunks, with the arrival
//    counts in flags[]; all arrivals in this launch at a node inside a superchunk come from that
//    superchunk's workgroup, so workgroup-scope fences suffice again.
// 3. k_bu_top: one workgroup climbs from the second list through the remaining nodes (those whose
//    range crosses a superchunk edge).
// Each node is finished exactly once, with its children combined in the fixed order (left, right).
constexpr int kBuChunk = 4096;
static_assert(kBuChunk <= 65536, "start offsets are 16-bit");

__device__ inline void wg_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup"); }
__device__ inline void wg_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }

constexpr int kBuLocalThreads = 512;  // 8 waves per workgroup, 4 workgroups (32 waves) per CU

template <int PREC>
__global__ __launch_bounds__(kBuLocalThreads) void k_bu_local(const uint4* __restrict__ parts, int64_t nI,
                                                              uint4* __restrict__ geo, const int2* __restrict__ rng,
                                                  const int2* __restrict__ kids, const int2* __restrict__ up,
                                                  int32_t* __restrict__ flags, NodeSum* __restrict__ sums, void* mono,
                                                  double qp, double qm, int32_t* __restrict__ gfront,
                                                  int32_t* __restrict__ ctrl) {
    __shared__ int32_t cnt[kBuChunk];    // arrival count of a local node, -1 if the node is not local
    __shared__ uint16_t start[kBuChunk];  // local nodes with two leaf children (offset from base)
    __shared__ int32_t nstart;
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kBuChunk;
    const int64_t lastLeaf = base + kBuChunk - 1;
    if (tid == 0) nstart = 0;
    __syncthreads();
    for (int k = tid; k < kBuChunk; k += blockDim.x) {
        const int64_t i = base + k;
        int32_t c = -1;
        if (i < nI) {
            // leaf children from the topology, not from flags[] (other chunks may already be arriving there)
            const int2 kc = kids[i];
            const int32_t leaves = (kc.x < 0 ? 1 : 0) + (kc.y < 0 ? 1 : 0);
            const int2 fl = rng[i];
            const int64_t first = fl.x, last = fl.y;
            if (first >= base && last <= lastLeaf) {
                c = leaves;
                if (leaves == 2) start[atomicAdd(&nstart, 1)] = (uint16_t)k;
            } else if (leaves == 2) {
                gfront[atomicAdd(&ctrl[0], 1)] = (int32_t)i;
            }
        }
        cnt[k] = c;
    }
    __syncthreads();
    const int ns = nstart;
    for (int t = tid; t < ns; t += blockDim.x) {
        int32_t p = (int32_t)(base + start[t]);
        ChildData cur = finish_start<PREC>(p, parts, geo, kids, sums, mono, qp, qm);
        while (true) {
            if (p == 0) {
                ctrl[kMaxRounds + 1] = 1;  // root complete
                break;
            }
            const int2 u = up[p];
            const int32_t par = u.x;
            if (par < 0 || par >= nI) break;
            const int64_t lp = (int64_t)par - base;
            if (lp >= 0 && lp < kBuChunk && cnt[lp] >= 0) {
                wg_release();
                if (atomicAdd(&cnt[lp], 1) != 1) break;
                wg_acquire();
                cur = finish_up<PREC>(par, p, u.y, cur, parts, geo, sums, mono, qp, qm);
                p = par;
            } else {
                if (atomicAdd(&flags[par], 1) == 1) gfront[atomicAdd(&ctrl[0], 1)] = par;
                break;
            }
        }
    }
}

// Middle level: one 1024-thread workgroup per superchunk of kBuSuper leaves climbs from the listed
// nodes whose leaf range lies inside its superchunk.  Such a node's descendants lie in the same
// superchunk and were completed either by the first launch or by this workgroup, so workgroup
// fences suffice; a finished node whose parent crosses the superchunk edge is handed on (second
// arrival) to the final list.
constexpr int64_t kBuSuper = (int64_t)kBuChunk * 64;

template <int PREC>
__global__ __launch_bounds__(1024) void k_bu_mid(const uint4* __restrict__ parts, int64_t nI, uint4* __restrict__ geo,
                                                 const int2* __restrict__ rng, const int2* __restrict__ kids,
                                                 const int2* __restrict__ up,
                                                 int32_t* __restrict__ flags, NodeSum* __restrict__ sums, void* mono,
                                                 double qp, double qm, const int32_t* __restrict__ starts,
                                                 int32_t* __restrict__ next, int32_t* __restrict__ ctrl) {
    const int64_t lo = (int64_t)blockIdx.x * kBuSuper, hi = lo + kBuSuper - 1;  // leaf range of the superchunk
    const int ns = ctrl[0];
    for (int t = threadIdx.x; t < ns; t += blockDim.x) {
        int32_t p = starts[t];
        if (p < lo || p > hi) continue;  // a node's index lies in its own range
        const int2 fl = rng[p];
        if (fl.x < lo || fl.y > hi) {
            next[atomicAdd(&ctrl[1], 1)] = p;  // ready, but it crosses the superchunk edge
            continue;
        }
        ChildData cur = finish_start<PREC>(p, parts, geo, kids, sums, mono, qp, qm);
        while (true) {
            if (p == 0) {
                ctrl[kMaxRounds + 1] = 1;
                break;
            }
            const int2 u = up[p];
            const int32_t par = u.x;
            if (par < 0 || par >= nI) break;
            const bool inside = par >= lo && par <= hi && rng[par].x >= lo && rng[par].y <= hi;
            if (inside) {
                wg_release();
                if (atomicAdd(&flags[par], 1) != 1) break;
                wg_acquire();
                cur = finish_up<PREC>(par, p, u.y, cur, parts, geo, sums, mono, qp, qm);
                p = par;
            } else {
                if (atomicAdd(&flags[par], 1) == 1) next[atomicAdd(&ctrl[1], 1)] = par;
                break;
            }
        }
    }
}

template <int PREC>
__global__ __launch_bounds__(1024) void k_bu_top(const uint4* __restrict__ parts, int64_t nI, uint4* __restrict__ geo,
                                                 const int2* __restrict__ kids, const int2* __restrict__ up,
                                                 int32_t* __restrict__ flags, NodeSum* __restrict__ sums, void* mono,
                                                 double qp, double qm, const int32_t* __restrict__ starts,
                                                 int32_t* __restrict__ ctrl) {
    const int ns = ctrl[1];
    for (int t = threadIdx.x; t < ns; t += blockDim.x) {
        int32_t p = starts[t];
        // a listed node's children are complete: both loaded
        ChildData cur = finish_start<PREC>(p, parts, geo, kids, sums, mono, qp, qm);
        while (true) {
            if (p == 0) {
                ctrl[kMaxRounds + 1] = 1;
                break;
            }
            const int2 u = up[p];
            const int32_t par = u.x;
            if (par < 0 || par >= nI) break;
            wg_release();
            if (atomicAdd(&flags[par], 1) != 1) break;
            wg_acquire();
            cur = finish_up<PREC>(par, p, u.y, cur, parts, geo, sums, mono, qp, qm);
            p = par;
        }
    }
}

// ------------------------------------------------------------------ cell offsets
// out[c] = first sorted index whose cell (key >> shift) is >= c, for c in [0, ncells].
// Every entry is written exactly once, by the particle that starts the run.
__global__ void k_cell_start(const uint64_t* __restrict__ keys, int64_t n, int shift, int64_t ncells,
                             int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t c = (int64_t)(keys[i] >> shift);
    const int64_t cp = (i == 0) ? -1 : (int64_t)(keys[i - 1] >> shift);
    for (int64_t cc = cp + 1; cc <= c; ++cc) out[cc] = (int32_t)i;
    if (i == n - 1)
        for (int64_t cc = c + 1; cc <= ncells; ++cc) out[cc] = (int32_t)n;
}
__global__ void k_fill_i32(int32_t* __restrict__ out, int64_t m, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = v;
}

hipError_t launch_fill_i32(int32_t* out, int64_t m, int32_t v, hipStream_t s) {
    if (m > 0) k_fill_i32<<<dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(out, m, v);
    return hipGetLastError();
}

hipError_t launch_cell_start(const uint64_t* keys, int64_t n, int shift, int64_t ncells, int32_t* out,
                             hipStream_t s) {
    if (n == 0) {
        k_fill_i32<<<dim3((unsigned)((ncells + 1 + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(out, ncells + 1, 0);
    } else {
        k_cell_start<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(keys, n, shift, ncells,
                                                                                       out);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ start cells of the walk
// The keys of one Morton cell of level `lvl` (8^lvl cells, edge L / 2^lvl) are a contiguous sorted
// range [s, e), and in a binary radix tree they form exactly one subtree: the node whose range is
// [s, e - 1] (a leaf if e - s == 1).  A Karras node's index is one of its range ends, so that node
// is s or e - 1.  cellnode[c] = that node (~leaf for a single particle, kNoNode if the cell is
// empty); the walk seeds each group's traversal with the 27 cells around the group's cell.
constexpr int32_t kNoNode = (int32_t)0x80000000;

__global__ void k_cell_nodes(const int32_t* __restrict__ cell4, int lvl, int64_t nI, const uint4* __restrict__ geo,
                             int32_t* __restrict__ cellnode, int32_t* __restrict__ counters) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncell = 1 << (3 * lvl);
    if (c >= ncell) return;
    const int f = 1 << (3 * (kGroupLevelBits - lvl));  // level-4 cells per cell of this level
    const int32_t s = cell4[c * f], e = cell4[(c + 1) * f];
    int32_t node = kNoNode;
    if (e - s == 1) {
        node = ~s;
    } else if (e - s >= 2) {
        if (s < nI && (int32_t)geo[2 * s].w == s && (int32_t)geo[2 * s + 1].w == e - 1) node = s;
        else if (e - 1 < nI && (int32_t)geo[2 * (e - 1)].w == s && (int32_t)geo[2 * (e - 1) + 1].w == e - 1) node = e - 1;
        else counters[kCtrErr] = 4;  // no node holds exactly this cell: malformed hierarchy
    }
    cellnode[c] = node;
}

// ------------------------------------------------------------------ walk groups
// Groups are runs of <= gs consecutive sorted particles inside one level-4 Morton cell
// (edge L/16).  goff[c] = index of the first group of cell c; counters[0] = number of groups.
constexpr int kNCell4 = 1 << (3 * kGroupLevelBits);

__global__ __launch_bounds__(256) void k_group_offsets(const int32_t* __restrict__ cell4, int gs,
                                                       int32_t* __restrict__ goff, int32_t* __restrict__ counters) {
    __shared__ uint32_t lds[4];
    constexpr int per = kNCell4 / 256;
    const int tid = threadIdx.x;
    uint32_t v[per], s = 0;
#pragma unroll
    for (int j = 0; j < per; ++j) {
        const int c = tid * per + j;
        const int cnt = cell4[c + 1] - cell4[c];
        v[j] = (uint32_t)((cnt + gs - 1) / gs);
        s += v[j];
    }
    uint32_t tot;
    uint32_t off = block256_excl_scan(s, lds, &tot);
#pragma unroll
    for (int j = 0; j < per; ++j) {
        goff[tid * per + j] = (int32_t)off;
        off += v[j];
    }
    if (tid == 0) {
        goff[kNCell4] = (int32_t)tot;
        counters[kCtrNGroups] = (int32_t)tot;
        counters[kCtrErr] = 0;
    }
    if (tid < kXcds) counters[kCtrXcd + tid] = 0;  // walk queue heads
}

__global__ void k_groups(const int32_t* __restrict__ cell4, const int32_t* __restrict__ goff, int gs,
                         int64_t gmax, int2* __restrict__ groups) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= gmax || g >= goff[kNCell4]) return;
    int lo = 0, hi = kNCell4;  // find c with goff[c] <= g < goff[c+1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (goff[mid] <= g) lo = mid; else hi = mid;
    }
    // skip empty cells that share the same offset
    int c = lo;
    const int64_t j = g - goff[c];
    const int32_t first = cell4[c] + (int32_t)(j * gs);
    const int32_t cnt = min((int32_t)gs, cell4[c + 1] - first);
    groups[g] = make_int2(first, cnt);
}

hipError_t build_tree(Ctx& c, int64_t n, hipStream_t s) {
    if (n >= 2) {
        const int64_t nI = n - 1;
        const unsigned nbI = (unsigned)((nI + kBlock - 1) / kBlock);
        int2* rng = (int2*)c.keys1;  // leaf ranges; the sort's second key buffer is free until the next sort
        k_karras<<<dim3(nbI), dim3(kBlock), 0, s>>>(c.keys0, n, rng, c.node_kids, c.up, c.flags);
        int32_t* ctrl = c.dev_counters + kCtrRounds;
        hipError_t e = launch_fill_i32(ctrl, kMaxRounds + 2, 0, s);
        if (e != hipSuccess) return e;
        const unsigned nbLocal = (unsigned)((n + kBuChunk - 1) / kBuChunk);
        const unsigned nbSuper = (unsigned)((n + kBuSuper - 1) / kBuSuper);
        const double qp = c.p.g_constant * c.m_plus, qm = c.p.g_constant * c.m_minus;
        NodeSum* sums = (NodeSum*)c.node_sum;
        if (c.p.precision == JTP_FP64) {
            k_bu_local<JTP_FP64><<<dim3(nbLocal), dim3(kBuLocalThreads), 0, s>>>(c.parts, nI, c.node_geo, rng, c.node_kids, c.up,
                                                                     c.flags, sums, c.node_mono, qp, qm, c.frontier0, ctrl);
            k_bu_mid<JTP_FP64><<<dim3(nbSuper), dim3(1024), 0, s>>>(c.parts, nI, c.node_geo, rng, c.node_kids, c.up,
                                                                    c.flags, sums, c.node_mono, qp, qm, c.frontier0,
                                                                    c.frontier1, ctrl);
            k_bu_top<JTP_FP64><<<dim3(1), dim3(1024), 0, s>>>(c.parts, nI, c.node_geo, c.node_kids, c.up, c.flags,
                                                              sums, c.node_mono, qp, qm, c.frontier1, ctrl);
        } else {
            k_bu_local<JTP_FP32><<<dim3(nbLocal), dim3(kBuLocalThreads), 0, s>>>(c.parts, nI, c.node_geo, rng, c.node_kids, c.up,
                                                                     c.flags, sums, c.node_mono, qp, qm, c.frontier0, ctrl);
            k_bu_mid<JTP_FP32><<<dim3(nbSuper), dim3(1024), 0, s>>>(c.parts, nI, c.node_geo, rng, c.node_kids, c.up,
                                                                    c.flags, sums, c.node_mono, qp, qm, c.frontier0,
                                                                    c.frontier1, ctrl);
            k_bu_top<JTP_FP32><<<dim3(1), dim3(1024), 0, s>>>(c.parts, nI, c.node_geo, c.node_kids, c.up, c.flags,
                                                              sums, c.node_mono, qp, qm, c.frontier1, ctrl);
        }
    }
    hipError_t e = launch_cell_start(c.keys0, n, 63 - 3 * kGroupLevelBits, kNCell4, c.cell4, s);
    if (e != hipSuccess) return e;
    int32_t* goff = c.cell4 + (kNCell4 + 1);
    k_group_offsets<<<dim3(1), dim3(256), 0, s>>>(c.cell4, c.p.group_size, goff, c.dev_counters);
    if (c.n_groups_max > 0)
        k_groups<<<dim3((unsigned)((c.n_groups_max + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(
            c.cell4, goff, c.p.group_size, c.n_groups_max, c.groups);
    if (n >= 2) {
        const int lvl = c.start_level;
        k_cell_nodes<<<dim3(((1u << (3 * lvl)) + 255) / 256), dim3(256), 0, s>>>(c.cell4, lvl, n - 1, c.node_geo,
                                                                              c.cellnode, c.dev_counters);
    }
    return hipGetLastError();
}

}  // namespace jtp
