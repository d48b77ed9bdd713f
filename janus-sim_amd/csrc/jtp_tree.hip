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

__global__ void k_karras(const uint64_t* __restrict__ keys, int64_t n, uint4* __restrict__ geo,
                         int2* __restrict__ kids, int32_t* __restrict__ parent_int,
                         int32_t* __restrict__ parent_leaf, int32_t* __restrict__ flags) {
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
    geo[2 * i].w = (uint32_t)first;
    geo[2 * i + 1].w = (uint32_t)last;
    if (lc >= 0) parent_int[lc] = (int32_t)i; else parent_leaf[~lc] = (int32_t)i;
    if (rc >= 0) parent_int[rc] = (int32_t)i; else parent_leaf[~rc] = (int32_t)i;
    flags[i] = 0;
}

struct ChildData {
    uint32_t mn[3], mx[3];
    uint64_t sp[3], sm[3];
    uint32_t np, nm;
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
    } else {
        const uint4 g0 = geo[2 * c], g1 = geo[2 * c + 1];
        o.mn[0] = g0.x; o.mn[1] = g0.y; o.mn[2] = g0.z;
        o.mx[0] = g1.x; o.mx[1] = g1.y; o.mx[2] = g1.z;
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

// Bottom-up pass as a sequence of frontiers, one kernel per tree level.
// Round 0 holds the nodes whose two children are leaves; a node joins the next
// frontier when its second child completes (relaxed arrival counter).  Kernel
// boundaries publish each round's writes to the next, so no device-scope fence
// is needed inside a kernel (the arrival-counter climb with a fence per level is
// what the reference does, nbody_gpu_twopass.rs:485-560).  Each node combines its
// children in the fixed order (left, right), so the result does not depend on
// which round or thread produced it.
__global__ __launch_bounds__(256) void k_bu_init(const int2* __restrict__ kids, int64_t nI, int32_t* __restrict__ cnt,
                                                 int32_t* __restrict__ f0, int32_t* __restrict__ ctrl) {
    __shared__ int32_t lds[6];
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nI; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        int leaves = 0;
        if (i < nI) {
            const int2 k = kids[i];
            leaves = (k.x < 0 ? 1 : 0) + (k.y < 0 ? 1 : 0);
            cnt[i] = leaves;
        }
        const int slot = block_append(leaves == 2, &ctrl[0], lds);
        if (slot >= 0) f0[slot] = (int32_t)i;
    }
}

__global__ __launch_bounds__(256) void k_bu_round(int r, const uint4* __restrict__ parts, int64_t nI, uint4* __restrict__ geo,
                           const int2* __restrict__ kids, const int32_t* __restrict__ parent_int,
                           int32_t* __restrict__ cnt, NodeSum* __restrict__ sums, const int32_t* __restrict__ fin,
                           int32_t* __restrict__ fout, int32_t* __restrict__ ctrl) {
    __shared__ int32_t lds[6];
    const int32_t size = ctrl[r];
    // uniform trip count per block so that the frontier append can use block_append
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < size; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t idx = base + threadIdx.x;
        int32_t par = -1;
        bool ready = false;
        const int32_t p = idx < size ? fin[idx] : -1;
        if (p >= 0 && p < nI) {
            const int2 k = kids[p];
            ChildData a, b;
            load_child(k.x, parts, geo, sums, a);
            load_child(k.y, parts, geo, sums, b);
            NodeSum s;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                s.sp[ax] = a.sp[ax] + b.sp[ax];
                s.sm[ax] = a.sm[ax] + b.sm[ax];
            }
            s.np = a.np + b.np;
            s.nm = a.nm + b.nm;
            s.pad[0] = s.pad[1] = 0;
            sums[p] = s;
            const uint4 g0 = geo[2 * p], g1 = geo[2 * p + 1];
            geo[2 * p] = make_uint4(min(a.mn[0], b.mn[0]), min(a.mn[1], b.mn[1]), min(a.mn[2], b.mn[2]), g0.w);
            geo[2 * p + 1] = make_uint4(max(a.mx[0], b.mx[0]), max(a.mx[1], b.mx[1]), max(a.mx[2], b.mx[2]), g1.w);
            if (p == 0) {
                ctrl[kMaxRounds + 1] = 1;  // root complete
            } else {
                par = parent_int[p];
                if (par >= 0 && par < nI) ready = atomicAdd(&cnt[par], 1) == 1;
            }
        }
        const int slot = block_append(ready, &ctrl[r + 1], lds);
        if (slot >= 0) fout[slot] = par;
    }
}

// Rounded integer mean of a fixed-point coordinate sum.
__device__ inline uint32_t mean_fixed(uint64_t s, uint32_t cnt) {
    return (uint32_t)((s + (uint64_t)(cnt >> 1)) / (uint64_t)cnt);
}

// FP32 walk record: {com+ xyz (fixed), q+ (float bits)}, {com- xyz, q-}; q = sigma G |m| n.
__global__ void k_mono_f32(const NodeSum* __restrict__ sums, const uint4* __restrict__ geo, int64_t nI,
                           float qp_unit, float qm_unit, uint4* __restrict__ mono) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nI) return;
    const NodeSum s = sums[i];
    const uint4 g0 = geo[2 * i], g1 = geo[2 * i + 1];
    uint32_t c[3] = {g0.x + (g1.x - g0.x) / 2, g0.y + (g1.y - g0.y) / 2, g0.z + (g1.z - g0.z) / 2};
    uint4 a, b;
    if (s.np) {
        a = make_uint4(mean_fixed(s.sp[0], s.np), mean_fixed(s.sp[1], s.np), mean_fixed(s.sp[2], s.np),
                       __float_as_uint(qp_unit * (float)s.np));
    } else {
        a = make_uint4(c[0], c[1], c[2], __float_as_uint(0.0f));
    }
    if (s.nm) {
        b = make_uint4(mean_fixed(s.sm[0], s.nm), mean_fixed(s.sm[1], s.nm), mean_fixed(s.sm[2], s.nm),
                       __float_as_uint(-qm_unit * (float)s.nm));
    } else {
        b = make_uint4(c[0], c[1], c[2], __float_as_uint(0.0f));
    }
    mono[2 * i] = a;
    mono[2 * i + 1] = b;
}

// FP64 walk record: {com+ xyz (fixed units, as double), q+}, {com- xyz, q-}.
__device__ inline double mean_fixed_f64(uint64_t s, uint32_t cnt) {
    const uint64_t q = s / cnt, r = s % cnt;
    return (double)q + (double)r / (double)cnt;
}
__global__ void k_mono_f64(const NodeSum* __restrict__ sums, const uint4* __restrict__ geo, int64_t nI,
                           double qp_unit, double qm_unit, double4* __restrict__ mono) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nI) return;
    const NodeSum s = sums[i];
    const uint4 g0 = geo[2 * i], g1 = geo[2 * i + 1];
    const double c[3] = {(double)g0.x + 0.5 * (double)(g1.x - g0.x), (double)g0.y + 0.5 * (double)(g1.y - g0.y),
                         (double)g0.z + 0.5 * (double)(g1.z - g0.z)};
    double4 a, b;
    if (s.np)
        a = make_double4(mean_fixed_f64(s.sp[0], s.np), mean_fixed_f64(s.sp[1], s.np),
                         mean_fixed_f64(s.sp[2], s.np), qp_unit * (double)s.np);
    else
        a = make_double4(c[0], c[1], c[2], 0.0);
    if (s.nm)
        b = make_double4(mean_fixed_f64(s.sm[0], s.nm), mean_fixed_f64(s.sm[1], s.nm),
                         mean_fixed_f64(s.sm[2], s.nm), -qm_unit * (double)s.nm);
    else
        b = make_double4(c[0], c[1], c[2], 0.0);
    mono[2 * i] = a;
    mono[2 * i + 1] = b;
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
    const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    if (n >= 2) {
        const int64_t nI = n - 1;
        const unsigned nbI = (unsigned)((nI + kBlock - 1) / kBlock);
        launch_fill_i32(c.parent_int, nI, -1, s);
        launch_fill_i32(c.parent_leaf, n, -1, s);
        k_karras<<<dim3(nbI), dim3(kBlock), 0, s>>>(c.keys0, n, c.node_geo, c.node_kids, c.parent_int,
                                                    c.parent_leaf, c.flags);
        int32_t* ctrl = c.dev_counters + kCtrRounds;
        launch_fill_i32(ctrl, kMaxRounds + 2, 0, s);
        const unsigned nbInit = nbI < 4096 ? nbI : 4096;
        k_bu_init<<<dim3(nbInit), dim3(kBlock), 0, s>>>(c.node_kids, nI, c.flags, c.frontier0, ctrl);
        int blocks = (int)((nI / 4 + kBlock - 1) / kBlock);
        blocks = blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
        for (int r = 0; r < kMaxRounds; ++r) {
            int32_t* fin = (r & 1) ? c.frontier1 : c.frontier0;
            int32_t* fout = (r & 1) ? c.frontier0 : c.frontier1;
            k_bu_round<<<dim3(blocks), dim3(kBlock), 0, s>>>(r, c.parts, nI, c.node_geo, c.node_kids, c.parent_int,
                                                             c.flags, (NodeSum*)c.node_sum, fin, fout, ctrl);
        }
        const double qp = c.p.g_constant * c.m_plus, qm = c.p.g_constant * c.m_minus;
        if (c.p.precision == JTP_FP64)
            k_mono_f64<<<dim3(nbI), dim3(kBlock), 0, s>>>((const NodeSum*)c.node_sum, c.node_geo, nI, qp, qm,
                                                          (double4*)c.node_mono);
        else
            k_mono_f32<<<dim3(nbI), dim3(kBlock), 0, s>>>((const NodeSum*)c.node_sum, c.node_geo, nI, (float)qp,
                                                          (float)qm, (uint4*)c.node_mono);
    }
    hipError_t e = launch_cell_start(c.keys0, n, 63 - 3 * kGroupLevelBits, kNCell4, c.cell4, s);
    if (e != hipSuccess) return e;
    int32_t* goff = c.cell4 + (kNCell4 + 1);
    k_group_offsets<<<dim3(1), dim3(256), 0, s>>>(c.cell4, c.p.group_size, goff, c.dev_counters);
    if (c.n_groups_max > 0)
        k_groups<<<dim3((unsigned)((c.n_groups_max + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(
            c.cell4, goff, c.p.group_size, c.n_groups_max, c.groups);
    return hipGetLastError();
}

}  // namespace jtp
