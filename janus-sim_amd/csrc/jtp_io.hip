// Snapshot I/O in the reference's two on-disk formats, with the records packed and unpacked on the
// GPU and the file streamed through pinned host buffers.
//
// Reference behaviour replaced:
//   save_snapshot, src/bin/janus_mu19_production.rs:199-216 (host loop writing 25-byte records from a
//     device-to-host copy of the particle arrays) -> JTP_SNAP_PRODUCTION.
//   write_snapshot_v3 / read_snapshot_v3 / read_header_only, src/snapshot_v3.rs:108-157, 159-255,
//     313-370, 376-463 (408-byte header, 36-byte records, one record at a time through BufWriter /
//     BufReader) -> JTP_SNAP_V3.
//
// A chunk of kIoChunk records is packed by k_pack_* into a device staging buffer as the exact file
// bytes, copied to a pinned host buffer and written with one fwrite; reading runs the same path
// backwards.  The 25-byte production records are not 4-byte aligned, so each 256-record block is
// assembled byte-wise in LDS and leaves (or arrives) as aligned 32-bit words.
#include "jtp_internal.h"

#include <cstdio>
#include <cstring>
#include <string>

namespace jtp {

constexpr int kRecProd = 25;   // 3 f32 position, 3 f32 velocity, i8 sign
constexpr int kRecV3 = 36;     // + f32 mass, f32 epsilon, u8 sign, split_level, is_star, flags
constexpr int kHeadProd = 16;  // u32 N, f32 L, u32 step, f32 z
constexpr int kHeadV3 = 408;
constexpr uint64_t kMagicV3 = 0x0A335653554E414Aull;  // "JANUSV3\n" little-endian
constexpr int64_t kIoChunk = 1 << 20;                 // records per staging chunk

__device__ inline float fixed_to_f32(uint32_t u, double L) {
    return (float)(((double)u * (1.0 / 4294967296.0) - 0.5) * L);
}

__device__ inline uint32_t f32_to_fixed(float x, double L) {
    double t = (double)x / L + 0.5;
    t -= floor(t);
    uint64_t w = (uint64_t)(t * 4294967296.0);
    if (w >= 4294967296ull) w -= 4294967296ull;
    return (uint32_t)w;
}

__device__ inline float load_vel(const void* v, int64_t k, int f64) {
    return f64 ? (float)((const double*)v)[k] : ((const float*)v)[k];
}

__device__ inline void store_vel(void* v, int64_t k, int f64, float x) {
    if (f64) ((double*)v)[k] = (double)x;
    else ((float*)v)[k] = x;
}

__device__ inline void put_f32(uint8_t* b, float x) {
    const uint32_t u = __float_as_uint(x);
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k] = (uint8_t)(u >> (8 * k));
}

__device__ inline float get_f32(const uint8_t* b) {
    return __uint_as_float((uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24));
}

// Production records of particles [first, first + m) into out (m * 25 bytes).
__global__ __launch_bounds__(256) void k_pack_prod(const uint32_t* __restrict__ pos, const void* __restrict__ vel,
                                                   int vel_f64, const int8_t* __restrict__ sp, int64_t first,
                                                   int64_t m, double L, float vscale, uint8_t* __restrict__ out) {
    __shared__ uint32_t words[256 * kRecProd / 4];
    uint8_t* lb = (uint8_t*)words;
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    const int64_t r = r0 + threadIdx.x;
    if (r < m) {
        const int64_t i = first + r;
        uint8_t* b = lb + threadIdx.x * kRecProd;
        for (int a = 0; a < 3; ++a) put_f32(b + 4 * a, fixed_to_f32(pos[3 * i + a], L));
        for (int a = 0; a < 3; ++a) put_f32(b + 12 + 4 * a, load_vel(vel, 3 * i + a, vel_f64) * vscale);
        b[24] = (uint8_t)(sp[i] > 0 ? 1 : -1);
    }
    __syncthreads();
    const int64_t nrec = m - r0 < 256 ? m - r0 : 256;
    const int64_t nbytes = nrec * kRecProd;
    uint8_t* dst = out + r0 * kRecProd;  // 256 * 25 is a multiple of 4: dst is 4-byte aligned
    for (int64_t w = threadIdx.x; w < nbytes / 4; w += 256) ((uint32_t*)dst)[w] = words[w];
    for (int64_t t = (nbytes / 4) * 4 + threadIdx.x; t < nbytes; t += 256) dst[t] = lb[t];
}

// The inverse; bad[0] counts sign bytes that are 0.
__global__ __launch_bounds__(256) void k_unpack_prod(const uint8_t* __restrict__ in, int64_t first, int64_t m,
                                                     double L, float inv_vscale, uint32_t* __restrict__ pos,
                                                     void* __restrict__ vel, int vel_f64, int8_t* __restrict__ sp,
                                                     int32_t* __restrict__ bad) {
    __shared__ uint32_t words[256 * kRecProd / 4];
    uint8_t* lb = (uint8_t*)words;
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    const int64_t nrec = m - r0 < 256 ? m - r0 : 256;
    const int64_t nbytes = nrec * kRecProd;
    const uint8_t* src = in + r0 * kRecProd;
    for (int64_t w = threadIdx.x; w < nbytes / 4; w += 256) words[w] = ((const uint32_t*)src)[w];
    for (int64_t t = (nbytes / 4) * 4 + threadIdx.x; t < nbytes; t += 256) lb[t] = src[t];
    __syncthreads();
    const int64_t r = r0 + threadIdx.x;
    if (r >= m) return;
    const int64_t i = first + r;
    const uint8_t* b = lb + threadIdx.x * kRecProd;
    for (int a = 0; a < 3; ++a) pos[3 * i + a] = f32_to_fixed(get_f32(b + 4 * a), L);
    for (int a = 0; a < 3; ++a) store_vel(vel, 3 * i + a, vel_f64, get_f32(b + 12 + 4 * a) * inv_vscale);
    const int8_t s = (int8_t)b[24];
    sp[i] = s > 0 ? 1 : -1;
    if (s == 0) atomicAdd(bad, 1);
}

// v3 records (36 bytes, 4-byte aligned): position, velocity, mass and epsilon per species, sign 1 / 255.
__global__ __launch_bounds__(256) void k_pack_v3(const uint32_t* __restrict__ pos, const void* __restrict__ vel,
                                                 int vel_f64, const int8_t* __restrict__ sp, int64_t first,
                                                 int64_t m, double L, float vscale, float mp, float mm, float ep,
                                                 float em, uint32_t* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    const int64_t i = first + r;
    const bool plus = sp[i] > 0;
    uint32_t* o = out + r * (kRecV3 / 4);
    for (int a = 0; a < 3; ++a) o[a] = __float_as_uint(fixed_to_f32(pos[3 * i + a], L));
    for (int a = 0; a < 3; ++a) o[3 + a] = __float_as_uint(load_vel(vel, 3 * i + a, vel_f64) * vscale);
    o[6] = __float_as_uint(plus ? mp : mm);
    o[7] = __float_as_uint(plus ? ep : em);
    o[8] = plus ? 1u : 255u;  // sign, split_level 0, is_star 0, flags 0
}

// The inverse; bad[0] counts sign bytes other than 1 and 255.
__global__ __launch_bounds__(256) void k_unpack_v3(const uint32_t* __restrict__ in, int64_t first, int64_t m,
                                                   double L, float inv_vscale, uint32_t* __restrict__ pos,
                                                   void* __restrict__ vel, int vel_f64, int8_t* __restrict__ sp,
                                                   int32_t* __restrict__ bad) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    const int64_t i = first + r;
    const uint32_t* w = in + r * (kRecV3 / 4);
    for (int a = 0; a < 3; ++a) pos[3 * i + a] = f32_to_fixed(__uint_as_float(w[a]), L);
    for (int a = 0; a < 3; ++a) store_vel(vel, 3 * i + a, vel_f64, __uint_as_float(w[3 + a]) * inv_vscale);
    const uint32_t s = w[8] & 0xffu;
    sp[i] = s == 1u ? 1 : -1;
    if (s != 1u && s != 255u) atomicAdd(bad, 1);
}

}  // namespace jtp

using namespace jtp;

namespace {

thread_local std::string g_io_err;

int io_fail(int code, const std::string& msg) {
    g_io_err = msg;
    return code;
}

template <typename T>
void put_le(uint8_t*& p, T v) {
    std::memcpy(p, &v, sizeof(T));  // gfx950 hosts (x86-64) are little-endian like the file
    p += sizeof(T);
}

template <typename T>
T get_le(const uint8_t*& p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return v;
}

// snapshot_v3.rs:108-157 field order
void encode_v3_header(const jtp_snap_meta* m, uint8_t* h) {
    uint8_t* p = h;
    put_le<uint64_t>(p, kMagicV3);
    put_le<uint32_t>(p, 3);
    put_le<uint32_t>(p, kHeadV3);
    put_le<uint64_t>(p, (uint64_t)m->n);
    put_le<double>(p, m->a);
    put_le<double>(p, m->t_gyr);
    put_le<double>(p, m->box_size);
    put_le<double>(p, m->h0);
    put_le<double>(p, m->mu);
    put_le<double>(p, m->omega_b);
    put_le<double>(p, m->m_part_plus_base);
    put_le<double>(p, m->m_part_minus_base);
    put_le<double>(p, m->eps_plus_base);
    put_le<double>(p, m->eps_minus_base);
    put_le<uint32_t>(p, m->n_split_max);
    put_le<uint32_t>(p, m->seed_ic);
    put_le<double>(p, m->z_init);
    put_le<uint64_t>(p, m->n_stars);
    put_le<double>(p, m->z_start_run);
    put_le<double>(p, m->sfr);
    put_le<double>(p, m->rho_max);
    std::memcpy(p, m->run_label, 256);
}

// snapshot_v3.rs:159-255; returns the header size stored in the file (>= 408) or a JTP_ERR code
int decode_v3_header(const uint8_t* h, jtp_snap_meta* m) {
    const uint8_t* p = h;
    const uint64_t magic = get_le<uint64_t>(p);
    const uint32_t version = get_le<uint32_t>(p);
    const uint32_t hsize = get_le<uint32_t>(p);
    if (magic != kMagicV3) return io_fail(JTP_ERR_ARG, "not a v3 snapshot (bad magic number)");
    if (version != 3) return io_fail(JTP_ERR_ARG, "unsupported v3 snapshot version " + std::to_string(version));
    if (hsize < (uint32_t)kHeadV3) return io_fail(JTP_ERR_ARG, "v3 header_size below 408");
    m->n = (int64_t)get_le<uint64_t>(p);
    m->a = get_le<double>(p);
    m->t_gyr = get_le<double>(p);
    m->box_size = get_le<double>(p);
    m->h0 = get_le<double>(p);
    m->mu = get_le<double>(p);
    m->omega_b = get_le<double>(p);
    m->m_part_plus_base = get_le<double>(p);
    m->m_part_minus_base = get_le<double>(p);
    m->eps_plus_base = get_le<double>(p);
    m->eps_minus_base = get_le<double>(p);
    m->n_split_max = get_le<uint32_t>(p);
    m->seed_ic = get_le<uint32_t>(p);
    m->z_init = get_le<double>(p);
    m->n_stars = get_le<uint64_t>(p);
    m->z_start_run = get_le<double>(p);
    m->sfr = get_le<double>(p);
    m->rho_max = get_le<double>(p);
    std::memcpy(m->run_label, p, 256);
    m->run_label[255] = 0;
    m->z = 1.0 / m->a - 1.0;
    return (int)hsize;
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

struct Staging {
    void* dev = nullptr;
    void* host = nullptr;
    int32_t* bad = nullptr;
    ~Staging() {
        if (dev) (void)hipFree(dev);
        if (bad) (void)hipFree(bad);
        if (host) (void)hipHostFree(host);
    }
};

int rec_size(int32_t format) { return format == JTP_SNAP_V3 ? kRecV3 : kRecProd; }

int check_format(int32_t format) {
    if (format != JTP_SNAP_PRODUCTION && format != JTP_SNAP_V3) return io_fail(JTP_ERR_ARG, "unknown snapshot format");
    return JTP_OK;
}

#define IO_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return io_fail(JTP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" {

void jtp_snap_meta_default(jtp_snap_meta* m, int32_t format, double box_size) {
    if (!m) return;
    std::memset(m, 0, sizeof(*m));
    m->format = format;
    m->box_size = box_size;
    m->velocity_scale = 1.0;
    // SnapshotHeaderV3::default, snapshot_v3.rs:48-76
    m->a = 1.0;
    m->h0 = 69.9;
    m->mu = 19.0;
    m->omega_b = 0.05;
    m->m_part_plus_base = 1e10;
    m->m_part_minus_base = 1e10;
    m->eps_plus_base = 0.5;
    m->eps_minus_base = 2.5;
    m->seed_ic = 42;
    m->z_init = 10.0;
    m->z_start_run = 10.0;
    m->mass_plus = 1.0f;
    m->mass_minus = 1.0f;
    m->eps_plus = 0.5f;
    m->eps_minus = 2.5f;
}

const char* jtp_io_last_error(void) { return g_io_err.c_str(); }

int jtp_snapshot_write(const char* path, const jtp_snap_meta* meta, const uint32_t* d_pos, const void* d_vel,
                       int32_t vel_f64, const int8_t* d_species, void* stream) {
    if (!path || !meta) return io_fail(JTP_ERR_ARG, "null path or metadata");
    int rc = check_format(meta->format);
    if (rc != JTP_OK) return rc;
    const int64_t n = meta->n;
    if (n < 0 || (meta->format == JTP_SNAP_PRODUCTION && n > (int64_t)0xffffffffu))
        return io_fail(JTP_ERR_ARG, "particle count out of range for the format");
    if (n > 0 && (!d_pos || !d_vel || !d_species)) return io_fail(JTP_ERR_ARG, "null device pointer");
    if (!(meta->box_size > 0) || !(meta->velocity_scale != 0)) return io_fail(JTP_ERR_ARG, "bad box size or velocity scale");
    hipStream_t s = (hipStream_t)stream;
    File f;
    f.f = std::fopen(path, "wb");
    if (!f.f) return io_fail(JTP_ERR_ARG, std::string("cannot create ") + path);
    if (meta->format == JTP_SNAP_PRODUCTION) {
        uint8_t h[kHeadProd];
        uint8_t* p = h;
        put_le<uint32_t>(p, (uint32_t)n);
        put_le<float>(p, (float)meta->box_size);
        put_le<uint32_t>(p, meta->step);
        put_le<float>(p, (float)meta->z);
        if (std::fwrite(h, 1, kHeadProd, f.f) != (size_t)kHeadProd) return io_fail(JTP_ERR_ARG, "write failed");
    } else {
        uint8_t h[kHeadV3];
        encode_v3_header(meta, h);
        if (std::fwrite(h, 1, kHeadV3, f.f) != (size_t)kHeadV3) return io_fail(JTP_ERR_ARG, "write failed");
    }
    const int rs = rec_size(meta->format);
    const int64_t chunk = n < kIoChunk ? n : kIoChunk;
    Staging st;
    if (chunk > 0) {
        IO_HIP(hipMalloc(&st.dev, (size_t)chunk * rs));
        IO_HIP(hipHostMalloc(&st.host, (size_t)chunk * rs, 0));
    }
    const float vs = (float)meta->velocity_scale;
    for (int64_t first = 0; first < n; first += chunk) {
        const int64_t m = n - first < chunk ? n - first : chunk;
        const unsigned nb = (unsigned)((m + 255) / 256);
        if (meta->format == JTP_SNAP_PRODUCTION)
            k_pack_prod<<<dim3(nb), dim3(256), 0, s>>>(d_pos, d_vel, vel_f64, d_species, first, m, meta->box_size, vs,
                                                       (uint8_t*)st.dev);
        else
            k_pack_v3<<<dim3(nb), dim3(256), 0, s>>>(d_pos, d_vel, vel_f64, d_species, first, m, meta->box_size, vs,
                                                     meta->mass_plus, meta->mass_minus, meta->eps_plus,
                                                     meta->eps_minus, (uint32_t*)st.dev);
        IO_HIP(hipGetLastError());
        IO_HIP(hipMemcpyAsync(st.host, st.dev, (size_t)m * rs, hipMemcpyDeviceToHost, s));
        IO_HIP(hipStreamSynchronize(s));
        if (std::fwrite(st.host, 1, (size_t)m * rs, f.f) != (size_t)m * rs) return io_fail(JTP_ERR_ARG, "write failed");
    }
    if (std::fflush(f.f) != 0) return io_fail(JTP_ERR_ARG, "write failed");
    return JTP_OK;
}

int jtp_snapshot_read_meta(const char* path, int32_t format, jtp_snap_meta* meta) {
    if (!path || !meta) return io_fail(JTP_ERR_ARG, "null path or metadata");
    int rc = check_format(format);
    if (rc != JTP_OK) return rc;
    jtp_snap_meta_default(meta, format, 0.0);
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return io_fail(JTP_ERR_ARG, std::string("cannot open ") + path);
    if (format == JTP_SNAP_PRODUCTION) {
        uint8_t h[kHeadProd];
        if (std::fread(h, 1, kHeadProd, f.f) != (size_t)kHeadProd) return io_fail(JTP_ERR_ARG, "short header");
        const uint8_t* p = h;
        meta->n = get_le<uint32_t>(p);
        meta->box_size = get_le<float>(p);
        meta->step = get_le<uint32_t>(p);
        meta->z = get_le<float>(p);
    } else {
        uint8_t h[kHeadV3];
        if (std::fread(h, 1, kHeadV3, f.f) != (size_t)kHeadV3) return io_fail(JTP_ERR_ARG, "short header");
        rc = decode_v3_header(h, meta);
        if (rc < 0) return rc;
    }
    return JTP_OK;
}

int jtp_snapshot_read(const char* path, int32_t format, int64_t n, uint32_t* d_pos, void* d_vel, int32_t vel_f64,
                      int8_t* d_species, double velocity_scale, void* stream) {
    if (!path) return io_fail(JTP_ERR_ARG, "null path");
    int rc = check_format(format);
    if (rc != JTP_OK) return rc;
    if (!(velocity_scale != 0)) return io_fail(JTP_ERR_ARG, "bad velocity scale");
    jtp_snap_meta meta;
    rc = jtp_snapshot_read_meta(path, format, &meta);
    if (rc != JTP_OK) return rc;
    if (n != meta.n) return io_fail(JTP_ERR_ARG, "n does not match the file's particle count " + std::to_string(meta.n));
    if (n > 0 && (!d_pos || !d_vel || !d_species)) return io_fail(JTP_ERR_ARG, "null device pointer");
    if (!(meta.box_size > 0)) return io_fail(JTP_ERR_ARG, "file holds a non-positive box size");
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return io_fail(JTP_ERR_ARG, std::string("cannot open ") + path);
    long skip = kHeadProd;
    if (format == JTP_SNAP_V3) {
        uint8_t h[kHeadV3];
        if (std::fread(h, 1, kHeadV3, f.f) != (size_t)kHeadV3) return io_fail(JTP_ERR_ARG, "short header");
        jtp_snap_meta tmp;
        const int hs = decode_v3_header(h, &tmp);
        if (hs < 0) return hs;
        skip = hs;  // header_size > 408: the extra bytes are skipped (snapshot_v3.rs:425-431)
    }
    if (std::fseek(f.f, skip, SEEK_SET) != 0) return io_fail(JTP_ERR_ARG, "seek failed");
    const int rs = rec_size(format);
    const int64_t chunk = n < kIoChunk ? n : kIoChunk;
    hipStream_t s = (hipStream_t)stream;
    Staging st;
    if (chunk > 0) {
        IO_HIP(hipMalloc(&st.dev, (size_t)chunk * rs));
        IO_HIP(hipHostMalloc(&st.host, (size_t)chunk * rs, 0));
        IO_HIP(hipMalloc(&st.bad, sizeof(int32_t)));
        IO_HIP(launch_fill_i32(st.bad, 1, 0, s));
    }
    const float inv_vs = (float)(1.0 / velocity_scale);
    for (int64_t first = 0; first < n; first += chunk) {
        const int64_t m = n - first < chunk ? n - first : chunk;
        IO_HIP(hipStreamSynchronize(s));  // the previous chunk's copy has left the pinned buffer
        if (std::fread(st.host, 1, (size_t)m * rs, f.f) != (size_t)m * rs)
            return io_fail(JTP_ERR_ARG, "file shorter than its particle count");
        IO_HIP(hipMemcpyAsync(st.dev, st.host, (size_t)m * rs, hipMemcpyHostToDevice, s));
        const unsigned nb = (unsigned)((m + 255) / 256);
        if (format == JTP_SNAP_PRODUCTION)
            k_unpack_prod<<<dim3(nb), dim3(256), 0, s>>>((const uint8_t*)st.dev, first, m, meta.box_size, inv_vs, d_pos,
                                                         d_vel, vel_f64, d_species, st.bad);
        else
            k_unpack_v3<<<dim3(nb), dim3(256), 0, s>>>((const uint32_t*)st.dev, first, m, meta.box_size, inv_vs, d_pos,
                                                       d_vel, vel_f64, d_species, st.bad);
        IO_HIP(hipGetLastError());
    }
    if (n > 0) {
        int32_t bad = 0;
        IO_HIP(hipMemcpyAsync(&bad, st.bad, sizeof(bad), hipMemcpyDeviceToHost, s));
        IO_HIP(hipStreamSynchronize(s));
        if (bad) return io_fail(JTP_ERR_ARG, std::to_string(bad) + " records carry an invalid sign byte");
    }
    return JTP_OK;
}

}  // extern "C"
