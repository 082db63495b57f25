// State storage, library context and the elementwise / reduction kernels of
// libqfe.  The state is one contiguous 2^n array of complex128 in HBM,
// interleaved (re, im), index bit q = qubit q: the layout of the reference
// Computer's coeff_ (computer.h:176), so host transfers are plain copies.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "qfe_device.h"
#include "qfe_internal.h"

namespace qfe {

namespace {
thread_local std::string g_last_error;
Context g_ctx;
std::mutex g_init_mutex;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? QFE_ERR_NOMEM : QFE_ERR_DEVICE;
}

Context& ctx() { return g_ctx; }

int ensure_init() {
    if (g_ctx.ready) return QFE_OK;
    return qfe_init(0);
}

int scratch_reserve(size_t bytes, void** out) {
    Context& c = g_ctx;
    if (c.scratch_bytes < bytes) {
        if (c.scratch) {
            QFE_HIP(hipStreamSynchronize(c.stream));
            QFE_HIP(hipFree(c.scratch));
            c.scratch = nullptr;
            c.scratch_bytes = 0;
        }
        size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
        QFE_HIP(hipMalloc(&c.scratch, want));
        c.scratch_bytes = want;
    }
    *out = c.scratch;
    return QFE_OK;
}

int pinned_reserve(size_t bytes, void** out) {
    Context& c = g_ctx;
    if (c.pinned_bytes < bytes) {
        if (c.pinned) {
            QFE_HIP(hipStreamSynchronize(c.stream));
            QFE_HIP(hipHostFree(c.pinned));
            c.pinned = nullptr;
            c.pinned_bytes = 0;
        }
        size_t want = bytes < (1u << 16) ? (1u << 16) : bytes;
        QFE_HIP(hipHostMalloc(&c.pinned, want, hipHostMallocDefault));
        c.pinned_bytes = want;
    }
    *out = c.pinned;
    return QFE_OK;
}

// Pinned, double-buffered host staging for small host-to-device uploads: the
// host fills one half while the copy out of the other may still be queued, so
// planning the next call overlaps the kernels of the previous one.
namespace {
struct Staging {
    void* buf[2] = {nullptr, nullptr};
    size_t cap[2] = {0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    int cur = 0;
};
Staging g_staging;
}  // namespace

int staging_begin(size_t bytes, void** host) {
    Staging& S = g_staging;
    const int h = S.cur;
    if (S.pending[h]) {
        QFE_HIP(hipEventSynchronize(S.ev[h]));
        S.pending[h] = false;
    }
    if (S.cap[h] < bytes) {
        if (S.buf[h]) QFE_HIP(hipHostFree(S.buf[h]));
        S.buf[h] = nullptr;
        S.cap[h] = 0;
        const size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
        QFE_HIP(hipHostMalloc(&S.buf[h], want, hipHostMallocDefault));
        S.cap[h] = want;
    }
    if (!S.ev[h]) QFE_HIP(hipEventCreateWithFlags(&S.ev[h], hipEventDisableTiming));
    *host = S.buf[h];
    return QFE_OK;
}

int staging_commit(void* dev, size_t bytes) {
    Staging& S = g_staging;
    const int h = S.cur;
    QFE_HIP(hipMemcpyAsync(dev, S.buf[h], bytes, hipMemcpyHostToDevice, g_ctx.stream));
    QFE_HIP(hipEventRecord(S.ev[h], g_ctx.stream));
    S.pending[h] = true;
    S.cur ^= 1;
    return QFE_OK;
}

// ---- elementwise kernels ---------------------------------------------------------

// Streaming elementwise kernels: kStreamItems amplitudes per thread, each
// instruction of a wave covering 1 KiB contiguous, workgroups spread over the
// XCDs in contiguous eighths; non-temporal for states beyond the caches.
constexpr int kStreamTPB = 256;
constexpr int kStreamItems = 4;

template <bool NT>
__global__ __launch_bounds__(kStreamTPB) void k_fill_zero(double2* __restrict__ d, uint64_t N) {
    const uint64_t base = xcd_block() * (kStreamTPB * kStreamItems) + threadIdx.x;
#pragma unroll
    for (int u = 0; u < kStreamItems; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamTPB;
        if (i < N) st_s<NT>(d + i, make_double2(0.0, 0.0));
    }
}

__global__ void k_set_one(double2* __restrict__ d, uint64_t idx) {
    if (threadIdx.x == 0 && blockIdx.x == 0) d[idx] = make_double2(1.0, 0.0);
}

__global__ void k_scatter(double2* __restrict__ d, const uint64_t* __restrict__ idx, const double2* __restrict__ v,
                          uint64_t count) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) d[idx[i]] = v[i];
}

template <bool NT>
__global__ __launch_bounds__(kStreamTPB) void k_scale(double2* __restrict__ d, uint64_t N, double2 a) {
    const uint64_t base = xcd_block() * (kStreamTPB * kStreamItems) + threadIdx.x;
    double2 v[kStreamItems];
#pragma unroll
    for (int u = 0; u < kStreamItems; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamTPB;
        if (i < N) v[u] = ld_s<NT>(d + i);
    }
#pragma unroll
    for (int u = 0; u < kStreamItems; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamTPB;
        if (i < N) st_s<NT>(d + i, cmul(a, v[u]));
    }
}

template <bool NT>
__global__ __launch_bounds__(kStreamTPB) void k_axpy(double2* __restrict__ y, const double2* __restrict__ x, uint64_t N,
                                                     double2 a) {
    const uint64_t base = xcd_block() * (kStreamTPB * kStreamItems) + threadIdx.x;
    double2 v[kStreamItems], w[kStreamItems];
#pragma unroll
    for (int u = 0; u < kStreamItems; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamTPB;
        if (i < N) {
            v[u] = ld_s<NT>(y + i);
            w[u] = ld_s<NT>(x + i);
        }
    }
#pragma unroll
    for (int u = 0; u < kStreamItems; ++u) {
        const uint64_t i = base + (uint64_t)u * kStreamTPB;
        if (i < N) st_s<NT>(y + i, cadd(v[u], cmul(a, w[u])));
    }
}

static unsigned stream_grid(uint64_t N) {
    const uint64_t g = (N + kStreamTPB * kStreamItems - 1) / (kStreamTPB * kStreamItems);
    return (unsigned)(g ? g : 1);
}
static bool stream_nt(uint64_t N) { return N && use_nt(63 - __builtin_clzll(N), 63u); }

// Deterministic two-stage reductions.  Stage 1: block b reduces the fixed
// contiguous slice [b*chunk, (b+1)*chunk) in a fixed order; stage 2: one block
// reduces the partials in index order.
constexpr int kRedBlock = 256;

__global__ __launch_bounds__(kRedBlock) void k_dot_stage1(const double2* __restrict__ a, const double2* __restrict__ b,
                                                          uint64_t N, uint64_t chunk, double2* __restrict__ partial) {
    __shared__ double2 lds[16];
    const uint64_t begin = (uint64_t)blockIdx.x * chunk;
    uint64_t end = begin + chunk;
    if (end > N) end = N;
    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t i = begin + threadIdx.x; i < end; i += kRedBlock) acc = cadd(acc, cconjmul(a[i], b[i]));
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kRedBlock) void k_norm_stage1(const double2* __restrict__ a, uint64_t N, uint64_t chunk,
                                                           double2* __restrict__ partial) {
    __shared__ double2 lds[16];
    const uint64_t begin = (uint64_t)blockIdx.x * chunk;
    uint64_t end = begin + chunk;
    if (end > N) end = N;
    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t i = begin + threadIdx.x; i < end; i += kRedBlock) {
        const double2 v = a[i];
        acc.x += v.x * v.x + v.y * v.y;
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kRedBlock) void k_neq_stage1(const double2* __restrict__ a, const double2* __restrict__ b,
                                                          uint64_t N, uint64_t chunk, double2* __restrict__ partial) {
    __shared__ double2 lds[16];
    const uint64_t begin = (uint64_t)blockIdx.x * chunk;
    uint64_t end = begin + chunk;
    if (end > N) end = N;
    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t i = begin + threadIdx.x; i < end; i += kRedBlock) {
        const double2 u = a[i], v = b[i];
        if (u.x != v.x || u.y != v.y) acc.x += 1.0;
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// Sums `count` rows of partials: out[r] = sum_b partial[r * stride + b], b < nb.
__global__ __launch_bounds__(kRedBlock) void k_sum_rows(const double2* __restrict__ partial, uint64_t stride,
                                                        uint32_t nb, double2* __restrict__ out) {
    __shared__ double2 lds[16];
    const double2* row = partial + (uint64_t)blockIdx.x * stride;
    double2 acc = make_double2(0.0, 0.0);
    for (uint32_t i = threadIdx.x; i < nb; i += kRedBlock) acc = cadd(acc, row[i]);
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

static unsigned grid_for(uint64_t N, int block = 256, uint64_t cap = 65536) {
    uint64_t g = (N + block - 1) / block;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

int launch_zero(double2* d, uint64_t N, hipStream_t st) {
    if (stream_nt(N))
        hipLaunchKernelGGL(k_fill_zero<true>, dim3(stream_grid(N)), dim3(kStreamTPB), 0, st, d, N);
    else
        hipLaunchKernelGGL(k_fill_zero<false>, dim3(stream_grid(N)), dim3(kStreamTPB), 0, st, d, N);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

int launch_basis(double2* d, uint64_t N, uint64_t idx, hipStream_t st) {
    int rc = launch_zero(d, N, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_set_one, dim3(1), dim3(64), 0, st, d, idx);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

// Reduction driver shared by dot / norm / equality.
enum RedKind { R_DOT, R_NORM, R_NEQ };

static int reduce_pair(RedKind kind, const double2* a, const double2* b, uint64_t N, double2* result) {
    Context& c = ctx();
    uint64_t nb = (N + 4095) / 4096;  // ~4096 amplitudes per block
    if (nb > 2048) nb = 2048;
    if (nb == 0) nb = 1;
    const uint64_t chunk = (N + nb - 1) / nb;
    void* scr = nullptr;
    int rc = scratch_reserve((nb + 1) * sizeof(double2), &scr);
    if (rc) return rc;
    double2* partial = static_cast<double2*>(scr);
    double2* out = partial + nb;
    switch (kind) {
        case R_DOT:
            hipLaunchKernelGGL(k_dot_stage1, dim3((unsigned)nb), dim3(kRedBlock), 0, c.stream, a, b, N, chunk, partial);
            break;
        case R_NORM:
            hipLaunchKernelGGL(k_norm_stage1, dim3((unsigned)nb), dim3(kRedBlock), 0, c.stream, a, N, chunk, partial);
            break;
        case R_NEQ:
            hipLaunchKernelGGL(k_neq_stage1, dim3((unsigned)nb), dim3(kRedBlock), 0, c.stream, a, b, N, chunk, partial);
            break;
    }
    QFE_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(kRedBlock), 0, c.stream, partial, (uint64_t)nb, (uint32_t)nb, out);
    QFE_HIP(hipGetLastError());
    void* pin = nullptr;
    rc = pinned_reserve(sizeof(double2), &pin);
    if (rc) return rc;
    QFE_HIP(hipMemcpyAsync(pin, out, sizeof(double2), hipMemcpyDeviceToHost, c.stream));
    QFE_HIP(hipStreamSynchronize(c.stream));
    std::memcpy(result, pin, sizeof(double2));
    c.counters.reduce++;
    c.counters.bytes += (kind == R_NORM ? 16.0 : 32.0) * (double)N;
    return QFE_OK;
}

static bool valid(qfe_state s) { return s != nullptr && s->d != nullptr; }

}  // namespace qfe

using namespace qfe;

extern "C" {

const char* qfe_last_error(void) { return g_last_error.c_str(); }

int qfe_device_count(int* count) {
    if (!count) return fail(QFE_ERR_ARG, "qfe_device_count: null pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return QFE_OK;
}

int qfe_init(int device) {
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (g_ctx.ready) {
        if (g_ctx.device != device) return fail(QFE_ERR_STATE, "qfe_init: library already bound to another device");
        return QFE_OK;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(QFE_ERR_DEVICE, "qfe_init: no GPU visible (the engine needs an MI355X)");
    if (device < 0 || device >= n) return fail(QFE_ERR_ARG, "qfe_init: device index out of range");
    QFE_HIP(hipSetDevice(device));
    QFE_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    int rc = noalloc_scratch_setup(device);
    if (rc) return rc;
    g_ctx.device = device;
    g_ctx.ready = true;
    return QFE_OK;
}

int qfe_get_stream(void** stream) {
    if (!stream) return fail(QFE_ERR_ARG, "qfe_get_stream: null pointer");
    int rc = ensure_init();
    if (rc) return rc;
    *stream = (void*)g_ctx.stream;
    return QFE_OK;
}

int qfe_synchronize(void) {
    int rc = ensure_init();
    if (rc) return rc;
    QFE_HIP(hipStreamSynchronize(g_ctx.stream));
    return QFE_OK;
}

int qfe_set_tuning(int tile_qubits, int low_qubits, int max_ops_per_pass) {
    Tuning t = g_ctx.tuning;
    if (tile_qubits == -1 || low_qubits == -1 || max_ops_per_pass == -1) {
        // the default: the tile shape follows the state size
        const Tuning d;
        t.tile_auto = true;
        t.tile_qubits = d.tile_qubits;
        t.low_qubits = d.low_qubits;
        t.max_ops = d.max_ops;
        g_ctx.tuning = t;
        return QFE_OK;
    }
    if (!tile_qubits && !low_qubits && !max_ops_per_pass) return QFE_OK;
    t.tile_auto = false;
    if (tile_qubits) t.tile_qubits = tile_qubits;
    if (low_qubits) t.low_qubits = low_qubits;
    if (max_ops_per_pass) t.max_ops = max_ops_per_pass;
    if (t.tile_qubits < 2 || t.tile_qubits > 13) return fail(QFE_ERR_ARG, "qfe_set_tuning: tile_qubits must be in [2, 13]");
    if (t.low_qubits < 0 || t.low_qubits > t.tile_qubits) return fail(QFE_ERR_ARG, "qfe_set_tuning: low_qubits must be in [0, tile_qubits]");
    if (t.max_ops < 1 || t.max_ops > 4096) return fail(QFE_ERR_ARG, "qfe_set_tuning: max_ops_per_pass must be in [1, 4096]");
    g_ctx.tuning = t;
    return QFE_OK;
}

int qfe_get_tuning(int* tile_qubits, int* low_qubits, int* max_ops_per_pass) {
    const bool a = g_ctx.tuning.tile_auto;
    if (tile_qubits) *tile_qubits = a ? -1 : g_ctx.tuning.tile_qubits;
    if (low_qubits) *low_qubits = a ? -1 : g_ctx.tuning.low_qubits;
    if (max_ops_per_pass) *max_ops_per_pass = a ? -1 : g_ctx.tuning.max_ops;
    return QFE_OK;
}

int qfe_set_sector_word_limit(uint64_t max_pair_words) {
    g_ctx.tuning.sector_words = max_pair_words;
    return QFE_OK;
}

int qfe_set_pass_tuning(int orbit_max_dims) {
    if (orbit_max_dims < 1 || orbit_max_dims > 7) return fail(QFE_ERR_ARG, "qfe_set_pass_tuning: orbit_max_dims must be in [1, 7]");
    g_ctx.tuning.orbit_dims = orbit_max_dims;
    return QFE_OK;
}

int qfe_get_pass_tuning(int* orbit_max_dims) {
    if (orbit_max_dims) *orbit_max_dims = g_ctx.tuning.orbit_dims;
    return QFE_OK;
}

int qfe_event_create(void** event) {
    if (!event) return fail(QFE_ERR_ARG, "qfe_event_create: null pointer");
    int rc = ensure_init();
    if (rc) return rc;
    hipEvent_t e;
    QFE_HIP(hipEventCreate(&e));
    *event = (void*)e;
    return QFE_OK;
}

int qfe_event_record(void* event) {
    if (!event) return fail(QFE_ERR_ARG, "qfe_event_record: null event");
    QFE_HIP(hipEventRecord((hipEvent_t)event, g_ctx.stream));
    return QFE_OK;
}

int qfe_event_elapsed_ms(void* start, void* stop, float* ms) {
    if (!start || !stop || !ms) return fail(QFE_ERR_ARG, "qfe_event_elapsed_ms: invalid argument");
    QFE_HIP(hipEventSynchronize((hipEvent_t)stop));
    QFE_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return QFE_OK;
}

int qfe_event_destroy(void* event) {
    if (!event) return QFE_OK;
    QFE_HIP(hipEventDestroy((hipEvent_t)event));
    return QFE_OK;
}

int qfe_get_counters(uint64_t* fused_passes, uint64_t* single_passes, uint64_t* reduce_passes, double* bytes) {
    if (fused_passes) *fused_passes = g_ctx.counters.fused;
    if (single_passes) *single_passes = g_ctx.counters.single;
    if (reduce_passes) *reduce_passes = g_ctx.counters.reduce;
    if (bytes) *bytes = g_ctx.counters.bytes;
    return QFE_OK;
}

int qfe_get_pass_counts(uint64_t* out, uint64_t count) {
    if (count && !out) return fail(QFE_ERR_ARG, "qfe_get_pass_counts: null output");
    const Counters& c = g_ctx.counters;
    const uint64_t v[7] = {c.fused, c.single, c.reduce, c.tile, c.orbit, c.lane_orbit, c.nt};
    for (uint64_t i = 0; i < count && i < 7; ++i) out[i] = v[i];
    return QFE_OK;
}

int qfe_reset_counters(void) {
    g_ctx.counters = Counters{};
    return QFE_OK;
}

int qfe_state_create(int nqubit, qfe_state* out) {
    if (!out) return fail(QFE_ERR_ARG, "qfe_state_create: null output pointer");
    *out = nullptr;
    if (nqubit < 1 || nqubit > 36) return fail(QFE_ERR_ARG, "qfe_state_create: nqubit must be in [1, 36]");
    int rc = ensure_init();
    if (rc) return rc;
    qfe_state_s* s = new qfe_state_s();
    s->n = nqubit;
    s->N = 1ull << nqubit;
    hipError_t e = hipMalloc(&s->d, s->N * sizeof(double2));
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "hipMalloc(state)");
    }
    rc = launch_basis(s->d, s->N, 0, g_ctx.stream);
    if (rc) {
        (void)hipFree(s->d);
        delete s;
        return rc;
    }
    *out = s;
    return QFE_OK;
}

int qfe_state_destroy(qfe_state s) {
    if (!s) return QFE_OK;
    if (s->d) {
        // queued kernels may still use the buffer
        hipError_t e = hipStreamSynchronize(g_ctx.stream);
        hipError_t f = hipFree(s->d);
        s->d = nullptr;
        delete s;
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        if (f != hipSuccess) return hip_fail(f, "hipFree(state)");
        return QFE_OK;
    }
    delete s;
    return QFE_OK;
}

int qfe_state_nqubit(qfe_state s, int* nqubit) {
    if (!valid(s) || !nqubit) return fail(QFE_ERR_ARG, "qfe_state_nqubit: invalid argument");
    *nqubit = s->n;
    return QFE_OK;
}

int qfe_state_copy(qfe_state dst, qfe_state src) {
    if (!valid(dst) || !valid(src)) return fail(QFE_ERR_ARG, "qfe_state_copy: invalid handle");
    if (dst->n != src->n) return fail(QFE_ERR_ARG, "qfe_state_copy: qubit counts differ");
    if (dst == src) return QFE_OK;
    QFE_HIP(hipMemcpyAsync(dst->d, src->d, src->N * sizeof(double2), hipMemcpyDeviceToDevice, g_ctx.stream));
    return QFE_OK;
}

int qfe_state_upload(qfe_state s, const qfe_c64* host, uint64_t namps) {
    if (!valid(s) || (!host && namps)) return fail(QFE_ERR_ARG, "qfe_state_upload: invalid argument");
    if (namps != s->N) return fail(QFE_ERR_ARG, "qfe_state_upload: expected 2^n amplitudes");
    QFE_HIP(hipMemcpyAsync(s->d, host, namps * sizeof(double2), hipMemcpyHostToDevice, g_ctx.stream));
    QFE_HIP(hipStreamSynchronize(g_ctx.stream));
    return QFE_OK;
}

int qfe_state_download(qfe_state s, qfe_c64* host, uint64_t namps) {
    if (!valid(s) || (!host && namps)) return fail(QFE_ERR_ARG, "qfe_state_download: invalid argument");
    if (namps != s->N) return fail(QFE_ERR_ARG, "qfe_state_download: expected 2^n amplitudes");
    QFE_HIP(hipMemcpyAsync(host, s->d, namps * sizeof(double2), hipMemcpyDeviceToHost, g_ctx.stream));
    QFE_HIP(hipStreamSynchronize(g_ctx.stream));
    return QFE_OK;
}

int qfe_state_reset(qfe_state s) {
    if (!valid(s)) return fail(QFE_ERR_ARG, "qfe_state_reset: invalid handle");
    return launch_basis(s->d, s->N, 0, g_ctx.stream);
}

int qfe_state_zero(qfe_state s) {
    if (!valid(s)) return fail(QFE_ERR_ARG, "qfe_state_zero: invalid handle");
    return launch_zero(s->d, s->N, g_ctx.stream);
}

int qfe_state_set_amps(qfe_state s, const uint64_t* idx, const qfe_c64* val, uint64_t count) {
    if (!valid(s) || (count && (!idx || !val))) return fail(QFE_ERR_ARG, "qfe_state_set_amps: invalid argument");
    for (uint64_t i = 0; i < count; ++i)
        if (idx[i] >= s->N) return fail(QFE_ERR_ARG, "qfe_state_set_amps: basis index out of range");
    // Later duplicates win, as the reference's sequential assignment does
    // (computer.cc:42-47): keep the last value per index before scattering.
    std::vector<uint64_t> uidx;
    std::vector<qfe_c64> uval;
    {
        std::unordered_map<uint64_t, size_t> pos;
        for (uint64_t i = 0; i < count; ++i) {
            auto it = pos.find(idx[i]);
            if (it == pos.end()) {
                pos.emplace(idx[i], uidx.size());
                uidx.push_back(idx[i]);
                uval.push_back(val[i]);
            } else {
                uval[it->second] = val[i];
            }
        }
    }
    idx = uidx.data();
    val = uval.data();
    count = uidx.size();
    int rc = launch_zero(s->d, s->N, g_ctx.stream);
    if (rc || count == 0) return rc;
    void* scr = nullptr;
    const size_t bytes = count * (sizeof(uint64_t) + sizeof(double2));
    rc = scratch_reserve(bytes, &scr);
    if (rc) return rc;
    uint64_t* didx = static_cast<uint64_t*>(scr);
    double2* dval = reinterpret_cast<double2*>(didx + count);
    QFE_HIP(hipMemcpyAsync(didx, idx, count * sizeof(uint64_t), hipMemcpyHostToDevice, g_ctx.stream));
    QFE_HIP(hipMemcpyAsync(dval, val, count * sizeof(double2), hipMemcpyHostToDevice, g_ctx.stream));
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, g_ctx.stream, s->d, didx, dval,
                       count);
    QFE_HIP(hipGetLastError());
    QFE_HIP(hipStreamSynchronize(g_ctx.stream));
    return QFE_OK;
}

int qfe_state_get_amp(qfe_state s, uint64_t idx, qfe_c64* out) {
    if (!valid(s) || !out) return fail(QFE_ERR_ARG, "qfe_state_get_amp: invalid argument");
    if (idx >= s->N) return fail(QFE_ERR_ARG, "qfe_state_get_amp: basis index out of range");
    void* pin = nullptr;
    int rc = pinned_reserve(sizeof(double2), &pin);
    if (rc) return rc;
    QFE_HIP(hipMemcpyAsync(pin, s->d + idx, sizeof(double2), hipMemcpyDeviceToHost, g_ctx.stream));
    QFE_HIP(hipStreamSynchronize(g_ctx.stream));
    std::memcpy(out, pin, sizeof(double2));
    return QFE_OK;
}

int qfe_state_scale(qfe_state s, qfe_c64 a) {
    if (!valid(s)) return fail(QFE_ERR_ARG, "qfe_state_scale: invalid handle");
    if (stream_nt(s->N))
        hipLaunchKernelGGL(k_scale<true>, dim3(stream_grid(s->N)), dim3(kStreamTPB), 0, g_ctx.stream, s->d, s->N,
                           make_double2(a.re, a.im));
    else
        hipLaunchKernelGGL(k_scale<false>, dim3(stream_grid(s->N)), dim3(kStreamTPB), 0, g_ctx.stream, s->d, s->N,
                           make_double2(a.re, a.im));
    QFE_HIP(hipGetLastError());
    g_ctx.counters.single++;
    g_ctx.counters.bytes += 32.0 * (double)s->N;
    return QFE_OK;
}

int qfe_state_axpy(qfe_state y, qfe_c64 a, qfe_state x) {
    if (!valid(y) || !valid(x)) return fail(QFE_ERR_ARG, "qfe_state_axpy: invalid handle");
    if (y->n != x->n) return fail(QFE_ERR_ARG, "qfe_state_axpy: qubit counts differ");
    if (stream_nt(y->N))
        hipLaunchKernelGGL(k_axpy<true>, dim3(stream_grid(y->N)), dim3(kStreamTPB), 0, g_ctx.stream, y->d, x->d, y->N,
                           make_double2(a.re, a.im));
    else
        hipLaunchKernelGGL(k_axpy<false>, dim3(stream_grid(y->N)), dim3(kStreamTPB), 0, g_ctx.stream, y->d, x->d, y->N,
                           make_double2(a.re, a.im));
    QFE_HIP(hipGetLastError());
    g_ctx.counters.single++;
    g_ctx.counters.bytes += 48.0 * (double)y->N;
    return QFE_OK;
}

int qfe_state_dot(qfe_state a, qfe_state b, qfe_c64* out) {
    if (!valid(a) || !valid(b) || !out) return fail(QFE_ERR_ARG, "qfe_state_dot: invalid argument");
    if (a->n != b->n) return fail(QFE_ERR_ARG, "qfe_state_dot: qubit counts differ");
    double2 r;
    int rc = reduce_pair(R_DOT, a->d, b->d, a->N, &r);
    if (rc) return rc;
    out->re = r.x;
    out->im = r.y;
    return QFE_OK;
}

int qfe_state_norm2(qfe_state s, double* out) {
    if (!valid(s) || !out) return fail(QFE_ERR_ARG, "qfe_state_norm2: invalid argument");
    double2 r;
    int rc = reduce_pair(R_NORM, s->d, s->d, s->N, &r);
    if (rc) return rc;
    *out = r.x;
    return QFE_OK;
}

int qfe_state_equal(qfe_state a, qfe_state b, int* equal) {
    if (!valid(a) || !valid(b) || !equal) return fail(QFE_ERR_ARG, "qfe_state_equal: invalid argument");
    if (a->n != b->n) {
        *equal = 0;
        return QFE_OK;
    }
    double2 r;
    int rc = reduce_pair(R_NEQ, a->d, b->d, a->N, &r);
    if (rc) return rc;
    *equal = (r.x == 0.0) ? 1 : 0;
    return QFE_OK;
}

}  // extern "C"
