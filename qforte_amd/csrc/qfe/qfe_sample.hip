// Computational-basis sampling from |psi|^2 on the device: the probability
// scan and discrete_distribution draw of Computer::measure_circuit,
// measure_z_readouts_fast and measure_readouts (computer.cc:166-188,
// 195-210, 247-262).
//
// Two passes and a draw:
//  1. k_sample_scan: one workgroup per chunk of kChunkSubs sub-blocks of 64
//     amplitudes.  Each wave reduces whole sub-blocks (one amplitude per lane,
//     coalesced 1 KiB reads), the workgroup writes the inclusive prefix of its
//     sub-block masses and the chunk mass.  This is the only pass over the
//     state (16 bytes per amplitude, HBM-bound).
//  2. k_sample_chunks: one workgroup turns the chunk masses into an inclusive
//     prefix (2^(n-14) values).
//  3. k_sample_guide: a guide table over the chunk prefix (one entry per chunk).
//  4. k_sample_draw: one wave per sample.  u = U[0,1) * total from a
//     counter-based generator (SplitMix64, Steele, Lea and Flood 2014, applied
//     to seed + (i + 1) * golden ratio), then a binary search over the few
//     chunks the guide table leaves, one over the chunk's sub-block prefix and
//     a wave scan over the sub-block's 64 probabilities.
// The first position whose prefix exceeds u always has positive mass; if
// rounding leaves u at or above a level's total, the last position with
// positive mass at that level is taken.

#include "qfe_device.h"
#include "qfe_internal.h"

namespace qfe {
namespace {

constexpr uint32_t kSubLen = 64;      // amplitudes per sub-block: one wave
constexpr uint32_t kChunkSubs = 256;  // sub-blocks per chunk
constexpr int kScanThreads = 256;     // 4 waves per chunk
constexpr int kScanUnroll = 4;        // sub-blocks in flight per wave
constexpr int kChunkThreads = 1024;
constexpr int kDrawWaves = 4;  // samples per workgroup

struct SampleShape {
    uint64_t N;
    uint32_t sub_len;     // min(N, 64)
    uint64_t nsub;        // N / sub_len
    uint32_t chunk_subs;  // min(nsub, 256)
    uint64_t nchunk;      // nsub / chunk_subs
};

SampleShape sample_shape(uint64_t N) {
    SampleShape s{};
    s.N = N;
    s.sub_len = (uint32_t)std::min<uint64_t>(N, kSubLen);
    s.nsub = N / s.sub_len;
    s.chunk_subs = (uint32_t)std::min<uint64_t>(s.nsub, kChunkSubs);
    s.nchunk = s.nsub / s.chunk_subs;
    return s;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Inclusive prefix over the 64 lanes of a wave (Hillis-Steele).
__device__ __forceinline__ double wave_scan_d(double v, uint32_t lane) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const double o = __shfl_up(v, off, kWave);
        if (lane >= (uint32_t)off) v += o;
    }
    return v;
}

__global__ __launch_bounds__(kScanThreads) void k_sample_scan(const double2* __restrict__ psi, SampleShape sh,
                                                               double* __restrict__ sub_prefix,
                                                               double* __restrict__ chunk_mass) {
    __shared__ double mass[kChunkSubs];
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const uint64_t chunk = blockIdx.x;
    const uint64_t sub0 = chunk * sh.chunk_subs;
    constexpr uint32_t kWaves = kScanThreads / kWave;
    for (uint32_t s = wave * kScanUnroll; s < sh.chunk_subs; s += kWaves * kScanUnroll) {
        double p[kScanUnroll];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) {
            p[u] = 0.0;
            const uint32_t si = s + u;
            if (si < sh.chunk_subs && lane < sh.sub_len) {
                const double2 a = psi[(sub0 + si) * sh.sub_len + lane];
                p[u] = a.x * a.x + a.y * a.y;
            }
        }
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) {
            const double t = wave_sum_d(p[u]);
            if (lane == 0 && s + u < sh.chunk_subs) mass[s + u] = t;
        }
    }
    __syncthreads();
    if (wave == 0) {
        // inclusive prefix of the chunk's sub-block masses, 64 at a time
        double carry = 0.0;
        for (uint32_t b = 0; b < sh.chunk_subs; b += kWave) {
            const double v = b + lane < sh.chunk_subs ? mass[b + lane] : 0.0;
            const double pre = wave_scan_d(v, lane) + carry;
            if (b + lane < sh.chunk_subs) sub_prefix[sub0 + b + lane] = pre;
            carry = __shfl(pre, kWave - 1, kWave);
        }
        if (lane == 0) chunk_mass[chunk] = carry;
    }
}

// In-place inclusive prefix of the chunk masses (one workgroup).
__global__ __launch_bounds__(kChunkThreads) void k_sample_chunks(double* __restrict__ v, uint64_t n) {
    __shared__ double part[kChunkThreads / kWave];
    const uint32_t tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    const uint64_t per = (n + kChunkThreads - 1) / kChunkThreads;
    const uint64_t b = std::min<uint64_t>(n, tid * per), e = std::min<uint64_t>(n, b + per);
    double own = 0.0;
    for (uint64_t i = b; i < e; ++i) own += v[i];
    // exclusive offset of this thread's segment: wave scan, then across waves
    const double incl = wave_scan_d(own, lane);
    if (lane == kWave - 1) part[wave] = incl;
    __syncthreads();
    double off = incl - own;
    for (uint32_t w = 0; w < wave; ++w) off += part[w];
    for (uint64_t i = b; i < e; ++i) {
        off += v[i];
        v[i] = off;
    }
}

__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

// First index in [lo, hi] of a non-decreasing prefix array whose value
// exceeds u; if none does, the first index that reaches the last value (the
// last position with positive mass).  The answer must lie in [lo, hi].
__device__ __forceinline__ uint64_t prefix_find_in(const double* __restrict__ v, uint64_t n, double u, uint64_t lo,
                                                   uint64_t hi) {
    const double top = v[n - 1];
    const bool over = !(u < top);
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (over ? !(v[mid] < top) : v[mid] > u)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}
__device__ __forceinline__ uint64_t prefix_find(const double* __restrict__ v, uint64_t n, double u) {
    return prefix_find_in(v, n, u, 0, n - 1);
}

// Guide table of the chunk prefix (indexed search, Chen and Asau 1974):
// guide[b] = first chunk whose prefix exceeds b * total / nguide.  A draw u
// with b = floor(u * nguide / total) has its chunk in [guide[b - 1],
// guide[b + 2]] even when that floor is off by one in rounding.
__global__ __launch_bounds__(256) void k_sample_guide(const double* __restrict__ chunk_prefix, uint64_t nchunk,
                                                      uint32_t* __restrict__ guide, uint32_t nguide) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nguide) return;
    const double target = chunk_prefix[nchunk - 1] * ((double)b / (double)nguide);
    uint64_t lo = 0, hi = nchunk - 1;
    while (lo < hi) {  // first c with prefix[c] > target (nchunk - 1 if none)
        const uint64_t mid = (lo + hi) >> 1;
        if (chunk_prefix[mid] > target)
            hi = mid;
        else
            lo = mid + 1;
    }
    guide[b] = (uint32_t)lo;
}

__global__ __launch_bounds__(kDrawWaves * kWave) void k_sample_draw(const double2* __restrict__ psi, SampleShape sh,
                                                                     const double* __restrict__ sub_prefix,
                                                                     const double* __restrict__ chunk_prefix,
                                                                     const uint32_t* __restrict__ guide, uint32_t nguide,
                                                                     uint64_t nsamples, uint64_t seed,
                                                                     uint64_t* __restrict__ out) {
    const uint32_t lane = threadIdx.x % kWave;
    const uint64_t i = (uint64_t)blockIdx.x * kDrawWaves + threadIdx.x / kWave;
    if (i >= nsamples) return;  // whole waves leave together
    const double r = uniform01(seed, i), total = chunk_prefix[sh.nchunk - 1];
    double u = r * total;
    const uint32_t b = min((uint32_t)(r * (double)nguide), nguide - 1);
    const uint64_t glo = guide[b > 0 ? b - 1 : 0];
    const uint64_t ghi = b + 2 < nguide ? guide[b + 2] : sh.nchunk - 1;
    const uint64_t c = u < total ? prefix_find_in(chunk_prefix, sh.nchunk, u, glo, ghi) : prefix_find(chunk_prefix, sh.nchunk, u);
    if (c) u -= chunk_prefix[c - 1];
    const double* sp = sub_prefix + c * sh.chunk_subs;
    const uint64_t s = prefix_find(sp, sh.chunk_subs, u);
    if (s) u -= sp[s - 1];
    const uint64_t base = (c * sh.chunk_subs + s) * sh.sub_len;
    double p = 0.0;
    if (lane < sh.sub_len) {
        const double2 a = psi[base + lane];
        p = a.x * a.x + a.y * a.y;
    }
    const double pre = wave_scan_d(p, lane);
    const uint64_t hit = __ballot(lane < sh.sub_len && pre > u);
    uint32_t pick;
    if (hit) {
        pick = (uint32_t)__builtin_ctzll(hit);
    } else {
        const uint64_t pos = __ballot(p > 0.0);
        pick = pos ? 63u - (uint32_t)__builtin_clzll(pos) : 0u;
    }
    if (lane == 0) out[i] = base + pick;
}

}  // namespace
}  // namespace qfe

using namespace qfe;

int qfe_sample(qfe_state s, uint64_t nsamples, uint64_t seed, uint64_t* out) {
    if (!s || !s->d || (nsamples && !out)) return fail(QFE_ERR_ARG, "qfe_sample: invalid argument");
    if (!nsamples) return QFE_OK;
    if (nsamples > (1ull << 32)) return fail(QFE_ERR_ARG, "qfe_sample: at most 2^32 samples per call");
    int rc = ensure_init();
    if (rc) return rc;
    const SampleShape sh = sample_shape(s->N);
    const uint32_t nguide = (uint32_t)sh.nchunk;
    const size_t o_chunk = (size_t)sh.nsub * 8;
    const size_t o_guide = o_chunk + ((size_t)sh.nchunk * 8 + 255) / 256 * 256;
    const size_t o_out = o_guide + ((size_t)nguide * 4 + 255) / 256 * 256;
    void* scratch = nullptr;
    rc = scratch_reserve(o_out + (size_t)nsamples * 8, &scratch);
    if (rc) return rc;
    double* sub_prefix = (double*)scratch;
    double* chunk_prefix = (double*)((char*)scratch + o_chunk);
    uint32_t* guide = (uint32_t*)((char*)scratch + o_guide);
    uint64_t* dout = (uint64_t*)((char*)scratch + o_out);
    hipStream_t st = ctx().stream;
    hipLaunchKernelGGL(k_sample_scan, dim3((unsigned)sh.nchunk), dim3(kScanThreads), 0, st, s->d, sh, sub_prefix,
                       chunk_prefix);
    QFE_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sample_chunks, dim3(1), dim3(kChunkThreads), 0, st, chunk_prefix, sh.nchunk);
    QFE_HIP(hipGetLastError());
    double total = 0.0;
    QFE_HIP(hipMemcpyAsync(&total, chunk_prefix + sh.nchunk - 1, 8, hipMemcpyDeviceToHost, st));
    QFE_HIP(hipStreamSynchronize(st));
    if (!(total > 0.0)) return fail(QFE_ERR_ARG, "qfe_sample: the state has zero norm");
    hipLaunchKernelGGL(k_sample_guide, dim3((nguide + 255) / 256), dim3(256), 0, st, chunk_prefix, sh.nchunk, guide, nguide);
    QFE_HIP(hipGetLastError());
    const uint64_t blocks = (nsamples + kDrawWaves - 1) / kDrawWaves;
    hipLaunchKernelGGL(k_sample_draw, dim3((unsigned)blocks), dim3(kDrawWaves * kWave), 0, st, s->d, sh, sub_prefix,
                       chunk_prefix, guide, nguide, nsamples, seed, dout);
    QFE_HIP(hipGetLastError());
    QFE_HIP(hipMemcpyAsync(out, dout, nsamples * 8, hipMemcpyDeviceToHost, st));
    QFE_HIP(hipStreamSynchronize(st));
    return QFE_OK;
}
