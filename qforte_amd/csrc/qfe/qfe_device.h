// Device helpers shared by the kernels of libqfe (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace qfe {

constexpr int kWave = 64;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// a*x + b*y
__device__ __forceinline__ double2 cmix(double2 a, double2 x, double2 b, double2 y) {
    return make_double2(a.x * x.x - a.y * x.y + b.x * y.x - b.y * y.y, a.x * x.y + a.y * x.x + b.x * y.y + b.y * y.x);
}

// Insert a zero bit at position h: the k-th index with bit h clear.
__device__ __forceinline__ uint64_t ins0(uint64_t k, uint32_t h) {
    uint64_t lo = k & ((1ull << h) - 1ull);
    return ((k ^ lo) << 1) | lo;
}
// 32-bit form for indices inside an LDS tile.
__device__ __forceinline__ uint32_t ins0_32(uint32_t k, uint32_t h) {
    const uint32_t lo = k & ((1u << h) - 1u);
    return ((k ^ lo) << 1) | lo;
}
// Insert zero bits at two distinct positions lo_b < hi_b.
__device__ __forceinline__ uint64_t ins00(uint64_t k, uint32_t lo_b, uint32_t hi_b) {
    return ins0(ins0(k, lo_b), hi_b);
}

__device__ __forceinline__ double parity_sign(uint64_t v) { return (__popcll(v) & 1) ? -1.0 : 1.0; }

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on XCD
// b mod 8).  A streaming kernel that takes data chunk xcd_block() instead of
// blockIdx.x gives each XCD one contiguous eighth of the chunks, so each XCD
// walks its own run of DRAM pages; on 16 GiB in-place sweeps this measured
// 3-6 % faster than the plain mapping.  Grids that are not a multiple of 8
// keep the plain mapping.
__device__ __forceinline__ uint64_t xcd_block() {
    const uint32_t b = blockIdx.x, n = gridDim.x;
    return (n & 7u) ? (uint64_t)b : (uint64_t)(b & 7u) * (n >> 3) + (b >> 3);
}

// Streaming loads and stores that bypass cache residency (each amplitude of a
// sweep is touched once).  They pay only when every load or store
// instruction of a wave covers whole 128-byte lines: when the lowest bit that
// pairs amplitudes (a gate target, an insertion or orbit pivot bit) is below
// kNtMinBit, consecutive lanes touch every other amplitude, the partner half
// of each line comes from another instruction, and non-temporal accesses then
// fetch and write partial lines (measured up to 5x slower than plain
// accesses).  States small enough to stay in the caches (below 2^24
// amplitudes = 256 MiB, the size of the last-level cache) lose by bypassing
// them, so the host asks for NT only from kNtMinQubits on.  Kernels take NT
// as a template flag or argument chosen on the host (use_nt).
constexpr uint32_t kNtMinBit = 3;
constexpr int kNtMinQubits = 24;
inline bool use_nt(int nqubit, uint32_t low_bit) { return nqubit >= kNtMinQubits && low_bit >= kNtMinBit; }

__device__ __forceinline__ double2 ld_nt(const double2* p) {
    return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}
template <bool NT>
__device__ __forceinline__ double2 ld_s(const double2* p) {
    if constexpr (NT)
        return ld_nt(p);
    else
        return *p;
}
template <bool NT>
__device__ __forceinline__ void st_s(double2* p, double2 v) {
    if constexpr (NT)
        st_nt(p, v);
    else
        *p = v;
}

// c_i <- c_{i ^ v}: a butterfly of conditional swaps, one stage per bit of v.
// Every array index is a compile-time constant (v only selects whether a
// stage swaps), so the array stays in registers.
template <int M>
__device__ __forceinline__ void xor_permute(double2 (&c)[1 << M], uint32_t v) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
        if ((v >> k) & 1u) {
#pragma unroll
            for (int i = 0; i < (1 << M); ++i) {
                if (i & (1 << k)) continue;
                const double2 t = c[i];
                c[i] = c[i | (1 << k)];
                c[i | (1 << k)] = t;
            }
        }
    }
}

// Call f(std::integral_constant<uint32_t, V>{}) for the V in [0, 2^M) equal to
// the uniform runtime value v: a chain of scalar compares, after which every
// array index in f that depends on V is a compile-time constant (a partner
// a[i ^ V] is a register name, not a data movement).
template <int M, uint32_t V = 0, typename F>
__device__ __forceinline__ void dispatch_mask(uint32_t v, F&& f) {
    if constexpr (V + 1 < (1u << M)) {
        if (v == V) {
            f(std::integral_constant<uint32_t, V>{});
            return;
        }
        dispatch_mask<M, V + 1>(v, static_cast<F&&>(f));
    } else {
        f(std::integral_constant<uint32_t, V>{});
    }
}

// Wave-level sum of a double2 (fixed butterfly order -> deterministic).
__device__ __forceinline__ double2 wave_sum(double2 v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, kWave);
        v.y += __shfl_xor(v.y, off, kWave);
    }
    return v;
}

// Block-level sum of a double2; blockDim.x must be a multiple of 64 and at most 1024.
// Result valid in thread 0.  Deterministic for a fixed block size.
__device__ __forceinline__ double2 block_sum(double2 v, double2* lds /* >= 16 entries */) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (threadIdx.x == 0) {
        const int nw = blockDim.x / kWave;
        for (int w = 0; w < nw; ++w) r = cadd(r, lds[w]);
    }
    __syncthreads();
    return r;
}

}  // namespace qfe
