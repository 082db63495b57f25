// Device helpers shared by the kernels of libqfe (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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
