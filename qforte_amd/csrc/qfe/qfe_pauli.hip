// Pauli sums: expectation values, cross expectation values and apply_operator.
//
// The reference evaluates <psi|H|psi> by copying psi, applying every term's
// Pauli gates as separate sweeps, scaling, accumulating and restoring the copy
// (computer.cc:80-91, 753-771): (w + 4) sweeps per term.  Here terms are
// grouped by X mask; every term of a group pairs the same amplitudes
// (j, j ^ x), so one read pass over the state evaluates a whole group:
//   sum_{t in g} c_t <bra|P_t|ket> = sum_{pairs} conj(bra_a) ket_b f(b) + conj(bra_b) ket_a f(a),
//   f(j) = sum_{t in g} c_t phase_t (-1)^popcount(j & z_t).
// Reductions are two-stage with a fixed order, so results are bitwise
// reproducible run to run and independent of what else runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "qfe_device.h"
#include "qfe_internal.h"

namespace qfe {

void group_terms(const qfe_pauli* terms, uint64_t count, std::vector<PauliGroup>& groups, std::vector<uint64_t>& zs,
                 std::vector<double2>& cs) {
    groups.clear();
    zs.clear();
    cs.clear();
    std::unordered_map<uint64_t, size_t> where;
    std::vector<std::vector<uint64_t>> members;
    for (uint64_t i = 0; i < count; ++i) {
        auto it = where.find(terms[i].x);
        if (it == where.end()) {
            where.emplace(terms[i].x, members.size());
            members.emplace_back();
            members.back().push_back(i);
        } else {
            members[it->second].push_back(i);
        }
    }
    for (auto& mem : members) {
        PauliGroup g{};
        g.x = terms[mem[0]].x;
        g.first = (uint32_t)zs.size();
        g.count = (uint32_t)mem.size();
        g.hb = g.x ? (uint32_t)(63 - __builtin_clzll(g.x)) : 0u;
        for (uint64_t i : mem) {
            const qfe_pauli& t = terms[i];
            zs.push_back(t.z);
            // fold the coefficient and the phase of the mask form
            cs.push_back(make_double2(t.coeff.re * t.phase.re - t.coeff.im * t.phase.im,
                                      t.coeff.re * t.phase.im + t.coeff.im * t.phase.re));
        }
        groups.push_back(g);
    }
}

constexpr int kEvTPB = 256;

// f(j) = sum_t c_t (-1)^popcount(j & z_t)
__device__ __forceinline__ double2 group_f(uint64_t j, const uint64_t* __restrict__ zs, const double2* __restrict__ cs,
                                           uint32_t count) {
    double2 f = make_double2(0.0, 0.0);
    for (uint32_t t = 0; t < count; ++t) {
        const double2 c = cs[t];
        if (__popcll(j & zs[t]) & 1) {
            f.x -= c.x;
            f.y -= c.y;
        } else {
            f.x += c.x;
            f.y += c.y;
        }
    }
    return f;
}

// One block = (group g, chunk ch).  Pairs of group g are split into nchunk
// equal contiguous ranges.
__global__ __launch_bounds__(kEvTPB) void k_expval_groups(const double2* __restrict__ bra,
                                                          const double2* __restrict__ ket, int n,
                                                          const PauliGroup* __restrict__ groups,
                                                          const uint64_t* __restrict__ zs,
                                                          const double2* __restrict__ cs, uint32_t nchunk,
                                                          double2* __restrict__ partial) {
    __shared__ double2 lds[16];
    const uint32_t g = blockIdx.x / nchunk;
    const uint32_t ch = blockIdx.x % nchunk;
    const PauliGroup G = groups[g];
    const uint64_t* z = zs + G.first;
    const double2* c = cs + G.first;
    double2 acc = make_double2(0.0, 0.0);
    if (G.x == 0) {
        const uint64_t N = 1ull << n;
        const uint64_t per = (N + nchunk - 1) / nchunk;
        const uint64_t b0 = (uint64_t)ch * per;
        const uint64_t e0 = b0 + per < N ? b0 + per : N;
        for (uint64_t j = b0 + threadIdx.x; j < e0; j += kEvTPB) {
            const double2 p = cconjmul(bra[j], ket[j]);
            acc = cadd(acc, cmul(p, group_f(j, z, c, G.count)));
        }
    } else {
        const uint64_t P = 1ull << (n - 1);
        const uint64_t per = (P + nchunk - 1) / nchunk;
        const uint64_t b0 = (uint64_t)ch * per;
        const uint64_t e0 = b0 + per < P ? b0 + per : P;
        for (uint64_t k = b0 + threadIdx.x; k < e0; k += kEvTPB) {
            const uint64_t ja = ins0(k, G.hb);
            const uint64_t jb = ja ^ G.x;
            const double2 ab = cconjmul(bra[ja], ket[jb]);
            const double2 ba = cconjmul(bra[jb], ket[ja]);
            acc = cadd(acc, cmul(ab, group_f(jb, z, c, G.count)));
            acc = cadd(acc, cmul(ba, group_f(ja, z, c, G.count)));
        }
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kEvTPB) void k_sum_groups(const double2* __restrict__ partial, uint32_t nchunk,
                                                       double2* __restrict__ out) {
    __shared__ double2 lds[16];
    const double2* row = partial + (uint64_t)blockIdx.x * nchunk;
    double2 acc = make_double2(0.0, 0.0);
    for (uint32_t i = threadIdx.x; i < nchunk; i += kEvTPB) acc = cadd(acc, row[i]);
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

__global__ void k_apply_sum(const double2* __restrict__ in, double2* __restrict__ out, uint64_t N,
                            const PauliGroup* __restrict__ groups, uint32_t ngroups, const uint64_t* __restrict__ zs,
                            const double2* __restrict__ cs);

// Upload groups/terms into device scratch after `reserved` bytes; returns pointers.
struct GroupDev {
    PauliGroup* groups;
    uint64_t* zs;
    double2* cs;
    char* tail;  // first free byte after the arrays
};

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

static int upload_groups(const std::vector<PauliGroup>& groups, const std::vector<uint64_t>& zs,
                         const std::vector<double2>& cs, size_t extra, GroupDev& out) {
    const size_t b_g = align256(groups.size() * sizeof(PauliGroup));
    const size_t b_z = align256(zs.size() * sizeof(uint64_t));
    const size_t b_c = align256(cs.size() * sizeof(double2));
    void* scr = nullptr;
    int rc = scratch_reserve(b_g + b_z + b_c + extra, &scr);
    if (rc) return rc;
    char* base = static_cast<char*>(scr);
    std::vector<char> host(b_g + b_z + b_c);
    std::memcpy(host.data(), groups.data(), groups.size() * sizeof(PauliGroup));
    std::memcpy(host.data() + b_g, zs.data(), zs.size() * sizeof(uint64_t));
    std::memcpy(host.data() + b_g + b_z, cs.data(), cs.size() * sizeof(double2));
    QFE_HIP(hipMemcpyAsync(base, host.data(), host.size(), hipMemcpyHostToDevice, ctx().stream));
    QFE_HIP(hipStreamSynchronize(ctx().stream));
    out.groups = reinterpret_cast<PauliGroup*>(base);
    out.zs = reinterpret_cast<uint64_t*>(base + b_g);
    out.cs = reinterpret_cast<double2*>(base + b_g + b_z);
    out.tail = base + b_g + b_z + b_c;
    return QFE_OK;
}

static uint32_t chunks_for(int n, size_t ngroups) {
    // aim for ~4096 pairs per block, at most 1024 chunks per group, and keep
    // the partials array bounded
    const uint64_t P = 1ull << (n > 0 ? n - 1 : 0);
    uint64_t ch = (P + 4095) / 4096;
    if (ch > 1024) ch = 1024;
    if (ch < 1) ch = 1;
    while (ch > 1 && ch * ngroups > (1ull << 22)) ch >>= 1;
    return (uint32_t)ch;
}

uint32_t expval_chunks(int n, size_t ngroups) { return chunks_for(n, ngroups); }

int launch_expval_async(const double2* bra, const double2* ket, int n, const PauliGroup* dgroups, uint32_t ngroups,
                        const uint64_t* dzs, const double2* dcs, double2* partial, uint32_t nchunk,
                        double2* dout) {
    Context& cx = ctx();
    if (ngroups == 0) return QFE_OK;
    hipLaunchKernelGGL(k_expval_groups, dim3((unsigned)((uint64_t)ngroups * nchunk)), dim3(kEvTPB), 0, cx.stream, bra,
                       ket, n, dgroups, dzs, dcs, nchunk, partial);
    QFE_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_groups, dim3(ngroups), dim3(kEvTPB), 0, cx.stream, partial, nchunk, dout);
    QFE_HIP(hipGetLastError());
    cx.counters.reduce += ngroups;
    cx.counters.bytes += (bra == ket ? 16.0 : 32.0) * (double)(1ull << n) * (double)ngroups;
    return QFE_OK;
}

int launch_apply_sum_async(const double2* in, double2* out, int n, const PauliGroup* dgroups, uint32_t ngroups,
                           const uint64_t* dzs, const double2* dcs) {
    Context& cx = ctx();
    const uint64_t N = 1ull << n;
    if (ngroups == 0) return launch_zero(out, N, cx.stream);
    hipLaunchKernelGGL(k_apply_sum, dim3((unsigned)((N + kEvTPB - 1) / kEvTPB)), dim3(kEvTPB), 0, cx.stream, in, out,
                       N, dgroups, ngroups, dzs, dcs);
    QFE_HIP(hipGetLastError());
    cx.counters.single++;
    cx.counters.bytes += 16.0 * (double)N * (double)(ngroups + 1);
    return QFE_OK;
}

// ---------------------------------------------------------------------------
// Orbit-fused expectation passes.  Groups whose X masks span <= 4 dimensions
// over GF(2) (5 for a plain expectation value) are evaluated together: every thread loads the 2^M amplitudes
// of its orbit {j ^ c_i} once and accumulates, for each group g with mask
// v_g in orbit coordinates, sum_i conj(bra_i) ket_{i^v} f_g(i^v).  One read
// pass of the state then serves up to 16 groups.
// ---------------------------------------------------------------------------
constexpr int kEvOrbitMax = 4;
// Same-state passes (bra == ket) go one dimension wider: the pair symmetry of
// their products leaves them memory-bound at 5 dimensions.
constexpr int kEvOrbitMaxSame = 5;
static_assert((1 << kEvOrbitMaxSame) <= 32, "OrbitTerm sign masks hold 32 elements");
// Smallest state that takes orbit passes: below it one launch over every
// group (a read of the state per group) is faster than the chain of orbit
// passes, whose workgroups then cover too little of the state (measured on
// MI355X, 200 and 1000 random terms: 0.25 vs 0.50 and 0.96 vs 1.91 ms at 20
// qubits, 1.25 vs 0.89 and 5.2 vs 3.5 ms at 22).
constexpr int kEvOrbitMinQubits = 21;
constexpr int kEvOrbitGroups = 16;

struct OrbitGroup {
    uint32_t v;      // X mask in orbit coordinates
    uint32_t first;  // first term (into the OrbitTerm array)
    uint32_t count;
    uint32_t slot;   // output row
};

struct OrbitTerm {
    uint64_t z;
    uint32_t smask;  // bit i = parity(c_i & z)
    uint32_t pmask;  // bit i = bit i ^ v of smask (v: the group's X mask in orbit coordinates)
    double2 c;       // coefficient * phase
};


// p_i = conj(b_i) a_{i ^ v} for every element of an orbit, v dispatched to a
// compile-time constant so that a_{i ^ v} names a register (no data movement).
template <int M, uint32_t V = 0>
__device__ __forceinline__ void partner_products(uint32_t v, double2 (&p)[1 << M], const double2 (&a)[1 << M],
                                                 const double2 (&b)[1 << M]) {
    if constexpr (V + 1 < (1u << M)) {
        if (v != V) {
            partner_products<M, V + 1>(v, p, a, b);
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < (1 << M); ++i) {
        // an empty volatile asm on the partner keeps the products inside this
        // branch: otherwise the compiler computes the products of every
        // branch up front and spills them
        double2 t = a[i ^ V];
        asm volatile("" : "+v"(t.x), "+v"(t.y));
        p[i] = cconjmul(b[i], t);
    }
}

// Cross expectation pass <bra|H|ket> (bra != ket; a plain expectation value
// takes k_orbit_expval_same).
template <int M, bool NT>
__global__ __launch_bounds__(kEvTPB) __attribute__((amdgpu_waves_per_eu(2))) void k_orbit_expval(const double2* __restrict__ bra,
                                                         const double2* __restrict__ ket, uint64_t norbit,
                                                         uint64_t b0, uint64_t b1, uint64_t b2, uint64_t b3,
                                                         uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3,
                                                         const OrbitGroup* __restrict__ groups, uint32_t ngroups,
                                                         const OrbitTerm* __restrict__ terms,
                                                         double2* __restrict__ partial, uint32_t nb) {
    // per-thread accumulator of each group (dynamic LDS: ngroups * kEvTPB
    // entries, so passes with few groups keep more workgroups per CU)
    extern __shared__ double2 acc_s[];
    __shared__ double2 lds[16];
    const uint64_t bb[4] = {b0, b1, b2, b3};
    const uint32_t pp[4] = {p0, p1, p2, p3};
    for (uint32_t g = 0; g < ngroups; ++g) acc_s[g * kEvTPB + threadIdx.x] = make_double2(0.0, 0.0);
    // chunk ch of nb contiguous orbit ranges (a multiple of the block size each)
    const uint64_t ch = xcd_block();
    const uint64_t per = ((norbit + nb - 1) / nb + kEvTPB - 1) / kEvTPB * kEvTPB;
    const uint64_t t1 = min(norbit, (ch + 1) * per);
    for (uint64_t t = ch * per + threadIdx.x; t < t1; t += kEvTPB) {
        uint64_t j = t;
#pragma unroll
        for (int k = 0; k < M; ++k) j = ins0(j, pp[k]);
        double2 a[1 << M], b[1 << M];
#pragma unroll
        for (int i = 0; i < (1 << M); ++i) {
            uint64_t c = 0;
#pragma unroll
            for (int k = 0; k < M; ++k)
                if ((i >> k) & 1) c ^= bb[k];
            a[i] = ld_s<NT>(ket + (j ^ c));
            b[i] = ld_s<NT>(bra + (j ^ c));
        }
        for (uint32_t g = 0; g < ngroups; ++g) {
            const OrbitGroup G = groups[g];
            double2 acc = make_double2(0.0, 0.0);
            // p_i = conj(bra_i) ket_{i ^ v}: the group's flip mask selects, at
            // compile time, which register holds the partner of each element
            // (no data movement)
            double2 p[1 << M];
            partner_products<M>(G.v, p, a, b);
            for (uint32_t tt = 0; tt < G.count; ++tt) {
                const OrbitTerm T = terms[G.first + tt];
                // bit i of pmask: the sign of the ket element i ^ v; the signs
                // are uniform, so each element is one FMA by +-1, and the
                // thread's own parity negates the sum
                double2 w = make_double2(0.0, 0.0);
#pragma unroll
                for (int i = 0; i < (1 << M); ++i) {
                    const double sg = ((T.pmask >> i) & 1u) ? -1.0 : 1.0;
                    w.x = fma(sg, p[i].x, w.x);
                    w.y = fma(sg, p[i].y, w.y);
                }
                if (__popcll(j & T.z) & 1) w = make_double2(-w.x, -w.y);
                acc = cadd(acc, cmul(T.c, w));
            }
            acc_s[g * kEvTPB + threadIdx.x] = cadd(acc_s[g * kEvTPB + threadIdx.x], acc);
        }
    }
    for (uint32_t g = 0; g < ngroups; ++g) {
        const double2 s = block_sum(acc_s[g * kEvTPB + threadIdx.x], lds);
        if (threadIdx.x == 0) partial[(uint64_t)groups[g].slot * nb + ch] = s;
    }
}

// Basis of an orbit pass: vectors and their pivots, ascending.
struct OrbitBasis {
    uint64_t b[kEvOrbitMaxSame];
    uint32_t p[kEvOrbitMaxSame];
};

// Sign vector of a term in orbit coordinates: bit k = bit 2^k of smask (the
// sign of element i is parity(i & zm), linear in i).
template <int M>
__device__ __forceinline__ uint32_t term_sign_vector(uint32_t smask) {
    uint32_t zm = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) zm |= ((smask >> (1u << k)) & 1u) << k;
    return zm;
}

// Same-state expectation pass (bra == ket).  For a group with flip v the
// products p_i = conj(a_i) a_{i ^ v} pair up as p_{i ^ v} = conj(p_i), and a
// term's signs s_i (bit i of pmask) satisfy s_{i ^ v} = sigma s_i with sigma
// = (-1)^(bit v of smask), because the sign pattern over an orbit is linear
// in the element index.  So, with h the highest bit of v,
//   sum_i s_i p_i = 2 sum_{i: bit h clear} s_i Re p_i    (sigma = +1)
//                 = 2i sum_{i: bit h clear} s_i Im p_i   (sigma = -1):
// two products and an FMA per pair of elements instead of a complex product
// and two FMAs per element.  A diagonal group (v = 0, the terms with no X or
// Y; one per operator; diag is the first in the pass, ngroups if none) sums
// s_i |a_i|^2 = H[zm], the Walsh-Hadamard transform of q_i = |a_i|^2
// at the term's sign vector zm (bit k of zm: bit 2^k of smask, since the
// sign pattern is linear in i): it runs after the other groups, transforms
// q in place of the amplitudes once, and then costs each of its terms (often
// the most numerous group) one register select instead of a sum over the
// orbit.  Further groups with v = 0 in the same pass (one per operator of
// qfe_expval_many) sum s_i |a_i|^2 directly.
template <int M, bool NT>
__global__ __launch_bounds__(kEvTPB) __attribute__((amdgpu_waves_per_eu(3))) void k_orbit_expval_same(
    const double2* __restrict__ psi, uint64_t norbit, OrbitBasis ob, const OrbitGroup* __restrict__ groups,
    uint32_t ngroups, uint32_t diag, const OrbitTerm* __restrict__ terms, double2* __restrict__ partial, uint32_t nb) {
    extern __shared__ double2 acc_s[];
    __shared__ double2 lds[16];
    for (uint32_t g = 0; g < ngroups; ++g) acc_s[g * kEvTPB + threadIdx.x] = make_double2(0.0, 0.0);
    const uint64_t ch = xcd_block();
    const uint64_t per = ((norbit + nb - 1) / nb + kEvTPB - 1) / kEvTPB * kEvTPB;
    const uint64_t t1 = min(norbit, (ch + 1) * per);
    for (uint64_t t = ch * per + threadIdx.x; t < t1; t += kEvTPB) {
        uint64_t j = t;
#pragma unroll
        for (int k = 0; k < M; ++k) j = ins0(j, ob.p[k]);
        double2 a[1 << M];
#pragma unroll
        for (int i = 0; i < (1 << M); ++i) {
            uint64_t c = 0;
#pragma unroll
            for (int k = 0; k < M; ++k)
                if ((i >> k) & 1) c ^= ob.b[k];
            a[i] = ld_s<NT>(psi + (j ^ c));
        }
        for (uint32_t g = 0; g < ngroups; ++g) {
            if (g == diag) continue;
            const OrbitGroup G = groups[g];
            double2 acc = make_double2(0.0, 0.0);
            dispatch_mask<M>(G.v, [&](auto vc) {
                constexpr uint32_t V = decltype(vc)::value;
                for (uint32_t tt = 0; tt < G.count; ++tt) {
                    const OrbitTerm T = terms[G.first + tt];
                    double w = 0.0;
                    bool imag = false;
                    // the empty volatile asm on each partner keeps its products
                    // inside this branch and this term: otherwise the compiler
                    // computes the products of every branch up front, out of the
                    // term loop, and spills them
                    if constexpr (V == 0) {
                        // a second diagonal group in one pass (one per operator
                        // of qfe_expval_many): s_i |a_i|^2 summed directly
#pragma unroll
                        for (int i = 0; i < (1 << M); ++i) {
                            double2 q = a[i];
                            asm volatile("" : "+v"(q.x), "+v"(q.y));
                            const double sg = ((T.pmask >> i) & 1u) ? -1.0 : 1.0;
                            w = fma(sg, fma(q.x, q.x, q.y * q.y), w);
                        }
                    } else {
                        constexpr uint32_t H = 1u << (31 - __builtin_clz(V));
                        imag = (T.smask >> V) & 1u;
                        if (!imag) {
#pragma unroll
                            for (int i = 0; i < (1 << M); ++i) {
                                if (i & H) continue;
                                double2 q = a[i ^ V];
                                asm volatile("" : "+v"(q.x), "+v"(q.y));
                                const double sg = ((T.pmask >> i) & 1u) ? -1.0 : 1.0;
                                w = fma(sg, fma(a[i].x, q.x, a[i].y * q.y), w);
                            }
                        } else {
#pragma unroll
                            for (int i = 0; i < (1 << M); ++i) {
                                if (i & H) continue;
                                double2 q = a[i ^ V];
                                asm volatile("" : "+v"(q.x), "+v"(q.y));
                                const double sg = ((T.pmask >> i) & 1u) ? -1.0 : 1.0;
                                w = fma(sg, fma(a[i].x, q.y, -(a[i].y * q.x)), w);
                            }
                        }
                        w *= 2.0;
                    }
                    if (__popcll(j & T.z) & 1) w = -w;
                    // c * w (sigma = +1) or c * i w (sigma = -1)
                    if (imag)
                        acc = make_double2(fma(-T.c.y, w, acc.x), fma(T.c.x, w, acc.y));
                    else
                        acc = make_double2(fma(T.c.x, w, acc.x), fma(T.c.y, w, acc.y));
                }
            });
            acc_s[g * kEvTPB + threadIdx.x] = cadd(acc_s[g * kEvTPB + threadIdx.x], acc);
        }
        if (diag < ngroups) {
            // q_i = |a_i|^2 over the real parts, then H[k] = sum_i (-1)^popcount(i & k) q_i in place
#pragma unroll
            for (int i = 0; i < (1 << M); ++i) a[i].x = fma(a[i].x, a[i].x, a[i].y * a[i].y);
#pragma unroll
            for (int k = 0; k < M; ++k) {
#pragma unroll
                for (int i = 0; i < (1 << M); ++i) {
                    if (i & (1 << k)) continue;
                    const double u = a[i].x, v = a[i | (1 << k)].x;
                    a[i].x = u + v;
                    a[i | (1 << k)].x = u - v;
                }
            }
            const OrbitGroup G = groups[diag];
            double2 acc = make_double2(0.0, 0.0);
            for (uint32_t tt = 0; tt < G.count; ++tt) {
                const OrbitTerm T = terms[G.first + tt];
                const uint32_t zm = term_sign_vector<M>(T.smask);
                double w = 0.0;
                // one branch per value of zm (an empty volatile asm keeps the
                // branches apart, so the select is never a variable index into
                // the register array, which would move it to scratch memory)
                dispatch_mask<M>(zm, [&](auto zc) {
                    double t = a[decltype(zc)::value].x;
                    asm volatile("" : "+v"(t));
                    w = t;
                });
                if (__popcll(j & T.z) & 1) w = -w;
                acc = make_double2(fma(T.c.x, w, acc.x), fma(T.c.y, w, acc.y));
            }
            acc_s[diag * kEvTPB + threadIdx.x] = cadd(acc_s[diag * kEvTPB + threadIdx.x], acc);
        }
    }
    for (uint32_t g = 0; g < ngroups; ++g) {
        const double2 s = block_sum(acc_s[g * kEvTPB + threadIdx.x], lds);
        if (threadIdx.x == 0) partial[(uint64_t)groups[g].slot * nb + ch] = s;
    }
}

// Pack groups into passes whose X masks span <= maxdim dimensions.  A
// pass grows its basis one vector at a time: each remaining mask is kept
// reduced by the basis so far (its coset representative, 0 once the mask
// lies in the span), and the next vector is the representative shared by the
// most remaining groups (fewest bits, then earliest group, on ties), which
// brings all of them into the span at once.  A representative whose highest
// bit, the new pivot, is below kNtMinBit is taken only when no other is left:
// a pass with such a pivot has every lane of a load touch every other
// amplitude and runs 1.7x longer (4.8 vs 2.8 ms at 30 qubits), so the masks
// that need one end up together.  The groups whose masks then lie in the span
// join the pass, in group order, up to kEvOrbitGroups.  On 1000 random terms
// of weight 1-6 over 30 qubits (647 X masks) this packs 92 passes of 5
// dimensions, one of them with a low pivot (91 passes, 13 with low pivots,
// without the pivot rule; 126 taking the groups in order, first fit).
void plan_expval(const std::vector<PauliGroup>& groups, int maxdim, std::vector<EvPass>& passes) {
    std::vector<uint32_t> rem(groups.size());
    for (size_t g = 0; g < groups.size(); ++g) rem[g] = (uint32_t)g;
    std::vector<uint64_t> red(groups.size());
    // open-addressing tally of representatives: key, count, first position,
    // and the step that wrote the slot (slots of older steps read as empty)
    size_t cap = 16;
    while (cap < 2 * groups.size()) cap <<= 1;
    std::vector<uint64_t> key(cap);
    std::vector<uint32_t> count(cap), first(cap), stamp(cap, 0);
    uint32_t step = 0;
    while (!rem.empty()) {
        EvPass p;
        for (size_t i = 0; i < rem.size(); ++i) red[i] = groups[rem[i]].x;
        for (;;) {
            size_t inside = 0;
            for (size_t i = 0; i < rem.size(); ++i) inside += red[i] == 0;
            if (inside >= (size_t)kEvOrbitGroups || (int)p.basis.size() >= maxdim) break;
            ++step;
            uint64_t best = 0;
            uint32_t bc = 0, bf = 0;
            for (size_t i = 0; i < rem.size(); ++i) {
                const uint64_t r = red[i];
                if (!r) continue;
                size_t h = (size_t)((r * 0x9E3779B97F4A7C15ull) >> 32) & (cap - 1);
                while (stamp[h] == step && key[h] != r) h = (h + 1) & (cap - 1);
                if (stamp[h] != step) {
                    stamp[h] = step;
                    key[h] = r;
                    count[h] = 0;
                    first[h] = (uint32_t)i;
                }
                // every representative passes through each count up to its
                // final one, so the running best ends as the maximum
                const uint32_t c = ++count[h];
                const int pc = __builtin_popcountll(r), bpc = __builtin_popcountll(best);
                const bool hi = (r >> kNtMinBit) != 0, bhi = (best >> kNtMinBit) != 0;
                if (!best || (hi && !bhi) ||
                    (hi == bhi && (c > bc || (c == bc && (pc < bpc || (pc == bpc && first[h] < bf)))))) {
                    best = r;
                    bc = c;
                    bf = first[h];
                }
            }
            if (!best) break;
            // best has no bit in a pivot column of the basis; its own pivot
            // is cleared from the other vectors and from every representative
            const int piv = 63 - __builtin_clzll(best);
            for (uint64_t& v : p.basis)
                if ((v >> piv) & 1ull) v ^= best;
            p.basis.push_back(best);
            for (size_t i = 0; i < rem.size(); ++i)
                if ((red[i] >> piv) & 1ull) red[i] ^= best;
        }
        size_t w = 0;
        for (size_t i = 0; i < rem.size(); ++i) {
            if (red[i] == 0 && p.groups.size() < (size_t)kEvOrbitGroups)
                p.groups.push_back(rem[i]);
            else
                rem[w++] = rem[i];
        }
        rem.resize(w);
        passes.push_back(std::move(p));
    }
}

// Plans of the last few operators, keyed by their X masks in group order:
// callers that evaluate one operator many times (a variational loop, a step
// of a benchmark) plan it once.  Planning 647 masks takes about 2 ms on the
// host, against 5-10 us per saved pass at 21 qubits.
namespace {
struct PlanEntry {
    int maxdim = 0;
    std::vector<uint64_t> masks;
    std::vector<EvPass> passes;
    uint64_t used = 0;
};
std::mutex g_plan_mu;
std::vector<PlanEntry> g_plans;
uint64_t g_plan_clock = 0;
constexpr size_t kPlanCacheEntries = 8;
}  // namespace

static void cached_plan(const std::vector<PauliGroup>& groups, int maxdim, std::vector<EvPass>& passes) {
    std::vector<uint64_t> masks(groups.size());
    for (size_t g = 0; g < groups.size(); ++g) masks[g] = groups[g].x;
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        for (PlanEntry& e : g_plans)
            if (e.maxdim == maxdim && e.masks == masks) {
                e.used = ++g_plan_clock;
                passes = e.passes;
                return;
            }
    }
    passes.clear();
    plan_expval(groups, maxdim, passes);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    if (g_plans.size() < kPlanCacheEntries) {
        g_plans.push_back(PlanEntry{maxdim, std::move(masks), passes, ++g_plan_clock});
        return;
    }
    PlanEntry* lru = &g_plans[0];
    for (PlanEntry& e : g_plans)
        if (e.used < lru->used) lru = &e;
    *lru = PlanEntry{maxdim, std::move(masks), passes, ++g_plan_clock};
}

template <int M>
static void launch_orbit_expval_k(const double2* bra, const double2* ket, uint64_t norbit, const OrbitBasis& ob,
                                  const OrbitGroup* dg, uint32_t ng, uint32_t diag, const OrbitTerm* dt,
                                  double2* partial, uint32_t nb, bool nt, hipStream_t st) {
    const size_t lds = (size_t)ng * kEvTPB * sizeof(double2);
    if (bra == ket) {
        if (nt)
            hipLaunchKernelGGL((k_orbit_expval_same<M, true>), dim3(nb), dim3(kEvTPB), lds, st, ket, norbit, ob, dg, ng,
                               diag, dt, partial, nb);
        else
            hipLaunchKernelGGL((k_orbit_expval_same<M, false>), dim3(nb), dim3(kEvTPB), lds, st, ket, norbit, ob, dg, ng,
                               diag, dt, partial, nb);
        return;
    }
    if constexpr (M <= kEvOrbitMax) {
        const uint64_t* b = ob.b;
        const uint32_t* piv = ob.p;
        if (nt)
            hipLaunchKernelGGL((k_orbit_expval<M, true>), dim3(nb), dim3(kEvTPB), lds, st, bra, ket, norbit, b[0], b[1],
                               b[2], b[3], piv[0], piv[1], piv[2], piv[3], dg, ng, dt, partial, nb);
        else
            hipLaunchKernelGGL((k_orbit_expval<M, false>), dim3(nb), dim3(kEvTPB), lds, st, bra, ket, norbit, b[0], b[1],
                               b[2], b[3], piv[0], piv[1], piv[2], piv[3], dg, ng, dt, partial, nb);
    }
}

// One orbit expectation pass as planned on the host: its basis (pivots
// ascending) and its run of OrbitGroup entries.
struct OrbitLaunch {
    int M;
    OrbitBasis ob;
    size_t g0, ng;
    uint32_t diag;  // position of the first group with v = 0 in the pass, ng if none
};

// Host tables of every pass: per pass its OrbitGroup run (slot = group index)
// and OrbitTerm run, in pass order.
static void build_orbit_tables(const std::vector<PauliGroup>& groups, const std::vector<uint64_t>& zs,
                               const std::vector<double2>& cs, int maxdim, std::vector<OrbitGroup>& og,
                               std::vector<OrbitTerm>& ot, std::vector<OrbitLaunch>& launches) {
    std::vector<EvPass> passes;
    cached_plan(groups, maxdim, passes);
    og.clear();
    ot.clear();
    launches.clear();
    for (const EvPass& p : passes) {
        std::vector<uint64_t> b = p.basis;
        std::sort(b.begin(), b.end(), [](uint64_t u, uint64_t v) { return __builtin_clzll(u) > __builtin_clzll(v); });
        OrbitLaunch L{};
        L.M = (int)b.size();
        for (int k = 0; k < L.M; ++k) {
            L.ob.b[k] = b[k];
            L.ob.p[k] = (uint32_t)(63 - __builtin_clzll(b[k]));
        }
        uint64_t coff[1 << kEvOrbitMaxSame];
        for (int i = 0; i < (1 << L.M); ++i) {
            uint64_t c = 0;
            for (int k = 0; k < L.M; ++k)
                if ((i >> k) & 1) c ^= b[k];
            coff[i] = c;
        }
        L.g0 = og.size();
        L.ng = p.groups.size();
        L.diag = (uint32_t)L.ng;
        for (uint32_t gi : p.groups) {
            const PauliGroup& pg = groups[gi];
            OrbitGroup o{};
            for (int k = 0; k < L.M; ++k)
                if ((pg.x >> L.ob.p[k]) & 1ull) o.v |= 1u << k;
            o.first = (uint32_t)ot.size();
            o.count = pg.count;
            o.slot = gi;
            if (o.v == 0 && L.diag == L.ng) L.diag = (uint32_t)(og.size() - L.g0);
            og.push_back(o);
            for (uint32_t t = 0; t < pg.count; ++t) {
                OrbitTerm T{};
                T.z = zs[pg.first + t];
                for (int i = 0; i < (1 << L.M); ++i)
                    if (__builtin_popcountll(coff[i] & T.z) & 1) T.smask |= 1u << i;
                for (int i = 0; i < (1 << L.M); ++i)
                    if ((T.smask >> (i ^ o.v)) & 1u) T.pmask |= 1u << i;
                T.c = cs[pg.first + t];
                ot.push_back(T);
            }
        }
        launches.push_back(L);
    }
}

// Workgroups of an orbit expectation pass over 2^n amplitudes.
static uint32_t orbit_expval_blocks(int n) {
    uint64_t nb64 = (1ull << n) / 4096;
    if (nb64 < 1) nb64 = 1;
    if (nb64 > 2048) nb64 = 2048;
    return (uint32_t)nb64;
}

static int launch_orbit_pass(const double2* bra, const double2* ket, int n, const OrbitLaunch& L,
                             const OrbitGroup* dg, const OrbitTerm* dt, double2* partial, uint32_t nb) {
    Context& cx = ctx();
    const uint64_t N = 1ull << n;
    const uint64_t norbit = N >> L.M;
    const OrbitGroup* g = dg + L.g0;
    const uint32_t ng = (uint32_t)L.ng;
    const bool nt = use_nt(n, L.M == 0 ? 63u : L.ob.p[0]);  // pivots ascending
    if (L.M > (bra == ket ? kEvOrbitMaxSame : kEvOrbitMax))
        return fail(QFE_ERR_STATE, "internal: expectation orbit dimension out of range");
    switch (L.M) {
        case 0: launch_orbit_expval_k<0>(bra, ket, norbit, L.ob, g, ng, L.diag, dt, partial, nb, nt, cx.stream); break;
        case 1: launch_orbit_expval_k<1>(bra, ket, norbit, L.ob, g, ng, L.diag, dt, partial, nb, nt, cx.stream); break;
        case 2: launch_orbit_expval_k<2>(bra, ket, norbit, L.ob, g, ng, L.diag, dt, partial, nb, nt, cx.stream); break;
        case 3: launch_orbit_expval_k<3>(bra, ket, norbit, L.ob, g, ng, L.diag, dt, partial, nb, nt, cx.stream); break;
        case 4: launch_orbit_expval_k<4>(bra, ket, norbit, L.ob, g, ng, L.diag, dt, partial, nb, nt, cx.stream); break;
        case 5: launch_orbit_expval_k<5>(bra, ket, norbit, L.ob, g, ng, L.diag, dt, partial, nb, nt, cx.stream); break;
        default: return fail(QFE_ERR_STATE, "internal: expectation orbit dimension out of range");
    }
    QFE_HIP(hipGetLastError());
    cx.counters.reduce++;
    cx.counters.nt += nt ? 1 : 0;
    cx.counters.bytes += (bra == ket ? 16.0 : 32.0) * (double)N;
    return QFE_OK;
}

static int expval_groups_orbit(const double2* bra, const double2* ket, int n, const std::vector<PauliGroup>& groups,
                               const std::vector<uint64_t>& zs, const std::vector<double2>& cs,
                               double2* out_per_group) {
    Context& cx = ctx();
    const size_t G = groups.size();
    const uint32_t nb = orbit_expval_blocks(n);
    std::vector<OrbitGroup> og;
    std::vector<OrbitTerm> ot;
    std::vector<OrbitLaunch> launches;
    build_orbit_tables(groups, zs, cs, bra == ket ? kEvOrbitMaxSame : kEvOrbitMax, og, ot, launches);
    const size_t b_g = align256(og.size() * sizeof(OrbitGroup));
    const size_t b_t = align256(ot.size() * sizeof(OrbitTerm));
    const size_t b_part = align256(G * (size_t)nb * sizeof(double2));
    const size_t b_out = align256(G * sizeof(double2));
    void* scr = nullptr;
    int rc = scratch_reserve(b_g + b_t + b_part + b_out, &scr);
    if (rc) return rc;
    char* base = static_cast<char*>(scr);
    char* host = nullptr;
    rc = staging_begin(b_g + b_t, reinterpret_cast<void**>(&host));
    if (rc) return rc;
    std::memcpy(host, og.data(), og.size() * sizeof(OrbitGroup));
    std::memcpy(host + b_g, ot.data(), ot.size() * sizeof(OrbitTerm));
    rc = staging_commit(base, b_g + b_t);
    if (rc) return rc;
    const OrbitGroup* dg = reinterpret_cast<const OrbitGroup*>(base);
    const OrbitTerm* dt = reinterpret_cast<const OrbitTerm*>(base + b_g);
    double2* partial = reinterpret_cast<double2*>(base + b_g + b_t);
    double2* dout = reinterpret_cast<double2*>(base + b_g + b_t + b_part);
    for (const OrbitLaunch& L : launches) {
        rc = launch_orbit_pass(bra, ket, n, L, dg, dt, partial, nb);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_sum_groups, dim3((unsigned)G), dim3(kEvTPB), 0, cx.stream, partial, nb, dout);
    QFE_HIP(hipGetLastError());
    void* pin = nullptr;
    rc = pinned_reserve(G * sizeof(double2), &pin);
    if (rc) return rc;
    QFE_HIP(hipMemcpyAsync(pin, dout, G * sizeof(double2), hipMemcpyDeviceToHost, cx.stream));
    QFE_HIP(hipStreamSynchronize(cx.stream));
    std::memcpy(out_per_group, pin, G * sizeof(double2));
    return QFE_OK;
}

int expval_groups(const double2* bra, const double2* ket, int n, const std::vector<PauliGroup>& groups,
                  const std::vector<uint64_t>& zs, const std::vector<double2>& cs, double2* out_per_group) {
    Context& cx = ctx();
    const size_t G = groups.size();
    if (G == 0) return QFE_OK;
    if (n >= kEvOrbitMinQubits) return expval_groups_orbit(bra, ket, n, groups, zs, cs, out_per_group);
    const uint32_t nchunk = chunks_for(n, G);
    const size_t b_part = align256(G * nchunk * sizeof(double2));
    const size_t b_out = align256(G * sizeof(double2));
    GroupDev gd;
    int rc = upload_groups(groups, zs, cs, b_part + b_out, gd);
    if (rc) return rc;
    double2* partial = reinterpret_cast<double2*>(gd.tail);
    double2* dout = reinterpret_cast<double2*>(gd.tail + b_part);
    // launch in slices to keep the grid below 2^31 blocks
    const uint64_t total_blocks = (uint64_t)G * nchunk;
    hipLaunchKernelGGL(k_expval_groups, dim3((unsigned)total_blocks), dim3(kEvTPB), 0, cx.stream, bra, ket, n,
                       gd.groups, gd.zs, gd.cs, nchunk, partial);
    QFE_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_groups, dim3((unsigned)G), dim3(kEvTPB), 0, cx.stream, partial, nchunk, dout);
    QFE_HIP(hipGetLastError());
    void* pin = nullptr;
    rc = pinned_reserve(G * sizeof(double2), &pin);
    if (rc) return rc;
    QFE_HIP(hipMemcpyAsync(pin, dout, G * sizeof(double2), hipMemcpyDeviceToHost, cx.stream));
    QFE_HIP(hipStreamSynchronize(cx.stream));
    std::memcpy(out_per_group, pin, G * sizeof(double2));
    cx.counters.reduce += G;
    const double per = (bra == ket ? 16.0 : 32.0) * (double)(1ull << n);
    cx.counters.bytes += per * (double)G;
    return QFE_OK;
}

// ---- expectation values with no allocation ------------------------------------------------
// qfe_expval_noalloc keeps everything beyond its input state and output in
// one fixed device array per GPU: the group and term tables, the per-block
// partial sums and the per-group results.  Orbit passes run in windows of
// whole passes whose partial sums fit the array; each window's groups are
// summed (k_sum_groups, the same fixed order) before the next window reuses
// the partials, so every group's value, and the host sum over groups in group
// order, is the one qfe_expval computes.  Calls that overlap on one GPU race
// on the array.
constexpr uint64_t kNoAllocMaxTerms = 1ull << 16;
constexpr size_t kNoAllocBytes = 6ull << 20;
static __device__ __align__(256) char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
static void* g_scratch_dev[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device (called by qfe_init with
// `device` current, once per device).
int noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return fail(QFE_ERR_ARG, "qfe_init: device index beyond the scratch table");
    if (g_scratch_dev[device]) return QFE_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(expectation scratch)");
    if (!ptr) return fail(QFE_ERR_DEVICE, "hipGetSymbolAddress(expectation scratch): null address");
    g_scratch_dev[device] = ptr;
    return QFE_OK;
}

// Bump layout inside the fixed array; false when the request does not fit.
struct FixedLayout {
    char* base;
    size_t cap, used = 0;
    char* take(size_t bytes) {
        bytes = align256(bytes);
        if (bytes > cap - used) return nullptr;
        char* p = base + used;
        used += bytes;
        return p;
    }
};

static int expval_groups_fixed(const double2* bra, const double2* ket, int n, const std::vector<PauliGroup>& groups,
                               const std::vector<uint64_t>& zs, const std::vector<double2>& cs, double2* out_per_group,
                               void* dev_scratch) {
    Context& cx = ctx();
    const size_t G = groups.size();
    if (G == 0) return QFE_OK;
    FixedLayout A{static_cast<char*>(dev_scratch), kNoAllocBytes};
    std::vector<double2> res(G);
    if (n < kEvOrbitMinQubits) {
        const uint32_t nchunk = chunks_for(n, G);
        char* dg = A.take(G * sizeof(PauliGroup));
        char* dz = A.take(zs.size() * sizeof(uint64_t));
        char* dc = A.take(cs.size() * sizeof(double2));
        char* part = A.take(G * nchunk * sizeof(double2));
        char* dout = A.take(G * sizeof(double2));
        if (!dg || !dz || !dc || !part || !dout) return fail(QFE_ERR_ARG, "qfe_expval_noalloc: operator exceeds the fixed scratch");
        QFE_HIP(hipMemcpyAsync(dg, groups.data(), G * sizeof(PauliGroup), hipMemcpyHostToDevice, cx.stream));
        QFE_HIP(hipMemcpyAsync(dz, zs.data(), zs.size() * sizeof(uint64_t), hipMemcpyHostToDevice, cx.stream));
        QFE_HIP(hipMemcpyAsync(dc, cs.data(), cs.size() * sizeof(double2), hipMemcpyHostToDevice, cx.stream));
        hipLaunchKernelGGL(k_expval_groups, dim3((unsigned)((uint64_t)G * nchunk)), dim3(kEvTPB), 0, cx.stream, bra, ket,
                           n, (const PauliGroup*)dg, (const uint64_t*)dz, (const double2*)dc, nchunk, (double2*)part);
        QFE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sum_groups, dim3((unsigned)G), dim3(kEvTPB), 0, cx.stream, (const double2*)part, nchunk,
                           (double2*)dout);
        QFE_HIP(hipGetLastError());
        QFE_HIP(hipMemcpyAsync(res.data(), dout, G * sizeof(double2), hipMemcpyDeviceToHost, cx.stream));
        QFE_HIP(hipStreamSynchronize(cx.stream));
        std::memcpy(out_per_group, res.data(), G * sizeof(double2));
        cx.counters.reduce += G;
        cx.counters.bytes += (bra == ket ? 16.0 : 32.0) * (double)(1ull << n) * (double)G;
        return QFE_OK;
    }
    const uint32_t nb = orbit_expval_blocks(n);
    std::vector<OrbitGroup> og;
    std::vector<OrbitTerm> ot;
    std::vector<OrbitLaunch> launches;
    build_orbit_tables(groups, zs, cs, bra == ket ? kEvOrbitMaxSame : kEvOrbitMax, og, ot, launches);
    char* dg = A.take(og.size() * sizeof(OrbitGroup));
    char* dt = A.take(ot.size() * sizeof(OrbitTerm));
    char* dout = A.take(G * sizeof(double2));
    const size_t row = (size_t)nb * sizeof(double2);
    const size_t window = (A.cap - A.used) / row;  // groups whose partials fit at once
    char* part = A.take(window * row);
    if (!dg || !dt || !dout || !part || window < (size_t)kEvOrbitGroups)
        return fail(QFE_ERR_ARG, "qfe_expval_noalloc: operator exceeds the fixed scratch");
    // window-local output rows: whole passes in order, slot = position in the
    // window (build_orbit_tables left the group index in slot)
    std::vector<uint32_t> gid(og.size());
    for (size_t i = 0; i < og.size(); ++i) gid[i] = og[i].slot;
    std::vector<size_t> win_start;  // first launch of each window
    size_t w0 = 0;
    for (size_t l = 0; l < launches.size(); ++l) {
        if (l == 0 || launches[l].g0 + launches[l].ng - w0 > window) {
            win_start.push_back(l);
            w0 = launches[l].g0;
        }
        for (size_t i = 0; i < launches[l].ng; ++i) og[launches[l].g0 + i].slot = (uint32_t)(launches[l].g0 + i - w0);
    }
    win_start.push_back(launches.size());
    QFE_HIP(hipMemcpyAsync(dg, og.data(), og.size() * sizeof(OrbitGroup), hipMemcpyHostToDevice, cx.stream));
    QFE_HIP(hipMemcpyAsync(dt, ot.data(), ot.size() * sizeof(OrbitTerm), hipMemcpyHostToDevice, cx.stream));
    for (size_t w = 0; w + 1 < win_start.size(); ++w) {
        const size_t l0 = win_start[w], l1 = win_start[w + 1];
        const size_t g0 = launches[l0].g0, g1 = launches[l1 - 1].g0 + launches[l1 - 1].ng;
        for (size_t l = l0; l < l1; ++l) {
            int rc = launch_orbit_pass(bra, ket, n, launches[l], (const OrbitGroup*)dg, (const OrbitTerm*)dt,
                                       (double2*)part, nb);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_sum_groups, dim3((unsigned)(g1 - g0)), dim3(kEvTPB), 0, cx.stream, (const double2*)part,
                           nb, (double2*)dout + g0);
        QFE_HIP(hipGetLastError());
    }
    // rows of dout are in pass order
    QFE_HIP(hipMemcpyAsync(res.data(), dout, G * sizeof(double2), hipMemcpyDeviceToHost, cx.stream));
    QFE_HIP(hipStreamSynchronize(cx.stream));
    for (size_t i = 0; i < og.size(); ++i) out_per_group[gid[i]] = res[i];
    return QFE_OK;
}

// out_k = sum_g ket[k ^ x_g] f_g(k ^ x_g): every thread owns output amplitudes
// and walks the groups in order (fixed summation order per element).
__global__ void k_apply_sum(const double2* __restrict__ in, double2* __restrict__ out, uint64_t N,
                            const PauliGroup* __restrict__ groups, uint32_t ngroups, const uint64_t* __restrict__ zs,
                            const double2* __restrict__ cs) {
    const uint64_t k = (uint64_t)blockIdx.x * kEvTPB + threadIdx.x;
    if (k >= N) return;
    double2 acc = make_double2(0.0, 0.0);
    for (uint32_t g = 0; g < ngroups; ++g) {
        const PauliGroup G = groups[g];
        const uint64_t src = k ^ G.x;
        const double2 f = group_f(src, zs + G.first, cs + G.first, G.count);
        acc = cadd(acc, cmul(f, in[src]));
    }
    out[k] = acc;
}

int launch_apply_pauli_sum(const double2* in, double2* out, int n, const std::vector<PauliGroup>& groups,
                           const std::vector<uint64_t>& zs, const std::vector<double2>& cs) {
    Context& cx = ctx();
    const uint64_t N = 1ull << n;
    if (groups.empty()) return launch_zero(out, N, cx.stream);
    GroupDev gd;
    int rc = upload_groups(groups, zs, cs, 0, gd);
    if (rc) return rc;
    hipLaunchKernelGGL(k_apply_sum, dim3((unsigned)((N + kEvTPB - 1) / kEvTPB)), dim3(kEvTPB), 0, cx.stream, in, out,
                       N, gd.groups, (uint32_t)groups.size(), gd.zs, gd.cs);
    QFE_HIP(hipGetLastError());
    cx.counters.single++;
    cx.counters.bytes += 16.0 * (double)N * (double)(groups.size() + 1);
    return QFE_OK;
}

static bool valid(qfe_state s) { return s != nullptr && s->d != nullptr; }

static int check_terms(const qfe_pauli* terms, uint64_t count, int n) {
    if (count && !terms) return fail(QFE_ERR_ARG, "Pauli terms: null pointer");
    const uint64_t lim = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
    for (uint64_t i = 0; i < count; ++i)
        if ((terms[i].x & ~lim) || (terms[i].z & ~lim))
            return fail(QFE_ERR_ARG, "Pauli term acts on a qubit outside the register");
    return QFE_OK;
}

}  // namespace qfe

using namespace qfe;

extern "C" {

int qfe_expval(qfe_state s, const qfe_pauli* terms, uint64_t count, qfe_c64* out) {
    if (!valid(s) || !out) return fail(QFE_ERR_ARG, "qfe_expval: invalid argument");
    int rc = check_terms(terms, count, s->n);
    if (rc) return rc;
    std::vector<PauliGroup> groups;
    std::vector<uint64_t> zs;
    std::vector<double2> cs;
    group_terms(terms, count, groups, zs, cs);
    std::vector<double2> per(groups.size());
    rc = expval_groups(s->d, s->d, s->n, groups, zs, cs, per.data());
    if (rc) return rc;
    double2 tot = make_double2(0.0, 0.0);
    for (const double2& v : per) tot = make_double2(tot.x + v.x, tot.y + v.y);
    out->re = tot.x;
    out->im = tot.y;
    return QFE_OK;
}

int qfe_expval_noalloc(qfe_state s, const qfe_pauli* terms, uint64_t count, qfe_c64* out) {
    if (!valid(s) || !out) return fail(QFE_ERR_ARG, "qfe_expval_noalloc: invalid argument");
    if (count > kNoAllocMaxTerms)
        return fail(QFE_ERR_ARG, "qfe_expval_noalloc: at most 65536 Pauli terms per call (fixed 6 MiB device scratch)");
    int rc = check_terms(terms, count, s->n);
    if (rc) return rc;
    void* dev = g_scratch_dev[ctx().device];
    if (!dev) return fail(QFE_ERR_STATE, "qfe_expval_noalloc: library not initialised on this device");
    std::vector<PauliGroup> groups;
    std::vector<uint64_t> zs;
    std::vector<double2> cs;
    group_terms(terms, count, groups, zs, cs);
    std::vector<double2> per(groups.size());
    rc = expval_groups_fixed(s->d, s->d, s->n, groups, zs, cs, per.data(), dev);
    if (rc) return rc;
    double2 tot = make_double2(0.0, 0.0);
    for (const double2& v : per) tot = make_double2(tot.x + v.x, tot.y + v.y);
    out->re = tot.x;
    out->im = tot.y;
    return QFE_OK;
}

int qfe_expval_many(qfe_state s, const qfe_pauli* terms, const uint64_t* op_offsets, uint64_t nops, qfe_c64* out) {
    return qfe_cross_expval_many(s, s, terms, op_offsets, nops, out);
}

int qfe_cross_expval_many(qfe_state bra, qfe_state s, const qfe_pauli* terms, const uint64_t* op_offsets,
                          uint64_t nops, qfe_c64* out) {
    if (!valid(s) || !valid(bra) || (nops && (!op_offsets || !out)))
        return fail(QFE_ERR_ARG, "qfe_cross_expval_many: invalid argument");
    if (bra->n != s->n) return fail(QFE_ERR_ARG, "qfe_cross_expval_many: qubit counts differ");
    if (nops == 0) return QFE_OK;
    const uint64_t total = op_offsets[nops];
    int rc = check_terms(terms, total, s->n);
    if (rc) return rc;
    // one launch for all operators: groups never span operators
    std::vector<PauliGroup> groups, g1;
    std::vector<uint64_t> zs, z1;
    std::vector<double2> cs, c1;
    std::vector<uint64_t> op_of_group;
    for (uint64_t k = 0; k < nops; ++k) {
        if (op_offsets[k + 1] < op_offsets[k]) return fail(QFE_ERR_ARG, "qfe_expval_many: offsets must be non-decreasing");
        group_terms(terms + op_offsets[k], op_offsets[k + 1] - op_offsets[k], g1, z1, c1);
        const uint32_t shift = (uint32_t)zs.size();
        for (PauliGroup g : g1) {
            g.first += shift;
            groups.push_back(g);
            op_of_group.push_back(k);
        }
        zs.insert(zs.end(), z1.begin(), z1.end());
        cs.insert(cs.end(), c1.begin(), c1.end());
    }
    std::vector<double2> per(groups.size());
    rc = expval_groups(bra->d, s->d, s->n, groups, zs, cs, per.data());
    if (rc) return rc;
    for (uint64_t k = 0; k < nops; ++k) out[k] = qfe_c64{0.0, 0.0};
    for (size_t g = 0; g < groups.size(); ++g) {
        out[op_of_group[g]].re += per[g].x;
        out[op_of_group[g]].im += per[g].y;
    }
    return QFE_OK;
}

int qfe_cross_expval(qfe_state bra, qfe_state ket, const qfe_pauli* terms, uint64_t count, qfe_c64* out) {
    if (!valid(bra) || !valid(ket) || !out) return fail(QFE_ERR_ARG, "qfe_cross_expval: invalid argument");
    if (bra->n != ket->n) return fail(QFE_ERR_ARG, "qfe_cross_expval: qubit counts differ");
    int rc = check_terms(terms, count, bra->n);
    if (rc) return rc;
    std::vector<PauliGroup> groups;
    std::vector<uint64_t> zs;
    std::vector<double2> cs;
    group_terms(terms, count, groups, zs, cs);
    std::vector<double2> per(groups.size());
    rc = expval_groups(bra->d, ket->d, bra->n, groups, zs, cs, per.data());
    if (rc) return rc;
    double2 tot = make_double2(0.0, 0.0);
    for (const double2& v : per) tot = make_double2(tot.x + v.x, tot.y + v.y);
    out->re = tot.x;
    out->im = tot.y;
    return QFE_OK;
}

int qfe_apply_pauli_sum(qfe_state in, const qfe_pauli* terms, uint64_t count, qfe_state out) {
    if (!valid(in) || !valid(out)) return fail(QFE_ERR_ARG, "qfe_apply_pauli_sum: invalid handle");
    if (in == out) return fail(QFE_ERR_ARG, "qfe_apply_pauli_sum: output must differ from input");
    if (in->n != out->n) return fail(QFE_ERR_ARG, "qfe_apply_pauli_sum: qubit counts differ");
    int rc = check_terms(terms, count, in->n);
    if (rc) return rc;
    std::vector<PauliGroup> groups;
    std::vector<uint64_t> zs;
    std::vector<double2> cs;
    group_terms(terms, count, groups, zs, cs);
    return launch_apply_pauli_sum(in->d, out->d, in->n, groups, zs, cs);
}

}  // extern "C"
