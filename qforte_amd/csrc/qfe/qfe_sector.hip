// Determinant-sector engine: UCC energies, adjoint gradients and pool
// gradients on the (N_alpha, N_beta) sector of the reference determinant, and
// the continuous-batching serving engine built on them.
//
// A UCC generator K = sum_t kappa_t P_t built from a fermionic excitation
// (JW of tau - tau^dag, ansatz.py:25-82 feeding trotterize) is a set of Pauli
// strings that share one X mask and commute, so the product of their
// exponentials that the reference applies (trotterization.py:29-35, one
// exponentiate_pauli_string per term) equals exp(theta K) exactly.  K maps a
// determinant j to j ^ x with weight alpha(j) = sum_t kappa_t (-1)^{j.z_t};
// for anti-Hermitian K, alpha(j) alpha(j ^ x) = -|alpha(j)|^2, so on every
// pair (j, j ^ x) exp(theta K) = cos(theta w) + sin(theta w) / w * K with
// w = |alpha(j)|: a rotation of determinant pairs.  Pairs never leave the
// sector when K conserves N_alpha and N_beta, so the state needs only
// C(norb, N_alpha) * C(norb, N_beta) amplitudes instead of 2^n.
//
// Writing z_t = z_0 ^ d_t, alpha(j) = s(j) A(j & m) with s(j) = (-1)^{j.z_0},
// m = x | (union of the d_t) and A(p) = sum_t kappa_t (-1)^{p.d_t}: the host
// tabulates A over the patterns p of m, checks anti-Hermiticity and sector
// conservation pattern by pattern, and emits one "entry" per active pattern
// pair (p, p ^ x).  The determinants of an entry are a product of an alpha
// string list and a beta string list (with partner ranks precomputed).
//
// <H> = sum_groups sum_{j, j^x in sector} conj(psi_j) psi_{j^x} T_g(j^x),
// T_g(j') = sum_{t in g} c_t (-1)^{j'.z_t}   (computer.cc:850-911 on the sector).
// For x != 0 the pairs (j, j^x) and (j^x, j) contribute complex-conjugate
// values (H_g is Hermitian), so only the half with the top bit of x clear in j
// is visited, with weight 2.  The pair space of every group is cut into work
// items of similar cost that run as independent workgroups over the chip.
//
// Layout: rank(j) = rank_alpha(alpha string) * D_beta + rank_beta(beta string),
// strings ranked in increasing numeric order; alpha = even bits, beta = odd
// bits (state_prep.py:5-11 occupation convention).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "qfe_device.h"
#include "qfe_internal.h"

namespace qfe {
namespace {

struct SecStep {  // one run of commuting exponentials sharing an X mask
    uint32_t e_first, e_count;
    int32_t param;
    uint32_t gen;
    double scale;
};

struct SecEntry {  // one class of determinant pairs of a run
    uint64_t x, z0;
    uint32_t la_off, la_len, lb_off, lb_len;
    double w, inv_lb;
    double2 ulo, uhi;  // A(p_lo) / w, A(p_hi) / w
    uint32_t step;
    uint32_t w_off;  // first pair word of the entry (when the template has pair words)
};

struct SecGroup {  // Hamiltonian terms sharing an X mask, on sector pairs
    uint64_t x;
    uint32_t la_off, la_len, lb_off, lb_len;  // half lists (x != 0) or full lists (x == 0)
    uint32_t t_first, t_count;
    uint32_t xa, xb;                          // compressed flips
    double inv_lb, weight;
};

struct SecItem {  // groups [g0, g1) whole, or pairs [p0, p1) of group g0 when g1 == g0 + 1
    uint32_t g0, g1, p0, p1;
};

struct SecTables {  // device view of a template (lives in device memory)
    uint64_t D, ref_rank;
    uint32_t Da, Db;
    const uint64_t* afull;
    const uint64_t* bfull;
    const uint32_t* astr;
    const uint32_t* bstr;
    const int32_t* arank;  // 2^norb entries, -1 outside the sector
    const int32_t* brank;
    const int2* lists;
    const SecStep* steps;
    const SecEntry* entries;
    const SecGroup* groups;
    const SecItem* items;
    const uint32_t* gen_step;  // steps of generator k: [gen_step[k], gen_step[k+1])
    const uint64_t* tz;
    const double2* tph;    // phase of each term
    const double2* tcoef;  // phase * coeff (fixed Hamiltonian)
    const uint64_t* map_off;
    const uint32_t* map_idx;
    const double2* map_val;
    const uint2* hchunks;  // groups [x, y) staged together by the sigma kernel
    const uint2* words;    // pair words of every entry, in entry order (null: use the lists)
    uint32_t nsteps, nentries, ngroups, nterms, nitems, ngen, nhchunks;
};

struct SlotDesc {
    const SecTables* t;
    double2* psi;
    double2* sig;   // H psi (gradients only)
    double2* coef;  // per-request coefficients (null: t->tcoef)
    double2* cs;    // (cos, sin) per entry
    double* part;   // per work item energy partials
    double* gacc;   // per generator <sigma|K|psi> (gradients only)
    const double* params;
    const double* inputs;
    double* out;
};

struct PoolOp {  // entries [e0, e1) of one pool operator
    uint32_t e0, e1;
};

constexpr uint32_t kTableLds = 1024;  // alpha + beta string tables staged in LDS when they fit
constexpr uint32_t kTermLds = 1024;   // Hamiltonian terms of one group staged in LDS when they fit
constexpr int kUnroll = 4;            // independent pairs in flight per thread
constexpr int kItemThreads = 256;

__device__ __forceinline__ void divmod_u(uint32_t i, uint32_t d, double inv, uint32_t& q, uint32_t& r) {
    int64_t qq = (int64_t)((double)i * inv);
    int64_t rr = (int64_t)i - qq * (int64_t)d;
    if (rr < 0) {
        qq -= 1;
        rr += d;
    } else if (rr >= (int64_t)d) {
        qq += 1;
        rr -= d;
    }
    q = (uint32_t)qq;
    r = (uint32_t)rr;
}

// Largest s with prefix[s] <= b (prefix[0] = 0, n + 1 entries).
__device__ __forceinline__ uint32_t find_slot(const uint32_t* __restrict__ prefix, uint32_t n, uint32_t b) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ---- 1. prologue: per-request Hamiltonian coefficients and rotation angles --------------
__global__ __launch_bounds__(256) void k_sec_prologue(const SlotDesc* __restrict__ slots, int need_coef) {
    const SlotDesc sd = slots[blockIdx.x];
    const SecTables* __restrict__ T = sd.t;
    if (need_coef && sd.coef) {
        for (uint32_t k = threadIdx.x; k < T->nterms; k += blockDim.x) {
            double cr = 0.0, ci = 0.0;
            const uint64_t e1 = T->map_off[k + 1];
            for (uint64_t e = T->map_off[k]; e < e1; ++e) {
                const double v = sd.inputs[T->map_idx[e]];
                const double2 f = T->map_val[e];
                cr += f.x * v;
                ci += f.y * v;
            }
            sd.coef[k] = cmul(T->tph[k], make_double2(cr, ci));
        }
    }
    for (uint32_t e = threadIdx.x; e < T->nentries; e += blockDim.x) {
        const SecEntry& en = T->entries[e];
        const SecStep& st = T->steps[en.step];
        double s, c;
        sincos(st.scale * sd.params[st.param] * en.w, &s, &c);
        sd.cs[e] = make_double2(c, s);
    }
}

// One pair (r1, r2) of an entry with old values (a, b), j1 the determinant
// of r1: psi_1' = c a + s s(j2) uhi b, psi_2' = c b + s s(j1) ulo a.  Every
// kernel that rotates pairs calls this, so a pair's result does not depend on
// which kernel or launch shape applied it.
__device__ __forceinline__ void pair_rotate_signed(double2* __restrict__ psi, uint64_t r1, uint64_t r2, double2 a,
                                                   double2 b, double c, double s1, double s2, double2 ulo, double2 uhi) {
    const double2 hb = cmul(uhi, b), la = cmul(ulo, a);
    psi[r1] = make_double2(c * a.x + s2 * hb.x, c * a.y + s2 * hb.y);
    psi[r2] = make_double2(c * b.x + s1 * la.x, c * b.y + s1 * la.y);
}
__device__ __forceinline__ void pair_rotate(double2* __restrict__ psi, uint64_t r1, uint64_t r2, double2 a, double2 b,
                                            uint64_t j1, uint64_t x, uint64_t z0, double c, double s, double2 ulo,
                                            double2 uhi) {
    pair_rotate_signed(psi, r1, r2, a, b, c, s * parity_sign(j1 & z0), s * parity_sign((j1 ^ x) & z0), ulo, uhi);
}

// Pair word of a sector pair: .x = r1 | [parity of (j1 ^ x) & z0] << 31,
// .y = r2 | [parity of j1 & z0] << 31, j1 the determinant of rank r1.  One
// 8-byte load per pair instead of two list loads, two string-table loads and
// a division; ranks stay below 2^31 (at most C(16, 8)^2 determinants).
constexpr uint32_t kWordRank = 0x7fffffffu;
__device__ __forceinline__ double word_sign(uint32_t v, double s) { return (v >> 31) ? -s : s; }

// <sigma|K|psi> contribution of one pair (the entry_overlap formula), with
// s1 = w (-1)^{|j1 & z0|} and s2 = w (-1)^{|(j1 ^ x) & z0|}.
__device__ __forceinline__ double pair_overlap_signed(const SecEntry& en, double s1, double s2, double2 pa, double2 pb,
                                                      double2 sa, double2 sb) {
    const double2 k1 = cmul(en.uhi, pb), k2 = cmul(en.ulo, pa);
    return s2 * (sa.x * k1.x + sa.y * k1.y) + s1 * (sb.x * k2.x + sb.y * k2.y);
}
__device__ __forceinline__ double pair_overlap(const SecEntry& en, uint64_t j1, double2 pa, double2 pb, double2 sa,
                                               double2 sb) {
    return pair_overlap_signed(en, en.w * parity_sign(j1 & en.z0), en.w * parity_sign((j1 ^ en.x) & en.z0), pa, pb, sa,
                               sb);
}

// v with its sign bit flipped when neg = 1u << 31: adding flip_sign(c, neg) is
// bit-identical to fma(+-1.0, c, acc), with one integer op for the sign.
// Determinant indices and Z masks fit 32 bits (at most 32 qubits).
__device__ __forceinline__ double flip_sign(double v, uint32_t neg) {
    return __hiloint2double(__double2hiint(v) ^ (int)neg, __double2loint(v));
}

// Rotate every pair of entry `en` by (c, sgn * s); pairs of one step are
// disjoint, so kUnroll of them are in flight per thread (loads first).
template <int THREADS>
__device__ __forceinline__ void rotate_entry(const SecTables* __restrict__ T, const SecEntry& en, double c, double s,
                                             double2* __restrict__ psi, const uint64_t* __restrict__ af,
                                             const uint64_t* __restrict__ bf, uint32_t tid) {
    const uint64_t Db = T->Db;
    const uint32_t lb = en.lb_len, total = en.la_len * lb;
    const uint64_t x = en.x, z0 = en.z0;
    const double2 ulo = en.ulo, uhi = en.uhi;
    const double inv = en.inv_lb;
    for (uint32_t i0 = tid; i0 < total; i0 += kUnroll * THREADS) {
        uint64_t r1[kUnroll], r2[kUnroll], j1[kUnroll];
        double2 a[kUnroll], b[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const uint32_t i = i0 + u * THREADS;
            if (i < total) {
                uint32_t ia, ib;
                divmod_u(i, lb, inv, ia, ib);
                const int2 pa = T->lists[en.la_off + ia];
                const int2 pb = T->lists[en.lb_off + ib];
                r1[u] = (uint64_t)pa.x * Db + (uint64_t)pb.x;
                r2[u] = (uint64_t)pa.y * Db + (uint64_t)pb.y;
                j1[u] = af[pa.x] | bf[pb.x];
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (i0 + u * THREADS < total) {
                a[u] = psi[r1[u]];
                b[u] = psi[r2[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (i0 + u * THREADS < total) pair_rotate(psi, r1[u], r2[u], a[u], b[u], j1[u], x, z0, c, s, ulo, uhi);
        }
    }
}

// Stage the string tables into LDS when they fit; returns the table pointers.
__device__ __forceinline__ void stage_tables(const SecTables* __restrict__ T, uint64_t* tab, const uint64_t*& af,
                                             const uint64_t*& bf, uint32_t tid, uint32_t nthreads) {
    if (T->Da + T->Db <= kTableLds) {
        for (uint32_t k = tid; k < T->Da; k += nthreads) tab[k] = T->afull[k];
        for (uint32_t k = tid; k < T->Db; k += nthreads) tab[T->Da + k] = T->bfull[k];
        af = tab;
        bf = tab + T->Da;
    } else {
        af = T->afull;
        bf = T->bfull;
    }
}

// ---- 2. ansatz: U|ref> by one workgroup per request ------------------------------------
// With PSI_LDS the state lives in LDS (dynamic LDS = D * 16 bytes + string
// tables) and is written to the request's buffer at the end.
// Pair words of chunk [base, base + K * THREADS) of an entry with P pairs.
template <int K, int THREADS>
__device__ __forceinline__ void load_words(uint2 (&w)[K], const uint2* __restrict__ src, uint32_t P, uint32_t base,
                                           uint32_t tid) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t i = base + k * THREADS + tid;
        w[k] = i < P ? src[i] : make_uint2(0u, 0u);
    }
}

// Position of a chunk of K * THREADS pairs: entry e, first pair base (e ==
// ne: past the last entry).
struct WordCursor {
    uint32_t e, base;
};

template <int K, int THREADS>
__device__ __forceinline__ WordCursor next_chunk(const SecEntry* __restrict__ ent, uint32_t ne, WordCursor c) {
    if (c.e >= ne) return c;
    c.base += K * THREADS;
    if (c.base >= ent[c.e].la_len * ent[c.e].lb_len) {
        ++c.e;
        c.base = 0;
    }
    return c;
}

template <int K, int THREADS>
__device__ __forceinline__ void load_chunk(uint2 (&w)[K], const SecEntry* __restrict__ ent, const uint2* __restrict__ words,
                                           uint32_t ne, WordCursor c, uint32_t tid) {
    if (c.e < ne) load_words<K, THREADS>(w, words + ent[c.e].w_off, ent[c.e].la_len * ent[c.e].lb_len, c.base, tid);
}

// The entries in order from pair words, K pairs per thread per chunk.  The
// words of the chunk DEPTH ahead are loaded while the current one rotates, so
// a step's dependent chain is the amplitude loads and stores (DEPTH = 2 when
// the state is in LDS and a step is shorter than a load from memory).
template <int K, int THREADS, int DEPTH>
__device__ __forceinline__ void ansatz_words(const SlotDesc& sd, const SecTables* __restrict__ T, double2* __restrict__ psi,
                                             uint32_t tid) {
    static_assert(DEPTH == 1 || DEPTH == 2, "prefetch depth 1 or 2");
    const SecEntry* __restrict__ ent = T->entries;
    const uint2* __restrict__ words = T->words;
    const uint32_t ne = T->nentries;
    WordCursor cur{0u, 0u};
    WordCursor c1 = next_chunk<K, THREADS>(ent, ne, cur);
    uint2 w[K], w1[K], w2[K];
    load_chunk<K, THREADS>(w, ent, words, ne, cur, tid);
    if constexpr (DEPTH == 2) load_chunk<K, THREADS>(w1, ent, words, ne, c1, tid);
    WordCursor c2 = next_chunk<K, THREADS>(ent, ne, c1);
    __syncthreads();
    while (cur.e < ne) {
        const uint32_t e = cur.e, P = ent[e].la_len * ent[e].lb_len;
        const double2 cs = sd.cs[e], ulo = ent[e].ulo, uhi = ent[e].uhi;
        double2 a[K], b[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (cur.base + k * THREADS + tid < P) {
                a[k] = psi[w[k].x & kWordRank];
                b[k] = psi[w[k].y & kWordRank];
            }
        }
        if constexpr (DEPTH == 2)
            load_chunk<K, THREADS>(w2, ent, words, ne, c2, tid);
        else
            load_chunk<K, THREADS>(w1, ent, words, ne, c1, tid);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (cur.base + k * THREADS + tid < P)
                pair_rotate_signed(psi, w[k].x & kWordRank, w[k].y & kWordRank, a[k], b[k], cs.x,
                                   word_sign(w[k].y, cs.y), word_sign(w[k].x, cs.y), ulo, uhi);
            w[k] = w1[k];
            if constexpr (DEPTH == 2) w1[k] = w2[k];
        }
        if (cur.base + K * THREADS >= P && (e + 1 == ne || ent[e + 1].step != ent[e].step)) __syncthreads();
        cur = c1;
        c1 = c2;
        c2 = next_chunk<K, THREADS>(ent, ne, c2);
    }
}

template <bool PSI_LDS, int THREADS, int K>
__global__ __launch_bounds__(THREADS) void k_sec_ansatz(const SlotDesc* __restrict__ slots) {
    extern __shared__ double2 sh[];
    const SlotDesc sd = slots[blockIdx.x];
    const SecTables* __restrict__ T = sd.t;
    const uint32_t tid = threadIdx.x;
    const uint64_t D = T->D;
    double2* __restrict__ psi = PSI_LDS ? sh : sd.psi;
    for (uint64_t r = tid; r < D; r += THREADS) psi[r] = make_double2(r == T->ref_rank ? 1.0 : 0.0, 0.0);
    if (T->words) {
        ansatz_words<K, THREADS, PSI_LDS ? 2 : 1>(sd, T, psi, tid);
    } else {
        const uint64_t *af, *bf;
        stage_tables(T, (uint64_t*)(sh + (PSI_LDS ? D : 0)), af, bf, tid, THREADS);
        __syncthreads();
        for (uint32_t s = 0; s < T->nsteps; ++s) {
            const uint32_t e0 = T->steps[s].e_first, e1 = e0 + T->steps[s].e_count;
            for (uint32_t e = e0; e < e1; ++e) {
                const double2 cs = sd.cs[e];
                rotate_entry<THREADS>(T, T->entries[e], cs.x, cs.y, psi, af, bf, tid);
            }
            __syncthreads();
        }
    }
    if (PSI_LDS)
        for (uint64_t r = tid; r < D; r += THREADS) sd.psi[r] = psi[r];
}

// ---- 3. energy work items: one workgroup per (request, item) ---------------------------
__global__ __launch_bounds__(kItemThreads) void k_sec_energy(const SlotDesc* __restrict__ slots,
                                                             const uint32_t* __restrict__ prefix, uint32_t nslots) {
    __shared__ uint32_t tz_s[kTermLds];
    __shared__ double2 c_s[kTermLds];
    __shared__ double2 red[16];
    const uint32_t slot = find_slot(prefix, nslots, blockIdx.x);
    const SlotDesc sd = slots[slot];
    const SecTables* __restrict__ T = sd.t;
    const SecItem it = T->items[blockIdx.x - prefix[slot]];
    const double2* __restrict__ coef = sd.coef ? sd.coef : T->tcoef;
    const double2* __restrict__ psi = sd.psi;
    const uint32_t tid = threadIdx.x;
    const uint64_t Db = T->Db;
    const bool whole = it.g1 != it.g0 + 1;
    double acc = 0.0;
    for (uint32_t g = it.g0; g < it.g1; ++g) {
        const SecGroup& gr = T->groups[g];
        const uint32_t t0 = gr.t_first, nt = gr.t_count;
        const bool lds = nt <= kTermLds;
        if (lds) {
            __syncthreads();
            for (uint32_t k = tid; k < nt; k += kItemThreads) {
                tz_s[k] = (uint32_t)T->tz[t0 + k];
                c_s[k] = coef[t0 + k];
            }
            __syncthreads();
        }
        const uint32_t lb = gr.lb_len;
        const uint32_t p0 = whole ? 0u : it.p0, p1 = whole ? gr.la_len * lb : it.p1;
        const uint32_t x = (uint32_t)gr.x;
        double gacc = 0.0;
        for (uint32_t i0 = p0 + tid; i0 < p1; i0 += kUnroll * kItemThreads) {
            uint32_t j2[kUnroll];
            double2 v[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const uint32_t i = i0 + u * kItemThreads;
                j2[u] = 0;
                v[u] = make_double2(0.0, 0.0);
                if (i < p1) {
                    uint32_t ia, ib;
                    divmod_u(i, lb, gr.inv_lb, ia, ib);
                    const int2 pa = T->lists[gr.la_off + ia];
                    const int2 pb = T->lists[gr.lb_off + ib];
                    j2[u] = (uint32_t)(T->afull[pa.x] | T->bfull[pb.x]) ^ x;
                    v[u] = cconjmul(psi[(uint64_t)pa.x * Db + (uint64_t)pb.x], psi[(uint64_t)pa.y * Db + (uint64_t)pb.y]);
                }
            }
            double tr[kUnroll], ti[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) tr[u] = ti[u] = 0.0;
            for (uint32_t t = 0; t < nt; ++t) {
                const uint32_t z = lds ? tz_s[t] : (uint32_t)T->tz[t0 + t];
                const double2 c = lds ? c_s[t] : coef[t0 + t];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    const uint32_t neg = (uint32_t)__popc(j2[u] & z) << 31;
                    tr[u] += flip_sign(c.x, neg);
                    ti[u] += flip_sign(c.y, neg);
                }
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) gacc += v[u].x * tr[u] - v[u].y * ti[u];
        }
        acc += gr.weight * gacc;
    }
    const double2 tot = block_sum(make_double2(acc, 0.0), red);
    if (tid == 0) sd.part[blockIdx.x - prefix[slot]] = tot.x;
}

// ---- 4. per-request sum of the item partials in a fixed order ---------------------------
__global__ __launch_bounds__(256) void k_sec_sum(const SlotDesc* __restrict__ slots) {
    __shared__ double2 red[16];
    const SlotDesc sd = slots[blockIdx.x];
    const uint32_t n = sd.t->nitems;
    double acc = 0.0;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) acc += sd.part[k];
    const double2 tot = block_sum(make_double2(acc, 0.0), red);
    if (threadIdx.x == 0) *sd.out = tot.x;
}

// ---- gradients ---------------------------------------------------------------------------
// sigma = H psi, output-centric: each determinant gathers from every group
// whose flip keeps it in the sector.
constexpr uint32_t kHpsiGroups = 256;  // group headers staged per chunk
constexpr uint32_t kHpsiTerms = 1024;  // terms staged per chunk (larger chunks read terms from memory)

// PER determinants per thread: each staged term is read once for all of them
// (batched launches use 4; a single request keeps 1 so it fills the chip).
template <int PER>
__global__ __launch_bounds__(256) void k_sec_hpsi(const SlotDesc* __restrict__ slots) {
    constexpr int kHpsiPer = PER;
    constexpr uint32_t kHpsiSpan = 256 * PER;  // determinants per block
    __shared__ uint4 hdr[kHpsiGroups];  // xa, xb, first term, term count
    __shared__ uint32_t gx[kHpsiGroups];
    __shared__ uint32_t tzs[kHpsiTerms];
    __shared__ double2 cfs[kHpsiTerms];
    const SlotDesc sd = slots[blockIdx.y];
    const SecTables* __restrict__ T = sd.t;
    if ((uint64_t)blockIdx.x * kHpsiSpan >= T->D) return;  // the whole block: a smaller request of the launch
    uint32_t sa[kHpsiPer], sb[kHpsiPer];
    uint32_t j[kHpsiPer];
    bool live[kHpsiPer];
    double sr[kHpsiPer], si[kHpsiPer];
#pragma unroll
    for (int q = 0; q < kHpsiPer; ++q) {
        const uint64_t r = (uint64_t)blockIdx.x * kHpsiSpan + (uint64_t)q * 256 + threadIdx.x;
        live[q] = r < T->D;  // every thread stays for the staging barriers
        const uint32_t ra = live[q] ? (uint32_t)(r / T->Db) : 0u, rb = live[q] ? (uint32_t)(r % T->Db) : 0u;
        sa[q] = T->astr[ra];
        sb[q] = T->bstr[rb];
        j[q] = (uint32_t)(T->afull[ra] | T->bfull[rb]);
        sr[q] = si[q] = 0.0;
    }
    const double2* __restrict__ coef = sd.coef ? sd.coef : T->tcoef;
    for (uint32_t c = 0; c < T->nhchunks; ++c) {
        const uint2 ch = T->hchunks[c];
        const uint32_t t0 = T->groups[ch.x].t_first;
        const uint32_t t1 = T->groups[ch.y - 1].t_first + T->groups[ch.y - 1].t_count;
        const bool tl = t1 - t0 <= kHpsiTerms;
        __syncthreads();
        for (uint32_t g = ch.x + threadIdx.x; g < ch.y; g += blockDim.x) {
            const SecGroup& gr = T->groups[g];
            hdr[g - ch.x] = make_uint4(gr.xa, gr.xb, gr.t_first, gr.t_count);
            gx[g - ch.x] = (uint32_t)gr.x;
        }
        if (tl)
            for (uint32_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
                tzs[t - t0] = (uint32_t)T->tz[t];
                cfs[t - t0] = coef[t];
            }
        __syncthreads();
        for (uint32_t g = 0; g < ch.y - ch.x; ++g) {
            const uint4 h = hdr[g];
            const int pa = __popc(h.x), pb = __popc(h.y);
            bool on[kHpsiPer];
            bool any = false;
#pragma unroll
            for (int q = 0; q < kHpsiPer; ++q) {
                on[q] = live[q] && 2 * __popc(sa[q] & h.x) == pa && 2 * __popc(sb[q] & h.y) == pb;
                any |= on[q];
            }
            if (!any) continue;
            uint32_t j2[kHpsiPer];
            double tr[kHpsiPer], ti[kHpsiPer];
#pragma unroll
            for (int q = 0; q < kHpsiPer; ++q) {
                j2[q] = j[q] ^ gx[g];
                tr[q] = ti[q] = 0.0;
            }
            for (uint32_t t = h.z; t < h.z + h.w; ++t) {
                const uint32_t z = tl ? tzs[t - t0] : (uint32_t)T->tz[t];
                const double2 cf = tl ? cfs[t - t0] : coef[t];
#pragma unroll
                for (int q = 0; q < kHpsiPer; ++q) {
                    const uint32_t neg = (uint32_t)__popc(j2[q] & z) << 31;
                    tr[q] += flip_sign(cf.x, neg);
                    ti[q] += flip_sign(cf.y, neg);
                }
            }
#pragma unroll
            for (int q = 0; q < kHpsiPer; ++q) {
                if (!on[q]) continue;
                const uint64_t r2 = (uint64_t)T->arank[sa[q] ^ h.x] * T->Db + (uint64_t)T->brank[sb[q] ^ h.y];
                const double2 p = sd.psi[r2];
                sr[q] += tr[q] * p.x - ti[q] * p.y;
                si[q] += tr[q] * p.y + ti[q] * p.x;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kHpsiPer; ++q) {
        const uint64_t r = (uint64_t)blockIdx.x * kHpsiSpan + (uint64_t)q * 256 + threadIdx.x;
        if (live[q]) sd.sig[r] = make_double2(sr[q], si[q]);
    }
}

// Sum over the pairs of entry en of Re(conj(sig) K psi).
template <int THREADS>
__device__ __forceinline__ double entry_overlap(const SecTables* __restrict__ T, const SecEntry& en,
                                                const double2* __restrict__ sig, const double2* __restrict__ psi,
                                                uint32_t tid) {
    const uint64_t Db = T->Db;
    const uint32_t lb = en.lb_len, total = en.la_len * lb;
    double acc = 0.0;
    for (uint32_t i = tid; i < total; i += THREADS) {
        uint32_t ia, ib;
        divmod_u(i, lb, en.inv_lb, ia, ib);
        const int2 pa = T->lists[en.la_off + ia];
        const int2 pb = T->lists[en.lb_off + ib];
        const uint64_t r1 = (uint64_t)pa.x * Db + (uint64_t)pb.x, r2 = (uint64_t)pa.y * Db + (uint64_t)pb.y;
        const uint64_t j1 = T->afull[pa.x] | T->bfull[pb.x];
        const double s1 = en.w * parity_sign(j1 & en.z0), s2 = en.w * parity_sign((j1 ^ en.x) & en.z0);
        // (K psi)_1 = s(j2) w uhi psi_2, (K psi)_2 = s(j1) w ulo psi_1
        const double2 k1 = cmul(en.uhi, psi[r2]), k2 = cmul(en.ulo, psi[r1]);
        const double2 a = sig[r1], b = sig[r2];
        acc += s2 * (a.x * k1.x + a.y * k1.y) + s1 * (b.x * k2.x + b.y * k2.y);
    }
    return acc;
}

// Reverse sweep (one workgroup per request): E = Re<psi|sigma>, then for
// k = M-1 .. 0: <sigma_k|K_k|psi_k>, and un-apply generator k on psi and sigma.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_sec_reverse(const SlotDesc* __restrict__ slots) {
    __shared__ double2 red[16];
    const SlotDesc sd = slots[blockIdx.x];
    const SecTables* __restrict__ T = sd.t;
    const uint32_t tid = threadIdx.x;
    double2* __restrict__ psi = sd.psi;
    double2* __restrict__ sig = sd.sig;
    double e = 0.0;
    for (uint64_t r = tid; r < T->D; r += THREADS) e += psi[r].x * sig[r].x + psi[r].y * sig[r].y;
    const double2 et = block_sum(make_double2(e, 0.0), red);
    if (tid == 0) *sd.out = et.x;
    for (uint32_t k = T->ngen; k-- > 0;) {
        const uint32_t s0 = T->gen_step[k], s1 = T->gen_step[k + 1];
        double acc = 0.0;
        if (s1 == s0 + 1 && T->words) {
            // one step from pair words: each pair's four old values give its
            // overlap term and its un-rotated psi and sigma pair in one pass
            for (uint32_t en = T->steps[s0].e_first; en < T->steps[s0].e_first + T->steps[s0].e_count; ++en) {
                const SecEntry& E = T->entries[en];
                const double2 cs = sd.cs[en];
                const uint2* __restrict__ wp = T->words + E.w_off;
                const uint32_t total = E.la_len * E.lb_len;
                for (uint32_t i = tid; i < total; i += THREADS) {
                    const uint2 w = wp[i];
                    const uint64_t r1 = w.x & kWordRank, r2 = w.y & kWordRank;
                    const double2 p1 = psi[r1], p2 = psi[r2], g1 = sig[r1], g2 = sig[r2];
                    acc += pair_overlap_signed(E, word_sign(w.y, E.w), word_sign(w.x, E.w), p1, p2, g1, g2);
                    const double n1 = word_sign(w.y, -cs.y), n2 = word_sign(w.x, -cs.y);
                    pair_rotate_signed(psi, r1, r2, p1, p2, cs.x, n1, n2, E.ulo, E.uhi);
                    pair_rotate_signed(sig, r1, r2, g1, g2, cs.x, n1, n2, E.ulo, E.uhi);
                }
            }
            const double2 tot = block_sum(make_double2(acc, 0.0), red);  // its barriers order the writes above
            if (tid == 0) sd.gacc[k] = tot.x;
            continue;
        }
        if (s1 == s0 + 1) {
            // one step: each pair's four old values give its overlap term and its
            // un-rotated psi and sigma pair in one pass (pairs of a step are disjoint)
            const uint64_t Db = T->Db;
            for (uint32_t en = T->steps[s0].e_first; en < T->steps[s0].e_first + T->steps[s0].e_count; ++en) {
                const SecEntry& E = T->entries[en];
                const double2 cs = sd.cs[en];
                const uint32_t lb = E.lb_len, total = E.la_len * lb;
                for (uint32_t i = tid; i < total; i += THREADS) {
                    uint32_t ia, ib;
                    divmod_u(i, lb, E.inv_lb, ia, ib);
                    const int2 pa = T->lists[E.la_off + ia];
                    const int2 pb = T->lists[E.lb_off + ib];
                    const uint64_t r1 = (uint64_t)pa.x * Db + (uint64_t)pb.x, r2 = (uint64_t)pa.y * Db + (uint64_t)pb.y;
                    const uint64_t j1 = T->afull[pa.x] | T->bfull[pb.x];
                    const double2 p1 = psi[r1], p2 = psi[r2], g1 = sig[r1], g2 = sig[r2];
                    acc += pair_overlap(E, j1, p1, p2, g1, g2);
                    pair_rotate(psi, r1, r2, p1, p2, j1, E.x, E.z0, cs.x, -cs.y, E.ulo, E.uhi);
                    pair_rotate(sig, r1, r2, g1, g2, j1, E.x, E.z0, cs.x, -cs.y, E.ulo, E.uhi);
                }
            }
            const double2 tot = block_sum(make_double2(acc, 0.0), red);  // its barriers order the writes above
            if (tid == 0) sd.gacc[k] = tot.x;
            continue;
        }
        for (uint32_t s = s0; s < s1; ++s)
            for (uint32_t en = T->steps[s].e_first; en < T->steps[s].e_first + T->steps[s].e_count; ++en)
                acc += entry_overlap<THREADS>(T, T->entries[en], sig, psi, tid);
        const double2 tot = block_sum(make_double2(acc, 0.0), red);  // its barriers end every read above
        if (tid == 0) sd.gacc[k] = tot.x;
        if (k == 0) break;
        for (uint32_t s = s1; s-- > s0;) {
            const uint32_t e0 = T->steps[s].e_first, e1 = e0 + T->steps[s].e_count;
            for (uint32_t en = e0; en < e1; ++en) {
                const double2 cs = sd.cs[en];
                rotate_entry<THREADS>(T, T->entries[en], cs.x, -cs.y, psi, T->afull, T->bfull, tid);
                rotate_entry<THREADS>(T, T->entries[en], cs.x, -cs.y, sig, T->afull, T->bfull, tid);
            }
            __syncthreads();
        }
    }
}

// Pool gradients: 2 Re<H psi|A_k psi> for the operators of one batch, one
// workgroup per operator (entries and lists of the pool's own tables).
__global__ __launch_bounds__(256) void k_sec_pool(const SecTables* __restrict__ T, const SecEntry* __restrict__ pe,
                                                  const int2* __restrict__ plists, const PoolOp* __restrict__ ops,
                                                  const double2* __restrict__ sig, const double2* __restrict__ psi,
                                                  double* __restrict__ out) {
    __shared__ double2 red[16];
    const PoolOp op = ops[blockIdx.x];
    const uint64_t Db = T->Db;
    double acc = 0.0;
    for (uint32_t e = op.e0; e < op.e1; ++e) {
        const SecEntry en = pe[e];
        const uint32_t lb = en.lb_len, total = en.la_len * lb;
        for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
            uint32_t ia, ib;
            divmod_u(i, lb, en.inv_lb, ia, ib);
            const int2 pa = plists[en.la_off + ia];
            const int2 pb = plists[en.lb_off + ib];
            const uint64_t r1 = (uint64_t)pa.x * Db + (uint64_t)pb.x, r2 = (uint64_t)pa.y * Db + (uint64_t)pb.y;
            const uint64_t j1 = T->afull[pa.x] | T->bfull[pb.x];
            const double s1 = en.w * parity_sign(j1 & en.z0), s2 = en.w * parity_sign((j1 ^ en.x) & en.z0);
            const double2 k1 = cmul(en.uhi, psi[r2]), k2 = cmul(en.ulo, psi[r1]);
            const double2 a = sig[r1], b = sig[r2];
            acc += s2 * (a.x * k1.x + a.y * k1.y) + s1 * (b.x * k2.x + b.y * k2.y);
        }
    }
    const double2 tot = block_sum(make_double2(acc, 0.0), red);
    if (threadIdx.x == 0) out[blockIdx.x] = 2.0 * tot.x;
}

// ---- grid-wide single-request kernels (large sectors) ------------------------------------
// A lone request of a large sector would run its ansatz and reverse sweep in
// one workgroup.  These kernels spread one step (or one generator) over the
// whole grid instead; the host launches one per step, in order, so kernel
// boundaries order the steps.  Pair t of a step is taken in entry order.
constexpr int kWideTPB = 256;

__device__ __forceinline__ bool step_pair(const SecTables* __restrict__ T, uint32_t e, uint32_t e_end, uint32_t t,
                                          uint32_t& entry, uint32_t& idx) {
    for (; e < e_end; ++e) {
        const uint32_t cnt = T->entries[e].la_len * T->entries[e].lb_len;
        if (t < cnt) {
            entry = e;
            idx = t;
            return true;
        }
        t -= cnt;
    }
    return false;
}

__device__ __forceinline__ void pair_of(const SecTables* __restrict__ T, const SecEntry& en, uint32_t i, uint64_t& r1,
                                        uint64_t& r2, uint64_t& j1) {
    uint32_t ia, ib;
    divmod_u(i, en.lb_len, en.inv_lb, ia, ib);
    const int2 pa = T->lists[en.la_off + ia];
    const int2 pb = T->lists[en.lb_off + ib];
    r1 = (uint64_t)pa.x * T->Db + (uint64_t)pb.x;
    r2 = (uint64_t)pa.y * T->Db + (uint64_t)pb.y;
    j1 = T->afull[pa.x] | T->bfull[pb.x];
}

// Rotate the pairs of step `step` by (c, s) (inverse: (c, -s)); `both` also rotates sigma.
__global__ __launch_bounds__(kWideTPB) void k_sec_step(const SlotDesc* __restrict__ slots, uint32_t step, int inverse,
                                                       int both) {
    const SlotDesc sd = slots[0];
    const SecTables* __restrict__ T = sd.t;
    const SecStep st = T->steps[step];
    uint32_t e, i;
    if (!step_pair(T, st.e_first, st.e_first + st.e_count, blockIdx.x * kWideTPB + threadIdx.x, e, i)) return;
    const SecEntry& en = T->entries[e];
    const double2 cs = sd.cs[e];
    const double sn = inverse ? -cs.y : cs.y;
    uint64_t r1, r2, j1;
    pair_of(T, en, i, r1, r2, j1);
    pair_rotate(sd.psi, r1, r2, sd.psi[r1], sd.psi[r2], j1, en.x, en.z0, cs.x, sn, en.ulo, en.uhi);
    if (both) pair_rotate(sd.sig, r1, r2, sd.sig[r1], sd.sig[r2], j1, en.x, en.z0, cs.x, sn, en.ulo, en.uhi);
}


// Generator with one step: its overlap and its un-rotation of psi and sigma
// read the same four old values of each pair, so one pass does both.
// partial[blockIdx.x] receives the block's overlap sum.
__global__ __launch_bounds__(kWideTPB) void k_sec_genrev(const SlotDesc* __restrict__ slots, uint32_t step,
                                                         double* __restrict__ partial) {
    __shared__ double2 red[16];
    const SlotDesc sd = slots[0];
    const SecTables* __restrict__ T = sd.t;
    const SecStep st = T->steps[step];
    double acc = 0.0;
    uint32_t e, i;
    if (step_pair(T, st.e_first, st.e_first + st.e_count, blockIdx.x * kWideTPB + threadIdx.x, e, i)) {
        const SecEntry& en = T->entries[e];
        const double2 cs = sd.cs[e];
        uint64_t r1, r2, j1;
        pair_of(T, en, i, r1, r2, j1);
        const double2 pa = sd.psi[r1], pb = sd.psi[r2], sa = sd.sig[r1], sb = sd.sig[r2];
        acc = pair_overlap(en, j1, pa, pb, sa, sb);
        pair_rotate(sd.psi, r1, r2, pa, pb, j1, en.x, en.z0, cs.x, -cs.y, en.ulo, en.uhi);
        pair_rotate(sd.sig, r1, r2, sa, sb, j1, en.x, en.z0, cs.x, -cs.y, en.ulo, en.uhi);
    }
    const double2 tot = block_sum(make_double2(acc, 0.0), red);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot.x;
}

// Generator with several steps: the overlap over all of its entries first
// (steps [s0, s1), their entries contiguous), un-rotation by k_sec_step after.
__global__ __launch_bounds__(kWideTPB) void k_sec_genov(const SlotDesc* __restrict__ slots, uint32_t s0, uint32_t s1,
                                                        double* __restrict__ partial) {
    __shared__ double2 red[16];
    const SlotDesc sd = slots[0];
    const SecTables* __restrict__ T = sd.t;
    const uint32_t e0 = T->steps[s0].e_first, e1 = T->steps[s1 - 1].e_first + T->steps[s1 - 1].e_count;
    double acc = 0.0;
    uint32_t e, i;
    if (step_pair(T, e0, e1, blockIdx.x * kWideTPB + threadIdx.x, e, i)) {
        const SecEntry& en = T->entries[e];
        uint64_t r1, r2, j1;
        pair_of(T, en, i, r1, r2, j1);
        acc = pair_overlap(en, j1, sd.psi[r1], sd.psi[r2], sd.sig[r1], sd.sig[r2]);
    }
    const double2 tot = block_sum(make_double2(acc, 0.0), red);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot.x;
}

// Re<psi|sigma> block partials (the energy of the gradient path).
__global__ __launch_bounds__(kWideTPB) void k_sec_edot(const SlotDesc* __restrict__ slots, double* __restrict__ partial) {
    __shared__ double2 red[16];
    const SlotDesc sd = slots[0];
    const uint64_t D = sd.t->D;
    const uint64_t r = (uint64_t)blockIdx.x * kWideTPB + threadIdx.x;
    double e = 0.0;
    if (r < D) e = sd.psi[r].x * sd.sig[r].x + sd.psi[r].y * sd.sig[r].y;
    const double2 tot = block_sum(make_double2(e, 0.0), red);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot.x;
}

// Fixed-order sums: out = sum of partial[seg[0]], gacc[k] = sum of partial[seg[k + 1]].
__global__ __launch_bounds__(kWideTPB) void k_sec_wsum(const SlotDesc* __restrict__ slots, const double* __restrict__ partial,
                                                       const uint2* __restrict__ seg, uint32_t nseg) {
    const uint32_t k = blockIdx.x * kWideTPB + threadIdx.x;
    if (k >= nseg) return;
    const uint2 sg = seg[k];
    double acc = 0.0;
    for (uint32_t q = 0; q < sg.y; ++q) acc += partial[sg.x + q];
    if (k == 0)
        *slots[0].out = acc;
    else
        slots[0].gacc[k - 1] = acc;
}

// The reference determinant (grid over the sector).
__global__ __launch_bounds__(kWideTPB) void k_sec_ref(const SlotDesc* __restrict__ slots) {
    const SlotDesc sd = slots[0];
    const uint64_t r = (uint64_t)blockIdx.x * kWideTPB + threadIdx.x;
    if (r < sd.t->D) sd.psi[r] = make_double2(r == sd.t->ref_rank ? 1.0 : 0.0, 0.0);
}

// out[i] = sigma[rank[i]] (PQE residual amplitudes).
__global__ void k_sec_gather(const SlotDesc* __restrict__ slots, const uint64_t* __restrict__ rank, uint64_t count,
                             double2* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = slots[0].sig[rank[i]];
}

// Scatter a sector state into a zeroed full state.
__global__ void k_sector_scatter(const double2* __restrict__ psi, uint64_t D, uint32_t Db, const uint64_t* __restrict__ afull,
                                 const uint64_t* __restrict__ bfull, double2* __restrict__ out) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= D) return;
    out[afull[r / Db] | bfull[r % Db]] = psi[r];
}

// Ansatz launch shapes by size tier: 0 = state in LDS, 256 threads; 1 = state
// in LDS, 512 threads; 2 = state in device memory, 1024 threads.
constexpr size_t kTier0Lds = 40 << 10;
constexpr int kTier2Threads = 1024;

int launch_ansatz(int tier, unsigned nblocks, size_t lds, hipStream_t st, const SlotDesc* d) {
    if (tier == 0)
        hipLaunchKernelGGL((k_sec_ansatz<true, 256, 4>), dim3(nblocks), dim3(256), lds, st, d);
    else if (tier == 1)
        hipLaunchKernelGGL((k_sec_ansatz<true, 512, 4>), dim3(nblocks), dim3(512), lds, st, d);
    else
        hipLaunchKernelGGL((k_sec_ansatz<false, kTier2Threads, 5>), dim3(nblocks), dim3(kTier2Threads), lds, st, d);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

// Reverse sweep launch shape follows the ansatz tier (256 / 512 / 1024 threads).
int launch_reverse(int tier, unsigned nblocks, hipStream_t st, const SlotDesc* d) {
    if (tier == 0)
        hipLaunchKernelGGL((k_sec_reverse<256>), dim3(nblocks), dim3(256), 0, st, d);
    else if (tier == 1)
        hipLaunchKernelGGL((k_sec_reverse<512>), dim3(nblocks), dim3(512), 0, st, d);
    else
        hipLaunchKernelGGL((k_sec_reverse<1024>), dim3(nblocks), dim3(1024), 0, st, d);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

// ---- host-side construction -------------------------------------------------------------

inline uint32_t compress(uint64_t v, int parity, int norb) {
    uint32_t out = 0;
    for (int i = 0; i < norb; ++i) out |= (uint32_t)((v >> (2 * i + parity)) & 1ull) << i;
    return out;
}
inline uint64_t spread(uint32_t v, int parity, int norb) {
    uint64_t out = 0;
    for (int i = 0; i < norb; ++i) out |= (uint64_t)((v >> i) & 1u) << (2 * i + parity);
    return out;
}
inline int popc32(uint32_t v) { return __builtin_popcount(v); }
inline int popc64(uint64_t v) { return __builtin_popcountll(v); }
inline double2 cplx_mul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
inline double cplx_abs(double2 a) { return std::hypot(a.x, a.y); }
inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Space {
    int norb = 0;
    std::vector<uint32_t> str[2];
    std::vector<int32_t> rank[2];
    void build(int nb, int na_, int nb_) {
        norb = nb;
        const int ne[2] = {na_, nb_};
        for (int s = 0; s < 2; ++s) {
            str[s].clear();
            rank[s].assign((size_t)1 << norb, -1);
            for (uint32_t v = 0; v < (1u << norb); ++v)
                if (popc32(v) == ne[s]) {
                    rank[s][v] = (int32_t)str[s].size();
                    str[s].push_back(v);
                }
        }
    }
};

// Lists of (rank, partner rank) string pairs.  kind 0: (a & mask) == pattern;
// kind 1: popcount(a & mask) == popcount(mask) / 2; kind 2: kind 1 and
// (a & clear) == 0.  The partner is a ^ flip.
struct ListPool {
    std::vector<int2> pool;
    std::map<std::tuple<int, int, uint32_t, uint32_t, uint32_t>, std::pair<uint32_t, uint32_t>> index;
    std::pair<uint32_t, uint32_t> get(const Space& sp, int s, int kind, uint32_t mask, uint32_t pattern, uint32_t flip) {
        auto key = std::make_tuple(s, kind, mask, pattern, flip);
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        const uint32_t off = (uint32_t)pool.size();
        for (size_t r = 0; r < sp.str[s].size(); ++r) {
            const uint32_t a = sp.str[s][r];
            bool ok;
            if (kind == 0)
                ok = (a & mask) == pattern;
            else
                ok = 2 * popc32(a & mask) == popc32(mask) && (kind == 1 || (a & pattern) == 0);
            if (!ok) continue;
            const int32_t partner = sp.rank[s][a ^ flip];
            if (partner < 0) continue;
            pool.push_back(make_int2((int)r, partner));
        }
        auto v = std::make_pair(off, (uint32_t)pool.size() - off);
        index.emplace(key, v);
        return v;
    }
};

// Split generator k (terms [b, e)) into runs and entries; appends to steps
// and entries.  Used for ansatz generators and for pool operators.
int build_runs(const Space& sp, ListPool& lp, int norb, uint64_t lim, const qfe_pauli* terms, uint64_t b, uint64_t e,
               int32_t param, double scale, uint32_t gen, std::vector<SecStep>& steps, std::vector<SecEntry>& entries,
               const char* who) {
    while (b < e) {
        const uint64_t x = terms[b].x, z0 = terms[b].z;
        uint64_t r = b;
        while (r < e && terms[r].x == x && (popc64(x & (terms[r].z ^ z0)) & 1) == 0) ++r;
        uint64_t m = x;
        for (uint64_t k = b; k < r; ++k) {
            if ((terms[k].x | terms[k].z) & ~lim) return fail(QFE_ERR_ARG, std::string(who) + ": term acts on a qubit out of range");
            m |= terms[k].z ^ z0;
        }
        if (popc64(m) > 20) return fail(QFE_ERR_ARG, std::string(who) + ": a run of commuting terms spans more than 20 qubits");
        double scale_abs = 0.0;
        std::vector<double2> kap(r - b);
        for (uint64_t k = b; k < r; ++k) {
            // exp(theta * coeff * P) with coeff purely imaginary, as the full-space program applies it
            kap[k - b] = cplx_mul(make_double2(0.0, terms[k].coeff.im), make_double2(terms[k].phase.re, terms[k].phase.im));
            scale_abs += cplx_abs(kap[k - b]);
        }
        const double tol = 1e-13 * (scale_abs > 0 ? scale_abs : 1.0);
        std::map<uint64_t, double2> A;  // A(p) over the submasks of m
        for (uint64_t p = m;; p = (p - 1) & m) {
            double2 s = make_double2(0.0, 0.0);
            for (uint64_t k = b; k < r; ++k) {
                const double sg = (popc64(p & (terms[k].z ^ z0)) & 1) ? -1.0 : 1.0;
                s.x += sg * kap[k - b].x;
                s.y += sg * kap[k - b].y;
            }
            if (cplx_abs(s) > tol) A.emplace(p, s);
            if (p == 0) break;
        }
        if (x == 0 && !A.empty()) return fail(QFE_ERR_ARG, std::string(who) + ": diagonal runs are not supported on the sector");
        SecStep st{};
        st.e_first = (uint32_t)entries.size();
        st.param = param;
        st.scale = scale;
        st.gen = gen;
        const double sx = (popc64(x & z0) & 1) ? -1.0 : 1.0;
        for (const auto& kv : A) {
            const uint64_t plo = kv.first, phi = plo ^ x;
            auto it = A.find(phi);
            if (it == A.end()) return fail(QFE_ERR_ARG, std::string(who) + ": a run is not anti-Hermitian");
            if (plo > phi) continue;
            const double2 alo = kv.second, ahi = it->second;
            const double w = cplx_abs(alo);
            const double2 prod = cplx_mul(alo, ahi);
            if (std::fabs(sx * prod.x + w * w) > 1e-12 * w * w || std::fabs(prod.y) > 1e-12 * w * w)
                return fail(QFE_ERR_ARG, std::string(who) + ": a run is not anti-Hermitian");
            const uint32_t pa = compress(plo, 0, norb), pb = compress(plo, 1, norb);
            const uint32_t qa = compress(phi, 0, norb), qb = compress(phi, 1, norb);
            if (popc32(pa) != popc32(qa) || popc32(pb) != popc32(qb))
                return fail(QFE_ERR_ARG, std::string(who) + ": a run does not conserve the alpha and beta particle numbers");
            const uint32_t ma = compress(m, 0, norb), mb = compress(m, 1, norb);
            const uint32_t xa = compress(x, 0, norb), xb = compress(x, 1, norb);
            auto la = lp.get(sp, 0, 0, ma, pa, xa);
            auto lb = lp.get(sp, 1, 0, mb, pb, xb);
            if (!la.second || !lb.second) continue;
            SecEntry en{};
            en.x = x;
            en.z0 = z0;
            en.la_off = la.first;
            en.la_len = la.second;
            en.lb_off = lb.first;
            en.lb_len = lb.second;
            en.w = w;
            en.inv_lb = 1.0 / (double)lb.second;
            en.ulo = make_double2(alo.x / w, alo.y / w);
            en.uhi = make_double2(ahi.x / w, ahi.y / w);
            en.step = (uint32_t)steps.size();
            entries.push_back(en);
        }
        st.e_count = (uint32_t)entries.size() - st.e_first;
        if (st.e_count) steps.push_back(st);
        b = r;
    }
    return QFE_OK;
}

// Packs host arrays into one device allocation at 256-byte offsets.
struct Packer {
    struct Piece {
        const void* src;
        size_t bytes, off;
    };
    std::vector<Piece> pieces;
    size_t total;
    explicit Packer(size_t head) : total(align256(head)) {}
    size_t add(const void* src, size_t bytes) {
        pieces.push_back({src, bytes, total});
        total += align256(bytes ? bytes : 1);
        return pieces.size() - 1;
    }
    void fill(std::vector<char>& img) const {
        img.assign(total, 0);
        for (const auto& p : pieces)
            if (p.bytes) std::memcpy(img.data() + p.off, p.src, p.bytes);
    }
};

}  // namespace
}  // namespace qfe

using namespace qfe;

struct qfe_template_s {
    int n = 0, norb = 0;
    uint64_t D = 0;
    uint32_t Da = 0, Db = 0;
    uint64_t nparam = 0, ninputs = 0, ngen = 0;
    std::vector<int32_t> gparam;
    std::vector<double> gscale;
    uint32_t nterms = 0, nentries = 0, nitems = 0;
    bool has_map = false;
    uint64_t cost = 0;     // work estimate for ordering slots in a launch
    int tier = 2;          // ansatz launch shape
    size_t lds_bytes = 0;  // dynamic LDS of the ansatz kernel
    Space sp;              // host copy of the string space (pool operators)
    void* dev = nullptr;   // all tables
    SecTables* dt = nullptr;
    const uint64_t* afull = nullptr;
    const uint64_t* bfull = nullptr;
    // per-request device bytes: psi, coef, cs, part (energy); + sig, gacc (gradient)
    size_t off_psi = 0, off_coef = 0, off_cs = 0, off_part = 0, off_sig = 0, off_gacc = 0;
    size_t work_bytes = 0, grad_bytes = 0, io_count = 0, request_bytes = 0, grad_request_bytes = 0;
    void* arena = nullptr;  // single-request arena: [desc][out][io][prefix][work incl. gradient parts]
    // grid-wide single-request path (large sectors): host step structure and
    // device [segments (uint2 x (ngen + 1))][block partials]
    std::vector<uint32_t> h_gen_step, step_pairs;
    std::vector<uint2> wide_seg;  // (first partial, count): the energy, then generator k
    uint32_t wide_parts = 0;
    void* wide = nullptr;
    // captured single-request sequences (energy, gradient) and their pinned
    // host block: [request head as uploaded][energy][generator overlaps]
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    void* gpin = nullptr;
    // host-side audit of each captured sequence: nodes, kernel nodes, kernel
    // nodes whose arguments were compared, memcpy nodes whose parameters were compared
    uint64_t audit[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
};

namespace qfe {
namespace {

int template_build(qfe_template t, int n, uint64_t ref, const qfe_pauli* gen_terms, const uint64_t* gen_off, uint64_t ngen,
                   const int32_t* gen_param, const double* gen_scale, uint64_t nparam, const qfe_pauli* ham, uint64_t nham,
                   const uint64_t* map_off, const uint32_t* map_idx, const qfe_c64* map_val, uint64_t ninputs) {
    if (n < 2 || n > 32 || (n & 1)) return fail(QFE_ERR_ARG, "qfe_template_create: nqubit must be even and in [2, 32]");
    const int norb = n / 2;
    const uint64_t lim = (1ull << n) - 1ull;
    if (ref & ~lim) return fail(QFE_ERR_ARG, "qfe_template_create: reference index out of range");
    const uint32_t refa = compress(ref, 0, norb), refb = compress(ref, 1, norb);
    Space& sp = t->sp;
    sp.build(norb, popc32(refa), popc32(refb));
    ListPool lp;
    t->n = n;
    t->norb = norb;
    t->Da = (uint32_t)sp.str[0].size();
    t->Db = (uint32_t)sp.str[1].size();
    t->D = (uint64_t)t->Da * t->Db;
    t->nparam = nparam;
    t->ninputs = ninputs;
    t->ngen = ngen;
    t->gparam.assign(gen_param, gen_param + ngen);
    t->gscale.assign(gen_scale, gen_scale + ngen);
    const uint64_t ref_rank = (uint64_t)sp.rank[0][refa] * t->Db + (uint64_t)sp.rank[1][refb];

    std::vector<SecStep> steps;
    std::vector<SecEntry> entries;
    std::vector<uint32_t> gen_step{0};
    for (uint64_t g = 0; g < ngen; ++g) {
        if (gen_off[g + 1] < gen_off[g]) return fail(QFE_ERR_ARG, "qfe_template_create: generator offsets must not decrease");
        if (gen_param[g] < 0 || (uint64_t)gen_param[g] >= nparam)
            return fail(QFE_ERR_ARG, "qfe_template_create: generator parameter index out of range");
        int rc = build_runs(sp, lp, norb, lim, gen_terms, gen_off[g], gen_off[g + 1], gen_param[g], gen_scale[g],
                            (uint32_t)g, steps, entries, "qfe_template_create");
        if (rc) return rc;
        gen_step.push_back((uint32_t)steps.size());
    }

    // Hamiltonian: groups by X mask (first appearance), sector pairs only
    std::vector<SecGroup> groups;
    std::vector<uint64_t> tz;
    std::vector<double2> tph, tcoef;
    std::vector<uint64_t> moff{0};
    std::vector<uint32_t> midx;
    std::vector<double2> mval;
    {
        std::map<uint64_t, std::vector<uint64_t>> byx;
        std::vector<uint64_t> order;
        for (uint64_t k = 0; k < nham; ++k) {
            if ((ham[k].x | ham[k].z) & ~lim)
                return fail(QFE_ERR_ARG, "qfe_template_create: Hamiltonian term acts on a qubit out of range");
            auto& v = byx[ham[k].x];
            if (v.empty()) order.push_back(ham[k].x);
            v.push_back(k);
        }
        for (uint64_t x : order) {
            const uint32_t xa = compress(x, 0, norb), xb = compress(x, 1, norb);
            if ((popc32(xa) & 1) || (popc32(xb) & 1)) continue;  // never connects two sector determinants
            SecGroup gr{};
            gr.x = x;
            gr.xa = xa;
            gr.xb = xb;
            std::pair<uint32_t, uint32_t> la, lb;
            if (x == 0) {
                la = lp.get(sp, 0, 1, 0, 0, 0);
                lb = lp.get(sp, 1, 1, 0, 0, 0);
                gr.weight = 1.0;
            } else {
                // visit the half of the pairs with the top bit of x clear, weight 2
                const int hb = 63 - __builtin_clzll(x);
                const uint32_t clear = 1u << (hb / 2);
                if ((hb & 1) == 0) {
                    la = lp.get(sp, 0, 2, xa, clear, xa);
                    lb = lp.get(sp, 1, 1, xb, 0, xb);
                } else {
                    la = lp.get(sp, 0, 1, xa, 0, xa);
                    lb = lp.get(sp, 1, 2, xb, clear, xb);
                }
                gr.weight = 2.0;
            }
            if (!la.second || !lb.second) continue;
            gr.la_off = la.first;
            gr.la_len = la.second;
            gr.lb_off = lb.first;
            gr.lb_len = lb.second;
            gr.inv_lb = 1.0 / (double)lb.second;
            gr.t_first = (uint32_t)tz.size();
            for (uint64_t k : byx[x]) {
                tz.push_back(ham[k].z);
                const double2 ph = make_double2(ham[k].phase.re, ham[k].phase.im);
                tph.push_back(ph);
                tcoef.push_back(cplx_mul(ph, make_double2(ham[k].coeff.re, ham[k].coeff.im)));
                if (map_off) {
                    if (map_off[k + 1] < map_off[k]) return fail(QFE_ERR_ARG, "qfe_template_create: map offsets must not decrease");
                    for (uint64_t e = map_off[k]; e < map_off[k + 1]; ++e) {
                        if (map_idx[e] >= ninputs) return fail(QFE_ERR_ARG, "qfe_template_create: map index out of range");
                        midx.push_back(map_idx[e]);
                        mval.push_back(make_double2(map_val[e].re, map_val[e].im));
                    }
                    moff.push_back(midx.size());
                }
            }
            gr.t_count = (uint32_t)tz.size() - gr.t_first;
            groups.push_back(gr);
        }
    }
    // work items of similar cost (pairs x (terms + 4)): small groups merge,
    // large groups split into pair ranges
    std::vector<SecItem> items;
    {
        const uint64_t target = 1ull << 18;
        uint32_t g = 0;
        while (g < groups.size()) {
            const uint64_t pairs = (uint64_t)groups[g].la_len * groups[g].lb_len;
            const uint64_t per = groups[g].t_count + 4;
            if (pairs * per >= target) {
                uint64_t chunk = std::max<uint64_t>(2048, target / per);
                chunk = (chunk + 255) & ~255ull;
                for (uint64_t p = 0; p < pairs; p += chunk)
                    items.push_back(SecItem{g, g + 1, (uint32_t)p, (uint32_t)std::min(pairs, p + chunk)});
                ++g;
                continue;
            }
            uint32_t g1 = g;
            uint64_t cost = 0;
            while (g1 < groups.size()) {
                const uint64_t c = (uint64_t)groups[g1].la_len * groups[g1].lb_len * (groups[g1].t_count + 4);
                if (g1 > g && (cost + c > target || c >= target)) break;
                cost += c;
                ++g1;
            }
            // a single small group is "whole" only when g1 != g + 1; give it its full pair range instead
            if (g1 == g + 1)
                items.push_back(SecItem{g, g + 1, 0, (uint32_t)((uint64_t)groups[g].la_len * groups[g].lb_len)});
            else
                items.push_back(SecItem{g, g1, 0, 0});
            g = g1;
        }
    }
    t->h_gen_step = gen_step;
    t->step_pairs.assign(steps.size(), 0);
    for (size_t k = 0; k < steps.size(); ++k)
        for (uint32_t e = steps[k].e_first; e < steps[k].e_first + steps[k].e_count; ++e)
            t->step_pairs[k] += entries[e].la_len * entries[e].lb_len;
    // sigma-kernel chunks: up to kHpsiGroups groups and kHpsiTerms terms each
    std::vector<uint2> hchunks;
    for (uint32_t g = 0; g < groups.size();) {
        uint32_t g1 = g + 1;
        uint32_t terms = groups[g].t_count;
        while (g1 < groups.size() && g1 - g < kHpsiGroups && terms + groups[g1].t_count <= kHpsiTerms)
            terms += groups[g1++].t_count;
        hchunks.push_back(make_uint2(g, g1));
        g = g1;
    }
    t->has_map = map_off != nullptr;
    t->nterms = (uint32_t)tz.size();
    t->nentries = (uint32_t)entries.size();
    t->nitems = (uint32_t)items.size();

    uint64_t cost = t->D;
    for (const auto& en : entries) cost += (uint64_t)en.la_len * en.lb_len * 4;
    for (const auto& gr : groups) cost += (uint64_t)gr.la_len * gr.lb_len * (2 + gr.t_count);
    t->cost = cost;

    {  // ansatz tier: the state in LDS when it fits a workgroup's share
        int dev = 0, max_lds = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) max_lds = 65536;
        const size_t tab = ((uint64_t)t->Da + t->Db) <= kTableLds ? ((uint64_t)t->Da + t->Db) * 8 : 0;
        const size_t with_psi = t->D * sizeof(double2) + tab;
        if (with_psi <= kTier0Lds) {
            t->tier = 0;
            t->lds_bytes = with_psi;
        } else if (with_psi + 1024 <= (size_t)max_lds) {
            t->tier = 1;
            t->lds_bytes = with_psi;
        } else {
            t->tier = 2;
            t->lds_bytes = tab;
        }
    }

    std::vector<uint64_t> af(t->Da), bf(t->Db);
    for (uint32_t a = 0; a < t->Da; ++a) af[a] = spread(sp.str[0][a], 0, norb);
    for (uint32_t b2 = 0; b2 < t->Db; ++b2) bf[b2] = spread(sp.str[1][b2], 1, norb);
    // pair words of every entry, in entry order, when they fit the cap
    std::vector<uint2> words;
    {
        uint64_t total = 0;
        for (const auto& en : entries) total += (uint64_t)en.la_len * en.lb_len;
        if (total <= ctx().tuning.sector_words) {
            words.reserve(total);
            for (auto& en : entries) {
                en.w_off = (uint32_t)words.size();
                for (uint32_t ia = 0; ia < en.la_len; ++ia) {
                    const int2 pa = lp.pool[en.la_off + ia];
                    for (uint32_t ib = 0; ib < en.lb_len; ++ib) {
                        const int2 pb = lp.pool[en.lb_off + ib];
                        const uint64_t j1 = af[pa.x] | bf[pb.x];
                        const uint32_t r1 = (uint32_t)((uint64_t)pa.x * t->Db + pb.x);
                        const uint32_t r2 = (uint32_t)((uint64_t)pa.y * t->Db + pb.y);
                        const uint32_t g1 = popc64(j1 & en.z0) & 1, g2 = popc64((j1 ^ en.x) & en.z0) & 1;
                        words.push_back(make_uint2(r1 | (g2 << 31), r2 | (g1 << 31)));
                    }
                }
            }
        }
    }
    Packer pk(sizeof(SecTables));
    const size_t i_af = pk.add(af.data(), af.size() * 8), i_bf = pk.add(bf.data(), bf.size() * 8);
    const size_t i_as = pk.add(sp.str[0].data(), sp.str[0].size() * 4), i_bs = pk.add(sp.str[1].data(), sp.str[1].size() * 4);
    const size_t i_ar = pk.add(sp.rank[0].data(), sp.rank[0].size() * 4), i_br = pk.add(sp.rank[1].data(), sp.rank[1].size() * 4);
    const size_t i_l = pk.add(lp.pool.data(), lp.pool.size() * sizeof(int2));
    const size_t i_st = pk.add(steps.data(), steps.size() * sizeof(SecStep));
    const size_t i_en = pk.add(entries.data(), entries.size() * sizeof(SecEntry));
    const size_t i_gr = pk.add(groups.data(), groups.size() * sizeof(SecGroup));
    const size_t i_it = pk.add(items.data(), items.size() * sizeof(SecItem));
    const size_t i_gs = pk.add(gen_step.data(), gen_step.size() * 4);
    const size_t i_tz = pk.add(tz.data(), tz.size() * 8);
    const size_t i_tph = pk.add(tph.data(), tph.size() * sizeof(double2));
    const size_t i_tc = pk.add(tcoef.data(), tcoef.size() * sizeof(double2));
    const size_t i_mo = pk.add(moff.data(), moff.size() * 8);
    const size_t i_mi = pk.add(midx.data(), midx.size() * 4);
    const size_t i_mv = pk.add(mval.data(), mval.size() * sizeof(double2));
    const size_t i_hc = pk.add(hchunks.data(), hchunks.size() * sizeof(uint2));
    const size_t i_w = pk.add(words.data(), words.size() * sizeof(uint2));
    QFE_HIP(hipMalloc(&t->dev, pk.total));
    char* base = (char*)t->dev;
    auto at = [&](size_t i) { return base + pk.pieces[i].off; };
    SecTables h{};
    h.D = t->D;
    h.ref_rank = ref_rank;
    h.Da = t->Da;
    h.Db = t->Db;
    h.afull = (const uint64_t*)at(i_af);
    h.bfull = (const uint64_t*)at(i_bf);
    h.astr = (const uint32_t*)at(i_as);
    h.bstr = (const uint32_t*)at(i_bs);
    h.arank = (const int32_t*)at(i_ar);
    h.brank = (const int32_t*)at(i_br);
    h.lists = (const int2*)at(i_l);
    h.steps = (const SecStep*)at(i_st);
    h.entries = (const SecEntry*)at(i_en);
    h.groups = (const SecGroup*)at(i_gr);
    h.items = (const SecItem*)at(i_it);
    h.gen_step = (const uint32_t*)at(i_gs);
    h.tz = (const uint64_t*)at(i_tz);
    h.tph = (const double2*)at(i_tph);
    h.tcoef = (const double2*)at(i_tc);
    h.map_off = t->has_map ? (const uint64_t*)at(i_mo) : nullptr;
    h.map_idx = (const uint32_t*)at(i_mi);
    h.map_val = (const double2*)at(i_mv);
    h.hchunks = (const uint2*)at(i_hc);
    h.nhchunks = (uint32_t)hchunks.size();
    h.words = entries.empty() || words.empty() ? nullptr : (const uint2*)at(i_w);
    h.nsteps = (uint32_t)steps.size();
    h.nentries = (uint32_t)entries.size();
    h.ngroups = (uint32_t)groups.size();
    h.nterms = (uint32_t)tz.size();
    h.nitems = (uint32_t)items.size();
    h.ngen = (uint32_t)ngen;
    std::vector<char> img;
    pk.fill(img);
    std::memcpy(img.data(), &h, sizeof(h));
    QFE_HIP(hipMemcpy(t->dev, img.data(), pk.total, hipMemcpyHostToDevice));
    t->dt = (SecTables*)t->dev;
    t->afull = h.afull;
    t->bfull = h.bfull;

    // per-request layout: energy parts first, gradient parts after
    size_t o = 0;
    t->off_psi = o;
    o += align256(t->D * sizeof(double2));
    t->off_coef = o;
    if (t->has_map) o += align256((size_t)t->nterms * sizeof(double2));
    t->off_cs = o;
    o += align256((size_t)t->nentries * sizeof(double2) + 16);
    t->off_part = o;
    o += align256((size_t)t->nitems * 8 + 8);
    t->work_bytes = o;
    t->off_sig = o;
    o += align256(t->D * sizeof(double2));
    t->off_gacc = o;
    o += align256(ngen * 8 + 8);
    t->grad_bytes = o;
    t->io_count = nparam + ninputs;
    t->request_bytes = t->work_bytes + align256(t->io_count * 8 + 8) + sizeof(SlotDesc) + 16;
    if (t->tier == 2) {  // partial-sum segments of the grid-wide reverse sweep
        auto blocks = [](uint64_t items) { return (uint32_t)((items + kWideTPB - 1) / kWideTPB); };
        t->wide_seg.push_back(make_uint2(0, blocks(t->D)));
        uint32_t at = blocks(t->D);
        for (uint64_t k = 0; k < ngen; ++k) {
            uint64_t pairs = 0;
            for (uint32_t sdx = gen_step[k]; sdx < gen_step[k + 1]; ++sdx) pairs += t->step_pairs[sdx];
            t->wide_seg.push_back(make_uint2(at, blocks(pairs)));
            at += blocks(pairs);
        }
        t->wide_parts = at;
        const size_t seg_bytes = align256(t->wide_seg.size() * sizeof(uint2));
        QFE_HIP(hipMalloc(&t->wide, seg_bytes + (size_t)t->wide_parts * 8 + 8));
        QFE_HIP(hipMemcpy(t->wide, t->wide_seg.data(), t->wide_seg.size() * sizeof(uint2), hipMemcpyHostToDevice));
    }
    t->grad_request_bytes = t->grad_bytes + align256(t->io_count * 8 + 8) + sizeof(SlotDesc) + 16 + ngen * 8;
    return QFE_OK;
}

SlotDesc make_slot(qfe_template t, char* work, const double* io, double* out, bool grad) {
    SlotDesc d{};
    d.t = t->dt;
    d.psi = (double2*)(work + t->off_psi);
    d.coef = t->has_map ? (double2*)(work + t->off_coef) : nullptr;
    d.cs = (double2*)(work + t->off_cs);
    d.part = (double*)(work + t->off_part);
    d.sig = grad ? (double2*)(work + t->off_sig) : nullptr;
    d.gacc = grad ? (double*)(work + t->off_gacc) : nullptr;
    d.params = io;
    d.inputs = io + t->nparam;
    d.out = out;
    return d;
}

// Energy of k requests of one tier: prologue, ansatz, items, sums.
int launch_energy(const std::vector<qfe_template>& ts, int tier, size_t lds, const SlotDesc* ddesc, const uint32_t* dprefix,
                  uint32_t total_items, hipStream_t st) {
    const unsigned k = (unsigned)ts.size();
    hipLaunchKernelGGL(k_sec_prologue, dim3(k), dim3(256), 0, st, ddesc, 1);
    QFE_HIP(hipGetLastError());
    int rc = launch_ansatz(tier, k, lds, st, ddesc);
    if (rc) return rc;
    if (total_items) {
        hipLaunchKernelGGL(k_sec_energy, dim3(total_items), dim3(kItemThreads), 0, st, ddesc, dprefix, (uint32_t)k);
        QFE_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_sec_sum, dim3(k), dim3(256), 0, st, ddesc);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

// Energy and adjoint gradient of k requests of one tier: prologue, ansatz,
// sigma = H psi (grid y = request), reverse sweep.
int launch_gradient(const std::vector<qfe_template>& ts, int tier, size_t lds, const SlotDesc* ddesc, hipStream_t st) {
    const unsigned k = (unsigned)ts.size();
    uint64_t chunks = 1;
    for (const auto& t : ts) chunks = std::max<uint64_t>(chunks, (t->D + 1023) / 1024);
    hipLaunchKernelGGL(k_sec_prologue, dim3(k), dim3(256), 0, st, ddesc, 1);
    QFE_HIP(hipGetLastError());
    int rc = launch_ansatz(tier, k, lds, st, ddesc);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sec_hpsi<4>, dim3((unsigned)chunks, k), dim3(256), 0, st, ddesc);
    QFE_HIP(hipGetLastError());
    return launch_reverse(tier, k, st, ddesc);
}

// ---- grid-wide single-request sequences (tier 2) -----------------------------------------
uint32_t wide_blocks(uint64_t items) { return (uint32_t)((items + kWideTPB - 1) / kWideTPB); }

// U|ref>: the reference determinant, then one launch per step.
int wide_ansatz(qfe_template t, const SlotDesc* d, hipStream_t st) {
    hipLaunchKernelGGL(k_sec_ref, dim3(wide_blocks(t->D)), dim3(kWideTPB), 0, st, d);
    QFE_HIP(hipGetLastError());
    for (uint32_t k = 0; k < (uint32_t)t->step_pairs.size(); ++k) {
        if (!t->step_pairs[k]) continue;
        hipLaunchKernelGGL(k_sec_step, dim3(wide_blocks(t->step_pairs[k])), dim3(kWideTPB), 0, st, d, k, 0, 0);
        QFE_HIP(hipGetLastError());
    }
    return QFE_OK;
}

// Energy: prologue, ansatz, work items, sum (the items and sum kernels of the batched path).
int wide_energy(qfe_template t, const SlotDesc* d, const uint32_t* dprefix, hipStream_t st) {
    hipLaunchKernelGGL(k_sec_prologue, dim3(1), dim3(256), 0, st, d, 1);
    QFE_HIP(hipGetLastError());
    int rc = wide_ansatz(t, d, st);
    if (rc) return rc;
    if (t->nitems) {
        hipLaunchKernelGGL(k_sec_energy, dim3(t->nitems), dim3(kItemThreads), 0, st, d, dprefix, 1u);
        QFE_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_sec_sum, dim3(1), dim3(256), 0, st, d);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

// psi and sigma = H psi (pool gradients and the gradient below).
int wide_psi_sigma(qfe_template t, const SlotDesc* d, hipStream_t st) {
    hipLaunchKernelGGL(k_sec_prologue, dim3(1), dim3(256), 0, st, d, 1);
    QFE_HIP(hipGetLastError());
    int rc = wide_ansatz(t, d, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sec_hpsi<1>, dim3((unsigned)((t->D + 255) / 256), 1), dim3(256), 0, st, d);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

// Energy and generator overlaps: E = Re<psi|sigma>, then for k = M-1 .. 0 the
// overlap of generator k and its un-rotation on psi and sigma; fixed-order sums.
int wide_gradient(qfe_template t, const SlotDesc* d, hipStream_t st) {
    int rc = wide_psi_sigma(t, d, st);
    if (rc) return rc;
    const size_t seg_bytes = align256(t->wide_seg.size() * sizeof(uint2));
    double* part = (double*)((char*)t->wide + seg_bytes);
    hipLaunchKernelGGL(k_sec_edot, dim3(t->wide_seg[0].y), dim3(kWideTPB), 0, st, d, part + t->wide_seg[0].x);
    QFE_HIP(hipGetLastError());
    for (uint32_t k = (uint32_t)t->ngen; k-- > 0;) {
        const uint2 sg = t->wide_seg[k + 1];
        if (!sg.y) continue;
        const uint32_t s0 = t->h_gen_step[k], s1 = t->h_gen_step[k + 1];
        if (s1 == s0 + 1) {
            hipLaunchKernelGGL(k_sec_genrev, dim3(sg.y), dim3(kWideTPB), 0, st, d, s0, part + sg.x);
            QFE_HIP(hipGetLastError());
            continue;
        }
        hipLaunchKernelGGL(k_sec_genov, dim3(sg.y), dim3(kWideTPB), 0, st, d, s0, s1, part + sg.x);
        QFE_HIP(hipGetLastError());
        for (uint32_t sdx = s1; sdx-- > s0;) {
            if (!t->step_pairs[sdx]) continue;
            hipLaunchKernelGGL(k_sec_step, dim3(wide_blocks(t->step_pairs[sdx])), dim3(kWideTPB), 0, st, d, sdx, 1, 1);
            QFE_HIP(hipGetLastError());
        }
    }
    const uint32_t nseg = (uint32_t)t->wide_seg.size();
    hipLaunchKernelGGL(k_sec_wsum, dim3(wide_blocks(nseg)), dim3(kWideTPB), 0, st, d, (const double*)part,
                       (const uint2*)t->wide, nseg);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

// Single request on the library stream.  Arena: [desc 256][out 256][prefix 256][io][work + gradient parts].
struct Single {
    SlotDesc* ddesc;
    double* dout;
    uint32_t* dprefix;
    char* work;
};

int single_setup(qfe_template t, const double* params, const double* inputs, bool grad, Single& S) {
    const size_t io_bytes = align256(t->io_count * 8 + 8);
    const size_t head = 768 + io_bytes;
    if (!t->arena) QFE_HIP(hipMalloc(&t->arena, head + t->grad_bytes));
    char* dev = (char*)t->arena;
    S.ddesc = (SlotDesc*)dev;
    S.dout = (double*)(dev + 256);
    S.dprefix = (uint32_t*)(dev + 512);
    S.work = dev + head;
    void* host = nullptr;
    int rc = staging_begin(head, &host);
    if (rc) return rc;
    char* h = (char*)host;
    SlotDesc d = make_slot(t, S.work, (const double*)(dev + 768), S.dout, grad);
    std::memcpy(h, &d, sizeof(d));
    const uint32_t pre[2] = {0, t->nitems};
    std::memcpy(h + 512, pre, sizeof(pre));
    if (t->nparam) std::memcpy(h + 768, params, t->nparam * 8);
    if (t->ninputs) std::memcpy(h + 768 + t->nparam * 8, inputs, t->ninputs * 8);
    return staging_commit(dev, head);
}

int fetch_doubles(const double* dev, size_t count, double* host) {
    void* pin = nullptr;
    int rc = pinned_reserve(count * 8 + 8, &pin);
    if (rc) return rc;
    QFE_HIP(hipMemcpyAsync(pin, dev, count * 8, hipMemcpyDeviceToHost, ctx().stream));
    QFE_HIP(hipStreamSynchronize(ctx().stream));
    std::memcpy(host, pin, count * 8);
    return QFE_OK;
}

// grad[p] = sum over generators k with parameter p of 2 scale_k <sigma_k|K_k|psi_k>.
void gradient_from_overlaps(qfe_template t, const double* acc, double* grad) {
    for (uint64_t i = 0; i < t->nparam; ++i) grad[i] = 0.0;
    for (uint64_t k = 0; k < t->ngen; ++k) grad[t->gparam[k]] += 2.0 * t->gscale[k] * acc[k];
}

// psi = U|ref> and sigma = H psi for one request (pool gradients).
int single_psi_sigma(qfe_template t, const double* params, const double* inputs, Single& S) {
    int rc = single_setup(t, params, inputs, true, S);
    if (rc) return rc;
    hipStream_t st = ctx().stream;
    if (t->tier == 2) return wide_psi_sigma(t, S.ddesc, st);
    hipLaunchKernelGGL(k_sec_prologue, dim3(1), dim3(256), 0, st, S.ddesc, 1);
    QFE_HIP(hipGetLastError());
    rc = launch_ansatz(t->tier, 1, t->lds_bytes, st, S.ddesc);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sec_hpsi<1>, dim3((unsigned)((t->D + 255) / 256), 1), dim3(256), 0, st, S.ddesc);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

// ---- captured single-request sequences (large sectors) ----------------------------------
// The grid-wide sequences launch one kernel per step and per generator; as a
// hipGraph per template they replay without per-launch host work.  The first
// node uploads the whole request head (descriptor, prefix, amplitudes,
// inputs) from the template's pinned block, so a replay does not depend on
// what other calls left in the arena; the last nodes copy the energy (and the
// generator overlaps) back into the same block.
size_t single_head(qfe_template t) { return 768 + align256(t->io_count * 8 + 8); }

// ---- host-side audit of a captured sequence ---------------------------------------------
// After instantiation every node of a template graph is compared with the
// sequence the template's structure fixes, computed here independently of the
// wide_* launchers: the graph must be one chain from its root; each node must
// be the expected memcpy (addresses and size inside the arena and the pinned
// block) or the expected kernel with a non-zero grid and a block of at most
// 1024 threads, every scalar argument equal to the expected one and every
// pointer argument equal to the expected address (partial-sum pointers with
// their whole grid inside the partial array); the uploaded descriptor's
// pointers must lie inside the arena.  A mismatch fails graph_build with
// QFE_ERR_STATE and the graph is not kept, so nothing malformed is launched.
struct GraphAudit {
    uint64_t nodes = 0, kernels = 0, kernels_checked = 0, copies_checked = 0;
};

struct NodeExpect {
    bool copy = false;
    const void* func = nullptr;
    unsigned grid = 0, block = 0;
    std::vector<std::vector<char>> args;  // expected bytes of each argument
    const void* dst = nullptr;
    const void* src = nullptr;
    size_t bytes = 0;
};

template <typename... A>
NodeExpect expect_kernel(const void* func, uint64_t grid, unsigned block, A... a) {
    NodeExpect e;
    e.func = func;
    e.grid = grid > 0x7fffffffull ? 0u : (unsigned)grid;  // out of range: fails the non-zero grid check
    e.block = block;
    (e.args.push_back(std::vector<char>((const char*)&a, (const char*)&a + sizeof(A))), ...);
    return e;
}

NodeExpect expect_copy(const void* dst, const void* src, size_t bytes) {
    NodeExpect e;
    e.copy = true;
    e.dst = dst;
    e.src = src;
    e.bytes = bytes;
    return e;
}

void expected_nodes(qfe_template t, int grad, std::vector<NodeExpect>& out) {
    const size_t head = single_head(t);
    char* dev = (char*)t->arena;
    char* pin = (char*)t->gpin;
    const SlotDesc* d = (const SlotDesc*)dev;
    const auto KB = [](uint64_t items) { return (items + kWideTPB - 1) / kWideTPB; };
    out.clear();
    out.push_back(expect_copy(dev, pin, head));
    out.push_back(expect_kernel((const void*)&k_sec_prologue, 1, 256, d, (int)1));
    out.push_back(expect_kernel((const void*)&k_sec_ref, KB(t->D), kWideTPB, d));
    for (uint32_t k = 0; k < (uint32_t)t->step_pairs.size(); ++k)
        if (t->step_pairs[k])
            out.push_back(expect_kernel((const void*)&k_sec_step, KB(t->step_pairs[k]), kWideTPB, d, k, (int)0, (int)0));
    if (!grad) {
        if (t->nitems)
            out.push_back(expect_kernel((const void*)&k_sec_energy, t->nitems, kItemThreads, d,
                                        (const uint32_t*)(dev + 512), (uint32_t)1));
        out.push_back(expect_kernel((const void*)&k_sec_sum, 1, 256, d));
        out.push_back(expect_copy(pin + head, dev + 256, 8));
        return;
    }
    out.push_back(expect_kernel((const void*)&k_sec_hpsi<1>, (t->D + 255) / 256, 256, d));
    double* part = (double*)((char*)t->wide + align256(t->wide_seg.size() * sizeof(uint2)));
    out.push_back(expect_kernel((const void*)&k_sec_edot, t->wide_seg[0].y, kWideTPB, d, part + t->wide_seg[0].x));
    for (uint32_t k = (uint32_t)t->ngen; k-- > 0;) {
        const uint2 sg = t->wide_seg[k + 1];
        if (!sg.y) continue;
        const uint32_t s0 = t->h_gen_step[k], s1 = t->h_gen_step[k + 1];
        if (s1 == s0 + 1) {
            out.push_back(expect_kernel((const void*)&k_sec_genrev, sg.y, kWideTPB, d, s0, part + sg.x));
            continue;
        }
        out.push_back(expect_kernel((const void*)&k_sec_genov, sg.y, kWideTPB, d, s0, s1, part + sg.x));
        for (uint32_t sdx = s1; sdx-- > s0;)
            if (t->step_pairs[sdx])
                out.push_back(expect_kernel((const void*)&k_sec_step, KB(t->step_pairs[sdx]), kWideTPB, d, sdx, (int)1, (int)1));
    }
    const uint32_t nseg = (uint32_t)t->wide_seg.size();
    out.push_back(expect_kernel((const void*)&k_sec_wsum, KB(nseg), kWideTPB, d, (const double*)part,
                                (const uint2*)t->wide, nseg));
    out.push_back(expect_copy(pin + head, dev + 256, 8));
    if (t->ngen) out.push_back(expect_copy(pin + head + 8, dev + head + t->off_gacc, t->ngen * 8));
}

bool inside(const void* p, size_t bytes, const void* base, size_t cap) {
    const uintptr_t a = (uintptr_t)p, b = (uintptr_t)base;
    return a >= b && bytes <= cap && a - b <= cap - bytes;
}

int audit_fail(int grad, size_t node, const std::string& what) {
    return fail(QFE_ERR_STATE, std::string("qfe_template: captured ") + (grad ? "gradient" : "energy") + " sequence, node " +
                                   std::to_string(node) + ": " + what);
}

int audit_graph(qfe_template t, hipGraph_t g, int grad, GraphAudit& A) {
    const size_t head = single_head(t);
    const size_t arena_bytes = head + t->grad_bytes, pin_bytes = head + 8 * (t->ngen + 2);
    const char* dev = (const char*)t->arena;
    // descriptor as the first node uploads it
    SlotDesc d;
    std::memcpy(&d, t->gpin, sizeof(d));
    const char* work = dev + head;
    auto in_work = [&](const void* p, size_t bytes) { return p && inside(p, bytes, work, t->grad_bytes); };
    if (d.t != t->dt || d.out != (double*)(dev + 256) || !in_work(d.psi, t->D * sizeof(double2)) ||
        !in_work(d.sig, t->D * sizeof(double2)) || !in_work(d.cs, (size_t)t->nentries * sizeof(double2)) ||
        !in_work(d.part, (size_t)t->nitems * 8) || !in_work(d.gacc, t->ngen * 8) ||
        (t->has_map && !in_work(d.coef, (size_t)t->nterms * sizeof(double2))) ||
        !inside(d.params, t->io_count * 8, dev + 768, head - 768))
        return audit_fail(grad, 0, "descriptor pointer outside the arena");
    std::vector<NodeExpect> want;
    expected_nodes(t, grad, want);
    size_t nroot = 0;
    QFE_HIP(hipGraphGetRootNodes(g, nullptr, &nroot));
    if (nroot != 1) return audit_fail(grad, 0, std::to_string(nroot) + " root nodes (expected one chain)");
    hipGraphNode_t node = nullptr;
    QFE_HIP(hipGraphGetRootNodes(g, &node, &nroot));
    const double* part_lo = (const double*)((const char*)t->wide + align256(t->wide_seg.size() * sizeof(uint2)));
    const double* part_hi = part_lo + t->wide_parts;
    A = GraphAudit{};
    for (size_t i = 0;; ++i) {
        if (i >= want.size()) return audit_fail(grad, i, "more nodes than the sequence has launches");
        const NodeExpect& w = want[i];
        hipGraphNodeType type;
        QFE_HIP(hipGraphNodeGetType(node, &type));
        if (w.copy) {
            if (type != hipGraphNodeTypeMemcpy) return audit_fail(grad, i, "expected a memcpy node");
            // the expected copy itself must stay inside the arena and the pinned block
            const bool up = w.dst == t->arena;
            if (up ? !(inside(w.dst, w.bytes, t->arena, arena_bytes) && inside(w.src, w.bytes, t->gpin, pin_bytes))
                   : !(inside(w.src, w.bytes, t->arena, arena_bytes) && inside(w.dst, w.bytes, t->gpin, pin_bytes)))
                return audit_fail(grad, i, "memcpy outside the arena or the pinned block");
            hipMemcpy3DParms p{};
            // a runtime that does not read back captured 1D copies returns null pointers: shape unchecked
            if (hipGraphMemcpyNodeGetParams(node, &p) == hipSuccess && (p.dstPtr.ptr || p.srcPtr.ptr)) {
                // a captured 1D copy reads back as a width-only extent of `bytes`
                const size_t bytes = p.extent.width * std::max<size_t>(1, p.extent.height) * std::max<size_t>(1, p.extent.depth);
                if (p.dstPtr.ptr != w.dst || p.srcPtr.ptr != w.src || bytes != w.bytes)
                    return audit_fail(grad, i, "memcpy addresses or size differ from the sequence");
                ++A.copies_checked;
            }
        } else {
            if (type != hipGraphNodeTypeKernel) return audit_fail(grad, i, "expected a kernel node");
            hipKernelNodeParams p{};
            QFE_HIP(hipGraphKernelNodeGetParams(node, &p));
            if (p.func != w.func) return audit_fail(grad, i, "kernel differs from the sequence");
            if (p.gridDim.x == 0 || p.gridDim.y != 1 || p.gridDim.z != 1 || w.grid == 0)
                return audit_fail(grad, i, "zero or out-of-range grid");
            if (p.blockDim.x == 0 || p.blockDim.x > 1024 || p.blockDim.y != 1 || p.blockDim.z != 1)
                return audit_fail(grad, i, "block outside 1..1024 threads");
            if (p.gridDim.x != w.grid || p.blockDim.x != w.block)
                return audit_fail(grad, i, "launch shape differs from the sequence (grid " + std::to_string(p.gridDim.x) +
                                               ", expected " + std::to_string(w.grid) + ")");
            ++A.kernels;
            // a runtime that keeps no argument copies leaves kernelParams null: shape checked only
            for (size_t a = 0; p.kernelParams && a < w.args.size(); ++a) {
                if (!p.kernelParams[a] || std::memcmp(p.kernelParams[a], w.args[a].data(), w.args[a].size()))
                    return audit_fail(grad, i, "argument " + std::to_string(a) + " differs from the sequence");
                if (w.args[a].size() == sizeof(void*) && w.func != (const void*)&k_sec_wsum) {
                    const double* q;
                    std::memcpy(&q, w.args[a].data(), sizeof(q));
                    // a partial-sum pointer covers one double per workgroup of its launch
                    if (q >= part_lo && q <= part_hi && q + p.gridDim.x > part_hi)
                        return audit_fail(grad, i, "partial sums past the end of the partial array");
                }
            }
            if (p.kernelParams) ++A.kernels_checked;
        }
        size_t ndep = 0;
        QFE_HIP(hipGraphNodeGetDependentNodes(node, nullptr, &ndep));
        if (ndep == 0) {
            if (i + 1 != want.size()) return audit_fail(grad, i, "chain ends before the sequence does");
            A.nodes = i + 1;
            break;
        }
        if (ndep != 1) return audit_fail(grad, i, "node with " + std::to_string(ndep) + " dependents (expected one chain)");
        QFE_HIP(hipGraphNodeGetDependentNodes(node, &node, &ndep));
    }
    return QFE_OK;
}

int graph_build(qfe_template t, int grad) {
    const size_t head = single_head(t);
    if (!t->arena) QFE_HIP(hipMalloc(&t->arena, head + t->grad_bytes));
    if (!t->gpin) QFE_HIP(hipHostMalloc(&t->gpin, head + 8 * (t->ngen + 2), hipHostMallocDefault));
    char* dev = (char*)t->arena;
    char* pin = (char*)t->gpin;
    const SlotDesc* ddesc = (const SlotDesc*)dev;
    double* dout = (double*)(dev + 256);
    // descriptor with the gradient parts (the energy sequence ignores them) and the one-slot prefix
    SlotDesc d = make_slot(t, dev + head, (const double*)(dev + 768), dout, true);
    std::memcpy(pin, &d, sizeof(d));
    const uint32_t pre[2] = {0, t->nitems};
    std::memcpy(pin + 512, pre, sizeof(pre));
    hipStream_t st = ctx().stream;
    QFE_HIP(hipStreamSynchronize(st));
    QFE_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    auto record = [&]() -> int {
        QFE_HIP(hipMemcpyAsync(dev, pin, head, hipMemcpyHostToDevice, st));
        int r = grad ? wide_gradient(t, ddesc, st) : wide_energy(t, ddesc, (const uint32_t*)(dev + 512), st);
        if (r) return r;
        QFE_HIP(hipMemcpyAsync(pin + head, dout, 8, hipMemcpyDeviceToHost, st));
        if (grad && t->ngen)
            QFE_HIP(hipMemcpyAsync(pin + head + 8, dev + head + t->off_gacc, t->ngen * 8, hipMemcpyDeviceToHost, st));
        return QFE_OK;
    };
    int rc = record();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "qfe_template: end capture");
    if (!rc) {
        const hipError_t ie = hipGraphInstantiate(&t->gexec[grad], g, nullptr, nullptr, 0);
        if (ie != hipSuccess) rc = hip_fail(ie, "qfe_template: instantiate");
    }
    if (!rc) {
        GraphAudit A;
        rc = audit_graph(t, g, grad, A);
        if (!rc) {
            t->audit[grad][0] = A.nodes;
            t->audit[grad][1] = A.kernels;
            t->audit[grad][2] = A.kernels_checked;
            t->audit[grad][3] = A.copies_checked;
        }
    }
    if (rc) {
        if (t->gexec[grad]) (void)hipGraphExecDestroy(t->gexec[grad]);
        if (g) (void)hipGraphDestroy(g);
        t->gexec[grad] = nullptr;
        return rc;
    }
    t->graph[grad] = g;
    return QFE_OK;
}

// Replay: energy into *energy, and with grad the ngen generator overlaps into acc.
int graph_run(qfe_template t, int grad, const double* params, const double* inputs, double* energy, double* acc) {
    if (!t->gexec[grad]) {
        int rc = graph_build(t, grad);
        if (rc) return rc;
    }
    char* pin = (char*)t->gpin;
    const size_t head = single_head(t);
    // every call synchronizes before returning, so no replay is reading the block
    if (t->nparam) std::memcpy(pin + 768, params, t->nparam * 8);
    if (t->ninputs) std::memcpy(pin + 768 + t->nparam * 8, inputs, t->ninputs * 8);
    hipStream_t st = ctx().stream;
    QFE_HIP(hipGraphLaunch(t->gexec[grad], st));
    QFE_HIP(hipStreamSynchronize(st));
    std::memcpy(energy, pin + head, 8);
    if (grad && t->ngen) std::memcpy(acc, pin + head + 8, t->ngen * 8);
    return QFE_OK;
}

}  // namespace
}  // namespace qfe

// ---- serving engine -----------------------------------------------------------------------

namespace {

struct Request {
    uint64_t id;
    qfe_template t;
    std::vector<double> params, inputs;
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    void* dev = nullptr;
    size_t dev_cap = 0;
    void* up = nullptr;
    size_t up_cap = 0;
    double* out = nullptr;
    size_t out_cap = 0;
    std::vector<uint64_t> ids;  // in slot order
    std::vector<qfe_template> ts;
    std::vector<size_t> goff;  // gradient launches: each request's first overlap after the k energies
    bool busy = false;
    bool grad = false;
    int tier = 0;
    uint64_t seq = 0;  // launch order
};

}  // namespace

struct qfe_batch_s {
    uint64_t budget = 0;
    int max_active = 0;
    uint64_t next_id = 1;
    uint64_t active = 0;
    uint64_t seq = 0;
    std::deque<Request> queue[3];   // energy requests, one queue per ansatz tier
    std::deque<Request> gqueue[3];  // energy + gradient requests
    std::vector<Lane> lanes;
    std::deque<std::pair<uint64_t, double>> done;
    std::unordered_map<uint64_t, std::vector<double>> grads;  // completed gradients not yet taken
};

namespace {

// Requests of different tiers never share a launch, so small requests do not
// wait for the longest request of a mixed launch (nor, in tier 2, requests of
// different sector dimensions).  One lane per tier: tier 0
// on the library stream, tiers 1 and 2 on two streams the library creates once
// and every engine shares.  With the null stream (synchronous copies, torch)
// these are four streams, HIP's default four hardware queues per process, so
// no lane shares a queue with another.  Streams that share a queue run one
// kernel at a time between them: with a third lane stream of its own (five
// streams on four queues) the 16-qubit requests of tier 1 waited behind the
// tier-2 launches (p50 13.6 ms at half load, 1.2 ms without), and two lanes
// per tier (six lane streams) gave 13.4 ms p50 over all classes.  Letting two
// lanes take tier-1 and tier-2 launches by age instead (one dimension never on
// both) brought the 18-qubit p50 from 7.7 to 3.6 ms but the 16-qubit one from
// 1.2 to 2.5 ms, and two concurrent tier-2 launches slowed the tier-0 lane:
// 21.2K instead of 21.7K energies/s, p50 45 instead of 3 ms at 1.5x load.
constexpr int kLanesPerTier = 1;
constexpr int kLanes = 3 * kLanesPerTier;

hipStream_t g_lane_stream[kLanes] = {};

int lane_stream(int i, hipStream_t* out) {
    if (i == 0) {  // the tier-0 lane runs on the library stream
        *out = ctx().stream;
        return QFE_OK;
    }
    if (!g_lane_stream[i]) QFE_HIP(hipStreamCreateWithFlags(&g_lane_stream[i], hipStreamNonBlocking));
    *out = g_lane_stream[i];
    return QFE_OK;
}
// A launch admits requests up to this footprint even under a larger budget:
// the requests' states then stay in the 256 MB last-level cache while their
// per-step gathers run (4096 mixed 8-20 qubit requests: 19.5K -> 21.8K
// energies/s and 6.5K -> 7.0K gradients/s against a 64 GiB share).
constexpr uint64_t kLaunchFootprint = 256ull << 20;

template <typename T>
int grow_device(void*& p, size_t& cap, size_t need) {
    if (cap >= need) return QFE_OK;
    if (p) QFE_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    QFE_HIP(hipMalloc(&p, need + need / 4));
    cap = need + need / 4;
    return QFE_OK;
}

int grow_pinned(void*& p, size_t& cap, size_t need) {
    if (cap >= need) return QFE_OK;
    if (p) QFE_HIP(hipHostFree(p));
    p = nullptr;
    cap = 0;
    QFE_HIP(hipHostMalloc(&p, need + need / 4, hipHostMallocDefault));
    cap = need + need / 4;
    return QFE_OK;
}

// Largest launch of gradient requests (the sigma kernel's grid y is the request).
constexpr size_t kMaxGradLaunch = 65535;

int lane_launch(qfe_batch b, Lane& L) {
    // a launch holds one kind: the kind whose oldest queued request arrived first
    std::deque<Request>& qe = b->queue[L.tier];
    std::deque<Request>& qg = b->gqueue[L.tier];
    const bool grad = qe.empty() || (!qg.empty() && qg.front().id < qe.front().id);
    std::deque<Request>& queue = grad ? qg : qe;
    // admit from the front of the queue while the footprint fits this lane's
    // share.  A tier-2 launch holds requests of one sector dimension, that of
    // its oldest request: a tier-2 request runs its whole ansatz chain in one
    // workgroup, and the launch ends with its longest chain (alone, 1.7 ms for
    // an 18-qubit and 6.6 ms for a 20-qubit UCCSD energy), so shorter requests
    // would otherwise wait for the longest; requests of other dimensions keep
    // their place in the queue.
    const uint64_t share = std::min<uint64_t>(b->budget / kLanes, kLaunchFootprint);
    const bool one_dim = L.tier == 2;
    const uint64_t dim0 = queue.front().t->D;
    std::vector<Request> take;
    std::vector<Request> skipped;
    uint64_t foot = 0;
    while (!queue.empty() && (int64_t)(b->active + take.size()) < b->max_active &&
           (!grad || take.size() < kMaxGradLaunch)) {
        if (one_dim && queue.front().t->D != dim0) {
            skipped.push_back(std::move(queue.front()));
            queue.pop_front();
            continue;
        }
        const uint64_t f = grad ? queue.front().t->grad_request_bytes : queue.front().t->request_bytes;
        if (!take.empty() && foot + f > share) break;
        foot += f;
        take.push_back(std::move(queue.front()));
        queue.pop_front();
    }
    for (size_t i = skipped.size(); i-- > 0;) queue.push_front(std::move(skipped[i]));
    if (take.empty()) return QFE_OK;
    // heaviest requests first, so their workgroups start first
    std::stable_sort(take.begin(), take.end(), [](const Request& x, const Request& y) { return x.t->cost > y.t->cost; });
    const size_t k = take.size();
    size_t io_total = 0, work_total = 0, lds = 0, gtotal = 0;
    uint64_t items = 0;
    std::vector<qfe_template> ts(k);
    for (size_t i = 0; i < k; ++i) {
        ts[i] = take[i].t;
        io_total += ts[i]->io_count;
        work_total += grad ? ts[i]->grad_bytes : ts[i]->work_bytes;
        lds = std::max(lds, ts[i]->lds_bytes);
        items += ts[i]->nitems;
        if (grad) gtotal += ts[i]->ngen;
    }
    if (items >= (1ull << 31)) return fail(QFE_ERR_ARG, "qfe_batch: too many work items in one launch");
    // results: k energies, then the generator overlaps of every gradient request
    const size_t nres = k + gtotal;
    const size_t o_out = align256(k * sizeof(SlotDesc));
    const size_t o_pre = o_out + align256(nres * 8);
    const size_t o_io = o_pre + align256((k + 1) * 4);
    const size_t head = o_io + align256(io_total * 8 + 8);
    const size_t need = head + work_total;
    int rc = grow_device<char>(L.dev, L.dev_cap, need);
    if (!rc) rc = grow_pinned(L.up, L.up_cap, head);
    if (!rc) {
        void* p = L.out;
        size_t cap = L.out_cap * 8;
        rc = grow_pinned(p, cap, nres * 8 + 8);
        L.out = (double*)p;
        L.out_cap = cap / 8;
    }
    if (rc) return rc;
    char* dev = (char*)L.dev;
    char* up = (char*)L.up;
    double* dout = (double*)(dev + o_out);
    uint32_t* hpre = (uint32_t*)(up + o_pre);
    size_t io_pos = 0, work_pos = head, gpos = k;
    uint32_t acc_items = 0;
    L.ids.clear();
    L.goff.clear();
    for (size_t i = 0; i < k; ++i) {
        qfe_template t = ts[i];
        const double* dio = (const double*)(dev + o_io) + io_pos;
        double* hio = (double*)(up + o_io) + io_pos;
        SlotDesc d = make_slot(t, dev + work_pos, dio, dout + i, grad);
        if (grad) {
            d.gacc = dout + gpos;
            L.goff.push_back(gpos);
            gpos += t->ngen;
        }
        std::memcpy(up + i * sizeof(SlotDesc), &d, sizeof(d));
        if (t->nparam) std::memcpy(hio, take[i].params.data(), t->nparam * 8);
        if (t->ninputs) std::memcpy(hio + t->nparam, take[i].inputs.data(), t->ninputs * 8);
        hpre[i] = acc_items;
        acc_items += t->nitems;
        io_pos += t->io_count;
        work_pos += grad ? t->grad_bytes : t->work_bytes;
        L.ids.push_back(take[i].id);
    }
    hpre[k] = acc_items;
    QFE_HIP(hipMemcpyAsync(dev, up, head, hipMemcpyHostToDevice, L.stream));
    if (grad)
        rc = launch_gradient(ts, L.tier, lds, (const SlotDesc*)dev, L.stream);
    else
        rc = launch_energy(ts, L.tier, lds, (const SlotDesc*)dev, (const uint32_t*)(dev + o_pre), acc_items, L.stream);
    if (rc) return rc;
    QFE_HIP(hipMemcpyAsync(L.out, dout, nres * 8, hipMemcpyDeviceToHost, L.stream));
    QFE_HIP(hipEventRecord(L.done, L.stream));
    L.ts = std::move(ts);
    L.grad = grad;
    L.busy = true;
    L.seq = ++b->seq;
    b->active += k;
    return QFE_OK;
}

void retire(qfe_batch b, Lane& L) {
    for (size_t i = 0; i < L.ids.size(); ++i) b->done.emplace_back(L.ids[i], L.out[i]);
    if (L.grad)
        for (size_t i = 0; i < L.ids.size(); ++i) {
            std::vector<double> g(L.ts[i]->nparam);
            gradient_from_overlaps(L.ts[i], L.out + L.goff[i], g.data());
            b->grads[L.ids[i]] = std::move(g);
        }
    b->active -= L.ids.size();
    L.ids.clear();
    L.busy = false;
}

// Retire every finished launch; with wait_oldest, first block on the oldest
// launch still running.
int collect(qfe_batch b, bool wait_oldest) {
    if (wait_oldest) {
        Lane* oldest = nullptr;
        for (auto& L : b->lanes)
            if (L.busy && (!oldest || L.seq < oldest->seq)) oldest = &L;
        if (oldest) {
            QFE_HIP(hipEventSynchronize(oldest->done));
            retire(b, *oldest);
        }
    }
    for (auto& L : b->lanes) {
        if (!L.busy) continue;
        const hipError_t q = hipEventQuery(L.done);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return hip_fail(q, "qfe_batch: launch completion");
        retire(b, L);
    }
    return QFE_OK;
}

Request make_request(qfe_batch b, qfe_template t, const double* params, const double* inputs) {
    Request r;
    r.id = b->next_id++;
    r.t = t;
    r.params.assign(params, params + t->nparam);
    if (t->ninputs) r.inputs.assign(inputs, inputs + t->ninputs);
    return r;
}

bool any_busy(qfe_batch b) {
    for (const auto& L : b->lanes)
        if (L.busy) return true;
    return false;
}

int fill_lanes(qfe_batch b) {
    for (auto& L : b->lanes) {
        if (L.busy || (b->queue[L.tier].empty() && b->gqueue[L.tier].empty())) continue;
        int rc = lane_launch(b, L);
        if (rc) return rc;
    }
    return QFE_OK;
}

int hand_out(qfe_batch b, uint64_t* ids, double* es, uint64_t cap, uint64_t* nd) {
    uint64_t k = 0;
    while (k < cap && !b->done.empty()) {
        ids[k] = b->done.front().first;
        es[k] = b->done.front().second;
        b->done.pop_front();
        ++k;
    }
    *nd = k;
    return QFE_OK;
}

}  // namespace

extern "C" {

int qfe_template_create(int nqubit, uint64_t ref_index, const qfe_pauli* gen_terms, const uint64_t* gen_offsets,
                        uint64_t ngen, const int32_t* gen_param, const double* gen_scale, uint64_t nparam,
                        const qfe_pauli* ham_terms, uint64_t nham, const uint64_t* map_offsets,
                        const uint32_t* map_index, const qfe_c64* map_value, uint64_t ninputs, qfe_template* out) {
    if (!out || !gen_offsets || (ngen && (!gen_terms || !gen_param || !gen_scale)) || (nham && !ham_terms))
        return fail(QFE_ERR_ARG, "qfe_template_create: null argument");
    if (map_offsets && (!map_index || !map_value) && map_offsets[nham] > 0)
        return fail(QFE_ERR_ARG, "qfe_template_create: null map arrays");
    if (!map_offsets && ninputs) return fail(QFE_ERR_ARG, "qfe_template_create: inputs need a coefficient map");
    int rc = ensure_init();
    if (rc) return rc;
    qfe_template t = new qfe_template_s();
    rc = template_build(t, nqubit, ref_index, gen_terms, gen_offsets, ngen, gen_param, gen_scale, nparam, ham_terms, nham,
                        map_offsets, map_index, map_value, ninputs);
    if (rc) {
        qfe_template_destroy(t);
        return rc;
    }
    *out = t;
    return QFE_OK;
}

int qfe_template_destroy(qfe_template t) {
    if (!t) return QFE_OK;
    if (ctx().ready) (void)hipStreamSynchronize(ctx().stream);
    if (t->dev) (void)hipFree(t->dev);
    if (t->arena) (void)hipFree(t->arena);
    if (t->wide) (void)hipFree(t->wide);
    for (int k = 0; k < 2; ++k) {
        if (t->gexec[k]) (void)hipGraphExecDestroy(t->gexec[k]);
        if (t->graph[k]) (void)hipGraphDestroy(t->graph[k]);
    }
    if (t->gpin) (void)hipHostFree(t->gpin);
    delete t;
    return QFE_OK;
}

int qfe_template_info(qfe_template t, uint64_t* sector_dim, uint64_t* request_bytes) {
    if (!t) return fail(QFE_ERR_ARG, "qfe_template_info: invalid handle");
    if (sector_dim) *sector_dim = t->D;
    if (request_bytes) *request_bytes = t->request_bytes;
    return QFE_OK;
}

int qfe_template_graph_audit(qfe_template t, int gradient, uint64_t* out4) {
    if (!t || !out4 || (gradient != 0 && gradient != 1)) return fail(QFE_ERR_ARG, "qfe_template_graph_audit: invalid argument");
    for (int k = 0; k < 4; ++k) out4[k] = 0;
    if (t->tier != 2) return QFE_OK;  // one-workgroup sequences are launched directly, never captured
    if (!t->gexec[gradient]) {
        int rc = graph_build(t, gradient);
        if (rc) return rc;
    }
    for (int k = 0; k < 4; ++k) out4[k] = t->audit[gradient][k];
    return QFE_OK;
}

int qfe_template_energy(qfe_template t, const double* params, const double* inputs, double* energy) {
    if (!t || !energy || (t->nparam && !params) || (t->ninputs && !inputs))
        return fail(QFE_ERR_ARG, "qfe_template_energy: null argument");
    if (t->tier == 2) return graph_run(t, 0, params, inputs, energy, nullptr);
    Single S;
    int rc = single_setup(t, params, inputs, false, S);
    if (rc) return rc;
    rc = launch_energy({t}, t->tier, t->lds_bytes, S.ddesc, S.dprefix, t->nitems, ctx().stream);
    if (rc) return rc;
    return fetch_doubles(S.dout, 1, energy);
}

int qfe_template_prepare(qfe_template t, const double* params, qfe_state out) {
    if (!t || !out || (t->nparam && !params)) return fail(QFE_ERR_ARG, "qfe_template_prepare: null argument");
    if (out->n != t->n) return fail(QFE_ERR_ARG, "qfe_template_prepare: state has the wrong number of qubits");
    std::vector<double> zeros(t->ninputs + 1, 0.0);
    Single S;
    int rc = single_setup(t, params, zeros.data(), false, S);
    if (rc) return rc;
    hipStream_t st = ctx().stream;
    hipLaunchKernelGGL(k_sec_prologue, dim3(1), dim3(256), 0, st, S.ddesc, 0);
    QFE_HIP(hipGetLastError());
    rc = t->tier == 2 ? wide_ansatz(t, S.ddesc, st) : launch_ansatz(t->tier, 1, t->lds_bytes, st, S.ddesc);
    if (!rc) rc = launch_zero(out->d, out->N, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sector_scatter, dim3((unsigned)((t->D + 255) / 256)), dim3(256), 0, st,
                       (const double2*)(S.work + t->off_psi), t->D, t->Db, t->afull, t->bfull, out->d);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

int qfe_template_gradient(qfe_template t, const double* params, const double* inputs, double* grad, double* energy) {
    if (!t || !grad || (t->nparam && !params) || (t->ninputs && !inputs))
        return fail(QFE_ERR_ARG, "qfe_template_gradient: null argument");
    std::vector<double> acc(t->ngen + 1);
    if (t->tier == 2) {
        int rc = graph_run(t, 1, params, inputs, &acc[t->ngen], acc.data());
        if (rc) return rc;
        gradient_from_overlaps(t, acc.data(), grad);
        if (energy) *energy = acc[t->ngen];
        return QFE_OK;
    }
    Single S;
    int rc = single_setup(t, params, inputs, true, S);
    if (!rc) rc = launch_gradient({t}, t->tier, t->lds_bytes, S.ddesc, ctx().stream);
    if (rc) return rc;
    rc = fetch_doubles((const double*)(S.work + t->off_gacc), t->ngen, acc.data());
    if (!rc) rc = fetch_doubles(S.dout, 1, &acc[t->ngen]);
    if (rc) return rc;
    gradient_from_overlaps(t, acc.data(), grad);
    if (energy) *energy = acc[t->ngen];
    return QFE_OK;
}

int qfe_template_residuals(qfe_template t, const double* params, const double* inputs, const uint64_t* idx,
                           uint64_t count, qfe_c64* out) {
    if (!t || (count && (!idx || !out)) || (t->nparam && !params) || (t->ninputs && !inputs))
        return fail(QFE_ERR_ARG, "qfe_template_residuals: null argument");
    std::vector<uint64_t> rank(count);
    for (uint64_t i = 0; i < count; ++i) {
        if (idx[i] >> t->n) return fail(QFE_ERR_ARG, "qfe_template_residuals: basis index out of range");
        const int32_t ra = t->sp.rank[0][compress(idx[i], 0, t->norb)];
        const int32_t rb = t->sp.rank[1][compress(idx[i], 1, t->norb)];
        if (ra < 0 || rb < 0) return fail(QFE_ERR_ARG, "qfe_template_residuals: basis index outside the sector");
        rank[i] = (uint64_t)ra * t->Db + (uint64_t)rb;
    }
    // sigma = H U|ref>, then U^dag on sigma: every step inverted, last step
    // first (psi is carried back to |ref> alongside)
    Single S;
    int rc = single_psi_sigma(t, params, inputs, S);
    if (rc) return rc;
    hipStream_t st = ctx().stream;
    for (uint32_t sdx = (uint32_t)t->step_pairs.size(); sdx-- > 0;) {
        if (!t->step_pairs[sdx]) continue;
        hipLaunchKernelGGL(k_sec_step, dim3(wide_blocks(t->step_pairs[sdx])), dim3(kWideTPB), 0, st, S.ddesc, sdx, 1, 1);
        QFE_HIP(hipGetLastError());
    }
    if (!count) return QFE_OK;
    void* scratch = nullptr;
    const size_t o_out = align256(count * 8);
    rc = scratch_reserve(o_out + count * sizeof(double2), &scratch);
    if (rc) return rc;
    QFE_HIP(hipMemcpyAsync(scratch, rank.data(), count * 8, hipMemcpyHostToDevice, st));
    double2* dres = (double2*)((char*)scratch + o_out);
    hipLaunchKernelGGL(k_sec_gather, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, S.ddesc,
                       (const uint64_t*)scratch, count, dres);
    QFE_HIP(hipGetLastError());
    QFE_HIP(hipMemcpyAsync(out, dres, count * sizeof(double2), hipMemcpyDeviceToHost, st));
    QFE_HIP(hipStreamSynchronize(st));
    return QFE_OK;
}

int qfe_template_pool_gradients(qfe_template t, const double* params, const double* inputs, const qfe_pauli* pool_terms,
                                const uint64_t* op_offsets, uint64_t nops, uint64_t batch, double* grads) {
    if (!t || (nops && (!pool_terms || !op_offsets || !grads)) || (t->nparam && !params) || (t->ninputs && !inputs))
        return fail(QFE_ERR_ARG, "qfe_template_pool_gradients: null argument");
    if (batch == 0) batch = 64;
    // pool operators as pair rotations on the template's sector
    ListPool lp;
    std::vector<SecStep> steps;
    std::vector<SecEntry> entries;
    std::vector<PoolOp> ops(nops);
    const uint64_t lim = (1ull << t->n) - 1ull;
    for (uint64_t o = 0; o < nops; ++o) {
        if (op_offsets[o + 1] < op_offsets[o]) return fail(QFE_ERR_ARG, "qfe_template_pool_gradients: offsets must not decrease");
        ops[o].e0 = (uint32_t)entries.size();
        int rc = build_runs(t->sp, lp, t->norb, lim, pool_terms, op_offsets[o], op_offsets[o + 1], 0, 1.0, (uint32_t)o, steps,
                            entries, "qfe_template_pool_gradients");
        if (rc) return rc;
        ops[o].e1 = (uint32_t)entries.size();
    }
    Single S;
    int rc = single_psi_sigma(t, params, inputs, S);
    if (rc) return rc;
    Packer pk(0);
    const size_t i_l = pk.add(lp.pool.data(), lp.pool.size() * sizeof(int2));
    const size_t i_e = pk.add(entries.data(), entries.size() * sizeof(SecEntry));
    const size_t i_o = pk.add(ops.data(), ops.size() * sizeof(PoolOp));
    const size_t o_out = pk.total;
    const size_t total = o_out + align256(nops * 8 + 8);
    void* dev = nullptr;
    QFE_HIP(hipMalloc(&dev, total));
    std::vector<char> img;
    pk.fill(img);
    hipError_t e = hipMemcpy(dev, img.data(), pk.total, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(dev);
        return hip_fail(e, "qfe_template_pool_gradients: upload");
    }
    char* base = (char*)dev;
    const double2* psi = (const double2*)(S.work + t->off_psi);
    const double2* sig = (const double2*)(S.work + t->off_sig);
    // batches of `batch` operators per launch
    for (uint64_t b0 = 0; b0 < nops; b0 += batch) {
        const uint64_t b1 = std::min(nops, b0 + batch);
        hipLaunchKernelGGL(k_sec_pool, dim3((unsigned)(b1 - b0)), dim3(256), 0, ctx().stream, t->dt,
                           (const SecEntry*)(base + pk.pieces[i_e].off), (const int2*)(base + pk.pieces[i_l].off),
                           (const PoolOp*)(base + pk.pieces[i_o].off) + b0, sig, psi, (double*)(base + o_out) + b0);
        e = hipGetLastError();
        if (e != hipSuccess) break;
    }
    if (e == hipSuccess && nops) rc = fetch_doubles((const double*)(base + o_out), nops, grads);
    (void)hipStreamSynchronize(ctx().stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(e, "qfe_template_pool_gradients: launch");
    return rc;
}

int qfe_batch_create(uint64_t hbm_budget_bytes, int max_active, qfe_batch* out) {
    if (!out || max_active < 1 || hbm_budget_bytes == 0) return fail(QFE_ERR_ARG, "qfe_batch_create: invalid argument");
    int rc = ensure_init();
    if (rc) return rc;
    qfe_batch b = new qfe_batch_s();
    b->budget = hbm_budget_bytes;
    b->max_active = max_active;
    b->lanes.resize(kLanes);
    for (int i = 0; i < kLanes; ++i) {
        Lane& L = b->lanes[i];
        L.tier = i / kLanesPerTier;
        rc = lane_stream(i, &L.stream);
        if (rc) {
            qfe_batch_destroy(b);
            return rc;
        }
        const hipError_t e = hipEventCreateWithFlags(&L.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            qfe_batch_destroy(b);
            return hip_fail(e, "qfe_batch_create");
        }
    }
    *out = b;
    return QFE_OK;
}

int qfe_batch_destroy(qfe_batch b) {
    if (!b) return QFE_OK;
    for (auto& L : b->lanes) {
        // the lane streams are shared by every engine: wait for this engine's own last launch only
        if (L.busy && L.done) (void)hipEventSynchronize(L.done);
        if (L.dev) (void)hipFree(L.dev);
        if (L.up) (void)hipHostFree(L.up);
        if (L.out) (void)hipHostFree(L.out);
        if (L.done) (void)hipEventDestroy(L.done);
    }
    delete b;
    return QFE_OK;
}

int qfe_batch_submit(qfe_batch b, qfe_template t, const double* params, const double* inputs, uint64_t* request_id) {
    if (!b || !t || !request_id) return fail(QFE_ERR_ARG, "qfe_batch_submit: invalid argument");
    if ((t->nparam && !params) || (t->ninputs && !inputs)) return fail(QFE_ERR_ARG, "qfe_batch_submit: null parameters");
    if (t->request_bytes > b->budget / kLanes)
        return fail(QFE_ERR_ARG, "qfe_batch_submit: one request exceeds the per-launch share of the HBM budget");
    b->queue[t->tier].push_back(make_request(b, t, params, inputs));
    *request_id = b->queue[t->tier].back().id;
    return QFE_OK;
}

int qfe_batch_submit_gradient(qfe_batch b, qfe_template t, const double* params, const double* inputs,
                              uint64_t* request_id) {
    if (!b || !t || !request_id) return fail(QFE_ERR_ARG, "qfe_batch_submit_gradient: invalid argument");
    if ((t->nparam && !params) || (t->ninputs && !inputs))
        return fail(QFE_ERR_ARG, "qfe_batch_submit_gradient: null parameters");
    if (t->grad_request_bytes > b->budget / kLanes)
        return fail(QFE_ERR_ARG, "qfe_batch_submit_gradient: one request exceeds the per-launch share of the HBM budget");
    b->gqueue[t->tier].push_back(make_request(b, t, params, inputs));
    *request_id = b->gqueue[t->tier].back().id;
    return QFE_OK;
}

int qfe_batch_take_gradient(qfe_batch b, uint64_t request_id, double* grad, uint64_t nparam) {
    if (!b || (nparam && !grad)) return fail(QFE_ERR_ARG, "qfe_batch_take_gradient: invalid argument");
    auto it = b->grads.find(request_id);
    if (it == b->grads.end())
        return fail(QFE_ERR_ARG, "qfe_batch_take_gradient: no completed gradient request with this id");
    if (it->second.size() != nparam) return fail(QFE_ERR_ARG, "qfe_batch_take_gradient: wrong gradient length");
    std::copy(it->second.begin(), it->second.end(), grad);
    b->grads.erase(it);
    return QFE_OK;
}

int qfe_batch_poll(qfe_batch b, uint64_t* done_ids, double* done_energies, uint64_t cap, uint64_t* ndone) {
    if (!b || !ndone || (cap && (!done_ids || !done_energies))) return fail(QFE_ERR_ARG, "qfe_batch_poll: invalid argument");
    int rc = collect(b, false);
    if (!rc) rc = fill_lanes(b);
    if (rc) return rc;
    return hand_out(b, done_ids, done_energies, cap, ndone);
}

int qfe_batch_step(qfe_batch b, uint64_t* done_ids, double* done_energies, uint64_t cap, uint64_t* ndone) {
    if (!b || !ndone || (cap && (!done_ids || !done_energies))) return fail(QFE_ERR_ARG, "qfe_batch_step: invalid argument");
    int rc = collect(b, false);
    if (!rc) rc = fill_lanes(b);
    if (!rc && b->done.empty() && any_busy(b)) {
        rc = collect(b, true);
        if (!rc) rc = fill_lanes(b);
    }
    if (rc) return rc;
    return hand_out(b, done_ids, done_energies, cap, ndone);
}

int qfe_batch_pending(qfe_batch b, uint64_t* queued, uint64_t* active) {
    if (!b) return fail(QFE_ERR_ARG, "qfe_batch_pending: invalid handle");
    if (queued) {
        *queued = 0;
        for (int i = 0; i < 3; ++i) *queued += b->queue[i].size() + b->gqueue[i].size();
    }
    if (active) *active = b->active + b->done.size();
    return QFE_OK;
}

}  // extern "C"
