// Gates and Pauli exponentials: lowering, pass planning and the sweep kernels.
//
// The reference applies every gate as its own full-state sweep
// (computer.cc:340-751) and every Pauli exponential as the 2*n_XY + 2(w-1) + 1
// gates exponentiate_pauli_string builds (exponentiate.py:9-87).  Here each
// operation is lowered to one of four device ops (K_U1, K_DIAG1, K_U2,
// K_PEXP) and a planner groups commuting-safe runs of ops into fused passes:
// a workgroup loads a tile of 2^k amplitudes spanning a chosen set of qubits
// into LDS, applies every op of the pass there, and writes the tile back once.
// A pass holding a single op uses a streaming kernel without LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "qfe_device.h"
#include "qfe_internal.h"

namespace qfe {

// ---------------------------------------------------------------------------
// Lowering (host)
// ---------------------------------------------------------------------------

static inline double cabs2(const qfe_c64& a) { return a.re * a.re + a.im * a.im; }
static inline double cabs(const qfe_c64& a) { return std::sqrt(cabs2(a)); }
static inline bool is_one(const qfe_c64& a) { return a.re == 1.0 && a.im == 0.0; }
static inline double2 d2(const qfe_c64& a) { return make_double2(a.re, a.im); }

// The reference's element threshold (computer.h:172).
constexpr double kComputeThreshold = 1.0e-16;

// A U1 whose matrix is exactly [[0, 1], [1, 0]] (X, the target half of cX)
// moves amplitudes without arithmetic; DevOp::hb flags it for the kernels.
static inline uint32_t is_swap(const double2* m) {
    return m[0].x == 0.0 && m[0].y == 0.0 && m[3].x == 0.0 && m[3].y == 0.0 && m[1].x == 1.0 && m[1].y == 0.0 &&
                   m[2].x == 1.0 && m[2].y == 0.0
               ? 1u
               : 0u;
}

// The two-qubit swap in the |c t> basis (rows and columns 2c + t): exactly
// 1 at (0,0), (1,2), (2,1), (3,3), 0 elsewhere.
static inline uint32_t is_qubit_swap(const double2* m) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            const bool one = (r == c && (r == 0 || r == 3)) || (r == 1 && c == 2) || (r == 2 && c == 1);
            const double2 v = m[4 * r + c];
            if (v.y != 0.0 || v.x != (one ? 1.0 : 0.0)) return 0u;
        }
    return 1u;
}

static void add_u1(int t, const qfe_c64 u[4], uint64_t cmask, std::vector<HostOp>& out) {
    // u = {m00, m01, m10, m11}; classified like computer.cc:353-422
    const bool diag_nz = cabs(u[0]) + cabs(u[3]) > kComputeThreshold;
    const bool off_nz = cabs(u[1]) + cabs(u[2]) > kComputeThreshold;
    HostOp h{};
    h.op.t = (uint32_t)t;
    h.op.cmask = cmask;
    h.op.cval = cmask;
    h.support = (1ull << t) | cmask;
    if (diag_nz && !off_nz) {
        if (is_one(u[0]) && is_one(u[3])) return;  // identity: the reference touches nothing
        h.op.kind = K_DIAG1;
        h.op.m[0] = d2(u[0]);
        h.op.m[1] = d2(u[3]);
        h.mix = 0;
    } else {
        h.op.kind = K_U1;
        for (int i = 0; i < 4; ++i) h.op.m[i] = d2(u[i]);
        h.op.hb = is_swap(h.op.m);
        h.mix = 1ull << t;
    }
    out.push_back(h);
}

int lower_gate(const qfe_gate& g, int n, std::vector<HostOp>& out) {
    const int t = g.target, c = g.control;
    if (t < 0 || c < 0) return fail(QFE_ERR_ARG, "apply_gate: negative qubit index");
    if (t == c) {
        if (t >= n) return fail(QFE_ERR_ARG, "apply_gate: target qubit " + std::to_string(t) + " out of range for " + std::to_string(n) + " qubits");
        const qfe_c64 u[4] = {g.m[0], g.m[1], g.m[4], g.m[5]};
        add_u1(t, u, 0, out);
        return QFE_OK;
    }
    // Controlled-U detection exactly as computer.cc:469-470.
    const bool controlled = (cabs(g.m[1]) + cabs(g.m[4]) < kComputeThreshold) && is_one(g.m[0]) && is_one(g.m[5]);
    if (controlled) {
        if (c >= n) return QFE_OK;  // the control bit reads 0 on every basis state: no-op
        if (t >= n) return fail(QFE_ERR_ARG, "apply_gate: target qubit " + std::to_string(t) + " out of range for " + std::to_string(n) + " qubits");
        const qfe_c64 u[4] = {g.m[10], g.m[11], g.m[14], g.m[15]};
        add_u1(t, u, 1ull << c, out);
        return QFE_OK;
    }
    if (t >= n || c >= n)
        return fail(QFE_ERR_ARG, "apply_gate: qubit index out of range for a 2-qubit gate on " + std::to_string(n) + " qubits");
    HostOp h{};
    h.op.kind = K_U2;
    h.op.t = (uint32_t)t;
    h.op.c = (uint32_t)c;
    for (int i = 0; i < 16; ++i) h.op.m[i] = d2(g.m[i]);
    h.op.hb = is_qubit_swap(h.op.m);  // a permutation a tile's index map absorbs
    h.mix = (1ull << t) | (1ull << c);
    h.support = h.mix;
    out.push_back(h);
    return QFE_OK;
}

int lower_pexp(const qfe_pauli& p, double angle, int n, std::vector<HostOp>& out) {
    const uint64_t lim = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
    if ((p.x & ~lim) || (p.z & ~lim)) return fail(QFE_ERR_ARG, "apply_pauli_exp: Pauli string acts on a qubit out of range");
    // Hermitian P needs conj(phase) = phase * (-1)^popcount(x & z).
    const bool odd = __builtin_popcountll(p.x & p.z) & 1;
    const double er = odd ? -p.phase.re : p.phase.re, ei = odd ? -p.phase.im : p.phase.im;
    if (std::fabs(er - p.phase.re) > 1e-15 || std::fabs(-ei - p.phase.im) > 1e-15)
        return fail(QFE_ERR_ARG, "apply_pauli_exp: the Pauli operator is not Hermitian");
    if (std::fabs(cabs2(p.phase) - 1.0) > 1e-12) return fail(QFE_ERR_ARG, "apply_pauli_exp: phase must have modulus 1");
    if (angle == 0.0) return QFE_OK;
    HostOp h{};
    h.op.kind = K_PEXP;
    h.op.x = p.x;
    h.op.z = p.z;
    h.op.hb = p.x ? (uint32_t)(63 - __builtin_clzll(p.x)) : 0u;
    h.op.m[0] = make_double2(std::cos(angle), std::sin(angle));
    h.op.m[1] = d2(p.phase);
    h.mix = p.x;
    h.support = p.x | p.z;
    out.push_back(h);
    return QFE_OK;
}

// ---------------------------------------------------------------------------
// Device: single-op streaming kernels
// ---------------------------------------------------------------------------

constexpr int kTPB = 256;  // threads per block for streaming kernels
constexpr int kUnroll = 4; // work items per thread

// A 2x2 on (a, b): a' = m00 a + m01 b, b' = m10 a + m11 b.
__device__ __forceinline__ void u2x2(double2& a, double2& b, const double2* m) {
    const double2 na = cmix(m[0], a, m[1], b);
    const double2 nb = cmix(m[2], a, m[3], b);
    a = na;
    b = nb;
}

// exp(i th P) on the pair (a at ja, b at jb = ja ^ x):
// a' = cos a + i sin ph (-1)^{jb.z} b ; b' = cos b + i sin ph (-1)^{ja.z} a.
__device__ __forceinline__ void pexp_pair(double2& a, double2& b, double c, double s, double2 ph, bool neg_a, bool neg_b) {
    const double2 iph = make_double2(-ph.y, ph.x);  // i * phase
    const double sa = neg_a ? -s : s;               // sign for the term feeding b (from a's index)
    const double sb = neg_b ? -s : s;               // sign for the term feeding a (from b's index)
    const double2 ta = cmul(iph, b);
    const double2 tb = cmul(iph, a);
    const double2 na = make_double2(c * a.x + sb * ta.x, c * a.y + sb * ta.y);
    const double2 nb = make_double2(c * b.x + sa * tb.x, c * b.y + sa * tb.y);
    a = na;
    b = nb;
}

// U1 on bit t with an optional single control bit c (cbit = 1<<c or 0).
template <bool NT>
__global__ __launch_bounds__(kTPB) void k_u1(double2* __restrict__ s, uint64_t nitems, uint32_t t, uint32_t c,
                                             uint64_t cbit, double2 m00, double2 m01, double2 m10, double2 m11) {
    const double2 m[4] = {m00, m01, m10, m11};
    const uint64_t base = xcd_block() * (kTPB * kUnroll) + threadIdx.x;
    const uint32_t lo = t < c ? t : c, hi = t < c ? c : t;
    double2 a[kUnroll], b[kUnroll];
    uint64_t j0[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t k = base + (uint64_t)u * kTPB;
        if (k < nitems) {
            j0[u] = cbit ? (ins00(k, lo, hi) | cbit) : ins0(k, t);
            a[u] = ld_s<NT>(s + j0[u]);
            b[u] = ld_s<NT>(s + (j0[u] | (1ull << t)));
        }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t k = base + (uint64_t)u * kTPB;
        if (k < nitems) {
            u2x2(a[u], b[u], m);
            st_s<NT>(s + j0[u], a[u]);
            st_s<NT>(s + (j0[u] | (1ull << t)), b[u]);
        }
    }
}

// Diagonal on bit t: multiply the bit-t = 0 half by d0 and the bit-t = 1 half
// by d1, skipping a half whose factor is exactly 1 (computer.cc:370-395).
// which: 0 = both halves, 1 = only bit=0 half, 2 = only bit=1 half.
__global__ __launch_bounds__(kTPB) void k_diag1(double2* __restrict__ s, uint64_t nitems, uint32_t t, uint32_t c,
                                                uint64_t cbit, double2 d0, double2 d1, int which) {
    const uint64_t base = xcd_block() * (kTPB * kUnroll) + threadIdx.x;
    const uint32_t lo = t < c ? t : c, hi = t < c ? c : t;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t k = base + (uint64_t)u * kTPB;
        if (k < nitems) {
            const uint64_t j0 = cbit ? (ins00(k, lo, hi) | cbit) : ins0(k, t);
            if (which != 2) s[j0] = cmul(d0, s[j0]);
            if (which != 1) {
                const uint64_t j1 = j0 | (1ull << t);
                s[j1] = cmul(d1, s[j1]);
            }
        }
    }
}

// General 4x4 on bits (t, c), basis 2*c_bit + t_bit.
__global__ __launch_bounds__(kTPB) void k_u2(double2* __restrict__ s, uint64_t nitems, uint32_t t, uint32_t c,
                                             const double2* __restrict__ mat) {
    const uint64_t base = xcd_block() * (kTPB * kUnroll) + threadIdx.x;
    const uint32_t lo = t < c ? t : c, hi = t < c ? c : t;
    double2 m[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = mat[i];
    const uint64_t tb = 1ull << t, cb = 1ull << c;
#pragma unroll 2
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t k = base + (uint64_t)u * kTPB;
        if (k < nitems) {
            const uint64_t j = ins00(k, lo, hi);
            const uint64_t idx[4] = {j, j | tb, j | cb, j | cb | tb};
            double2 v[4], r[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = s[idx[q]];
#pragma unroll
            for (int row = 0; row < 4; ++row) {
                double2 acc = make_double2(0.0, 0.0);
#pragma unroll
                for (int col = 0; col < 4; ++col) acc = cadd(acc, cmul(m[row * 4 + col], v[col]));
                r[row] = acc;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) s[idx[q]] = r[q];
        }
    }
}

// exp(i th P) with x != 0: pairs (ja, ja ^ x), ja with bit hb clear.
template <bool NT>
__global__ __launch_bounds__(kTPB) void k_pexp(double2* __restrict__ s, uint64_t nitems, uint32_t hb, uint64_t x,
                                               uint64_t z, double c, double sn, double2 ph) {
    const uint64_t base = xcd_block() * (kTPB * kUnroll) + threadIdx.x;
    double2 a[kUnroll], b[kUnroll];
    uint64_t ja[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t k = base + (uint64_t)u * kTPB;
        if (k < nitems) {
            ja[u] = ins0(k, hb);
            a[u] = ld_s<NT>(s + ja[u]);
            b[u] = ld_s<NT>(s + (ja[u] ^ x));
        }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t k = base + (uint64_t)u * kTPB;
        if (k < nitems) {
            const uint64_t jb = ja[u] ^ x;
            pexp_pair(a[u], b[u], c, sn, ph, __popcll(ja[u] & z) & 1, __popcll(jb & z) & 1);
            st_s<NT>(s + ja[u], a[u]);
            st_s<NT>(s + jb, b[u]);
        }
    }
}

// exp(i th P) with x == 0: diagonal phase cos + i sin ph (-1)^{j.z}.
__global__ __launch_bounds__(kTPB) void k_pexp_diag(double2* __restrict__ s, uint64_t N, uint64_t z, double c,
                                                    double sn, double2 ph) {
    const uint64_t base = xcd_block() * (kTPB * kUnroll) + threadIdx.x;
    const double2 fp = make_double2(c - sn * ph.y, sn * ph.x);  // cos + i sin ph
    const double2 fm = make_double2(c + sn * ph.y, -sn * ph.x); // cos - i sin ph
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t j = base + (uint64_t)u * kTPB;
        if (j < N) s[j] = cmul((__popcll(j & z) & 1) ? fm : fp, s[j]);
    }
}

static unsigned blocks_for(uint64_t nitems) {
    const uint64_t per = (uint64_t)kTPB * kUnroll;
    return (unsigned)((nitems + per - 1) / per);
}

static int launch_single(qfe_state_s* st, const HostOp& h, double2* dev_mat) {
    Context& cx = ctx();
    const DevOp& o = h.op;
    const uint64_t N = st->N;
    const uint64_t cbit = o.cmask;  // lowering produces at most one control bit
    const uint32_t cpos = cbit ? (uint32_t)(63 - __builtin_clzll(cbit)) : o.t;
    double bytes = 0.0;
    switch (o.kind) {
        case K_U1: {
            const uint64_t items = cbit ? N / 4 : N / 2;
            const uint32_t low = cbit ? std::min(o.t, cpos) : o.t;
            if (use_nt(st->n, low))
                hipLaunchKernelGGL(k_u1<true>, dim3(blocks_for(items)), dim3(kTPB), 0, cx.stream, st->d, items, o.t, cpos,
                                   cbit, o.m[0], o.m[1], o.m[2], o.m[3]);
            else
                hipLaunchKernelGGL(k_u1<false>, dim3(blocks_for(items)), dim3(kTPB), 0, cx.stream, st->d, items, o.t, cpos,
                                   cbit, o.m[0], o.m[1], o.m[2], o.m[3]);
            bytes = 64.0 * (double)items;
            break;
        }
        case K_DIAG1: {
            const uint64_t items = cbit ? N / 4 : N / 2;
            const bool one0 = o.m[0].x == 1.0 && o.m[0].y == 0.0;
            const bool one1 = o.m[1].x == 1.0 && o.m[1].y == 0.0;
            const int which = one0 ? 2 : (one1 ? 1 : 0);
            hipLaunchKernelGGL(k_diag1, dim3(blocks_for(items)), dim3(kTPB), 0, cx.stream, st->d, items, o.t, cpos, cbit,
                               o.m[0], o.m[1], which);
            bytes = (which ? 32.0 : 64.0) * (double)items;
            break;
        }
        case K_U2: {
            const uint64_t items = N / 4;
            hipLaunchKernelGGL(k_u2, dim3(blocks_for(items)), dim3(kTPB), 0, cx.stream, st->d, items, o.t, o.c,
                               (const double2*)dev_mat);
            bytes = 128.0 * (double)items;
            break;
        }
        case K_PEXP: {
            if (o.x) {
                const uint64_t items = N / 2;
                if (use_nt(st->n, o.hb))
                    hipLaunchKernelGGL(k_pexp<true>, dim3(blocks_for(items)), dim3(kTPB), 0, cx.stream, st->d, items,
                                       o.hb, o.x, o.z, o.m[0].x, o.m[0].y, o.m[1]);
                else
                    hipLaunchKernelGGL(k_pexp<false>, dim3(blocks_for(items)), dim3(kTPB), 0, cx.stream, st->d, items,
                                       o.hb, o.x, o.z, o.m[0].x, o.m[0].y, o.m[1]);
                bytes = 64.0 * (double)items;
            } else {
                hipLaunchKernelGGL(k_pexp_diag, dim3(blocks_for(N)), dim3(kTPB), 0, cx.stream, st->d, N, o.z, o.m[0].x,
                                   o.m[0].y, o.m[1]);
                bytes = 32.0 * (double)N;
            }
            break;
        }
        default:
            return fail(QFE_ERR_ARG, "internal: unknown op kind");
    }
    QFE_HIP(hipGetLastError());
    cx.counters.single++;
    cx.counters.bytes += bytes;
    return QFE_OK;
}

// ---------------------------------------------------------------------------
// Device: fused LDS tile pass
// ---------------------------------------------------------------------------

constexpr int kFusedTPB = 256;
// DevOp::pad of the head op of a group of one-qubit ops inside a tile pass:
// the group length, and kGroupPlain when every op of the group is a dense
// U1 with no condition.
constexpr uint32_t kGroupLen = 0xffu;
constexpr uint32_t kGroupPlain = 0x100u;
// A plain group whose ops are all diag(u0, u1) * R with R real (an Rz after an
// Ry, for one): every op carries R in m[0] = (r00, r01), m[1] = (r10, r11),
// and the head op carries in m[4 + i] the product of the ops' row phases for
// element i of the group (bit k of i = the value of op k's target).
constexpr uint32_t kGroupRot = 0x200u;
constexpr int kWarmOpsMaxTile = 10;  // tiles up to this many qubits touch their ops first
constexpr int kMaxTileQubits = 13;  // 2^13 * 16 B = 128 KiB of LDS

struct PassDesc {
    uint64_t bmask;          // tile qubits (global mask)
    uint32_t k;              // popcount(bmask)
    uint32_t nlow;           // trailing contiguous tile bits 0..nlow-1
    uint32_t nops;
    uint32_t nrest;          // number of non-tile qubits
    uint8_t hipos[kMaxTileQubits];   // global positions of tile bits nlow..k-1
    uint8_t restpos[64];             // global positions of non-tile bits, ascending
};

// Apply U1 op `o` along item axis AX (pairs i, i | 2^AX of the 8 amplitudes
// a thread holds); e[i] are the tile indices (conditions on other bits).
template <int AX>
__device__ __forceinline__ void group_axis(double2 (&a)[8], const uint32_t (&e)[8], const DevOp& o, uint64_t base) {
    if ((base & o.omask) != o.oval) return;  // uniform across the block
    const uint32_t cm = (uint32_t)o.cmask, cv = (uint32_t)o.cval;
    const double2 m0 = o.m[0], m1 = o.m[1], m2 = o.m[2], m3 = o.m[3];
    const bool swap = o.hb != 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i & (1 << AX)) continue;
        const int j = i | (1 << AX);
        if ((e[i] & cm) != cv) continue;
        if (swap) {
            const double2 t = a[i];
            a[i] = a[j];
            a[j] = t;
        } else {
            const double2 x = a[i], y = a[j];
            a[i] = cmix(m0, x, m1, y);
            a[j] = cmix(m2, x, m3, y);
        }
    }
}

// Logical -> physical index map of an LDS tile: P(l) = M l ^ off over GF(2),
// M given by its columns col[b] = M e_b.  CNOT, anti-controlled CNOT and X
// gates on tile bits permute amplitudes, so instead of moving data they
// update the map (P' = P o G: a CNOT c -> t adds col[t] to col[c], an X on t
// adds col[t] to off); every other op reads and writes tile[P(l)] for its
// logical indices l, and the write-back reads tile[P(l)] for amplitude l.
// P is evaluated per element from two LDS tables (index bits 0-5 and 6-12),
// rebuilt once after each run of map updates: two reads and an XOR instead
// of a GF(2) product over every tile bit.
// threadIdx.x behind an empty volatile asm: conditions on it inside the op
// loop of k_fused are evaluated where they are used instead of being hoisted
// out of the loop as lane masks, each holding two scalar registers for the
// whole kernel (on 11-qubit tiles those masks pushed the kernel past its
// register budget into scratch spills).
__device__ __forceinline__ uint32_t tid_here() {
    uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

template <int KQ>
struct TileMap {
    // column b of M in lane b of one vector register (every wave keeps the
    // same copy): a column at a runtime bit is one v_readlane and an update
    // a lane compare and select, where columns in scalar registers need a
    // compare chain over every tile bit
    uint32_t colv;
    uint32_t lane;
    uint32_t off;
    bool perm;   // false: P is the identity (no translation)
    bool dirty;  // the lookup tables are older than the map
    const uint32_t* tlo;  // LDS: P on index bits 0-5, with off
    const uint32_t* thi;  // LDS: M on index bits 6 and up
    __device__ __forceinline__ void init(const uint32_t* lo, const uint32_t* hi) {
        lane = threadIdx.x & (kWave - 1);
        colv = lane < (uint32_t)KQ ? 1u << lane : 0u;
        off = 0;
        perm = false;
        dirty = false;
        tlo = lo;
        thi = hi;
    }
    // column b (b uniform)
    __device__ __forceinline__ uint32_t column(uint32_t b) const {
        return (uint32_t)__builtin_amdgcn_readlane((int)colv, (int)b);
    }
    // M v from the columns, bit by bit (table builds)
    __device__ __forceinline__ uint32_t lin_cols(uint32_t v) const {
        uint32_t r = 0;
#pragma unroll
        for (int b = 0; b < KQ; ++b) r ^= ((v >> b) & 1u) ? column((uint32_t)b) : 0u;
        return r;
    }
    // M v = P(v) ^ P(0) from the lookup tables (after refresh): two uniform
    // LDS reads instead of a scalar loop over every tile bit
    __device__ __forceinline__ uint32_t lin(uint32_t v) const { return perm ? (tlo[v & 63u] ^ thi[v >> 6] ^ tlo[0]) : v; }
    // P(l) = M l ^ off: linear over GF(2), so it splits into two lookups
    __device__ __forceinline__ uint32_t operator()(uint32_t l) const { return perm ? (tlo[l & 63u] ^ thi[l >> 6]) : l; }
    // G = SWAP of tile bits a and b: M' e_a = M e_b and M' e_b = M e_a
    __device__ __forceinline__ void swap_cols(uint32_t a, uint32_t b) {
        const uint32_t ca = column(a), cb = column(b);
        colv = lane == a ? cb : (lane == b ? ca : colv);
        perm = true;
        dirty = true;
    }
    // G = X on t, controlled by tile bit c when cmask != 0 (cval: required value)
    __device__ __forceinline__ void apply_swap(uint32_t t, uint32_t cmask, uint32_t cval) {
        const uint32_t ct = column(t);
        if (cmask) {
            const uint32_t c = (uint32_t)__builtin_ctz(cmask);
            const uint32_t nc = column(c) ^ ct;
            colv = lane == c ? nc : colv;
            if (!(cval & cmask)) off ^= ct;
        } else {
            off ^= ct;
        }
        perm = true;
        dirty = true;
    }
    // Rebuild the lookup tables after map changes (uniform; every thread calls it).
    __device__ __forceinline__ void refresh(uint32_t* lo, uint32_t* hi) {
        if (!dirty) return;
        const uint32_t t = tid_here();
        if (t < 64) lo[t] = lin_cols(t) ^ off;
        if (t < (1u << (KQ > 6 ? KQ - 6 : 0))) hi[t] = lin_cols(t << 6);
        __syncthreads();
        dirty = false;
    }
};

// 2 or 3 consecutive U1 ops with distinct tile targets t0, t1 (, t2): every
// thread holds the 8 amplitudes spanned by the three bits (a free bit pads a
// pair) and applies the ops in order, partners at compile-time positions.
template <uint32_t T, int KQ>
__device__ __forceinline__ void fused_group(double2* tile, const TileMap<KQ>& P, uint64_t base,
                                            const DevOp* __restrict__ g, uint32_t cnt) {
    const uint32_t t0 = g[0].t, t1 = g[1].t;
    uint32_t t2 = 0;
    if (cnt > 2) {
        t2 = g[2].t;
    } else {
        while (t2 == t0 || t2 == t1) ++t2;
    }
    // ascending positions for the zero insertions
    const uint32_t lo = min(t0, min(t1, t2)), hi = max(t0, max(t1, t2));
    const uint32_t mid = t0 ^ t1 ^ t2 ^ lo ^ hi;
    const uint32_t b0 = 1u << t0, b1 = 1u << t1, b2 = 1u << t2;
    const uint32_t p0 = P.lin(b0), p1 = P.lin(b1), p2 = P.lin(b2);
    for (uint32_t it = tid_here(); it < T / 8; it += kFusedTPB) {
        const uint32_t eb = ins0_32(ins0_32(ins0_32(it, lo), mid), hi);
        const uint32_t pb = P(eb);
        uint32_t e[8], pe[8];
        double2 a[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            e[i] = eb | ((i & 1) ? b0 : 0u) | ((i & 2) ? b1 : 0u) | ((i & 4) ? b2 : 0u);
            pe[i] = pb ^ ((i & 1) ? p0 : 0u) ^ ((i & 2) ? p1 : 0u) ^ ((i & 4) ? p2 : 0u);
            a[i] = tile[pe[i]];
        }
        group_axis<0>(a, e, g[0], base);
        group_axis<1>(a, e, g[1], base);
        if (cnt > 2) group_axis<2>(a, e, g[2], base);
#pragma unroll
        for (int i = 0; i < 8; ++i) tile[pe[i]] = a[i];
    }
}

// Dense U1 along item axis AX: every pair of the 8 amplitudes.
template <int AX>
__device__ __forceinline__ void plain_axis(double2 (&a)[8], const DevOp& o) {
    const double2 m0 = o.m[0], m1 = o.m[1], m2 = o.m[2], m3 = o.m[3];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i & (1 << AX)) continue;
        const int j = i | (1 << AX);
        const double2 x = a[i], y = a[j];
        a[i] = cmix(m0, x, m1, y);
        a[j] = cmix(m2, x, m3, y);
    }
}

// A group of 2 or 3 dense U1 ops with no condition at all (kGroupPlain):
// fused_group without the per-element condition tests and swap selects, as
// straight-line code (no loop-carried amplitude array, so no register copies).
template <uint32_t T, int KQ>
__device__ __forceinline__ void plain_group(double2* tile, const TileMap<KQ>& P, const DevOp* __restrict__ g,
                                            uint32_t cnt) {
    const uint32_t t0 = g[0].t, t1 = g[1].t;
    uint32_t t2 = 0;
    if (cnt > 2) {
        t2 = g[2].t;
    } else {
        while (t2 == t0 || t2 == t1) ++t2;
    }
    const uint32_t lo = min(t0, min(t1, t2)), hi = max(t0, max(t1, t2));
    const uint32_t mid = t0 ^ t1 ^ t2 ^ lo ^ hi;
    const uint32_t p0 = P.lin(1u << t0), p1 = P.lin(1u << t1), p2 = P.lin(1u << t2);
    for (uint32_t it = tid_here(); it < T / 8; it += kFusedTPB) {
        const uint32_t pb = P(ins0_32(ins0_32(ins0_32(it, lo), mid), hi));
        uint32_t pe[8];
        double2 a[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            pe[i] = pb ^ ((i & 1) ? p0 : 0u) ^ ((i & 2) ? p1 : 0u) ^ ((i & 4) ? p2 : 0u);
            a[i] = tile[pe[i]];
        }
        plain_axis<0>(a, g[0]);
        plain_axis<1>(a, g[1]);
        if (cnt > 2) plain_axis<2>(a, g[2]);
#pragma unroll
        for (int i = 0; i < 8; ++i) tile[pe[i]] = a[i];
    }
}

// Real rotation R of a kGroupRot op along item axis AX of N elements: one
// multiply and one FMA per output component.
template <int AX, int N>
__device__ __forceinline__ void rot_axis(double2 (&a)[N], const DevOp& o) {
    const double r00 = o.m[0].x, r01 = o.m[0].y, r10 = o.m[1].x, r11 = o.m[1].y;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i & (1 << AX)) continue;
        const int j = i | (1 << AX);
        const double2 x = a[i], y = a[j];
        a[i] = make_double2(fma(r01, y.x, r00 * x.x), fma(r01, y.y, r00 * x.y));
        a[j] = make_double2(fma(r11, y.x, r10 * x.x), fma(r11, y.y, r10 * x.y));
    }
}

// A kGroupRot group of cnt <= G ops on G item axes (free tile bits pad the
// axes of a shorter group): the ops' real rotations, then every element times
// the product of the ops' row phases (the phases of op k act on axis k only,
// so they commute past the rotations of the other axes).  For three ops, 16
// FP64 instructions per element against 24 for three dense complex matrices.
// Phase of element i: head m[4 + i] for i < 12, else m[i - 8] of the second op.
template <uint32_t T, int KQ, int G>
__device__ __forceinline__ void rot_group(double2* tile, const TileMap<KQ>& P, const DevOp* __restrict__ g,
                                          uint32_t cnt) {
    constexpr int N = 1 << G;
    uint32_t t[G];
    uint32_t used = 0;
#pragma unroll
    for (int k = 0; k < G; ++k)
        if ((uint32_t)k < cnt) used |= 1u << g[k].t;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if ((uint32_t)k < cnt) {
            t[k] = g[k].t;
        } else {
            t[k] = (uint32_t)__builtin_ctz(~used);
            used |= 1u << t[k];
        }
    }
    // insertion positions, ascending
    uint32_t srt[G];
#pragma unroll
    for (int k = 0; k < G; ++k) srt[k] = t[k];
#pragma unroll
    for (int x = 0; x < G; ++x)
#pragma unroll
        for (int y = 0; y + 1 < G - x; ++y) {
            const uint32_t lo = min(srt[y], srt[y + 1]), hi = max(srt[y], srt[y + 1]);
            srt[y] = lo;
            srt[y + 1] = hi;
        }
    uint32_t pk[G];
#pragma unroll
    for (int k = 0; k < G; ++k) pk[k] = P.lin(1u << t[k]);
    for (uint32_t it = tid_here(); it < T / N; it += kFusedTPB) {
        uint32_t e = it;
#pragma unroll
        for (int k = 0; k < G; ++k) e = ins0_32(e, srt[k]);
        const uint32_t pb = P(e);
        double2 a[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            uint32_t q = pb;
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (i & (1 << k)) q ^= pk[k];
            a[i] = tile[q];
        }
        rot_axis<0, N>(a, g[0]);
        rot_axis<1, N>(a, g[1]);
        if (cnt > 2) rot_axis<2, N>(a, g[2]);
        if constexpr (G > 3) {
            if (cnt > 3) rot_axis<3, N>(a, g[3]);
        }
        // the store offsets are recomputed from an opaque copy of pb rather
        // than kept live across the rotations (N registers fewer)
        uint32_t pw = pb;
        asm volatile("" : "+v"(pw));
#pragma unroll
        for (int i = 0; i < N; ++i) {
            uint32_t q = pw;
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (i & (1 << k)) q ^= pk[k];
            tile[q] = cmul(i < 12 ? g[0].m[4 + i] : g[1].m[i - 8], a[i]);
        }
    }
}

template <int KQ>
__global__ __launch_bounds__(kFusedTPB) __attribute__((amdgpu_waves_per_eu(KQ <= 11 ? 4 : 1))) void k_fused(double2* __restrict__ s, uint64_t bmask, uint32_t nlow,
                                                     uint32_t nt, const DevOp* __restrict__ ops, uint32_t nops) {
    extern __shared__ double2 tile[];
    __shared__ uint64_t hioff[1 << 9];
    __shared__ uint32_t ptab_lo[64], ptab_hi[1 << (kMaxTileQubits - 6)];
    constexpr uint32_t T = 1u << KQ;
    const uint32_t nhi = KQ - nlow;
    const uint64_t himask = bmask & ~((1ull << nlow) - 1ull);
    // tile base: the tile number with a zero bit inserted at every tile qubit
    // (the low nlow tile qubits are a plain shift); pure scalar ALU
    auto tile_base = [&](uint64_t tn) {
        uint64_t b = tn << nlow;
        for (uint64_t m = himask; m; m &= m - 1) {
            const uint32_t p = (uint32_t)__builtin_ctzll(m);
            const uint64_t lo = b & ((1ull << p) - 1ull);
            b = ((b ^ lo) << 1) | lo;
        }
        return b;
    };
    // offsets of the high tile qubits (2^nhi entries, nhi <= 9)
    for (uint32_t h = threadIdx.x; h < (1u << nhi); h += kFusedTPB) {
        uint64_t off = 0;
        uint32_t i = 0;
        for (uint64_t m = himask; m; m &= m - 1, ++i)
            if ((h >> i) & 1u) off |= 1ull << __builtin_ctzll(m);
        hioff[h] = off;
    }
    __syncthreads();
    const uint32_t lowmask = (1u << nlow) - 1u;
    const uint64_t base = tile_base(xcd_block());
    for (uint32_t l = threadIdx.x; l < T; l += kFusedTPB) {
        const double2* src = s + (base | hioff[l >> nlow] | (l & lowmask));
        tile[l] = nt ? ld_nt(src) : *src;
    }
    if constexpr (KQ <= kWarmOpsMaxTile) {
        // Small tiles: a workgroup's life is its chain of ops, and each op's
        // first scalar load would miss the scalar cache (the ops of a pass are
        // read once per workgroup; 500 gates of any kind took 0.9 us each at
        // 12 qubits).  Touch the first 144 bytes of every op (its head and
        // m[0..3]) while the tile loads are in flight, eight ops per wait, so
        // the chain then hits the cache.
        uint32_t touch = 0;
#pragma unroll 8
        for (uint32_t oi = 0; oi < nops; ++oi) {
            const uint32_t* w = reinterpret_cast<const uint32_t*>(ops + oi);
            touch ^= w[0] ^ w[16] ^ w[32];
        }
        asm volatile("" ::"s"(touch));
    }
    TileMap<KQ> P;
    P.init(ptab_lo, ptab_hi);
    for (uint32_t oi = 0; oi < nops; ++oi) {
        const DevOp& o = ops[oi];
        if (o.kind == K_U1 && o.hb && __popc((uint32_t)o.cmask) <= 1) {
            // a permutation of the tile: the map absorbs it (registers only,
            // so no barrier)
            if ((base & o.omask) == o.oval) P.apply_swap(o.t, (uint32_t)o.cmask, (uint32_t)o.cval);
            continue;
        }
        if (o.kind == K_U2 && o.hb) {  // a qubit swap, absorbed the same way
            P.swap_cols(o.t, o.c);
            continue;
        }
        __syncthreads();
        P.refresh(ptab_lo, ptab_hi);
        if (o.kind == K_U1 && (o.pad & kGroupLen) >= 2) {
            // a group of one-qubit ops on distinct tile bits: one LDS round
            // trip for the whole group
            const uint32_t cnt = o.pad & kGroupLen;
            if ((o.pad & kGroupRot) && cnt > 3)
                rot_group<T, KQ, 4>(tile, P, ops + oi, cnt);
            else if (o.pad & kGroupRot)
                rot_group<T, KQ, 3>(tile, P, ops + oi, cnt);
            else if (o.pad & kGroupPlain)
                plain_group<T, KQ>(tile, P, ops + oi, cnt);
            else
                fused_group<T, KQ>(tile, P, base, ops + oi, cnt);
            oi += cnt - 1;
            continue;
        }
        if ((base & o.omask) != o.oval) continue;  // uniform across the block
        switch (o.kind) {
            case K_U1: {
                const uint32_t tb = 1u << o.t, ptb = P.lin(tb);
                if (o.hb) {  // X on the target with two or more conditions
                    for (uint32_t p = tid_here(); p < T / 2; p += kFusedTPB) {
                        const uint32_t l0 = ins0_32(p, o.t);
                        if ((l0 & (uint32_t)o.cmask) != (uint32_t)o.cval) continue;
                        const uint32_t q0 = P(l0);
                        const double2 a = tile[q0];
                        tile[q0] = tile[q0 ^ ptb];
                        tile[q0 ^ ptb] = a;
                    }
                    break;
                }
                const double2 m[4] = {o.m[0], o.m[1], o.m[2], o.m[3]};
                // the pairs of one op are disjoint: a thread reads a batch of
                // its pairs before writing any, so the LDS reads overlap
                // (batches of two keep the registers of 11-qubit tiles within
                // four waves per SIMD without spilling)
                constexpr uint32_t PP = (T / 2 + kFusedTPB - 1) / kFusedTPB;
                constexpr uint32_t PB = PP < 2 ? PP : 2;
                for (uint32_t r0 = 0; r0 < PP; r0 += PB) {
                    uint32_t q0[PB];
                    bool on[PB];
                    double2 a[PB], b[PB];
#pragma unroll
                    for (uint32_t r = 0; r < PB; ++r) {
                        const uint32_t p = tid_here() + (r0 + r) * kFusedTPB;
                        const uint32_t l0 = ins0_32(p, o.t);
                        q0[r] = P(l0);
                        on[r] = p < T / 2 && (l0 & (uint32_t)o.cmask) == (uint32_t)o.cval;
                        if (on[r]) {
                            a[r] = tile[q0[r]];
                            b[r] = tile[q0[r] ^ ptb];
                        }
                    }
#pragma unroll
                    for (uint32_t r = 0; r < PB; ++r) {
                        if (on[r]) {
                            u2x2(a[r], b[r], m);
                            tile[q0[r]] = a[r];
                            tile[q0[r] ^ ptb] = b[r];
                        }
                    }
                }
                break;
            }
            case K_DIAG1: {
                if (o.tin) {
                    const uint32_t ptb = P.lin(1u << o.t);
                    for (uint32_t p = tid_here(); p < T / 2; p += kFusedTPB) {
                        const uint32_t l0 = ins0_32(p, o.t);
                        if ((l0 & (uint32_t)o.cmask) != (uint32_t)o.cval) continue;
                        const uint32_t q0 = P(l0);
                        tile[q0] = cmul(o.m[0], tile[q0]);
                        tile[q0 ^ ptb] = cmul(o.m[1], tile[q0 ^ ptb]);
                    }
                } else {
                    const double2 f = ((base >> o.t) & 1ull) ? o.m[1] : o.m[0];
                    for (uint32_t l = tid_here(); l < T; l += kFusedTPB) {
                        if ((l & (uint32_t)o.cmask) != (uint32_t)o.cval) continue;
                        const uint32_t q = P(l);
                        tile[q] = cmul(f, tile[q]);
                    }
                }
                break;
            }
            case K_U2: {
                const uint32_t lo = o.t < o.c ? o.t : o.c, hi = o.t < o.c ? o.c : o.t;
                const uint32_t ptb = P.lin(1u << o.t), pcb = P.lin(1u << o.c);
                for (uint32_t p = tid_here(); p < T / 4; p += kFusedTPB) {
                    const uint32_t q = P(ins0_32(ins0_32(p, lo), hi));
                    const uint32_t idx[4] = {q, q ^ ptb, q ^ pcb, q ^ pcb ^ ptb};
                    double2 v[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) v[k] = tile[idx[k]];
#pragma unroll
                    for (int row = 0; row < 4; ++row) {
                        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
                        for (int col = 0; col < 4; ++col) acc = cadd(acc, cmul(o.m[row * 4 + col], v[col]));
                        tile[idx[row]] = acc;
                    }
                }
                break;
            }
            case K_PEXP: {
                const double c = o.m[0].x, sn = o.m[0].y;
                const double2 ph = o.m[1];
                const bool bneg = __popcll(base & o.ozmask) & 1;  // sign from out-of-tile z bits
                const uint32_t zl = (uint32_t)o.z;
                if (o.x == 0) {
                    const double2 fp = make_double2(c - sn * ph.y, sn * ph.x);
                    const double2 fm = make_double2(c + sn * ph.y, -sn * ph.x);
                    for (uint32_t l = tid_here(); l < T; l += kFusedTPB) {
                        const bool neg = ((__popc(l & zl) & 1) != 0) != bneg;
                        const uint32_t q = P(l);
                        tile[q] = cmul(neg ? fm : fp, tile[q]);
                    }
                } else {
                    const uint32_t xl = (uint32_t)o.x, pxl = P.lin(xl);
                    for (uint32_t p = tid_here(); p < T / 2; p += kFusedTPB) {
                        const uint32_t la = ins0_32(p, o.hb);
                        const uint32_t lb = la ^ xl;
                        const uint32_t qa = P(la), qb = qa ^ pxl;
                        double2 a = tile[qa], b = tile[qb];
                        const bool na = ((__popc(la & zl) & 1) != 0) != bneg;
                        const bool nb = ((__popc(lb & zl) & 1) != 0) != bneg;
                        pexp_pair(a, b, c, sn, ph, na, nb);
                        tile[qa] = a;
                        tile[qb] = b;
                    }
                }
                break;
            }
            default:
                break;
        }
    }
    __syncthreads();
    P.refresh(ptab_lo, ptab_hi);
    for (uint32_t l = threadIdx.x; l < T; l += kFusedTPB) {
        double2* dst = s + (base | hioff[l >> nlow] | (l & lowmask));
        const double2 v = tile[P(l)];
        if (nt)
            st_nt(dst, v);
        else
            *dst = v;
    }
}

// Map a global mask to tile-local bit positions.
static uint64_t to_local(uint64_t m, const int* local_of) {
    uint64_t r = 0;
    while (m) {
        const int b = __builtin_ctzll(m);
        m &= m - 1;
        r |= 1ull << local_of[b];
    }
    return r;
}

// Rewrite a global op for a tile with qubit mask bmask.
static DevOp localize(const DevOp& g, uint64_t bmask, const int* local_of) {
    DevOp o = g;
    o.omask = g.cmask & ~bmask;
    o.oval = g.cval & ~bmask;
    o.cmask = to_local(g.cmask & bmask, local_of);
    o.cval = to_local(g.cval & bmask, local_of);
    o.ozmask = 0;
    o.tin = 0;
    switch (g.kind) {
        case K_U1:
            o.t = (uint32_t)local_of[g.t];
            break;
        case K_DIAG1:
            if ((bmask >> g.t) & 1ull) {
                o.tin = 1;
                o.t = (uint32_t)local_of[g.t];
            } else {
                o.tin = 0;  // o.t stays global: read from the tile base
            }
            break;
        case K_U2:
            o.t = (uint32_t)local_of[g.t];
            o.c = (uint32_t)local_of[g.c];
            break;
        case K_PEXP:
            o.x = to_local(g.x, local_of);
            o.hb = o.x ? (uint32_t)(63 - __builtin_clzll(o.x)) : 0u;
            o.z = to_local(g.z & bmask, local_of);
            o.ozmask = g.z & ~bmask;
            break;
    }
    return o;
}

// ---------------------------------------------------------------------------
// Device: register-orbit pass for runs of Pauli exponentials
//
// The flip masks x_1..x_r of a run of exponentials span a space of dimension
// M <= 4 over GF(2).  With a reduced basis b_0..b_{M-1} (pivot = highest bit,
// each pivot bit present in one basis vector only), every index j with all
// pivot bits clear owns the orbit {j ^ c_i}, c_i = XOR of b_k over the bits k
// of i.  One thread loads its 2^M amplitudes into registers, applies every
// exponential of the run there (exponential o pairs i with i ^ v_o, v_o its
// mask in orbit coordinates) and stores them: one streaming sweep for the
// whole run, no LDS and no barriers.
// ---------------------------------------------------------------------------

constexpr int kOrbitMax = 4;
constexpr int kLaneOrbitMax = 7;  // lane orbits: 16 amplitudes per lane over 2, 4 or 8 lanes
constexpr int kOrbitTPB = 256;

struct OrbitOp {
    uint32_t v;      // mask in orbit coordinates (0: diagonal exponential)
    uint32_t smask;  // bit i = parity(c_i & z)
    uint64_t z;
    double c, sn;
    double2 ph;
};

struct OrbitDesc {
    uint64_t c[1 << kOrbitMax];  // orbit element offsets (host planning)
    uint64_t b[kOrbitMax];       // basis vectors, pivots ascending
    uint32_t piv[kOrbitMax];     // pivot bits, ascending
};

// exp(i th P) inside an orbit: every element i takes
//   a_i' = cos a_i + i sin ph (-1)^{s(i ^ v)} a_{i ^ v},
// the same formula for both members of each pair; the partner array comes
// from the xor butterfly, so no array index depends on v.
template <int M>
__device__ __forceinline__ void orbit_pexp(double2 (&a)[1 << M], uint32_t v, uint32_t sm, double c, double sn,
                                           double2 ph) {
    const double2 iph = make_double2(-ph.y, ph.x);  // i * phase
    dispatch_mask<M>(v, [&](auto vc) {
        constexpr int V = (int)decltype(vc)::value;
        // pair by pair (i, k = i ^ V), both updated from the old values
#pragma unroll
        for (int i = 0; i < (1 << M); ++i) {
            const int k = i ^ V;
            if (k < i) continue;
            const double2 x = a[i], y = a[k];
            const double si = ((sm >> k) & 1u) ? -sn : sn;  // the sign belongs to the partner
            const double2 ty = cmul(iph, y);
            a[i] = make_double2(c * x.x + si * ty.x, c * x.y + si * ty.y);
            if (k != i) {
                const double sk = ((sm >> i) & 1u) ? -sn : sn;
                const double2 tx = cmul(iph, x);
                a[k] = make_double2(c * y.x + sk * tx.x, c * y.y + sk * tx.y);
            }
        }
    });
}

// The basis travels as scalar kernel arguments (a by-value struct argument
// indexed in the unrolled loops was copied to scratch memory).
template <int M, bool NT>
__global__ __launch_bounds__(kOrbitTPB) void k_orbit(double2* __restrict__ s, uint64_t norbit, uint64_t b0,
                                                     uint64_t b1, uint64_t b2, uint64_t b3, uint32_t p0, uint32_t p1,
                                                     uint32_t p2, uint32_t p3, const OrbitOp* __restrict__ ops,
                                                     uint32_t nops) {
    const uint64_t t = xcd_block() * kOrbitTPB + threadIdx.x;
    if (t >= norbit) return;
    const uint64_t bb[4] = {b0, b1, b2, b3};
    const uint32_t pp[4] = {p0, p1, p2, p3};
    uint64_t j = t;
#pragma unroll
    for (int k = 0; k < M; ++k) j = ins0(j, pp[k]);
    uint64_t off[1 << M];
#pragma unroll
    for (int i = 0; i < (1 << M); ++i) {
        uint64_t c = 0;
#pragma unroll
        for (int k = 0; k < M; ++k)
            if ((i >> k) & 1) c ^= bb[k];
        off[i] = j ^ c;
    }
    double2 a[1 << M];
#pragma unroll
    for (int i = 0; i < (1 << M); ++i) a[i] = ld_s<NT>(s + off[i]);
    for (uint32_t o = 0; o < nops; ++o) {
        const OrbitOp op = ops[o];
        const bool pj = __popcll(j & op.z) & 1;
        if (op.v == 0) {
            const double2 fp = make_double2(op.c - op.sn * op.ph.y, op.sn * op.ph.x);
            const double2 fm = make_double2(op.c + op.sn * op.ph.y, -op.sn * op.ph.x);
#pragma unroll
            for (int i = 0; i < (1 << M); ++i) {
                const bool neg = pj != (((op.smask >> i) & 1u) != 0);
                a[i] = cmul(neg ? fm : fp, a[i]);
            }
        } else {
            orbit_pexp<M>(a, op.v, op.smask ^ (pj ? ~0u : 0u), op.c, op.sn, op.ph);
        }
    }
#pragma unroll
    for (int i = 0; i < (1 << M); ++i) st_s<NT>(s + off[i], a[i]);
}

// ---------------------------------------------------------------------------
// Device: register orbits of 4 + Q dimensions spread over 2^Q lanes
//
// An orbit of 2^(4+Q) amplitudes is held by 2^Q lanes of one wave, 16
// amplitudes each: lane group h owns the elements i = h * 16 + i'.  An
// exponential whose orbit flip v keeps the top Q bits (v >> 4 == 0) pairs
// elements inside a lane as in k_orbit; otherwise the partner of i' lives in
// the lane whose group is h ^ (v >> 4), at i' ^ (v & 15): the lane permutes
// its own 16 values by v & 15 (xor butterfly) and swaps them with that lane
// by a cross-lane shuffle.  Lanes of one group hold consecutive orbits, so a
// wave reads runs of 2^(6-Q) consecutive amplitudes per element.
// ---------------------------------------------------------------------------

struct OrbitOp2 {
    uint32_t v;
    uint32_t pad;
    uint64_t smask[2];  // bit i of the 128-bit word = parity(c_i & z), i over 2^(4+Q) elements
    uint64_t z;
    double c, sn;
    double2 ph;
};

// The E = 2^R sign bits of lane group g (elements g * E .. g * E + E - 1).
template <int R, int Q>
__device__ __forceinline__ uint32_t group_signs(const OrbitOp2& op, uint32_t g) {
    constexpr uint32_t WMASK = (R == 5) ? 0xFFFFFFFFu : ((1u << (1 << R)) - 1u);
    const uint32_t bit = g << R;
    if constexpr ((1 << (R + Q)) <= 64) {
        return (uint32_t)(op.smask[0] >> bit) & WMASK;
    } else {
        const uint64_t word = (bit & 64u) ? op.smask[1] : op.smask[0];
        return (uint32_t)(word >> (bit & 63u)) & WMASK;
    }
}

__device__ __forceinline__ double2 shfl_xor_d2(double2 v, int mask) {
    return make_double2(__shfl_xor(v.x, mask, kWave), __shfl_xor(v.y, mask, kWave));
}

// x with the signs of both components flipped when bit k of w is set (an
// exact sign change on the high words).
__device__ __forceinline__ double2 flip_by_bit(double2 x, uint32_t w, int k) {
    const int m = (int)((w << (31 - k)) & 0x80000000u);
    return make_double2(__hiloint2double(__double2hiint(x.x) ^ m, __double2loint(x.x)),
                        __hiloint2double(__double2hiint(x.y) ^ m, __double2loint(x.y)));
}

// (gx + i gy) q: one multiply and one FMA per component.
__device__ __forceinline__ double2 gmul(double gx, double gy, double2 q) {
    return make_double2(fma(gx, q.x, -gy * q.y), fma(gx, q.y, gy * q.x));
}

template <int R, int Q, bool NT>
__global__ __launch_bounds__(kOrbitTPB) void k_orbit_lanes(double2* __restrict__ s, uint64_t norbit,
                                                           const uint64_t* __restrict__ basis,  // R + Q, pivots ascending
                                                           const uint32_t* __restrict__ piv,
                                                           const OrbitOp2* __restrict__ ops, uint32_t nops) {
    constexpr int M = R + Q;
    constexpr int E = 1 << R;              // amplitudes per lane
    constexpr uint32_t OPW = kWave >> Q;  // orbits per wave
    static_assert(M <= 7, "the sign mask holds 128 elements");
    constexpr uint32_t WMASK = (E == 32) ? 0xFFFFFFFFu : ((1u << E) - 1u);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = ((uint64_t)blockIdx.x * kOrbitTPB + threadIdx.x) / kWave;
    const uint64_t o = wave * OPW + (lane & (OPW - 1));
    const uint32_t h = lane / OPW;
    // all lanes of a wave stay to the end (the shuffles need their partners);
    // lanes past the last orbit neither load nor store
    const bool live = o < norbit;
    uint64_t j = live ? o : 0;
    for (int k = 0; k < M; ++k) j = ins0(j, piv[k]);
    uint64_t hc = 0;
    for (int k = 0; k < Q; ++k)
        if ((h >> k) & 1u) hc ^= basis[R + k];
    uint64_t bl[R];
#pragma unroll
    for (int k = 0; k < R; ++k) bl[k] = basis[k];
    uint64_t off[E];
    double2 a[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        uint64_t x = j ^ hc;
#pragma unroll
        for (int k = 0; k < R; ++k)
            if ((i >> k) & 1) x ^= bl[k];
        off[i] = x;
        a[i] = live ? ld_s<NT>(s + off[i]) : make_double2(0.0, 0.0);
    }
    for (uint32_t oi = 0; oi < nops; ++oi) {
        const OrbitOp2 op = ops[oi];
        const uint32_t flip = (__popcll(j & op.z) & 1) ? WMASK : 0u;
        if (op.v == 0) {
            const uint32_t wo = group_signs<R, Q>(op, h) ^ flip;
            const double2 fp = make_double2(op.c - op.sn * op.ph.y, op.sn * op.ph.x);
            const double2 fm = make_double2(op.c + op.sn * op.ph.y, -op.sn * op.ph.x);
#pragma unroll
            for (int i = 0; i < E; ++i) a[i] = cmul(((wo >> i) & 1u) ? fm : fp, a[i]);
            continue;
        }
        // a_i' = cos a_i + (-1)^{s(partner)} g a_partner with g = sin * i * phase.  The partner of element
        // i = h * E + i' is (h ^ vq) * E + (i' ^ V), V = v mod E, vq = v / E; its sign bit is bit i' ^ V of
        // the E-bit word of group h ^ vq.  The sign goes onto the fresh product g a_partner.
        const uint32_t vq = op.v >> R;
        const uint32_t w = group_signs<R, Q>(op, h ^ vq) ^ flip;
        const int lm = (int)(vq * OPW);
        const double c = op.c, gx = -op.sn * op.ph.y, gy = op.sn * op.ph.x;
        dispatch_mask<R>(op.v & (uint32_t)(E - 1), [&](auto vc) {
            constexpr int V = (int)decltype(vc)::value;
            if (lm == 0) {
                // in-lane partners (V != 0 here): each pair updated together from its old values
#pragma unroll
                for (int i = 0; i < E; ++i) {
                    if (i < (i ^ V)) {
                        const double2 x = a[i], y = a[i ^ V];
                        const double2 ty = flip_by_bit(gmul(gx, gy, y), w, i ^ V);
                        const double2 tx = flip_by_bit(gmul(gx, gy, x), w, i);
                        a[i] = make_double2(fma(c, x.x, ty.x), fma(c, x.y, ty.y));
                        a[i ^ V] = make_double2(fma(c, y.x, tx.x), fma(c, y.y, tx.y));
                    }
                }
            } else {
                // partner of element i: element i ^ V of lane ^ lm
                double2 p[E];
#pragma unroll
                for (int i = 0; i < E; ++i) p[i] = shfl_xor_d2(a[i ^ V], lm);
#pragma unroll
                for (int i = 0; i < E; ++i) {
                    const double2 t = flip_by_bit(gmul(gx, gy, p[i]), w, i ^ V);
                    a[i] = make_double2(fma(c, a[i].x, t.x), fma(c, a[i].y, t.y));
                }
            }
        });
    }
    if (live) {
#pragma unroll
        for (int i = 0; i < E; ++i) st_s<NT>(s + off[i], a[i]);
    }
}

// Reduced GF(2) basis with pivot = highest bit; every pivot bit appears in
// exactly one basis vector.
struct Span {
    std::vector<uint64_t> b;
    uint64_t reduce(uint64_t x) const {
        for (uint64_t v : b)
            if ((x >> (63 - __builtin_clzll(v))) & 1ull) x ^= v;
        return x;
    }
    bool add(uint64_t x, int maxdim, int minpiv = 0) {
        const uint64_t r = reduce(x);
        if (!r) return true;
        if ((int)b.size() >= maxdim) return false;
        const int p = 63 - __builtin_clzll(r);
        if (p < minpiv) return false;
        for (uint64_t& v : b)
            if ((v >> p) & 1ull) v ^= r;
        b.push_back(r);
        return true;
    }
};

enum PassKind { P_SINGLE = 0, P_TILE = 1, P_ORBIT = 2 };

struct Pass {
    int kind = P_SINGLE;
    uint64_t bmask = 0;
    std::vector<int> ops;  // indices into the lowered op list
    Span span;             // P_ORBIT
};

// Greedy planner.  Ops may move ahead of an earlier op only when their qubit
// supports are disjoint (then the two commute exactly).  Each pass is grown
// two ways from the first pending op -- as an LDS tile (mixing qubits fit
// 2^k amplitudes) and as a register orbit (a run of exponentials whose flip
// masks span at most orbit_dims <= 7 dimensions) -- and the one covering more
// ops is kept.
static void grow_tile(const std::vector<HostOp>& ops, const std::vector<char>& done, size_t first, int K,
                      uint64_t lowmask, uint64_t all, int max_ops, Pass& p) {
    p = Pass{};
    p.kind = P_TILE;
    p.bmask = lowmask;
    uint64_t blocked = 0;
    for (size_t i = first; i < ops.size(); ++i) {
        if (done[i]) continue;
        const HostOp& h = ops[i];
        if (h.support & blocked) {
            blocked |= h.support;
            if ((blocked & all) == all) break;
            continue;
        }
        const uint64_t nb = p.bmask | h.mix;
        if (__builtin_popcountll(nb) <= K && (int)p.ops.size() < max_ops) {
            p.bmask = nb;
            p.ops.push_back((int)i);
        } else {
            blocked |= h.support;
            if ((blocked & all) == all || (int)p.ops.size() >= max_ops) break;
        }
    }
}

static void grow_orbit(const std::vector<HostOp>& ops, const std::vector<char>& done, size_t first, uint64_t all,
                       int max_ops, Pass& p, int kind = P_ORBIT, int maxdim = kOrbitMax, int minpiv = 0) {
    p = Pass{};
    p.kind = kind;
    uint64_t blocked = 0;
    for (size_t i = first; i < ops.size(); ++i) {
        if (done[i]) continue;
        const HostOp& h = ops[i];
        if (h.support & blocked) {
            blocked |= h.support;
            if ((blocked & all) == all) break;
            continue;
        }
        Span trial = p.span;
        if (h.op.kind == K_PEXP && (int)p.ops.size() < max_ops && trial.add(h.op.x, maxdim, minpiv)) {
            p.span = trial;
            p.ops.push_back((int)i);
        } else {
            blocked |= h.support;
            if ((blocked & all) == all || (int)p.ops.size() >= max_ops) break;
        }
    }
}

static void plan(const std::vector<HostOp>& ops, int n, const Tuning& tu, std::vector<Pass>& passes,
                 bool device_angles) {
    const TileShape shape = tile_shape(tu, n);
    const int K = std::min(shape.k, n);
    const int L = std::max(std::min(shape.low, K), K - 9);
    const uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
    const uint64_t lowmask = (L >= 64) ? ~0ull : ((1ull << L) - 1ull);
    std::vector<char> done(ops.size(), 0);
    size_t first = 0;
    while (first < ops.size()) {
        while (first < ops.size() && done[first]) ++first;
        if (first >= ops.size()) break;
        Pass tile, orbit;
        grow_tile(ops, done, first, K, lowmask, all, shape.max_ops, tile);
        grow_orbit(ops, done, first, all, shape.orbit_ops, orbit, P_ORBIT, std::max(1, std::min(tu.orbit_dims, kLaneOrbitMax)));
        Pass p;
        // an orbit pass at its op limit wins over a longer tile pass (tiles may
        // take more ops than orbits when the shape is chosen by state size)
        const size_t rival = std::min(tile.ops.size(), (size_t)shape.orbit_ops);
        if (orbit.ops.size() >= 2 && orbit.ops.size() >= rival && n > (int)orbit.span.b.size()) {
            p = std::move(orbit);
        } else if (tile.ops.size() >= 2) {
            p = std::move(tile);
        } else if (device_angles && ops[first].op.kind == K_PEXP && n > (int)orbit.span.b.size()) {
            p = std::move(orbit);  // a one-op orbit pass reads its angle from the descriptor image
        } else {
            // a standalone streaming pass
            p = Pass{};
            p.kind = P_SINGLE;
            p.ops.push_back((int)first);
        }
        for (int i : p.ops) done[i] = 1;
        passes.push_back(std::move(p));
    }
}

template <int M>
static void launch_orbit_k(double2* s, uint64_t norbit, const OrbitDesc& d, const OrbitOp* ops, uint32_t nops,
                           bool nt, hipStream_t st) {
    const unsigned blocks = (unsigned)((norbit + kOrbitTPB - 1) / kOrbitTPB);
    if (nt)
        hipLaunchKernelGGL((k_orbit<M, true>), dim3(blocks), dim3(kOrbitTPB), 0, st, s, norbit, d.b[0], d.b[1], d.b[2],
                           d.b[3], d.piv[0], d.piv[1], d.piv[2], d.piv[3], ops, nops);
    else
        hipLaunchKernelGGL((k_orbit<M, false>), dim3(blocks), dim3(kOrbitTPB), 0, st, s, norbit, d.b[0], d.b[1], d.b[2],
                           d.b[3], d.piv[0], d.piv[1], d.piv[2], d.piv[3], ops, nops);
}

static int launch_orbit(qfe_state_s* st, int M, const OrbitDesc& d, const OrbitOp* ops, uint32_t nops) {
    Context& cx = ctx();
    const uint64_t norbit = st->N >> M;
    const bool nt = use_nt(st->n, M == 0 ? 63u : d.piv[0]);  // pivots ascending: piv[0] is the lowest
    switch (M) {
        case 0: launch_orbit_k<0>(st->d, norbit, d, ops, nops, nt, cx.stream); break;  // only diagonal exponentials
        case 1: launch_orbit_k<1>(st->d, norbit, d, ops, nops, nt, cx.stream); break;
        case 2: launch_orbit_k<2>(st->d, norbit, d, ops, nops, nt, cx.stream); break;
        case 3: launch_orbit_k<3>(st->d, norbit, d, ops, nops, nt, cx.stream); break;
        case 4: launch_orbit_k<4>(st->d, norbit, d, ops, nops, nt, cx.stream); break;
        default: return fail(QFE_ERR_STATE, "internal: orbit dimension out of range");
    }
    QFE_HIP(hipGetLastError());
    cx.counters.fused++;
    cx.counters.orbit++;
    cx.counters.nt += nt ? 1 : 0;
    cx.counters.bytes += 32.0 * (double)st->N;
    return QFE_OK;
}

// Orbit descriptor and per-op orbit data for a finished orbit pass.
static void build_orbit(const Pass& p, const std::vector<HostOp>& ops, OrbitDesc& d, std::vector<OrbitOp>& out) {
    std::vector<uint64_t> b = p.span.b;
    std::sort(b.begin(), b.end(), [](uint64_t u, uint64_t v) { return __builtin_clzll(u) > __builtin_clzll(v); });
    const int M = (int)b.size();
    d = OrbitDesc{};
    for (int k = 0; k < M; ++k) {
        d.piv[k] = (uint32_t)(63 - __builtin_clzll(b[k]));
        d.b[k] = b[k];
    }
    for (int i = 0; i < (1 << M); ++i) {
        uint64_t c = 0;
        for (int k = 0; k < M; ++k)
            if ((i >> k) & 1) c ^= b[k];
        d.c[i] = c;
    }
    for (int oi : p.ops) {
        const DevOp& g = ops[oi].op;
        OrbitOp o{};
        uint32_t v = 0;
        for (int k = 0; k < M; ++k)
            if ((g.x >> d.piv[k]) & 1ull) v |= 1u << k;
        o.v = v;
        uint32_t sm = 0;
        for (int i = 0; i < (1 << M); ++i)
            if (__builtin_popcountll(d.c[i] & g.z) & 1) sm |= 1u << i;
        o.smask = sm;
        o.z = g.z;
        o.c = g.m[0].x;
        o.sn = g.m[0].y;
        o.ph = g.m[1];
        out.push_back(o);
    }
}

// Lane-orbit descriptor (basis sorted by pivot) and per-op data for an orbit
// pass of 5, 6 or 7 dimensions.
struct LaneDesc {
    uint64_t b[kLaneOrbitMax];
    uint32_t piv[kLaneOrbitMax];
};

static void build_orbit_lanes(const Pass& p, const std::vector<HostOp>& ops, LaneDesc& d, std::vector<OrbitOp2>& out) {
    std::vector<uint64_t> b = p.span.b;
    std::sort(b.begin(), b.end(), [](uint64_t u, uint64_t v) { return __builtin_clzll(u) > __builtin_clzll(v); });
    const int M = (int)b.size();
    d = LaneDesc{};
    for (int k = 0; k < M; ++k) {
        d.b[k] = b[k];
        d.piv[k] = (uint32_t)(63 - __builtin_clzll(b[k]));
    }
    std::vector<uint64_t> c(1u << M);
    for (int i = 0; i < (1 << M); ++i) {
        uint64_t x = 0;
        for (int k = 0; k < M; ++k)
            if ((i >> k) & 1) x ^= b[k];
        c[i] = x;
    }
    for (int oi : p.ops) {
        const DevOp& g = ops[oi].op;
        OrbitOp2 o{};
        for (int k = 0; k < M; ++k)
            if ((g.x >> d.piv[k]) & 1ull) o.v |= 1u << k;
        for (int i = 0; i < (1 << M); ++i)
            if (__builtin_popcountll(c[i] & g.z) & 1) o.smask[i >> 6] |= 1ull << (i & 63);
        o.z = g.z;
        o.c = g.m[0].x;
        o.sn = g.m[0].y;
        o.ph = g.m[1];
        out.push_back(o);
    }
}

// In-lane dimensions of a lane orbit: 2^kLaneBits amplitudes per lane, the
// other M - kLaneBits dimensions spread over lanes.
constexpr int kLaneBits = 4;

template <int M, bool NT>
static void launch_orbit_lanes_k(double2* s, uint64_t norbit, const uint64_t* b, const uint32_t* piv,
                                 const OrbitOp2* ops, uint32_t nops, hipStream_t st) {
    constexpr int Q = M - kLaneBits;
    const uint64_t threads = norbit << Q;
    const unsigned blocks = (unsigned)((threads + kOrbitTPB - 1) / kOrbitTPB);
    hipLaunchKernelGGL((k_orbit_lanes<kLaneBits, Q, NT>), dim3(blocks), dim3(kOrbitTPB), 0, st, s, norbit, b, piv, ops,
                       nops);
}

static int launch_orbit_lanes(qfe_state_s* st, int M, uint32_t low_pivot, const LaneDesc* dd, const OrbitOp2* ops,
                              uint32_t nops) {
    Context& cx = ctx();
    const uint64_t norbit = st->N >> M;
    const uint64_t* b = reinterpret_cast<const uint64_t*>(dd);
    const uint32_t* piv = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(dd) + offsetof(LaneDesc, piv));
    const bool nt = use_nt(st->n, low_pivot);
    if (M == 5 && nt)
        launch_orbit_lanes_k<5, true>(st->d, norbit, b, piv, ops, nops, cx.stream);
    else if (M == 5)
        launch_orbit_lanes_k<5, false>(st->d, norbit, b, piv, ops, nops, cx.stream);
    else if (M == 6 && nt)
        launch_orbit_lanes_k<6, true>(st->d, norbit, b, piv, ops, nops, cx.stream);
    else if (M == 6)
        launch_orbit_lanes_k<6, false>(st->d, norbit, b, piv, ops, nops, cx.stream);
    else if (M == 7 && nt)
        launch_orbit_lanes_k<7, true>(st->d, norbit, b, piv, ops, nops, cx.stream);
    else if (M == 7)
        launch_orbit_lanes_k<7, false>(st->d, norbit, b, piv, ops, nops, cx.stream);
    else
        return fail(QFE_ERR_STATE, "internal: lane orbit dimension out of range");
    QFE_HIP(hipGetLastError());
    cx.counters.fused++;
    cx.counters.lane_orbit++;
    cx.counters.nt += nt ? 1 : 0;
    cx.counters.bytes += 32.0 * (double)st->N;
    return QFE_OK;
}

template <int KQ>
static void launch_fused_k(double2* s, uint64_t nblocks, const PassDesc& pd, const DevOp* ops, uint32_t nt,
                           hipStream_t st) {
    const size_t lds = (size_t)sizeof(double2) << KQ;
    static bool attr_set = false;  // large tiles need the dynamic-LDS limit raised once
    if (!attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused<KQ>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr_set = true;
    }
    hipLaunchKernelGGL(k_fused<KQ>, dim3((unsigned)nblocks), dim3(kFusedTPB), lds, st, s, pd.bmask, pd.nlow, nt, ops,
                       pd.nops);
}

static int launch_fused(qfe_state_s* st, const PassDesc& pd, const DevOp* dops) {
    Context& cx = ctx();
    const int k = (int)pd.k;
    const uint64_t nblocks = st->N >> k;
    const uint32_t nt = use_nt(st->n, pd.nlow) ? 1u : 0u;  // tile loads and stores run over 2^nlow contiguous amplitudes
    switch (k) {
        case 2: launch_fused_k<2>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 3: launch_fused_k<3>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 4: launch_fused_k<4>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 5: launch_fused_k<5>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 6: launch_fused_k<6>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 7: launch_fused_k<7>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 8: launch_fused_k<8>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 9: launch_fused_k<9>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 10: launch_fused_k<10>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 11: launch_fused_k<11>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 12: launch_fused_k<12>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        case 13: launch_fused_k<13>(st->d, nblocks, pd, dops, nt, cx.stream); break;
        default: return fail(QFE_ERR_ARG, "internal: fused tile size out of range");
    }
    QFE_HIP(hipGetLastError());
    cx.counters.fused++;
    cx.counters.tile++;
    cx.counters.nt += nt;
    cx.counters.bytes += 32.0 * (double)st->N;
    return QFE_OK;
}

// Adjacent one-qubit ops on the same target under the same condition become
// one 2x2 (U = U_second * U_first): Ry then Rz on a qubit is one sweep op.
static void merge_single_qubit_runs(std::vector<HostOp>& ops) {
    auto as2x2 = [](const DevOp& o, double2 m[4]) {
        if (o.kind == K_U1) {
            for (int i = 0; i < 4; ++i) m[i] = o.m[i];
        } else {
            m[0] = o.m[0];
            m[1] = make_double2(0.0, 0.0);
            m[2] = make_double2(0.0, 0.0);
            m[3] = o.m[1];
        }
    };
    auto mul = [](double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
    auto add = [](double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); };
    std::vector<HostOp> out;
    out.reserve(ops.size());
    for (const HostOp& h : ops) {
        const bool one_q = h.op.kind == K_U1 || h.op.kind == K_DIAG1;
        if (one_q && !out.empty()) {
            HostOp& p = out.back();
            const bool prev_one_q = p.op.kind == K_U1 || p.op.kind == K_DIAG1;
            if (prev_one_q && p.op.t == h.op.t && p.op.cmask == h.op.cmask && p.op.cval == h.op.cval) {
                double2 a[4], b[4], r[4];
                as2x2(p.op, a);
                as2x2(h.op, b);
                r[0] = add(mul(b[0], a[0]), mul(b[1], a[2]));
                r[1] = add(mul(b[0], a[1]), mul(b[1], a[3]));
                r[2] = add(mul(b[2], a[0]), mul(b[3], a[2]));
                r[3] = add(mul(b[2], a[1]), mul(b[3], a[3]));
                const bool diag = r[1].x == 0.0 && r[1].y == 0.0 && r[2].x == 0.0 && r[2].y == 0.0;
                if (diag) {
                    p.op.kind = K_DIAG1;
                    p.op.m[0] = r[0];
                    p.op.m[1] = r[3];
                    p.mix = 0;
                } else {
                    p.op.kind = K_U1;
                    for (int i = 0; i < 4; ++i) p.op.m[i] = r[i];
                    p.op.hb = is_swap(p.op.m);
                    p.mix = 1ull << p.op.t;
                }
                continue;
            }
        }
        out.push_back(h);
    }
    ops.swap(out);
}

// A dense one-qubit gate with no condition (a plain group member).
static inline bool plain_u1(const DevOp& o) { return o.kind == K_U1 && !o.hb && o.cmask == 0; }
// m = diag(u0, u1) * R with R real, when each row of m is real up to a phase
// (to 1e-14 relative): r = R row-major, u = the row phases.
static bool split_row_phases(const double2* m, double r[4], double2 u[2]) {
    for (int row = 0; row < 2; ++row) {
        const double2 a = m[2 * row], b = m[2 * row + 1];
        const double na = std::hypot(a.x, a.y), nb = std::hypot(b.x, b.y);
        const double2 big = na >= nb ? a : b;
        const double nbig = std::max(na, nb);
        if (!(nbig > 0.0)) return false;
        const double2 ph = make_double2(big.x / nbig, big.y / nbig);
        const double ra = ph.x * a.x + ph.y * a.y, ia = ph.x * a.y - ph.y * a.x;  // conj(ph) * a
        const double rb = ph.x * b.x + ph.y * b.y, ib = ph.x * b.y - ph.y * b.x;
        if (std::fabs(ia) > 1e-14 * nbig || std::fabs(ib) > 1e-14 * nbig) return false;
        r[2 * row] = ra;
        r[2 * row + 1] = rb;
        u[row] = ph;
    }
    return true;
}

// A plain U1 whose matrix splits into row phases times a real rotation.
static inline bool rot_u1(const DevOp& o) {
    double r[4];
    double2 u[2];
    return plain_u1(o) && split_row_phases(o.m, r, u);
}
// An X with at most one condition bit, or a qubit swap: the tile's index map
// absorbs it (cX, anti-controlled X, SWAP), no amplitude moves.
static inline bool map_op(const HostOp& h) {
    return (h.op.kind == K_U1 && h.op.hb && __builtin_popcountll(h.op.cmask) <= 1) || (h.op.kind == K_U2 && h.op.hb);
}

// Order the ops of a tile pass for the kernel: any order in which every op
// follows each earlier op whose support it shares (ops with disjoint supports
// commute), chosen greedily so that index-map ops come as early as they can
// and dense unconditioned one-qubit gates on distinct qubits come in groups
// (one LDS round trip each): four when four ready gates each split into row
// phases times a real rotation, else up to three.  Returns the new order and
// the plain groups as (position, length).
static void schedule_tile_ops(const std::vector<HostOp>& ops, std::vector<int>& order,
                              std::vector<std::pair<size_t, uint32_t>>& groups) {
    std::vector<int> rem(order), out, ready, take, plain, other, rot;
    out.reserve(order.size());
    groups.clear();
    while (!rem.empty()) {
        // ready: no earlier remaining op shares a qubit with it
        ready.clear();
        uint64_t seen = 0;
        for (size_t i = 0; i < rem.size(); ++i) {
            const uint64_t sup = ops[rem[i]].support;
            if (!(sup & seen)) ready.push_back((int)i);
            seen |= sup;
        }
        take.clear();
        for (int i : ready)
            if (map_op(ops[rem[i]])) take.push_back(i);
        if (take.empty()) {
            plain.clear();
            other.clear();
            for (int i : ready) (plain_u1(ops[rem[i]].op) ? plain : other).push_back(i);
            rot.clear();
            for (int i : plain)
                if (rot_u1(ops[rem[i]].op)) rot.push_back(i);
            if (rot.size() >= 4) {
                // four rotations share one round trip of 16 elements per thread
                groups.push_back({out.size(), 4u});
                take.assign(rot.begin(), rot.begin() + 4);
            } else if (plain.size() >= 3 || (other.empty() && plain.size() >= 2)) {
                const size_t len = std::min<size_t>(3, plain.size());
                groups.push_back({out.size(), (uint32_t)len});
                take.assign(plain.begin(), plain.begin() + len);
            } else {
                take.push_back(other.empty() ? plain[0] : other[0]);
            }
        }
        for (int i : take) out.push_back(rem[i]);
        std::sort(take.begin(), take.end());
        for (size_t k = take.size(); k-- > 0;) rem.erase(rem.begin() + take[k]);
    }
    order.swap(out);
}

// Mark the groups of a tile pass in its localized op array: the plain groups
// the schedule formed (kGroupPlain; kGroupRot as well when every op splits
// into row phases times a real rotation, the only form a group of four
// takes), and runs of 2 or 3 other consecutive U1 ops with distinct tile
// targets (the head op's pad holds the run length); the fused kernel applies
// each group with one LDS round trip.  Index-map ops stay alone (absorbed
// without moving amplitudes).
static void group_u1_runs(std::vector<DevOp>& dops, size_t first, uint32_t count, int k,
                          const std::vector<std::pair<size_t, uint32_t>>& plain) {
    for (uint32_t j = 0; j < count; ++j) dops[first + j].pad = 0;
    if (k < 3) return;
    std::vector<char> in_plain(count, 0);
    for (const auto& g : plain) {
        dops[first + g.first].pad = g.second | kGroupPlain;
        for (uint32_t r = 0; r < g.second; ++r) in_plain[g.first + r] = 1;
    }
    auto groupable = [&](uint32_t j) {
        const DevOp& o = dops[first + j];
        return !in_plain[j] && o.kind == K_U1 && !(o.hb && __builtin_popcountll(o.cmask) <= 1);
    };
    uint32_t j = 0;
    while (j < count) {
        if (!groupable(j)) {
            ++j;
            continue;
        }
        uint32_t len = 1;
        uint32_t used = 1u << dops[first + j].t;
        while (len < 3 && j + len < count && groupable(j + len)) {
            const DevOp& o = dops[first + j + len];
            if ((used >> o.t) & 1u) break;
            used |= 1u << o.t;
            ++len;
        }
        bool allp = len >= 2;
        for (uint32_t r = 0; r < len; ++r) {
            const DevOp& o = dops[first + j + r];
            allp = allp && !o.hb && o.cmask == 0 && o.omask == 0;
        }
        dops[first + j].pad = len >= 2 ? (len | (allp ? kGroupPlain : 0u)) : 0u;
        j += len;
    }
    // plain groups whose ops all split into row phases times a real rotation
    for (uint32_t g = 0; g < count;) {
        DevOp& head = dops[first + g];
        const uint32_t len = head.pad & kGroupLen;
        if (!(head.pad & kGroupPlain) || len < 2) {
            ++g;
            continue;
        }
        double r[4][4];
        double2 u[4][2];
        bool ok = true;
        for (uint32_t k = 0; k < len && ok; ++k) ok = split_row_phases(dops[first + g + k].m, r[k], u[k]);
        if (!ok && len > 3) {
            // the dense group path takes at most three ops: the fourth runs alone
            head.pad = 3u | kGroupPlain;
            dops[first + g + 3].pad = 0;
            g += 3;
            continue;
        }
        if (ok) {
            for (uint32_t k = 0; k < len; ++k) {
                DevOp& o = dops[first + g + k];
                o.m[0] = make_double2(r[k][0], r[k][1]);
                o.m[1] = make_double2(r[k][2], r[k][3]);
            }
            for (int i = 0; i < (len > 3 ? 16 : 8); ++i) {
                double2 f = make_double2(1.0, 0.0);
                for (uint32_t k = 0; k < len; ++k) {
                    const double2 v = u[k][(i >> k) & 1];
                    f = make_double2(f.x * v.x - f.y * v.y, f.x * v.y + f.y * v.x);
                }
                if (i < 12)
                    head.m[4 + i] = f;
                else
                    dops[first + g + 1].m[i - 8] = f;
            }
            head.pad |= kGroupRot;
        }
        g += len;
    }
}

static inline size_t align_up(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

struct OpPlan {
    std::vector<HostOp> ops;
    std::vector<Pass> passes;
    std::vector<PassDesc> pdescs;
    std::vector<size_t> pass_pd, pass_op0, single_mat, span_d;
    std::vector<OrbitDesc> odescs;
    size_t off_ops = 0, off_mat = 0, off_orb = 0, off_ld = 0, off_lo = 0, total = 0;
};

int plan_build(std::vector<HostOp>& ops_in, int n, bool device_angles, OpPlan** out, std::vector<char>& image,
               std::vector<AngleRef>& angles) {
    *out = nullptr;
    Context& cx = ctx();
    OpPlan* P = new OpPlan();
    P->ops = std::move(ops_in);
    std::vector<HostOp>& ops = P->ops;
    merge_single_qubit_runs(ops);
    std::vector<Pass>& passes = P->passes;
    plan(ops, n, cx.tuning, passes, device_angles);

    // Descriptor image layout: [PassDesc per tile pass][DevOp arrays][U2 matrices for single passes]
    // [OrbitOp arrays][LaneDesc per lane-orbit pass][OrbitOp2 arrays]
    std::vector<DevOp> dops;
    std::vector<double2> mats;
    std::vector<OrbitOp> oops;
    std::vector<LaneDesc> ldescs;
    std::vector<OrbitOp2> lops;
    // angle refs by region (0: DevOp, 1: OrbitOp, 2: OrbitOp2) and index, resolved to bytes below
    struct Ref {
        int region;
        size_t index;
        int32_t src;
    };
    std::vector<Ref> refs;
    P->pass_pd.assign(passes.size(), (size_t)-1);
    P->pass_op0.assign(passes.size(), 0);
    P->single_mat.assign(passes.size(), (size_t)-1);
    P->span_d.assign(passes.size(), 0);
    P->odescs.resize(passes.size());
    auto fail_plan = [&](int code, const char* msg) {
        delete P;
        return fail(code, msg);
    };
    for (size_t pi = 0; pi < passes.size(); ++pi) {
        const Pass& p = passes[pi];
        if (p.kind == P_SINGLE) {
            const HostOp& h = ops[p.ops[0]];
            if (device_angles && h.op.kind == K_PEXP && h.src >= 0)
                return fail_plan(QFE_ERR_ARG, "plan: an exponential on every qubit of a 1-qubit register has no replayable pass");
            if (h.op.kind == K_U2) {
                P->single_mat[pi] = mats.size();
                for (int i = 0; i < 16; ++i) mats.push_back(h.op.m[i]);
            }
            continue;
        }
        if (p.kind == P_ORBIT && p.span.b.size() > (size_t)kOrbitMax) {
            P->span_d[pi] = ldescs.size();
            P->pass_op0[pi] = lops.size();
            ldescs.emplace_back();
            build_orbit_lanes(p, ops, ldescs.back(), lops);
            for (size_t j = 0; j < p.ops.size(); ++j)
                if (ops[p.ops[j]].src >= 0) refs.push_back({2, P->pass_op0[pi] + j, ops[p.ops[j]].src});
            continue;
        }
        if (p.kind == P_ORBIT) {
            P->pass_op0[pi] = oops.size();
            build_orbit(p, ops, P->odescs[pi], oops);
            for (size_t j = 0; j < p.ops.size(); ++j)
                if (ops[p.ops[j]].src >= 0) refs.push_back({1, P->pass_op0[pi] + j, ops[p.ops[j]].src});
            continue;
        }
        // extend the tile with the lowest free qubits up to K bits
        const int K = std::min(tile_shape(cx.tuning, n).k, n);
        uint64_t b = p.bmask;
        for (int q = 0; q < n && __builtin_popcountll(b) < K; ++q) b |= 1ull << q;
        PassDesc pd{};
        pd.bmask = b;
        pd.k = (uint32_t)__builtin_popcountll(b);
        uint32_t nlow = 0;
        while (nlow < 64 && ((b >> nlow) & 1ull)) ++nlow;
        // the planner keeps at least k - 9 low bits in every tile, so the
        // offset table of the high tile bits never exceeds 2^9 entries
        if (pd.k - nlow > 9) return fail_plan(QFE_ERR_STATE, "internal: fused tile has more than 9 high bits");
        pd.nlow = nlow;
        int local_of[64];
        for (int q = 0; q < 64; ++q) local_of[q] = -1;
        int li = 0, hi_i = 0, ri = 0;
        for (int q = 0; q < n; ++q) {
            if ((b >> q) & 1ull) {
                if ((uint32_t)li >= nlow) pd.hipos[hi_i++] = (uint8_t)q;
                local_of[q] = li++;
            } else {
                pd.restpos[ri++] = (uint8_t)q;
            }
        }
        pd.nrest = (uint32_t)ri;
        pd.nops = (uint32_t)p.ops.size();
        P->pass_op0[pi] = dops.size();
        std::vector<int> order(p.ops);
        std::vector<std::pair<size_t, uint32_t>> plain;
        if (pd.k >= 3) schedule_tile_ops(ops, order, plain);
        for (int oi : order) {
            if (ops[oi].src >= 0) refs.push_back({0, dops.size(), ops[oi].src});
            dops.push_back(localize(ops[oi].op, b, local_of));
        }
        group_u1_runs(dops, P->pass_op0[pi], pd.nops, (int)pd.k, plain);
        P->pass_pd[pi] = P->pdescs.size();
        P->pdescs.push_back(pd);
    }
    const size_t bytes_pd = P->pdescs.size() * sizeof(PassDesc);
    P->off_ops = align_up(bytes_pd, 256);
    const size_t bytes_ops = dops.size() * sizeof(DevOp);
    P->off_mat = align_up(P->off_ops + bytes_ops, 256);
    const size_t bytes_mat = mats.size() * sizeof(double2);
    P->off_orb = align_up(P->off_mat + bytes_mat, 256);
    const size_t bytes_orb = oops.size() * sizeof(OrbitOp);
    P->off_ld = align_up(P->off_orb + bytes_orb, 256);
    const size_t bytes_ld = ldescs.size() * sizeof(LaneDesc);
    P->off_lo = align_up(P->off_ld + bytes_ld, 256);
    const size_t bytes_lo = lops.size() * sizeof(OrbitOp2);
    P->total = (bytes_ops || bytes_mat || bytes_orb || bytes_ld) ? P->off_lo + bytes_lo : 0;
    image.assign(P->total, 0);
    char* host = image.data();
    if (bytes_pd) std::memcpy(host, P->pdescs.data(), bytes_pd);
    if (bytes_ops) std::memcpy(host + P->off_ops, dops.data(), bytes_ops);
    if (bytes_mat) std::memcpy(host + P->off_mat, mats.data(), bytes_mat);
    if (bytes_orb) std::memcpy(host + P->off_orb, oops.data(), bytes_orb);
    if (bytes_ld) std::memcpy(host + P->off_ld, ldescs.data(), bytes_ld);
    if (bytes_lo) std::memcpy(host + P->off_lo, lops.data(), bytes_lo);
    angles.clear();
    for (const Ref& r : refs) {
        uint64_t off;
        if (r.region == 0)
            off = P->off_ops + r.index * sizeof(DevOp) + offsetof(DevOp, m);
        else if (r.region == 1)
            off = P->off_orb + r.index * sizeof(OrbitOp) + offsetof(OrbitOp, c);
        else
            off = P->off_lo + r.index * sizeof(OrbitOp2) + offsetof(OrbitOp2, c);
        angles.push_back(AngleRef{off, r.src, 0});
    }
    *out = P;
    return QFE_OK;
}

int plan_launch(const OpPlan* P, qfe_state_s* st, char* dev) {
    for (size_t pi = 0; pi < P->passes.size(); ++pi) {
        const Pass& p = P->passes[pi];
        int rc;
        if (p.kind == P_SINGLE) {
            double2* m = P->single_mat[pi] == (size_t)-1 ? nullptr
                                                         : reinterpret_cast<double2*>(dev + P->off_mat) + P->single_mat[pi];
            rc = launch_single(st, P->ops[p.ops[0]], m);
        } else if (p.kind == P_ORBIT && p.span.b.size() > (size_t)kOrbitMax) {
            uint32_t low = 63;
            for (uint64_t v : p.span.b) low = std::min(low, (uint32_t)(63 - __builtin_clzll(v)));
            rc = launch_orbit_lanes(st, (int)p.span.b.size(), low,
                                    reinterpret_cast<const LaneDesc*>(dev + P->off_ld) + P->span_d[pi],
                                    reinterpret_cast<const OrbitOp2*>(dev + P->off_lo) + P->pass_op0[pi],
                                    (uint32_t)p.ops.size());
        } else if (p.kind == P_ORBIT) {
            const OrbitOp* oo = reinterpret_cast<const OrbitOp*>(dev + P->off_orb) + P->pass_op0[pi];
            rc = launch_orbit(st, (int)p.span.b.size(), P->odescs[pi], oo, (uint32_t)p.ops.size());
        } else {
            const DevOp* dop = reinterpret_cast<const DevOp*>(dev + P->off_ops) + P->pass_op0[pi];
            rc = launch_fused(st, P->pdescs[P->pass_pd[pi]], dop);
        }
        if (rc) return rc;
    }
    return QFE_OK;
}

void plan_free(OpPlan* p) { delete p; }

int run_ops(qfe_state_s* st, std::vector<HostOp>& ops) {
    if (ops.empty()) return QFE_OK;
    OpPlan* P = nullptr;
    std::vector<char> image;
    std::vector<AngleRef> angles;
    int rc = plan_build(ops, st->n, false, &P, image, angles);
    if (rc) return rc;
    char* dev = nullptr;
    if (P->total) {
        void* scr = nullptr;
        rc = scratch_reserve(P->total, &scr);
        char* host = nullptr;
        if (!rc) rc = staging_begin(P->total, reinterpret_cast<void**>(&host));
        if (!rc) {
            dev = static_cast<char*>(scr);
            std::memcpy(host, image.data(), P->total);
            rc = staging_commit(dev, P->total);
        }
    }
    if (!rc) rc = plan_launch(P, st, dev);
    plan_free(P);
    return rc;
}

}  // namespace qfe

using namespace qfe;

extern "C" {

int qfe_apply_gates(qfe_state s, const qfe_gate* gates, uint64_t count) {
    if (!s || !s->d || (count && !gates)) return fail(QFE_ERR_ARG, "qfe_apply_gates: invalid argument");
    std::vector<HostOp> ops;
    ops.reserve(count);
    for (uint64_t i = 0; i < count; ++i) {
        int rc = lower_gate(gates[i], s->n, ops);
        if (rc) return rc;
    }
    return run_ops(s, ops);
}

int qfe_apply_ops(qfe_state s, const qfe_op* ops_in, uint64_t count) {
    if (!s || !s->d || (count && !ops_in)) return fail(QFE_ERR_ARG, "qfe_apply_ops: invalid argument");
    std::vector<HostOp> ops;
    ops.reserve(count);
    for (uint64_t i = 0; i < count; ++i) {
        int rc;
        if (ops_in[i].kind == QFE_OP_GATE)
            rc = lower_gate(ops_in[i].gate, s->n, ops);
        else if (ops_in[i].kind == QFE_OP_PEXP)
            rc = lower_pexp(ops_in[i].pauli, ops_in[i].angle, s->n, ops);
        else
            rc = fail(QFE_ERR_ARG, "qfe_apply_ops: unknown op kind");
        if (rc) return rc;
    }
    return run_ops(s, ops);
}

}  // extern "C"
