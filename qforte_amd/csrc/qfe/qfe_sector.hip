// Determinant-sector engine: UCC energies on the (N_alpha, N_beta) sector of
// the reference determinant, and the continuous-batching serving engine.
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
// <H> = sum_groups sum_{j, j^x in sector} conj(psi_j) psi_{j^x}
//       sum_t c_t (-1)^{(j^x).z_t}   (computer.cc:850-911 restated on the sector)
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
#include <vector>

#include "qfe_common.cuh"
#include "qfe_internal.h"

namespace qfe {
namespace {

constexpr int kSecThreads = 512;

struct SecStep {  // one run of commuting exponentials sharing an X mask
    uint32_t e_first, e_count;
    int32_t param;
    uint32_t pad;
    double scale;
};

struct SecEntry {  // one class of determinant pairs of a run
    uint64_t x, z0;
    uint32_t la_off, la_len, lb_off, lb_len;
    double w, inv_lb;
    double2 ulo, uhi;  // A(p_lo) / w, A(p_hi) / w
    uint32_t step, pad;
};

struct SecGroup {  // Hamiltonian terms sharing an X mask, on sector pairs
    uint64_t x;
    uint32_t la_off, la_len, lb_off, lb_len;
    uint32_t t_first, t_count;
    double inv_lb;
};

struct SecTables {  // device view of a template (lives in device memory)
    uint64_t D, ref_rank;
    uint32_t Da, Db;
    const uint64_t* afull;
    const uint64_t* bfull;
    const int2* lists;
    const SecStep* steps;
    const SecEntry* entries;
    const SecGroup* groups;
    const uint64_t* tz;
    const double2* tph;    // phase of each term
    const double2* tcoef;  // phase * coeff (fixed Hamiltonian) or null
    const uint64_t* map_off;
    const uint32_t* map_idx;
    const double2* map_val;
    uint32_t nsteps, nentries, ngroups, nterms;
};

struct SlotDesc {
    const SecTables* t;
    double2* psi;
    double2* coef;  // per-request coefficients (null: use t->tcoef)
    double2* cs;    // (cos, sin) per entry
    const double* params;
    const double* inputs;
    double* out;
    uint32_t mode;  // 0: energy, 1: state only
    uint32_t pad;
};

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

// One workgroup per request: coefficients, angles, U|ref>, then <H>.
__global__ __launch_bounds__(kSecThreads) void k_sector(const SlotDesc* __restrict__ slots) {
    __shared__ double2 red[16];
    const SlotDesc sd = slots[blockIdx.x];
    const SecTables* __restrict__ T = sd.t;
    const uint32_t tid = threadIdx.x, B = blockDim.x;
    const uint64_t Db = T->Db;

    const double2* coef = sd.coef ? sd.coef : T->tcoef;
    if (sd.coef && sd.mode == 0) {
        for (uint32_t k = tid; k < T->nterms; k += B) {
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
    for (uint32_t e = tid; e < T->nentries; e += B) {
        const SecEntry& en = T->entries[e];
        const SecStep& st = T->steps[en.step];
        const double th = st.scale * sd.params[st.param] * en.w;
        double s, c;
        sincos(th, &s, &c);
        sd.cs[e] = make_double2(c, s);
    }
    double2* __restrict__ psi = sd.psi;
    for (uint64_t r = tid; r < T->D; r += B) psi[r] = make_double2(r == T->ref_rank ? 1.0 : 0.0, 0.0);
    __syncthreads();

    for (uint32_t s = 0; s < T->nsteps; ++s) {
        const uint32_t e0 = T->steps[s].e_first, e1 = e0 + T->steps[s].e_count;
        for (uint32_t e = e0; e < e1; ++e) {
            const SecEntry& en = T->entries[e];
            const double2 cs = sd.cs[e];
            const uint32_t lb = en.lb_len, total = en.la_len * lb;
            const uint64_t x = en.x, z0 = en.z0;
            const double2 ulo = en.ulo, uhi = en.uhi;
            const double inv = en.inv_lb;
            for (uint32_t i = tid; i < total; i += B) {
                uint32_t ia, ib;
                divmod_u(i, lb, inv, ia, ib);
                const int2 pa = T->lists[en.la_off + ia];
                const int2 pb = T->lists[en.lb_off + ib];
                const uint64_t r1 = (uint64_t)pa.x * Db + (uint64_t)pb.x;
                const uint64_t r2 = (uint64_t)pa.y * Db + (uint64_t)pb.y;
                const uint64_t j1 = T->afull[pa.x] | T->bfull[pb.x];
                const uint64_t j2 = j1 ^ x;
                const double s1 = cs.y * parity_sign(j1 & z0), s2 = cs.y * parity_sign(j2 & z0);
                const double2 a = psi[r1], b = psi[r2];
                // psi_1' = c a + s s(j2) uhi b ; psi_2' = c b + s s(j1) ulo a
                const double2 hb = cmul(uhi, b), la = cmul(ulo, a);
                psi[r1] = make_double2(cs.x * a.x + s2 * hb.x, cs.x * a.y + s2 * hb.y);
                psi[r2] = make_double2(cs.x * b.x + s1 * la.x, cs.x * b.y + s1 * la.y);
            }
        }
        __syncthreads();
    }
    if (sd.mode != 0) return;

    double acc = 0.0;
    for (uint32_t g = 0; g < T->ngroups; ++g) {
        const SecGroup& gr = T->groups[g];
        const uint32_t lb = gr.lb_len, total = gr.la_len * lb;
        const uint64_t x = gr.x;
        const uint32_t t0 = gr.t_first, t1 = gr.t_first + gr.t_count;
        const double inv = gr.inv_lb;
        for (uint32_t i = tid; i < total; i += B) {
            uint32_t ia, ib;
            divmod_u(i, lb, inv, ia, ib);
            const int2 pa = T->lists[gr.la_off + ia];
            const int2 pb = T->lists[gr.lb_off + ib];
            const uint64_t r1 = (uint64_t)pa.x * Db + (uint64_t)pb.x;
            const uint64_t r2 = (uint64_t)pa.y * Db + (uint64_t)pb.y;
            const uint64_t j2 = (T->afull[pa.x] | T->bfull[pb.x]) ^ x;
            double tr = 0.0, ti = 0.0;
            for (uint32_t t = t0; t < t1; ++t) {
                const double sg = parity_sign(j2 & T->tz[t]);
                tr += sg * coef[t].x;
                ti += sg * coef[t].y;
            }
            const double2 v = cconjmul(psi[r1], psi[r2]);
            acc += v.x * tr - v.y * ti;
        }
    }
    const double2 tot = block_sum(make_double2(acc, 0.0), red);
    if (tid == 0) *sd.out = tot.x;
}

// Scatter a sector state into a zeroed full state.
__global__ void k_sector_scatter(const double2* __restrict__ psi, uint64_t D, uint32_t Db, const uint64_t* __restrict__ afull,
                                 const uint64_t* __restrict__ bfull, double2* __restrict__ out) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= D) return;
    out[afull[r / Db] | bfull[r % Db]] = psi[r];
}

// ---- host-side construction -------------------------------------------------------

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

struct ListPool {
    std::vector<int2> pool;
    std::map<std::tuple<int, uint32_t, uint32_t, uint32_t>, std::pair<uint32_t, uint32_t>> index;
    // strings a of side s with (a & mask) == pattern (pattern == ~0u: popcount(a & mask) == popcount(mask) / 2),
    // each paired with the rank of a ^ flip.
    std::pair<uint32_t, uint32_t> get(const Space& sp, int s, uint32_t mask, uint32_t pattern, uint32_t flip) {
        auto key = std::make_tuple(s, mask, pattern, flip);
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        const uint32_t off = (uint32_t)pool.size();
        for (size_t r = 0; r < sp.str[s].size(); ++r) {
            const uint32_t a = sp.str[s][r];
            const bool ok = pattern == ~0u ? (2 * popc32(a & mask) == popc32(mask)) : ((a & mask) == pattern);
            if (!ok) continue;
            const int32_t partner = sp.rank[s][a ^ flip];
            if (partner < 0) continue;  // cannot happen for conserving entries; keep lists exact
            pool.push_back(make_int2((int)r, partner));
        }
        auto v = std::make_pair(off, (uint32_t)pool.size() - off);
        index.emplace(key, v);
        return v;
    }
};

inline double2 cplx_mul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
inline double cplx_abs(double2 a) { return std::hypot(a.x, a.y); }

}  // namespace
}  // namespace qfe

using namespace qfe;

struct qfe_template_s {
    int n = 0, norb = 0;
    uint64_t D = 0;
    uint32_t Da = 0, Db = 0;
    uint64_t nparam = 0, ninputs = 0;
    uint32_t nterms = 0, nentries = 0;
    bool has_map = false;
    uint64_t cost = 0;           // work estimate for ordering slots in a launch
    void* dev = nullptr;         // all tables
    SecTables* dt = nullptr;     // device struct inside dev
    const uint64_t* afull = nullptr;
    const uint64_t* bfull = nullptr;
    // per-request device bytes: work (psi, coef, cs) + io (params, inputs)
    size_t off_psi = 0, off_coef = 0, off_cs = 0, work_bytes = 0, io_count = 0, request_bytes = 0;
    // single-request arena
    void* arena = nullptr;
};

namespace qfe {
namespace {

inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

int template_build(qfe_template t, int n, uint64_t ref, const qfe_pauli* gen_terms, const uint64_t* gen_off, uint64_t ngen,
                   const int32_t* gen_param, const double* gen_scale, uint64_t nparam, const qfe_pauli* ham, uint64_t nham,
                   const uint64_t* map_off, const uint32_t* map_idx, const qfe_c64* map_val, uint64_t ninputs) {
    if (n < 2 || n > 32 || (n & 1)) return fail(QFE_ERR_ARG, "qfe_template_create: nqubit must be even and in [2, 32]");
    const int norb = n / 2;
    const uint64_t lim = (1ull << n) - 1ull;
    if (ref & ~lim) return fail(QFE_ERR_ARG, "qfe_template_create: reference index out of range");
    const uint32_t refa = compress(ref, 0, norb), refb = compress(ref, 1, norb);
    Space sp;
    sp.build(norb, popc32(refa), popc32(refb));
    ListPool lp;
    t->n = n;
    t->norb = norb;
    t->Da = (uint32_t)sp.str[0].size();
    t->Db = (uint32_t)sp.str[1].size();
    t->D = (uint64_t)t->Da * t->Db;
    t->nparam = nparam;
    t->ninputs = ninputs;
    const uint64_t ref_rank = (uint64_t)sp.rank[0][refa] * t->Db + (uint64_t)sp.rank[1][refb];

    // generators -> runs -> entries
    std::vector<SecStep> steps;
    std::vector<SecEntry> entries;
    for (uint64_t g = 0; g < ngen; ++g) {
        if (gen_off[g + 1] < gen_off[g]) return fail(QFE_ERR_ARG, "qfe_template_create: generator offsets must not decrease");
        if (gen_param[g] < 0 || (uint64_t)gen_param[g] >= nparam)
            return fail(QFE_ERR_ARG, "qfe_template_create: generator parameter index out of range");
        uint64_t b = gen_off[g];
        const uint64_t e = gen_off[g + 1];
        while (b < e) {
            const uint64_t x = gen_terms[b].x, z0 = gen_terms[b].z;
            uint64_t r = b;
            while (r < e && gen_terms[r].x == x && (popc64(x & (gen_terms[r].z ^ z0)) & 1) == 0) ++r;
            // run [b, r)
            uint64_t m = x;
            for (uint64_t k = b; k < r; ++k) {
                if ((gen_terms[k].x | gen_terms[k].z) & ~lim)
                    return fail(QFE_ERR_ARG, "qfe_template_create: generator term acts on a qubit out of range");
                m |= gen_terms[k].z ^ z0;
            }
            if (popc64(m) > 20) return fail(QFE_ERR_ARG, "qfe_template_create: generator run spans more than 20 qubits");
            double scale_abs = 0.0;
            std::vector<double2> kap(r - b);
            for (uint64_t k = b; k < r; ++k) {
                // exp(theta * coeff * P) with coeff purely imaginary, as the full-space program applies it
                kap[k - b] = cplx_mul(make_double2(0.0, gen_terms[k].coeff.im),
                                      make_double2(gen_terms[k].phase.re, gen_terms[k].phase.im));
                scale_abs += cplx_abs(kap[k - b]);
            }
            const double tol = 1e-13 * (scale_abs > 0 ? scale_abs : 1.0);
            // A(p) over the submasks of m
            std::map<uint64_t, double2> A;
            for (uint64_t p = m;; p = (p - 1) & m) {
                double2 s = make_double2(0.0, 0.0);
                for (uint64_t k = b; k < r; ++k) {
                    const double sg = (popc64(p & (gen_terms[k].z ^ z0)) & 1) ? -1.0 : 1.0;
                    s.x += sg * kap[k - b].x;
                    s.y += sg * kap[k - b].y;
                }
                if (cplx_abs(s) > tol) A.emplace(p, s);
                if (p == 0) break;
            }
            if (x == 0 && !A.empty())
                return fail(QFE_ERR_ARG, "qfe_template_create: diagonal generator runs are not supported on the sector");
            SecStep st{};
            st.e_first = (uint32_t)entries.size();
            st.param = gen_param[g];
            st.scale = gen_scale[g];
            const double sx = (popc64(x & z0) & 1) ? -1.0 : 1.0;
            for (const auto& kv : A) {
                const uint64_t plo = kv.first, phi = plo ^ x;
                auto it = A.find(phi);
                if (it == A.end())
                    return fail(QFE_ERR_ARG, "qfe_template_create: a generator run is not anti-Hermitian");
                if (plo > phi) continue;
                const double2 alo = kv.second, ahi = it->second;
                const double w = cplx_abs(alo);
                const double2 prod = cplx_mul(alo, ahi);
                if (std::fabs(sx * prod.x + w * w) > 1e-12 * w * w || std::fabs(prod.y) > 1e-12 * w * w)
                    return fail(QFE_ERR_ARG, "qfe_template_create: a generator run is not anti-Hermitian");
                const uint32_t pa = compress(plo, 0, norb), pb = compress(plo, 1, norb);
                const uint32_t qa = compress(phi, 0, norb), qb = compress(phi, 1, norb);
                if (popc32(pa) != popc32(qa) || popc32(pb) != popc32(qb))
                    return fail(QFE_ERR_ARG,
                                "qfe_template_create: a generator run does not conserve the alpha and beta particle numbers");
                const uint32_t ma = compress(m, 0, norb), mb = compress(m, 1, norb);
                const uint32_t xa = compress(x, 0, norb), xb = compress(x, 1, norb);
                auto la = lp.get(sp, 0, ma, pa, xa);
                auto lb = lp.get(sp, 1, mb, pb, xb);
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
            auto la = lp.get(sp, 0, xa, ~0u, xa);
            auto lb = lp.get(sp, 1, xb, ~0u, xb);
            if (!la.second || !lb.second) continue;
            SecGroup gr{};
            gr.x = x;
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
    t->has_map = map_off != nullptr;
    t->nterms = (uint32_t)tz.size();
    t->nentries = (uint32_t)entries.size();

    // cost estimate: pair updates + Hamiltonian pair-term evaluations
    uint64_t cost = t->D;
    for (const auto& en : entries) cost += (uint64_t)en.la_len * en.lb_len * 4;
    for (const auto& gr : groups) cost += (uint64_t)gr.la_len * gr.lb_len * (2 + gr.t_count);
    t->cost = cost;

    // device tables: one allocation
    std::vector<uint64_t> af(t->Da), bf(t->Db);
    for (uint32_t a = 0; a < t->Da; ++a) af[a] = spread(sp.str[0][a], 0, norb);
    for (uint32_t b2 = 0; b2 < t->Db; ++b2) bf[b2] = spread(sp.str[1][b2], 1, norb);
    struct Piece {
        const void* src;
        size_t bytes;
        size_t off;
    };
    std::vector<Piece> pieces;
    size_t total = align256(sizeof(SecTables));
    auto add = [&](const void* src, size_t bytes) {
        pieces.push_back({src, bytes, total});
        total += align256(bytes ? bytes : 1);
        return pieces.size() - 1;
    };
    const size_t i_af = add(af.data(), af.size() * 8), i_bf = add(bf.data(), bf.size() * 8);
    const size_t i_l = add(lp.pool.data(), lp.pool.size() * sizeof(int2));
    const size_t i_st = add(steps.data(), steps.size() * sizeof(SecStep));
    const size_t i_en = add(entries.data(), entries.size() * sizeof(SecEntry));
    const size_t i_gr = add(groups.data(), groups.size() * sizeof(SecGroup));
    const size_t i_tz = add(tz.data(), tz.size() * 8);
    const size_t i_tph = add(tph.data(), tph.size() * sizeof(double2));
    const size_t i_tc = add(tcoef.data(), tcoef.size() * sizeof(double2));
    const size_t i_mo = add(moff.data(), moff.size() * 8);
    const size_t i_mi = add(midx.data(), midx.size() * 4);
    const size_t i_mv = add(mval.data(), mval.size() * sizeof(double2));
    QFE_HIP(hipMalloc(&t->dev, total));
    char* base = (char*)t->dev;
    SecTables h{};
    h.D = t->D;
    h.ref_rank = ref_rank;
    h.Da = t->Da;
    h.Db = t->Db;
    h.afull = (const uint64_t*)(base + pieces[i_af].off);
    h.bfull = (const uint64_t*)(base + pieces[i_bf].off);
    h.lists = (const int2*)(base + pieces[i_l].off);
    h.steps = (const SecStep*)(base + pieces[i_st].off);
    h.entries = (const SecEntry*)(base + pieces[i_en].off);
    h.groups = (const SecGroup*)(base + pieces[i_gr].off);
    h.tz = (const uint64_t*)(base + pieces[i_tz].off);
    h.tph = (const double2*)(base + pieces[i_tph].off);
    h.tcoef = (const double2*)(base + pieces[i_tc].off);
    h.map_off = t->has_map ? (const uint64_t*)(base + pieces[i_mo].off) : nullptr;
    h.map_idx = (const uint32_t*)(base + pieces[i_mi].off);
    h.map_val = (const double2*)(base + pieces[i_mv].off);
    h.nsteps = (uint32_t)steps.size();
    h.nentries = (uint32_t)entries.size();
    h.ngroups = (uint32_t)groups.size();
    h.nterms = (uint32_t)tz.size();
    std::vector<char> img(total, 0);
    std::memcpy(img.data(), &h, sizeof(h));
    for (const auto& p : pieces)
        if (p.bytes) std::memcpy(img.data() + p.off, p.src, p.bytes);
    QFE_HIP(hipMemcpy(t->dev, img.data(), total, hipMemcpyHostToDevice));
    t->dt = (SecTables*)t->dev;
    t->afull = h.afull;
    t->bfull = h.bfull;

    // per-request device layout: work buffers (psi, coef, cs) and io (params, inputs)
    size_t o = 0;
    t->off_psi = o;
    o += align256(t->D * sizeof(double2));
    t->off_coef = o;
    if (t->has_map) o += align256((size_t)t->nterms * sizeof(double2));
    t->off_cs = o;
    o += align256((size_t)t->nentries * sizeof(double2) + 16);
    t->work_bytes = o;
    t->io_count = nparam + ninputs;
    t->request_bytes = t->work_bytes + align256(t->io_count * 8 + 8) + sizeof(SlotDesc) + 8;
    return QFE_OK;
}

SlotDesc make_slot(qfe_template t, char* work, const double* io, double* out, uint32_t mode) {
    SlotDesc d{};
    d.t = t->dt;
    d.psi = (double2*)(work + t->off_psi);
    d.coef = t->has_map ? (double2*)(work + t->off_coef) : nullptr;
    d.cs = (double2*)(work + t->off_cs);
    d.params = io;
    d.inputs = io + t->nparam;
    d.out = out;
    d.mode = mode;
    return d;
}

// One request on the library stream: arena = [desc][out][io][work]; the first
// three parts are uploaded through the pinned staging buffer.
int run_single(qfe_template t, const double* params, const double* inputs, uint32_t mode, double* energy) {
    Context& c = ctx();
    const size_t io_bytes = align256(t->io_count * 8 + 8);
    const size_t head = 512 + io_bytes;
    if (!t->arena) QFE_HIP(hipMalloc(&t->arena, head + t->work_bytes));
    char* dev = (char*)t->arena;
    double* dout = (double*)(dev + 256);
    const double* dio = (const double*)(dev + 512);
    void* host = nullptr;
    int rc = staging_begin(head, &host);
    if (rc) return rc;
    char* h = (char*)host;
    SlotDesc d = make_slot(t, dev + head, dio, dout, mode);
    std::memcpy(h, &d, sizeof(d));
    if (t->nparam) std::memcpy(h + 512, params, t->nparam * 8);
    if (t->ninputs) std::memcpy(h + 512 + t->nparam * 8, inputs, t->ninputs * 8);
    rc = staging_commit(dev, head);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sector, dim3(1), dim3(kSecThreads), 0, c.stream, (const SlotDesc*)dev);
    QFE_HIP(hipGetLastError());
    if (mode == 0) {
        void* pin = nullptr;
        rc = pinned_reserve(8, &pin);
        if (rc) return rc;
        QFE_HIP(hipMemcpyAsync(pin, dout, 8, hipMemcpyDeviceToHost, c.stream));
        QFE_HIP(hipStreamSynchronize(c.stream));
        *energy = *(double*)pin;
    }
    return QFE_OK;
}

}  // namespace
}  // namespace qfe

// ---- serving engine -------------------------------------------------------------------

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
    bool busy = false;
};

}  // namespace

struct qfe_batch_s {
    uint64_t budget = 0;
    int max_active = 0;
    uint64_t next_id = 1;
    uint64_t active = 0;
    std::deque<Request> queue;
    std::vector<Lane> lanes;
    std::deque<int> inflight;
    std::deque<std::pair<uint64_t, double>> done;
};

namespace {

constexpr int kLanes = 3;

int lane_launch(qfe_batch b, Lane& L) {
    // admit from the front of the queue while the footprint fits this lane's share
    const uint64_t share = b->budget / kLanes;
    std::vector<Request> take;
    uint64_t foot = 0;
    while (!b->queue.empty() && (int64_t)(b->active + take.size()) < b->max_active) {
        const Request& r = b->queue.front();
        const uint64_t f = r.t->request_bytes;
        if (!take.empty() && foot + f > share) break;
        foot += f;
        take.push_back(std::move(b->queue.front()));
        b->queue.pop_front();
    }
    if (take.empty()) return QFE_OK;
    // heaviest requests first, so their workgroups start first
    std::stable_sort(take.begin(), take.end(), [](const Request& x, const Request& y) { return x.t->cost > y.t->cost; });
    const size_t k = take.size();
    size_t io_total = 0, work_total = 0;
    for (const auto& r : take) {
        io_total += r.t->io_count;
        work_total += r.t->work_bytes;
    }
    const size_t head = align256(k * sizeof(SlotDesc)) + align256(k * 8) + align256(io_total * 8 + 8);
    const size_t need = head + work_total;
    if (L.dev_cap < need) {
        if (L.dev) QFE_HIP(hipFree(L.dev));
        L.dev = nullptr;
        L.dev_cap = 0;
        QFE_HIP(hipMalloc(&L.dev, need + need / 4));
        L.dev_cap = need + need / 4;
    }
    if (L.up_cap < head) {
        if (L.up) QFE_HIP(hipHostFree(L.up));
        L.up = nullptr;
        L.up_cap = 0;
        QFE_HIP(hipHostMalloc(&L.up, head + head / 4, hipHostMallocDefault));
        L.up_cap = head + head / 4;
    }
    if (L.out_cap < k) {
        if (L.out) QFE_HIP(hipHostFree(L.out));
        L.out = nullptr;
        L.out_cap = 0;
        QFE_HIP(hipHostMalloc((void**)&L.out, (k + 64) * 8, hipHostMallocDefault));
        L.out_cap = k + 64;
    }
    char* dev = (char*)L.dev;
    char* up = (char*)L.up;
    const size_t o_out = align256(k * sizeof(SlotDesc)), o_io = o_out + align256(k * 8);
    double* dout = (double*)(dev + o_out);
    size_t io_pos = 0, work_pos = head;
    L.ids.clear();
    for (size_t i = 0; i < k; ++i) {
        qfe_template t = take[i].t;
        const double* dio = (const double*)(dev + o_io) + io_pos;
        double* hio = (double*)(up + o_io) + io_pos;
        SlotDesc d = make_slot(t, dev + work_pos, dio, dout + i, 0);
        std::memcpy(up + i * sizeof(SlotDesc), &d, sizeof(d));
        if (t->nparam) std::memcpy(hio, take[i].params.data(), t->nparam * 8);
        if (t->ninputs) std::memcpy(hio + t->nparam, take[i].inputs.data(), t->ninputs * 8);
        io_pos += t->io_count;
        work_pos += t->work_bytes;
        L.ids.push_back(take[i].id);
    }
    QFE_HIP(hipMemcpyAsync(dev, up, head, hipMemcpyHostToDevice, L.stream));
    hipLaunchKernelGGL(k_sector, dim3((unsigned)k), dim3(kSecThreads), 0, L.stream, (const SlotDesc*)dev);
    QFE_HIP(hipGetLastError());
    QFE_HIP(hipMemcpyAsync(L.out, dout, k * 8, hipMemcpyDeviceToHost, L.stream));
    QFE_HIP(hipEventRecord(L.done, L.stream));
    L.busy = true;
    b->active += k;
    return QFE_OK;
}

int collect(qfe_batch b, bool wait_oldest) {
    bool waited = !wait_oldest;
    while (!b->inflight.empty()) {
        Lane& L = b->lanes[b->inflight.front()];
        hipError_t q;
        if (!waited) {
            q = hipEventSynchronize(L.done);
            waited = true;
        } else {
            q = hipEventQuery(L.done);
        }
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) return hip_fail(q, "qfe_batch: launch completion");
        for (size_t i = 0; i < L.ids.size(); ++i) b->done.emplace_back(L.ids[i], L.out[i]);
        b->active -= L.ids.size();
        L.ids.clear();
        L.busy = false;
        b->inflight.pop_front();
    }
    return QFE_OK;
}

int fill_lanes(qfe_batch b) {
    for (int i = 0; i < (int)b->lanes.size() && !b->queue.empty(); ++i) {
        Lane& L = b->lanes[i];
        if (L.busy) continue;
        int rc = lane_launch(b, L);
        if (rc) return rc;
        if (L.busy) b->inflight.push_back(i);
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
    delete t;
    return QFE_OK;
}

int qfe_template_info(qfe_template t, uint64_t* sector_dim, uint64_t* request_bytes) {
    if (!t) return fail(QFE_ERR_ARG, "qfe_template_info: invalid handle");
    if (sector_dim) *sector_dim = t->D;
    if (request_bytes) *request_bytes = t->request_bytes;
    return QFE_OK;
}

int qfe_template_energy(qfe_template t, const double* params, const double* inputs, double* energy) {
    if (!t || !energy || (t->nparam && !params) || (t->ninputs && !inputs))
        return fail(QFE_ERR_ARG, "qfe_template_energy: null argument");
    return run_single(t, params, inputs, 0, energy);
}

int qfe_template_prepare(qfe_template t, const double* params, qfe_state out) {
    if (!t || !out || (t->nparam && !params)) return fail(QFE_ERR_ARG, "qfe_template_prepare: null argument");
    if (out->n != t->n) return fail(QFE_ERR_ARG, "qfe_template_prepare: state has the wrong number of qubits");
    std::vector<double> zeros(t->ninputs + 1, 0.0);
    int rc = run_single(t, params, zeros.data(), 1, nullptr);
    if (rc) return rc;
    rc = launch_zero(out->d, out->N, ctx().stream);
    if (rc) return rc;
    const char* work = (const char*)t->arena + 512 + align256(t->io_count * 8 + 8);
    const unsigned grid = (unsigned)((t->D + 255) / 256);
    hipLaunchKernelGGL(k_sector_scatter, dim3(grid), dim3(256), 0, ctx().stream, (const double2*)(work + t->off_psi), t->D,
                       t->Db, t->afull, t->bfull, out->d);
    QFE_HIP(hipGetLastError());
    return QFE_OK;
}

int qfe_batch_create(uint64_t hbm_budget_bytes, int max_active, qfe_batch* out) {
    if (!out || max_active < 1 || hbm_budget_bytes == 0) return fail(QFE_ERR_ARG, "qfe_batch_create: invalid argument");
    int rc = ensure_init();
    if (rc) return rc;
    qfe_batch b = new qfe_batch_s();
    b->budget = hbm_budget_bytes;
    b->max_active = max_active;
    b->lanes.resize(kLanes);
    for (auto& L : b->lanes) {
        hipError_t e = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&L.done, hipEventDisableTiming);
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
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        if (L.dev) (void)hipFree(L.dev);
        if (L.up) (void)hipHostFree(L.up);
        if (L.out) (void)hipHostFree(L.out);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    delete b;
    return QFE_OK;
}

int qfe_batch_submit(qfe_batch b, qfe_template t, const double* params, const double* inputs, uint64_t* request_id) {
    if (!b || !t || !request_id) return fail(QFE_ERR_ARG, "qfe_batch_submit: invalid argument");
    if ((t->nparam && !params) || (t->ninputs && !inputs)) return fail(QFE_ERR_ARG, "qfe_batch_submit: null parameters");
    if (t->request_bytes > b->budget / kLanes)
        return fail(QFE_ERR_ARG, "qfe_batch_submit: one request exceeds the per-launch share of the HBM budget");
    Request r;
    r.id = b->next_id++;
    r.t = t;
    r.params.assign(params, params + t->nparam);
    if (t->ninputs) r.inputs.assign(inputs, inputs + t->ninputs);
    b->queue.push_back(std::move(r));
    *request_id = b->queue.back().id;
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
    if (!rc && b->done.empty() && !b->inflight.empty()) {
        rc = collect(b, true);
        if (!rc) rc = fill_lanes(b);
    }
    if (rc) return rc;
    return hand_out(b, done_ids, done_energies, cap, ndone);
}

int qfe_batch_pending(qfe_batch b, uint64_t* queued, uint64_t* active) {
    if (!b) return fail(QFE_ERR_ARG, "qfe_batch_pending: invalid handle");
    if (queued) *queued = b->queue.size();
    if (active) *active = b->active + b->done.size();
    return QFE_OK;
}

}  // extern "C"
