// Second quantisation: SQOperator, the Jordan-Wigner map and the operator
// pools (reference sq_operator.cc, sq_op_pool.cc, qubit_op_pool.cc).
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <set>

#include "core.h"

namespace qfa {

// ---- SQOperator ---------------------------------------------------------------------

void SQOperator::set_coeffs(const std::vector<cplx>& c) {
    if (c.size() != terms_.size()) throw std::invalid_argument("number of new coefficients for quantum operator must equal ");
    for (size_t i = 0; i < c.size(); ++i) std::get<0>(terms_[i]) = c[i];
}

void SQOperator::mult_coeffs(cplx m) {
    for (auto& t : terms_) std::get<0>(t) *= m;
}

namespace {

// Sort indices in descending order; return the sign of the permutation
// (sq_operator.cc:40-47, 88-104).  Equal indices keep their order.
int sort_descending(std::vector<size_t>& v) {
    std::vector<size_t> perm(v.size());
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return v[a] > v[b]; });
    std::vector<size_t> sorted(v.size());
    for (size_t i = 0; i < v.size(); ++i) sorted[i] = v[perm[i]];
    v.swap(sorted);
    // parity from the cycle decomposition of perm
    std::vector<char> seen(perm.size(), 0);
    size_t transpositions = 0;
    for (size_t i = 0; i < perm.size(); ++i) {
        if (seen[i]) continue;
        size_t len = 0;
        for (size_t j = i; !seen[j]; j = perm[j]) {
            seen[j] = 1;
            ++len;
        }
        transpositions += len - 1;
    }
    return (transpositions % 2) ? -1 : 1;
}

// a_p = 1/2 (X_p + i Y_p) Z_{<p},  a_p^dag = 1/2 (X_p - i Y_p) Z_{<p}
// (sq_operator.cc:106-136); qubit excitations drop the Z string.
QubitOperator jw_single(size_t p, bool creator, bool qubit_excitation) {
    Circuit xc, yc;
    if (!qubit_excitation) {
        for (size_t k = 0; k < p; ++k) {
            xc.add_gate(make_gate("Z", k, k));
            yc.add_gate(make_gate("Z", k, k));
        }
    }
    xc.add_gate(make_gate("X", p, p));
    yc.add_gate(make_gate("Y", p, p));
    QubitOperator op;
    op.add_term(0.5, xc);
    op.add_term(creator ? cplx(0.0, -0.5) : cplx(0.0, 0.5), yc);
    return op;
}

}  // namespace

void SQOperator::canonical_order() {
    for (auto& t : terms_) {
        std::get<0>(t) *= (double)sort_descending(std::get<1>(t));
        std::get<0>(t) *= (double)sort_descending(std::get<2>(t));
    }
}

void SQOperator::simplify() {
    canonical_order();
    std::map<std::pair<std::vector<size_t>, std::vector<size_t>>, cplx> uniq;
    for (const auto& t : terms_) uniq[{std::get<1>(t), std::get<2>(t)}] += std::get<0>(t);
    terms_.clear();
    for (const auto& kv : uniq)
        if (std::abs(kv.second) > 1.0e-12) terms_.emplace_back(kv.second, kv.first.first, kv.first.second);
}

QubitOperator SQOperator::jw_transform(bool qubit_excitation) {
    simplify();
    QubitOperator out;
    for (const auto& t : terms_) {
        const auto& cre = std::get<1>(t);
        const auto& ann = std::get<2>(t);
        if (cre.empty() && ann.empty()) {
            out.add_term(std::get<0>(t), Circuit());
            continue;
        }
        // an empty running product starts over from the next factor, as
        // sq_operator.cc:128-133 does
        QubitOperator prod;
        auto mul = [&](size_t p, bool creator) {
            QubitOperator f = jw_single(p, creator, qubit_excitation);
            if (prod.terms().empty()) {
                prod.add_op(f);
            } else {
                prod.operator_product(f);
            }
        };
        for (size_t p : cre) mul(p, true);
        for (size_t p : ann) mul(p, false);
        prod.mult_coeffs(std::get<0>(t));
        out.add_op(prod);
    }
    out.simplify();
    return out;
}

std::string SQOperator::str() const {
    std::vector<std::string> s;
    s.push_back("");
    for (const auto& t : terms_) {
        s.push_back(fmt_complex(std::get<0>(t)));
        s.push_back("(");
        for (size_t k : std::get<1>(t)) s.push_back(std::to_string(k) + "^");
        for (size_t k : std::get<2>(t)) s.push_back(std::to_string(k));
        s.push_back(")\n");
    }
    return join(s, " ");
}

// ---- QubitOpPool --------------------------------------------------------------------

void QubitOpPool::set_coeffs(const std::vector<cplx>& c) {
    if (c.size() != terms_.size())
        throw std::invalid_argument("Number of new coefficients for quantum op pool must equal number of terms in pool.");
    for (size_t i = 0; i < c.size(); ++i) terms_[i].first = c[i];
}

void QubitOpPool::set_op_coeffs(const std::vector<cplx>& c) {
    for (auto& t : terms_) {
        if (c.size() != t.second.terms().size())
            throw std::invalid_argument("Number of new coeficients for quantum operator must equal number of terim in operator.");
        t.second.set_coeffs(c);
    }
}

void QubitOpPool::square(bool upper_triangle_only) {
    std::vector<Term> out;
    for (size_t I = 0; I < terms_.size(); ++I)
        for (size_t J = upper_triangle_only ? I : 0; J < terms_.size(); ++J) {
            QubitOperator IJ;
            IJ.add_op(terms_[I].second);
            IJ.operator_product(terms_[J].second, false);
            IJ.simplify();
            out.emplace_back(std::conj(terms_[I].first) * terms_[J].first, IJ);
        }
    terms_ = std::move(out);
}

void QubitOpPool::join_op_from_right(const QubitOperator& q) {
    for (auto& t : terms_) {
        t.second.operator_product(q, false);
        t.second.simplify();
    }
}

void QubitOpPool::join_op_from_right_lazy(const QubitOperator& q) {
    for (auto& t : terms_) t.second.operator_product(q, false, false);
}

void QubitOpPool::join_op_from_left(const QubitOperator& q) {
    for (auto& t : terms_) {
        QubitOperator l;
        l.add_op(q);
        l.operator_product(t.second, false);
        l.simplify();
        t.second = l;
    }
}

void QubitOpPool::join_as_commutator(const QubitOperator& q) {
    // [q, A_k] = q A_k - A_k q (qubit_op_pool.cc:103-122)
    for (auto& t : terms_) {
        QubitOperator qa;
        qa.add_op(q);
        qa.operator_product(t.second, false);
        QubitOperator aq;
        aq.add_op(t.second);
        aq.operator_product(q, false);
        aq.mult_coeffs(-1.0);
        qa.add_op(aq);
        qa.simplify();
        t.second = qa;
    }
}

void QubitOpPool::fill_pool(const std::string& pool_type, size_t nqb) {
    // every Pauli string on nqb qubits, enumerated as base-4 numbers with the
    // most significant digit on qubit 0 (qubit_op_pool.cc:124-175, 195-209)
    const bool complete = pool_type == "complete_qubit";
    const bool odd_y = pool_type == "cqoy";
    if (!complete && !odd_y) throw std::invalid_argument("Invalid pool_type specified.");
    static const char* letter[4] = {"I", "X", "Y", "Z"};
    const uint64_t nterms = 1ull << (2 * nqb);
    for (uint64_t I = 0; I < nterms; ++I) {
        Circuit c;
        int ny = 0;
        for (size_t k = 0; k < nqb; ++k) {
            const int digit = (int)((I >> (2 * (nqb - 1 - k))) & 3ull);
            if (digit == 2) ++ny;
            if (digit) c.add_gate(make_gate(letter[digit], k, k));
        }
        if (complete || (ny % 2)) {
            QubitOperator op;
            op.add_term(1.0, c);
            add_term(1.0, op);
        }
    }
}

std::string QubitOpPool::str() const {
    std::vector<std::string> s;
    s.push_back("");
    for (size_t i = 0; i < terms_.size(); ++i) {
        s.push_back("----->");
        s.push_back(std::to_string(i));
        s.push_back("<-----\n");
        s.push_back(fmt_complex(terms_[i].first));
        s.push_back("[\n");
        s.push_back(terms_[i].second.str());
        s.push_back("\n");
        s.push_back("]\n\n");
    }
    return join(s, " ");
}

// ---- SQOpPool -------------------------------------------------------------------------

void SQOpPool::set_coeffs(const std::vector<cplx>& c) {
    if (c.size() != terms_.size()) throw std::invalid_argument("Number of new coefficients for quantum operator must equal.");
    for (size_t i = 0; i < c.size(); ++i) terms_[i].first = c[i];
}

void SQOpPool::set_orb_spaces(const std::vector<int>& ref, const std::vector<size_t>& orb_irreps_to_int) {
    ref_int_ = 0;
    for (size_t i = 0; i < ref.size(); ++i) ref_int_ += (uint64_t)ref[i] << i;
    n_spinorb_ = (int)ref.size();
    if (n_spinorb_ % 2 != 0) throw std::invalid_argument("The total number of spinorbitals must be even!");
    n_occ_alpha_ = n_occ_beta_ = n_vir_alpha_ = n_vir_beta_ = 0;
    for (size_t i = 0; i < ref.size(); ++i) {
        if (i % 2 == 0) {
            n_occ_alpha_ += ref[i];
            n_vir_alpha_ += ref[i] ^ 1;
        } else {
            n_occ_beta_ += ref[i];
            n_vir_beta_ += ref[i] ^ 1;
        }
    }
    orb_irreps_to_int_ = orb_irreps_to_int.empty() ? std::vector<size_t>(n_spinorb_ / 2, 0) : orb_irreps_to_int;
}

QubitOpPool SQOpPool::get_qubit_op_pool() {
    QubitOpPool out;
    for (auto& t : terms_) out.add_term(t.first, t.second.jw_transform());
    return out;
}

QubitOperator SQOpPool::get_qubit_operator(const std::string& order_type, bool combine_like_terms,
                                           bool qubit_excitations) {
    QubitOperator parent;
    if (order_type == "unique_lex") {
        for (auto& t : terms_) {
            QubitOperator child = t.second.jw_transform(qubit_excitations);
            child.mult_coeffs(t.first);
            parent.add_op(child);
        }
        parent.simplify();
        parent.order_terms();
    } else if (order_type == "commuting_grp_lex") {
        for (auto& t : terms_) {
            QubitOperator child = t.second.jw_transform(qubit_excitations);
            child.mult_coeffs(t.first);
            child.simplify(combine_like_terms);
            child.order_terms();
            parent.add_op(child);
        }
    } else {
        throw std::invalid_argument("Invalid order_type specified.");
    }
    return parent;
}

void SQOpPool::fill_pool(const std::string& pool_type) {
    if (pool_type == "GSD") return fill_gsd();
    static const std::map<std::string, int> ranks = {{"S", 1},    {"SD", 2},    {"SDT", 3},
                                                     {"SDTQ", 4}, {"SDTQP", 5}, {"SDTQPH", 6}};
    auto it = ranks.find(pool_type);
    if (it != ranks.end()) return fill_excitations(it->second);
    if (pool_type == "sa_SD") return fill_sa_sd();
    throw std::invalid_argument("Invalid pool_type specified.");
}

void SQOpPool::fill_excitations(int max_rank) {
    // Determinants with the reference's alpha and beta electron counts, built
    // from the permutations of each spin string that follow the reference's own
    // string in lexicographic order (MSB-first strings, sq_op_pool.cc:351-384),
    // then sorted as integers (sq_op_pool.cc:386-387).
    const int ns = n_spinorb_;
    const int norb = ns / 2;
    // spin strings, most significant orbital first: beta = odd qubits, alpha = even
    std::string a_str, b_str;
    for (int o = norb - 1; o >= 0; --o) {
        b_str += ((ref_int_ >> (2 * o + 1)) & 1ull) ? '1' : '0';
        a_str += ((ref_int_ >> (2 * o)) & 1ull) ? '1' : '0';
    }
    std::vector<std::string> a_perm, b_perm;
    do a_perm.push_back(a_str);
    while (std::next_permutation(a_str.begin(), a_str.end()));
    do b_perm.push_back(b_str);
    while (std::next_permutation(b_str.begin(), b_str.end()));
    std::vector<uint64_t> dets;
    dets.reserve(a_perm.size() * b_perm.size());
    for (const auto& sa : a_perm)
        for (const auto& sb : b_perm) {
            uint64_t d = 0;
            for (int k = 0; k < norb; ++k) {
                const int o = norb - 1 - k;
                if (sa[k] == '1') d |= 1ull << (2 * o);
                if (sb[k] == '1') d |= 1ull << (2 * o + 1);
            }
            dets.push_back(d);
        }
    std::sort(dets.begin(), dets.end());
    for (uint64_t d : dets) {
        const uint64_t ex = ref_int_ ^ d;
        const int nidx = __builtin_popcountll(ex);
        if (nidx % 2 != 0) throw std::invalid_argument("The number of excitation indices must be an even number!");
        const int rank = nidx / 2;
        if (rank == 0 || rank > max_rank) continue;
        // occupied / unoccupied indices in descending order (sq_op_pool.cc:403-413)
        std::vector<size_t> occ, unocc;
        for (int q = ns - 1; q >= 0; --q) {
            if (!((ex >> q) & 1ull)) continue;
            if ((ref_int_ >> q) & 1ull) {
                occ.push_back((size_t)q);
            } else {
                unocc.push_back((size_t)q);
            }
        }
        std::vector<size_t> all(occ);
        all.insert(all.end(), unocc.begin(), unocc.end());
        if (find_irrep(orb_irreps_to_int_, all)) continue;
        SQOperator t;
        t.add_term(+1.0, unocc, occ);
        t.add_term(-1.0, std::vector<size_t>(occ.rbegin(), occ.rend()), std::vector<size_t>(unocc.rbegin(), unocc.rend()));
        t.simplify();
        add_term(1.0, t);
    }
}

void SQOpPool::fill_gsd() {
    const size_t norb = (size_t)n_spinorb_ / 2;
    // generalized singles (sq_op_pool.cc:112-144)
    for (size_t i = 0; i < norb; ++i)
        for (size_t a = i; a < norb; ++a) {
            if (find_irrep(orb_irreps_to_int_, {2 * i, 2 * a})) continue;
            for (size_t spin = 0; spin < 2; ++spin) {
                const size_t pi = 2 * i + spin, pa = 2 * a + spin;
                if (pa == pi) continue;
                SQOperator t;
                t.add_term(+1.0, {pa}, {pi});
                t.add_term(-1.0, {pi}, {pa});
                t.simplify();
                if (!t.terms().empty()) add_term(1.0, t);
            }
        }
    // generalized doubles (sq_op_pool.cc:146-322); an operator is kept only if
    // neither it nor its adjoint was generated before.
    std::set<std::vector<size_t>> seen_op, seen_adj;
    auto consider = [&](size_t p, size_t q, size_t r, size_t s) {
        // p^ q^ r s - s^ r^ q p   (creators p, q; annihilators r, s)
        if (p == q || r == s) return;
        SQOperator t;
        t.add_term(+1.0, {p, q}, {r, s});
        t.add_term(-1.0, {s, r}, {q, p});
        t.simplify();
        if (t.terms().size() < 2) return;
        const auto& t0 = t.terms()[0];
        const auto& t1 = t.terms()[1];
        std::vector<size_t> v0 = {std::get<1>(t0)[0], std::get<1>(t0)[1], std::get<2>(t0)[0], std::get<2>(t0)[1]};
        std::vector<size_t> v1 = {std::get<1>(t1)[0], std::get<1>(t1)[1], std::get<2>(t1)[0], std::get<2>(t1)[1]};
        if (seen_op.count(v0) || seen_adj.count(v0)) return;
        seen_op.insert(v0);
        seen_adj.insert(v1);
        add_term(1.0, t);
    };
    for (size_t i = 0; i < norb; ++i)
        for (size_t j = i; j < norb; ++j)
            for (size_t a = 0; a < norb; ++a)
                for (size_t b = a; b < norb; ++b) {
                    if (find_irrep(orb_irreps_to_int_, {2 * i, 2 * j, 2 * a, 2 * b})) continue;
                    const size_t ia = 2 * i, ib = 2 * i + 1, ja = 2 * j, jb = 2 * j + 1;
                    const size_t aa = 2 * a, ab = 2 * a + 1, ba = 2 * b, bb = 2 * b + 1;
                    // same order of spin cases as the reference: aaaa, bbbb, abab, baba, abba, baab
                    if (aa != ba && ia != ja) consider(aa, ba, ia, ja);
                    if (ab != bb && ib != jb) consider(ab, bb, ib, jb);
                    if (aa != bb && ia != jb) consider(aa, bb, ia, jb);
                    if (ab != ba && ib != ja) consider(ab, ba, ib, ja);
                    if (aa != bb && ib != ja) consider(aa, bb, ib, ja);
                    if (ab != ba && ia != jb) consider(ab, ba, ia, jb);
                }
}

void SQOpPool::fill_sa_sd() {
    if (n_occ_alpha_ != n_occ_beta_)
        throw std::invalid_argument("sa_SD operator pool requires a closed-shell reference determinant!");
    const size_t nocc = (size_t)n_occ_alpha_, nvir = (size_t)n_vir_alpha_;
    const double r2 = 1.0 / std::sqrt(2.0), r12 = 1.0 / std::sqrt(12.0);
    for (size_t i = 0; i < nocc; ++i)
        for (size_t a = 0; a < nvir; ++a) {
            const size_t ia = 2 * i, ib = 2 * i + 1, aa = 2 * nocc + 2 * a, ab = aa + 1;
            if (find_irrep(orb_irreps_to_int_, {ia, aa})) continue;
            SQOperator t;
            t.add_term(+r2, {aa}, {ia});
            t.add_term(+r2, {ab}, {ib});
            t.add_term(-r2, {ia}, {aa});
            t.add_term(-r2, {ib}, {ab});
            t.simplify();
            add_term(1.0, t);
        }
    for (size_t i = 0; i < nocc; ++i)
        for (size_t j = i; j < nocc; ++j)
            for (size_t a = 0; a < nvir; ++a)
                for (size_t b = a; b < nvir; ++b) {
                    const size_t ia = 2 * i, ib = ia + 1, ja = 2 * j, jb = ja + 1;
                    const size_t aa = 2 * nocc + 2 * a, ab = aa + 1, ba = 2 * nocc + 2 * b, bb = ba + 1;
                    if (find_irrep(orb_irreps_to_int_, {ia, ja, aa, ba})) continue;
                    // singlet and triplet-coupled doubles (sq_op_pool.cc:473-565)
                    SQOperator s2, t2;
                    if (aa != ba && ia != ja) s2.add_term(2 * r12, {aa, ba}, {ia, ja});
                    if (ab != bb && ib != jb) s2.add_term(2 * r12, {ab, bb}, {ib, jb});
                    if (aa != bb && ia != jb) s2.add_term(r12, {aa, bb}, {ia, jb});
                    if (ab != ba && ib != ja) s2.add_term(r12, {ab, ba}, {ib, ja});
                    if (aa != bb && ib != ja) s2.add_term(r12, {aa, bb}, {ib, ja});
                    if (ab != ba && ia != jb) s2.add_term(r12, {ab, ba}, {ia, jb});
                    if (ja != ia && ba != aa) s2.add_term(-2 * r12, {ja, ia}, {ba, aa});
                    if (jb != ib && bb != ab) s2.add_term(-2 * r12, {jb, ib}, {bb, ab});
                    if (jb != ia && bb != aa) s2.add_term(-r12, {jb, ia}, {bb, aa});
                    if (ja != ib && ba != ab) s2.add_term(-r12, {ja, ib}, {ba, ab});
                    if (ja != ib && bb != aa) s2.add_term(-r12, {ja, ib}, {bb, aa});
                    if (jb != ia && ba != ab) s2.add_term(-r12, {jb, ia}, {ba, ab});

                    if (aa != bb && ia != jb) t2.add_term(0.5, {aa, bb}, {ia, jb});
                    if (ab != ba && ib != ja) t2.add_term(0.5, {ab, ba}, {ib, ja});
                    if (aa != bb && ib != ja) t2.add_term(-0.5, {aa, bb}, {ib, ja});
                    if (ab != ba && ia != jb) t2.add_term(-0.5, {ab, ba}, {ia, jb});
                    if (jb != ia && bb != aa) t2.add_term(-0.5, {jb, ia}, {bb, aa});
                    if (ja != ib && ba != ab) t2.add_term(-0.5, {ja, ib}, {ba, ab});
                    if (ja != ib && bb != aa) t2.add_term(0.5, {ja, ib}, {bb, aa});
                    if (jb != ia && ba != ab) t2.add_term(0.5, {jb, ia}, {ba, ab});

                    s2.simplify();
                    t2.simplify();
                    cplx ns(0.0, 0.0), nt(0.0, 0.0);
                    for (const auto& t : s2.terms()) ns += std::norm(std::get<0>(t));
                    for (const auto& t : t2.terms()) nt += std::norm(std::get<0>(t));
                    s2.mult_coeffs(std::sqrt(2.0 / ns));
                    t2.mult_coeffs(std::sqrt(2.0 / nt));
                    if (!s2.terms().empty()) add_term(1.0, s2);
                    if (!t2.terms().empty()) add_term(1.0, t2);
                }
}

std::string SQOpPool::str() const {
    std::vector<std::string> s;
    s.push_back("");
    for (size_t i = 0; i < terms_.size(); ++i) {
        s.push_back("----->");
        s.push_back(std::to_string(i));
        s.push_back("<-----\n");
        s.push_back(fmt_complex(terms_[i].first));
        s.push_back("[\n");
        s.push_back(terms_[i].second.str());
        s.push_back("]\n\n");
    }
    return join(s, " ");
}

// ---- molecular Hamiltonian as a linear map of the integrals ---------------------------

namespace {

// A product of ladder operators in mask form: sum of c * M(x, z).
struct MaskTerm {
    uint64_t x, z;
    double c;
};

// a_p^dag = 1/2 (X_p - i Y_p) Z_{<p} and a_p = 1/2 (X_p + i Y_p) Z_{<p}
// (sq_operator.cc:106-136); with Y = i X Z as operators these are
// 1/2 M(b, Z<p) +/- 1/2 M(b, Z<p | b), b = bit p.
void ladder(size_t p, bool creator, MaskTerm out[2]) {
    const uint64_t b = 1ull << p, lo = b - 1ull;
    out[0] = {b, lo, 0.5};
    out[1] = {b, lo | b, creator ? 0.5 : -0.5};
}

// M(x1, z1) M(x2, z2) = (-1)^{popcount(x2 & z1)} M(x1 ^ x2, z1 ^ z2)
std::vector<MaskTerm> ladder_product(const std::vector<size_t>& cre, const std::vector<size_t>& ann) {
    std::vector<MaskTerm> prod{{0, 0, 1.0}}, next;
    auto mul = [&](size_t p, bool creator) {
        MaskTerm f[2];
        ladder(p, creator, f);
        next.clear();
        for (const auto& a : prod)
            for (const auto& b : f) {
                const double sg = (__builtin_popcountll(b.x & a.z) & 1) ? -1.0 : 1.0;
                next.push_back({a.x ^ b.x, a.z ^ b.z, sg * a.c * b.c});
            }
        // combine equal masks (all coefficients are dyadic: the sums are exact)
        std::map<std::pair<uint64_t, uint64_t>, double> acc;
        for (const auto& t : next) acc[{t.x, t.z}] += t.c;
        prod.clear();
        for (const auto& kv : acc)
            if (kv.second != 0.0) prod.push_back({kv.first.first, kv.first.second, kv.second});
    };
    for (size_t p : cre) mul(p, true);
    for (size_t p : ann) mul(p, false);
    return prod;
}

inline size_t pair_index(size_t p, size_t q) { return p >= q ? p * (p + 1) / 2 + q : q * (q + 1) / 2 + p; }

}  // namespace

HamiltonianMap molecular_hamiltonian_map(size_t norb) {
    if (norb == 0 || norb > 32) throw std::invalid_argument("molecular_hamiltonian_map: norb must be in [1, 32]");
    const size_t P = norb * (norb + 1) / 2;
    const size_t h0 = 1, g0 = 1 + P;
    auto gidx = [&](size_t p, size_t q, size_t r, size_t s) { return g0 + pair_index(pair_index(p, q), pair_index(r, s)); };
    std::map<std::pair<uint64_t, uint64_t>, std::map<uint32_t, double>> acc;
    std::vector<std::pair<uint64_t, uint64_t>> order;
    auto add = [&](double factor, size_t input, const std::vector<size_t>& cre, const std::vector<size_t>& ann) {
        for (const auto& t : ladder_product(cre, ann)) {
            auto key = std::make_pair(t.x, t.z);
            auto it = acc.find(key);
            if (it == acc.end()) {
                it = acc.emplace(key, std::map<uint32_t, double>()).first;
                order.push_back(key);
            }
            it->second[(uint32_t)input] += factor * t.c;
        }
    };
    add(1.0, 0, {}, {});
    // the loop structure and spin cases of molecule_adapters.py:183-210
    for (size_t i = 0; i < norb; ++i) {
        const size_t ia = 2 * i, ib = 2 * i + 1;
        for (size_t j = 0; j < norb; ++j) {
            const size_t ja = 2 * j, jb = 2 * j + 1;
            add(1.0, h0 + pair_index(i, j), {ia}, {ja});
            add(1.0, h0 + pair_index(i, j), {ib}, {jb});
            for (size_t k = 0; k < norb; ++k) {
                const size_t ka = 2 * k, kb = 2 * k + 1;
                for (size_t l = 0; l < norb; ++l) {
                    const size_t la = 2 * l, lb = 2 * l + 1;
                    const size_t g = gidx(i, l, k, j);  // (il|kj) / 2
                    if (ia != jb && kb != la) add(0.5, g, {ia, jb}, {kb, la});
                    if (ib != ja && ka != lb) add(0.5, g, {ib, ja}, {ka, lb});
                    if (ia != ja && ka != la) add(0.5, g, {ia, ja}, {ka, la});
                    if (ib != jb && kb != lb) add(0.5, g, {ib, jb}, {kb, lb});
                }
            }
        }
    }
    HamiltonianMap m;
    m.norb = norb;
    m.ninputs = g0 + P * (P + 1) / 2;
    m.offsets.push_back(0);
    for (const auto& key : order) {
        const auto& row = acc[key];
        size_t kept = 0;
        for (const auto& kv : row)
            if (std::fabs(kv.second) > 1e-15) {
                m.index.push_back(kv.first);
                m.value.push_back(kv.second);
                ++kept;
            }
        if (!kept) continue;
        m.x.push_back(key.first);
        m.z.push_back(key.second);
        m.offsets.push_back(m.index.size());
    }
    return m;
}

std::vector<double> pack_integrals(size_t norb, double constant, const double* h, const double* g) {
    const size_t P = norb * (norb + 1) / 2;
    std::vector<double> in(1 + P + P * (P + 1) / 2, 0.0);
    in[0] = constant;
    for (size_t p = 0; p < norb; ++p)
        for (size_t q = 0; q <= p; ++q) in[1 + pair_index(p, q)] = h[p * norb + q];
    for (size_t p = 0; p < norb; ++p)
        for (size_t q = 0; q <= p; ++q)
            for (size_t r = 0; r < norb; ++r)
                for (size_t s = 0; s <= r; ++s) {
                    const size_t pq = pair_index(p, q), rs = pair_index(r, s);
                    if (pq < rs) continue;
                    in[1 + P + pair_index(pq, rs)] = g[((p * norb + q) * norb + r) * norb + s];
                }
    return in;
}

QubitOperator map_to_operator(const HamiltonianMap& m, const std::vector<double>& inputs) {
    if (inputs.size() != m.ninputs) throw std::invalid_argument("map_to_operator: wrong number of inputs");
    QubitOperator op;
    const cplx mi(0.0, -1.0);
    for (size_t k = 0; k < m.x.size(); ++k) {
        double c = 0.0;
        for (uint64_t e = m.offsets[k]; e < m.offsets[k + 1]; ++e) c += m.value[e] * inputs[m.index[e]];
        // canonical circuit C(x, z) = i^{popcount(x & z)} M(x, z)
        cplx coef = c;
        for (int y = __builtin_popcountll(m.x[k] & m.z[k]); y > 0; --y) coef *= mi;
        Circuit circ;
        for (size_t q = 0; q < 64; ++q) {
            const bool bx = (m.x[k] >> q) & 1ull, bz = (m.z[k] >> q) & 1ull;
            if (bx && bz) circ.add_gate(make_gate("Y", q, q));
            else if (bx) circ.add_gate(make_gate("X", q, q));
            else if (bz) circ.add_gate(make_gate("Z", q, q));
        }
        op.add_term(coef, circ);
    }
    return op;
}


}  // namespace qfa
