// Circuits and Pauli sums (reference circuit.cc, qubit_operator.cc).
#include <algorithm>
#include <cmath>
#include <map>

#include "core.h"

namespace qfa {

// ---- Circuit -------------------------------------------------------------------

void Circuit::add_circuit(const Circuit& c) {
    const size_t off = gates_.size();
    gates_.insert(gates_.end(), c.gates_.begin(), c.gates_.end());
    for (ExpSegment s : c.segs_) {
        s.begin += off;
        s.end += off;
        segs_.push_back(s);
    }
}

void Circuit::insert_gate(size_t pos, const Gate& g) {
    if (pos > gates_.size()) throw std::runtime_error("Circuit::insert_gate: position out of range");
    gates_.insert(gates_.begin() + pos, g);
    segs_.clear();
}

void Circuit::remove_gate(size_t pos) {
    if (pos >= gates_.size()) throw std::runtime_error("Circuit::remove_gate: position out of range");
    gates_.erase(gates_.begin() + pos);
    segs_.clear();
}

void Circuit::replace_gate(size_t pos, const Gate& g) {
    if (pos >= gates_.size()) throw std::runtime_error("Circuit::replace_gate: position out of range");
    gates_[pos] = g;
    segs_.clear();
}

void Circuit::swap_gates(size_t p1, size_t p2) {
    if (p1 >= gates_.size() || p2 >= gates_.size()) throw std::runtime_error("Circuit::swap_gates: position out of range");
    std::swap(gates_[p1], gates_[p2]);
    segs_.clear();
}

void Circuit::insert_circuit(size_t pos, const Circuit& c) {
    if (pos > gates_.size()) throw std::runtime_error("Circuit::insert_circuit: position out of range");
    gates_.insert(gates_.begin() + pos, c.gates_.begin(), c.gates_.end());
    segs_.clear();
}

void Circuit::remove_gates(size_t p1, size_t p2) {
    if (p1 >= gates_.size() || p2 >= gates_.size()) throw std::runtime_error("Circuit::remove_gates: position out of range");
    gates_.erase(gates_.begin() + p1, gates_.begin() + p2);
    segs_.clear();
}

Circuit Circuit::adjoint() const {
    Circuit out;
    out.gates_.reserve(gates_.size());
    for (auto it = gates_.rbegin(); it != gates_.rend(); ++it) out.gates_.push_back(it->adjoint());
    // exp(i a P)^dag = exp(-i a P) over the mirrored gate range
    const size_t n = gates_.size();
    for (auto it = segs_.rbegin(); it != segs_.rend(); ++it) {
        ExpSegment s = *it;
        s.begin = n - it->end;
        s.end = n - it->begin;
        s.angle = -it->angle;
        out.segs_.push_back(s);
    }
    return out;
}

void Circuit::set_parameters(const std::vector<double>& params) {
    size_t k = 0;
    for (auto& g : gates_) {
        if (!g.has_parameter()) continue;
        if (k < params.size() && std::fabs(*g.parameter() - params[k]) > 1e-12)
            g = make_gate(g.gate_id(), g.target(), g.control(), params[k]);
        ++k;
    }
    if (k != params.size())
        throw std::runtime_error("Circuit::set_parameters: number of parameters does not match number of gates with parameters");
    segs_.clear();
}

void Circuit::set_parameter(size_t pos, double param) {
    if (pos >= gates_.size()) throw std::runtime_error("Circuit::set_parameter: position out of range");
    if (!gates_[pos].has_parameter()) throw std::runtime_error("Circuit::set_parameter: gate does not have a parameter");
    gates_[pos] = make_gate(gates_[pos].gate_id(), gates_[pos].target(), gates_[pos].control(), param);
    segs_.clear();
}

std::vector<double> Circuit::get_parameters() const {
    std::vector<double> p;
    for (const auto& g : gates_)
        if (g.has_parameter()) p.push_back(*g.parameter());
    return p;
}

bool Circuit::is_pauli() const {
    for (const auto& g : gates_) {
        const std::string& id = g.gate_id();
        if (id != "X" && id != "Y" && id != "Z") return false;
    }
    return true;
}

namespace {
// Product of two single-qubit Paulis in written order (left * right).
std::pair<cplx, char> pauli_product(char l, char r) {
    const cplx i(0.0, 1.0);
    if (l == 'I') return {1.0, r};
    if (l == r) return {1.0, 'I'};
    if (l == 'X' && r == 'Y') return {i, 'Z'};
    if (l == 'Y' && r == 'X') return {-i, 'Z'};
    if (l == 'Y' && r == 'Z') return {i, 'X'};
    if (l == 'Z' && r == 'Y') return {-i, 'X'};
    if (l == 'Z' && r == 'X') return {i, 'Y'};
    /* l == 'X' && r == 'Z' */ return {-i, 'Y'};
}
}  // namespace

cplx Circuit::canonicalize_pauli_circuit() {
    if (gates_.empty()) return 1.0;
    if (!is_pauli())
        throw std::runtime_error(
            "Circuit::canonicalize_pauli_circuit is undefined for circuits with gates other than X, Y, or Z");
    std::stable_sort(gates_.begin(), gates_.end(), [](const Gate& a, const Gate& b) { return a.target() < b.target(); });
    std::vector<Gate> out;
    cplx coeff = 1.0;
    size_t i = 0;
    while (i < gates_.size()) {
        const size_t q = gates_[i].target();
        char acc = gates_[i].gate_id()[0];
        size_t j = i + 1;
        for (; j < gates_.size() && gates_[j].target() == q; ++j) {
            auto pr = pauli_product(acc, gates_[j].gate_id()[0]);
            coeff *= pr.first;
            acc = pr.second;
        }
        if (acc != 'I') out.push_back(make_gate(std::string(1, acc), q, q));
        i = j;
    }
    gates_ = std::move(out);
    segs_.clear();
    return coeff;
}

int Circuit::get_num_cnots() const {
    int n = 0;
    for (const auto& g : gates_) {
        const std::string& id = g.gate_id();
        if (id == "CNOT" || id == "cX" || id == "aCNOT" || id == "acX") {
            n += 1;
        } else if (id == "A" || id == "SWAP") {
            n += 3;
        }
    }
    return n;
}

std::string Circuit::str() const {
    std::vector<std::string> s;
    for (auto it = gates_.rbegin(); it != gates_.rend(); ++it) s.push_back(it->str());
    return "[" + join(s, " ") + "]";
}

size_t Circuit::num_qubits() const {
    size_t m = 0;
    for (const auto& g : gates_) m = std::max({m, g.target() + 1, g.control() + 1});
    return m;
}

bool operator==(const Circuit& a, const Circuit& b) { return a.gates() == b.gates(); }
bool operator<(const Circuit& a, const Circuit& b) {
    return std::lexicographical_compare(a.gates().begin(), a.gates().end(), b.gates().begin(), b.gates().end());
}

bool pauli_masks(const Circuit& c, uint64_t& x, uint64_t& z, cplx& phase) {
    // Left-multiply letter by letter in application order:
    //   X_q: flip; Z_q: sign (-1)^{(j^x)_q}; Y_q = i X Z: both, times i.
    x = z = 0;
    phase = 1.0;
    const cplx i(0.0, 1.0);
    for (const auto& g : c.gates()) {
        const char l = g.pauli_letter();
        if (!l) return false;
        const size_t q = g.target();
        if (q >= 64) return false;
        const uint64_t bit = 1ull << q;
        if (l == 'X') {
            x ^= bit;
        } else if (l == 'Z') {
            if (x & bit) phase = -phase;
            z ^= bit;
        } else if (l == 'Y') {
            phase *= i;
            if (x & bit) phase = -phase;
            z ^= bit;
            x ^= bit;
        }
    }
    return true;
}

// ---- QubitOperator ----------------------------------------------------------------

void QubitOperator::set_coeffs(const std::vector<cplx>& c) {
    if (c.size() != terms_.size()) throw std::invalid_argument("number of new coefficients for quantum operator must equal ");
    for (size_t i = 0; i < c.size(); ++i) terms_[i].first = c[i];
}

void QubitOperator::mult_coeffs(cplx m) {
    for (auto& t : terms_) t.first *= m;
}

void QubitOperator::order_terms() {
    simplify();
    // Comparator of qubit_operator.cc:53-71: targets position by position,
    // then gate ids position by position, else "not less" (the reference's
    // final size comparison compares a term with itself).
    std::sort(terms_.begin(), terms_.end(), [](const Term& a, const Term& b) {
        const auto& ga = a.second.gates();
        const auto& gb = b.second.gates();
        const size_t m = std::min(ga.size(), gb.size());
        for (size_t k = 0; k < m; ++k)
            if (ga[k].target() != gb[k].target()) return ga[k].target() < gb[k].target();
        for (size_t k = 0; k < m; ++k)
            if (ga[k].gate_id() != gb[k].gate_id()) return ga[k].gate_id() < gb[k].gate_id();
        return false;
    });
}

void QubitOperator::canonical_order() {
    for (auto& t : terms_) t.first *= t.second.canonicalize_pauli_circuit();
}

void QubitOperator::simplify(bool combine_like_terms) {
    canonical_order();
    std::map<Circuit, cplx> uniq;
    for (const auto& t : terms_) {
        auto it = uniq.find(t.second);
        if (it == uniq.end()) {
            uniq.emplace(t.second, t.first);
        } else {
            it->second += t.first;
        }
    }
    terms_.clear();
    for (const auto& kv : uniq)
        if (!combine_like_terms || std::abs(kv.second) > 1.0e-12) terms_.emplace_back(kv.second, kv.first);
}

void QubitOperator::operator_product(const QubitOperator& r, bool pre_simplify, bool post_simplify) {
    if (pre_simplify) simplify();
    std::vector<Term> prod;
    prod.reserve(terms_.size() * r.terms_.size());
    for (const auto& tl : terms_)
        for (const auto& tr : r.terms_) {
            Circuit c;
            for (const auto& g : tl.second.gates()) c.add_gate(g);
            for (const auto& g : tr.second.gates()) c.add_gate(g);
            prod.emplace_back(tl.first * tr.first, std::move(c));
        }
    terms_ = std::move(prod);
    if (post_simplify) {
        simplify();
    } else {
        canonical_order();
    }
}

bool QubitOperator::check_op_equivalence(QubitOperator other, bool reorder) {
    if (reorder) {
        order_terms();
        other.order_terms();
    }
    if (terms_.size() != other.terms_.size()) return false;
    for (size_t i = 0; i < terms_.size(); ++i) {
        if (std::abs(terms_[i].first - other.terms_[i].first) > 1.0e-10) return false;
        if (!(terms_[i].second == other.terms_[i].second)) return false;
    }
    return true;
}

std::string QubitOperator::str() const {
    std::vector<std::string> s;
    for (const auto& t : terms_) s.push_back(fmt_complex(t.first) + t.second.str());
    return join(s, "\n");
}

size_t QubitOperator::num_qubits() const {
    size_t m = 0;
    for (const auto& t : terms_) m = std::max(m, t.second.num_qubits());
    return m;
}

bool compile_pauli_sum(const QubitOperator& op, std::vector<qfe_pauli>& out, cplx mult) {
    out.clear();
    out.reserve(op.terms().size());
    for (const auto& t : op.terms()) {
        uint64_t x, z;
        cplx ph;
        if (!pauli_masks(t.second, x, z, ph)) return false;
        const cplx c = t.first * mult;
        out.push_back(qfe_pauli{x, z, qfe_c64{ph.real(), ph.imag()}, qfe_c64{c.real(), c.imag()}});
    }
    return true;
}

}  // namespace qfa
