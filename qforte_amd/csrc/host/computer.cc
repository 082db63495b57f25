// Computer on the qfe engine (reference computer.cc), plus the native
// exponentiate_pauli_string / trotterize (utils/exponentiate.py,
// utils/trotterization.py) that tag each exponential's gate run so the engine
// applies it as one pass.
#include <algorithm>
#include <cmath>
#include <random>

#include "core.h"

namespace qfa {

void check(int rc, const char* what) {
    if (rc == QFE_OK) return;
    std::string msg = std::string(what) + ": " + qfe_last_error();
    if (rc == QFE_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// ---- exponentials -----------------------------------------------------------------------

std::pair<Circuit, cplx> exponentiate_pauli_string(cplx coefficient, const Circuit& term, bool use_cRz,
                                                   long ancilla_idx, bool use_open_cRz) {
    // exponentiate.py:9-87: basis change (H for X, Rx(pi/2) for Y), CNOT
    // ladder cX(target=q_k, control=q_{k-1}), Rz(-2 Im c) on the last qubit,
    // then the ladder and the basis change undone.  Net operator:
    // exp(i Im(c) P) for a Pauli string on distinct qubits.
    if (std::fabs(coefficient.real()) > 1.0e-14) {
        char buf[128];
        std::snprintf(buf, sizeof buf, "exponentiate_pauli_string() called with a real coefficient (%g%+gj)",
                      coefficient.real(), coefficient.imag());
        throw std::invalid_argument(buf);
    }
    if (term.size() == 0) return {Circuit(), std::exp(coefficient)};
    Circuit to_z, ladder;
    long prev = -1, last = -1;
    for (const auto& g : term.gates()) {
        const std::string& id = g.gate_id();
        const size_t t = g.target(), c = g.control();
        if (id == "X") {
            to_z.add_gate(make_gate("H", t, c));
        } else if (id == "Y") {
            to_z.add_gate(make_gate("Rx", t, c, M_PI / 2));
        } else if (id == "I") {
            continue;
        }
        if (prev >= 0) ladder.add_gate(make_gate("cX", t, (size_t)prev));
        prev = (long)t;
        last = (long)t;
    }
    if (last < 0) throw std::invalid_argument("exponentiate_pauli_string(): the term has no X, Y or Z gate");
    const double theta = -2.0 * coefficient.imag();
    Circuit out;
    out.add_circuit(to_z);
    out.add_circuit(ladder);
    if (use_cRz) {
        if (ancilla_idx < 0) throw std::invalid_argument("exponentiate_pauli_string(): Use_cRz needs ancilla_idx");
        if (use_open_cRz) out.add_gate(make_gate("X", (size_t)ancilla_idx, (size_t)ancilla_idx));
        out.add_gate(make_gate("cRz", (size_t)last, (size_t)ancilla_idx, theta));
        if (use_open_cRz) out.add_gate(make_gate("X", (size_t)ancilla_idx, (size_t)ancilla_idx));
    } else {
        out.add_gate(make_gate("Rz", (size_t)last, (size_t)last, theta));
    }
    out.add_circuit(ladder.adjoint());
    out.add_circuit(to_z.adjoint());
    if (!use_cRz) {
        // tag the run when the term is a Pauli string on distinct qubits
        uint64_t x, z;
        cplx ph;
        bool distinct = true;
        uint64_t seen = 0;
        for (const auto& g : term.gates()) {
            if (g.gate_id() == "I") continue;
            if (g.target() >= 64 || (seen >> g.target()) & 1ull) distinct = false;
            else seen |= 1ull << g.target();
        }
        if (distinct && pauli_masks(term, x, z, ph)) {
            out.add_segment(ExpSegment{0, out.size(), x, z, ph, coefficient.imag()});
        }
    }
    return {out, 1.0};
}

std::pair<Circuit, cplx> trotterize(const QubitOperator& op, cplx factor, int trotter_number, int trotter_order) {
    // trotterization.py:10-57: product over terms in list order (term 0 first);
    // with trotter_number N > 1 every term repeats N times with coefficient / N.
    Circuit out;
    cplx phase = 1.0;
    if (trotter_order > 1) throw std::logic_error("Higher order trotterization is not yet implemented.");
    if (trotter_number == 1) {
        for (const auto& t : op.terms()) {
            auto r = exponentiate_pauli_string(factor * t.first, t.second);
            out.add_circuit(r.first);
            phase *= r.second;
        }
    } else {
        for (int k = 0; k < trotter_number; ++k)
            for (const auto& t : op.terms()) {
                auto r = exponentiate_pauli_string(factor * t.first / (double)trotter_number, t.second);
                out.add_circuit(r.first);
                phase *= r.second;
            }
    }
    return {out, phase};
}

std::pair<Circuit, cplx> trotterize_w_cRz(const QubitOperator& op, size_t ancilla, cplx factor, bool use_open_cRz,
                                          int trotter_number, int trotter_order) {
    // trotterization.py:60-128
    Circuit out;
    cplx phase = 1.0;
    if (trotter_order > 1) throw std::logic_error("Higher order trotterization is not yet implemented.");
    const int reps = trotter_number == 1 ? 1 : trotter_number;
    for (int k = 0; k < reps; ++k)
        for (const auto& t : op.terms()) {
            const cplx c = trotter_number == 1 ? factor * t.first : factor * t.first / (double)trotter_number;
            auto r = exponentiate_pauli_string(c, t.second, true, (long)ancilla, use_open_cRz);
            out.add_circuit(r.first);
            phase *= r.second;
        }
    return {out, phase};
}

// ---- Computer ------------------------------------------------------------------------------

Computer::Computer(int nqubit, double print_threshold) : n_(nqubit), thr_(print_threshold) {
    if (nqubit < 1) throw std::invalid_argument("Computer: the number of qubits must be positive");
    check(qfe_state_create(nqubit, &s_), "Computer");
}

Computer::Computer(const Computer& o) : n_(o.n_), thr_(o.thr_) {
    check(qfe_state_create(n_, &s_), "Computer(copy)");
    check(qfe_state_copy(s_, o.s_), "Computer(copy)");
}

Computer::~Computer() {
    if (s_) qfe_state_destroy(s_);
}

void Computer::lower(const Circuit& c, std::vector<qfe_op>& ops) const {
    const auto& gates = c.gates();
    std::vector<const ExpSegment*> seg_at(gates.size() + 1, nullptr);
    for (const auto& s : c.segments())
        if (s.begin < s.end && s.end <= gates.size()) seg_at[s.begin] = &s;
    ops.reserve(ops.size() + gates.size());
    size_t i = 0;
    while (i < gates.size()) {
        const ExpSegment* s = seg_at[i];
        if (s) {
            qfe_op op{};
            op.kind = QFE_OP_PEXP;
            op.angle = s->angle;
            op.pauli = qfe_pauli{s->x, s->z, qfe_c64{s->phase.real(), s->phase.imag()}, qfe_c64{1.0, 0.0}};
            ops.push_back(op);
            i = s->end;
            continue;
        }
        qfe_op op{};
        op.kind = QFE_OP_GATE;
        op.gate = gates[i].to_abi();
        ops.push_back(op);
        ++i;
    }
}

void Computer::apply_ops(const std::vector<qfe_op>& ops) {
    if (ops.empty()) return;
    check(qfe_apply_ops(s_, ops.data(), ops.size()), "apply_circuit");
}

void Computer::apply_circuit(const Circuit& c) {
    std::vector<qfe_op> ops;
    lower(c, ops);
    apply_ops(ops);
}

void Computer::apply_gate(const Gate& g) {
    const qfe_gate a = g.to_abi();
    check(qfe_apply_gates(s_, &a, 1), "apply_gate");
}

// The reference's naive scatter sweep (computer.cc:319-338, 427-456), on the
// host: a verification path independent of the device kernels.
void Computer::apply_gate_safe(const Gate& g) {
    std::vector<cplx> a = get_coeff_vec(), out(a.size(), 0.0);
    const Mat4& m = g.matrix();
    const uint64_t N = a.size();
    const double thr = 1.0e-16;
    if (g.nqubits() == 1) {
        const size_t t = g.target();
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                if (std::abs(m[i][j]) <= thr) continue;
                for (uint64_t J = 0; J < N; ++J) {
                    if ((int)((J >> t) & 1ull) != j) continue;
                    const uint64_t I = i ? (J | (1ull << t)) : (J & ~(1ull << t));
                    out[I] += m[i][j] * a[J];
                }
            }
    } else {
        const size_t t = g.target(), c = g.control();
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                if (std::abs(m[i][j]) <= thr) continue;
                const int jc = j >> 1, jt = j & 1, ic = i >> 1, it = i & 1;
                for (uint64_t J = 0; J < N; ++J) {
                    const int bc = c < 64 ? (int)((J >> c) & 1ull) : 0;
                    const int bt = t < 64 ? (int)((J >> t) & 1ull) : 0;
                    if (bc != jc || bt != jt) continue;
                    uint64_t I = J;
                    I = ic ? (I | (1ull << c)) : (I & ~(1ull << c));
                    I = it ? (I | (1ull << t)) : (I & ~(1ull << t));
                    if (I >= N) throw std::invalid_argument("apply_gate_safe: qubit index out of range");
                    out[I] += m[i][j] * a[J];
                }
            }
    }
    set_coeff_vec(out.data(), out.size());
}

void Computer::apply_circuit_safe(const Circuit& c) {
    for (const auto& g : c.gates()) apply_gate_safe(g);
}

void Computer::apply_operator(const QubitOperator& op) {
    std::vector<qfe_pauli> terms;
    if (compile_pauli_sum(op, terms)) {
        qfe_state out = nullptr;
        check(qfe_state_create(n_, &out), "apply_operator");
        int rc = qfe_apply_pauli_sum(s_, terms.data(), terms.size(), out);
        if (rc) {
            qfe_state_destroy(out);
            check(rc, "apply_operator");
        }
        qfe_state_destroy(s_);
        s_ = out;
        return;
    }
    // general circuits: per term copy, apply, scale, accumulate (computer.cc:80-91)
    Computer acc(*this);
    check(qfe_state_zero(acc.s_), "apply_operator");
    for (const auto& t : op.terms()) {
        Computer tmp(*this);
        tmp.apply_circuit(t.second);
        check(qfe_state_axpy(acc.s_, qfe_c64{t.first.real(), t.first.imag()}, tmp.s_), "apply_operator");
    }
    std::swap(s_, acc.s_);
}

void Computer::apply_constant(cplx a) { check(qfe_state_scale(s_, qfe_c64{a.real(), a.imag()}), "apply_constant"); }

cplx Computer::direct_op_exp_val(const QubitOperator& op) {
    std::vector<qfe_pauli> terms;
    if (compile_pauli_sum(op, terms)) {
        qfe_c64 r;
        check(qfe_expval(s_, terms.data(), terms.size(), &r), "direct_op_exp_val");
        return {r.re, r.im};
    }
    Computer tmp(*this);
    tmp.apply_operator(op);
    return dot(*this, tmp);
}

std::vector<cplx> Computer::direct_oppl_exp_val(const QubitOpPool& pool) {
    std::vector<int> idx(pool.terms().size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
    return direct_idxd_oppl_exp_val(pool, idx);
}

std::vector<cplx> Computer::direct_idxd_oppl_exp_val(const QubitOpPool& pool, const std::vector<int>& idxs) {
    std::vector<qfe_pauli> all, one;
    std::vector<uint64_t> off{0};
    bool pauli = true;
    for (int k : idxs) {
        if (k < 0 || (size_t)k >= pool.terms().size()) throw std::out_of_range("direct_idxd_oppl_exp_val: index out of range");
        if (!compile_pauli_sum(pool.terms()[k].second, one)) {
            pauli = false;
            break;
        }
        all.insert(all.end(), one.begin(), one.end());
        off.push_back(all.size());
    }
    std::vector<cplx> out;
    if (pauli) {
        std::vector<qfe_c64> r(idxs.size());
        check(qfe_expval_many(s_, all.data(), off.data(), idxs.size(), r.data()), "direct_oppl_exp_val");
        for (size_t i = 0; i < idxs.size(); ++i) out.push_back(cplx(r[i].re, r[i].im) * pool.terms()[idxs[i]].first);
        return out;
    }
    for (int k : idxs) out.push_back(direct_op_exp_val(pool.terms()[k].second) * pool.terms()[k].first);
    return out;
}

std::vector<cplx> Computer::direct_oppl_exp_val_w_mults(const QubitOpPool& pool, const std::vector<cplx>& mults) {
    // computer.cc:808-835 (default build): sum_l mults[l] c_pool c_l <circ_l>.
    // Pools of Pauli sums go to the device in one call: term l of operator k
    // carries the coefficient mults[l] c_l, and c_pool scales the operator's sum.
    std::vector<qfe_pauli> all, one;
    std::vector<uint64_t> off{0};
    bool pauli = true;
    for (const auto& pt : pool.terms()) {
        const auto& ts = pt.second.terms();
        if (ts.size() > mults.size()) throw std::out_of_range("direct_oppl_exp_val_w_mults: too few multipliers");
        if (!compile_pauli_sum(pt.second, one)) {
            pauli = false;
            break;
        }
        for (size_t l = 0; l < one.size(); ++l) {
            const cplx c = mults[l] * cplx(one[l].coeff.re, one[l].coeff.im);
            one[l].coeff = qfe_c64{c.real(), c.imag()};
        }
        all.insert(all.end(), one.begin(), one.end());
        off.push_back(all.size());
    }
    if (pauli) {
        const size_t nops = pool.terms().size();
        std::vector<qfe_c64> r(nops);
        if (nops) check(qfe_expval_many(s_, all.data(), off.data(), nops, r.data()), "direct_oppl_exp_val_w_mults");
        std::vector<cplx> out;
        for (size_t k = 0; k < nops; ++k) out.push_back(pool.terms()[k].first * cplx(r[k].re, r[k].im));
        return out;
    }
    std::vector<cplx> out;
    for (const auto& pt : pool.terms()) {
        cplx acc = 0.0;
        const auto& ts = pt.second.terms();
        for (size_t l = 0; l < ts.size(); ++l) {
            if (l >= mults.size()) throw std::out_of_range("direct_oppl_exp_val_w_mults: too few multipliers");
            acc += mults[l] * pt.first * ts[l].first * direct_circ_exp_val(ts[l].second);
        }
        out.push_back(acc);
    }
    return out;
}

cplx Computer::direct_circ_exp_val(const Circuit& c) {
    uint64_t x, z;
    cplx ph;
    if (pauli_masks(c, x, z, ph)) {
        qfe_pauli t{x, z, qfe_c64{ph.real(), ph.imag()}, qfe_c64{1.0, 0.0}};
        qfe_c64 r;
        check(qfe_expval(s_, &t, 1, &r), "direct_circ_exp_val");
        return {r.re, r.im};
    }
    Computer tmp(*this);
    tmp.apply_circuit(c);
    return dot(*this, tmp);
}

cplx Computer::direct_pauli_circ_exp_val(const Circuit& c) {
    // computer.cc:850-911 accepts X, Y, Z only; the value returned here is the
    // exact <psi|P|psi> (the reference's block formula negates it for an odd
    // number of Y factors).
    for (const auto& g : c.gates()) {
        const std::string& id = g.gate_id();
        if (id != "X" && id != "Y" && id != "Z") throw std::runtime_error("Not a valid pauli gate!");
    }
    return direct_circ_exp_val(c);
}

Computer Computer::basis_rotated(const Circuit& c) const {
    // measurement basis change of computer.cc:141-165: Z -> I, X -> H, Y -> Rx(pi/2)
    Computer tmp(*this);
    Circuit rot;
    for (const auto& g : c.gates()) {
        const std::string& id = g.gate_id();
        const size_t t = g.target();
        if (id == "Z") rot.add_gate(make_gate("I", t, t));
        else if (id == "X") rot.add_gate(make_gate("H", t, t));
        else if (id == "Y") rot.add_gate(make_gate("Rx", t, t, M_PI / 2));
    }
    tmp.apply_circuit(rot);
    return tmp;
}

// Basis indices drawn from |psi|^2 on the device, seeded from
// std::random_device as the reference's generator is (computer.cc:169-174).
std::vector<uint64_t> Computer::sample(size_t n_measurements) const {
    std::random_device rd;
    const uint64_t seed = ((uint64_t)rd() << 32) ^ (uint64_t)rd();
    std::vector<uint64_t> idx(n_measurements);
    check(qfe_sample(s_, n_measurements, seed, idx.data()), "measure");
    return idx;
}

std::vector<double> Computer::measure_circuit(const Circuit& c, size_t n_measurements) {
    const std::vector<uint64_t> idx = basis_rotated(c).sample(n_measurements);
    std::vector<double> out(n_measurements);
    for (size_t k = 0; k < n_measurements; ++k) {
        double v = 1.0;
        for (const auto& g : c.gates()) v *= 1.0 - 2.0 * (double)((idx[k] >> g.target()) & 1ull);
        out[k] = v;
    }
    return out;
}

std::vector<std::vector<int>> Computer::measure_z_readouts_fast(size_t na, size_t nb, size_t n_measurements) {
    const std::vector<uint64_t> idx = sample(n_measurements);
    std::vector<std::vector<int>> out(n_measurements);
    for (size_t k = 0; k < n_measurements; ++k)
        for (size_t l = na; l < nb + 1; ++l) out[k].push_back((int)((idx[k] >> l) & 1ull));
    return out;
}

std::vector<std::vector<int>> Computer::measure_readouts(const Circuit& c, size_t n_measurements) {
    const std::vector<uint64_t> idx = basis_rotated(c).sample(n_measurements);
    std::vector<std::vector<int>> out(n_measurements);
    for (size_t k = 0; k < n_measurements; ++k)
        for (const auto& g : c.gates()) out[k].push_back((int)((idx[k] >> g.target()) & 1ull));
    return out;
}

// sum_k prod_g (1 - 2 bit_g(k)) |a_k|^2 in the rotated basis (computer.cc:
// 275-315) = <psi'| Z(mask) |psi'>, mask the XOR of the gate targets.
double Computer::perfect_measure_circuit(const Circuit& c) {
    Computer tmp = basis_rotated(c);
    qfe_pauli t{};
    for (const auto& g : c.gates()) t.z ^= 1ull << g.target();
    t.phase.re = 1.0;
    t.coeff.re = 1.0;
    qfe_c64 out{};
    check(qfe_expval(tmp.s_, &t, 1, &out), "perfect_measure_circuit");
    return out.re;
}

std::vector<cplx> Computer::get_coeff_vec() const {
    std::vector<cplx> out(get_nbasis());
    get_coeff_into(out.data());
    return out;
}

void Computer::get_coeff_into(cplx* out) const {
    check(qfe_state_download(s_, reinterpret_cast<qfe_c64*>(out), get_nbasis()), "get_coeff_vec");
}

void Computer::set_coeff_vec(const cplx* data, size_t n) {
    if (n != get_nbasis()) throw std::invalid_argument("set_coeff_vec: expected 2^nqubit coefficients");
    check(qfe_state_upload(s_, reinterpret_cast<const qfe_c64*>(data), n), "set_coeff_vec");
}

cplx Computer::coeff(const QubitBasis& b) const {
    qfe_c64 r;
    check(qfe_state_get_amp(s_, b.index(), &r), "coeff");
    return {r.re, r.im};
}

void Computer::set_state(const std::vector<std::pair<QubitBasis, cplx>>& state) {
    std::vector<uint64_t> idx;
    std::vector<qfe_c64> val;
    for (const auto& bv : state) {
        idx.push_back(bv.first.index());
        val.push_back(qfe_c64{bv.second.real(), bv.second.imag()});
    }
    check(qfe_state_set_amps(s_, idx.data(), val.data(), idx.size()), "set_state");
}

void Computer::null_state() { check(qfe_state_zero(s_), "null_state"); }
void Computer::reset() { check(qfe_state_reset(s_), "reset"); }

std::string Computer::str() const {
    std::vector<cplx> a = get_coeff_vec();
    std::vector<std::string> s;
    s.push_back("Computer(");
    for (uint64_t i = 0; i < a.size(); ++i)
        if (std::abs(a[i]) >= thr_) s.push_back(fmt_complex(a[i]) + " " + QubitBasis(i).str((size_t)n_));
    s.push_back(")");
    return join(s, "\n");
}

bool operator==(const Computer& a, const Computer& b) {
    if (a.get_nqubit() != b.get_nqubit()) return false;
    int eq = 0;
    check(qfe_state_equal(a.handle(), b.handle(), &eq), "Computer.__eq__");
    return eq != 0;
}

cplx dot(const Computer& a, const Computer& b) {
    if (a.get_nqubit() != b.get_nqubit()) throw std::invalid_argument("inner_product: qubit counts differ");
    qfe_c64 r;
    check(qfe_state_dot(a.handle(), b.handle(), &r), "inner_product");
    return {r.re, r.im};
}

}  // namespace qfa
