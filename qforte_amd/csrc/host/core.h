// Host core of qforte_amd: the qforte data model (Gate, Circuit,
// QubitOperator, SQOperator, SQOpPool, QubitOpPool, QubitBasis) and the
// Computer, whose amplitudes live in HBM behind the qfe C ABI.
//
// The classes restate the behaviour of the reference's C++ core
// (src/qforte/*.h) -- term order, sign conventions, error messages -- so the
// qforte drivers run unchanged on top of them.  Each method cites the
// reference lines it follows.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <map>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../../include/qfe.h"

namespace qfa {

using cplx = std::complex<double>;
using Mat4 = std::array<std::array<cplx, 4>, 4>;

std::string fmt_complex(cplx v);  // helpers.cc:17-26 formatting ("{:+f}", "{:+f}j", "{:+f} {:+f}i")
std::string join(const std::vector<std::string>& parts, const std::string& sep);

// ---------------------------------------------------------------------------
// QubitBasis (qubit_basis.h:17-64)
// ---------------------------------------------------------------------------
class QubitBasis {
  public:
    explicit QubitBasis(uint64_t n = 0) : state_(n) {}
    bool get_bit(size_t pos) const { return (state_ >> pos) & 1ull; }
    void set_bit(size_t pos, bool val) { state_ = val ? (state_ | (1ull << pos)) : (state_ & ~(1ull << pos)); }
    void flip_bit(size_t pos) { state_ ^= 1ull << pos; }
    uint64_t index() const { return state_; }
    std::string str(size_t nqubit) const;  // qubit_basis.cc:3-14
  private:
    uint64_t state_;
};

// ---------------------------------------------------------------------------
// Gate (gate.h:52-154, make_gate.cc, gate.cc)
// ---------------------------------------------------------------------------
class Gate {
  public:
    Gate(std::string label, size_t target, size_t control, const Mat4& m,
         std::optional<std::pair<double, bool>> parameter = std::nullopt);
    size_t target() const { return target_; }
    size_t control() const { return control_; }
    const Mat4& matrix() const { return m_; }
    const std::string& gate_id() const { return label_; }
    size_t nqubits() const { return target_ == control_ ? 1 : 2; }
    bool has_parameter() const { return parameter_.has_value(); }
    std::optional<double> parameter() const {
        return parameter_ ? std::optional<double>(parameter_->first) : std::nullopt;
    }
    bool minus_parameter_on_adjoint() const { return parameter_ ? parameter_->second : false; }
    Gate adjoint() const;        // gate.cc:135-179
    std::string str() const;     // gate.cc:112-117
    std::string repr() const;    // gate.cc:119-130
    qfe_gate to_abi() const;
    // Pauli letter of this gate ('X','Y','Z','I') or 0 when it is not one
    char pauli_letter() const;

  private:
    std::string label_;
    size_t target_, control_;
    std::optional<std::pair<double, bool>> parameter_;
    Mat4 m_;
};

bool operator==(const Gate& a, const Gate& b);  // gate.cc:183-199
bool operator!=(const Gate& a, const Gate& b);
bool operator<(const Gate& a, const Gate& b);   // gate.cc:203-219

double normalize_angle(double angle, double period);  // make_gate.cc:7-16
Gate make_gate(const std::string& type, size_t target, size_t control, double parameter = 0.0);  // make_gate.cc:18-235
Gate make_control_gate(size_t control, const Gate& U);  // make_gate.cc:237-256

// ---------------------------------------------------------------------------
// Circuit (circuit.h:13-102, circuit.cc)
//
// Besides the gate list a circuit may carry "exponential segments": runs of
// gates [begin, end) known to equal exp(i angle P) for a Hermitian Pauli P
// (they were produced by exponentiate_pauli_string).  The engine applies such
// a run as one pass; the gate list itself is exactly the reference's.
// ---------------------------------------------------------------------------
struct ExpSegment {
    size_t begin, end;
    uint64_t x, z;
    cplx phase;
    double angle;
};

class Circuit {
  public:
    Circuit() = default;
    void add_gate(const Gate& g) { gates_.push_back(g); }
    void add_circuit(const Circuit& c);
    void insert_gate(size_t pos, const Gate& g);
    void remove_gate(size_t pos);
    void swap_gates(size_t p1, size_t p2);
    void insert_circuit(size_t pos, const Circuit& c);
    void remove_gates(size_t p1, size_t p2);
    void replace_gate(size_t pos, const Gate& g);
    const std::vector<Gate>& gates() const { return gates_; }
    size_t size() const { return gates_.size(); }
    Circuit adjoint() const;                               // circuit.cc:98-105
    void set_parameters(const std::vector<double>& params);  // circuit.cc:10-27
    std::vector<double> get_parameters() const;
    void set_parameter(size_t pos, double param);
    cplx canonicalize_pauli_circuit();                     // circuit.cc:107-169
    int get_num_cnots() const;                             // circuit.cc:171-183
    std::string str() const;                               // circuit.cc:185-192
    size_t num_qubits() const;                             // circuit.cc:194-200
    bool is_pauli() const;                                 // circuit.cc:222-230
    const std::vector<ExpSegment>& segments() const { return segs_; }
    void add_segment(const ExpSegment& s) { segs_.push_back(s); }

  private:
    std::vector<Gate> gates_;
    std::vector<ExpSegment> segs_;
};

bool operator==(const Circuit& a, const Circuit& b);
bool operator<(const Circuit& a, const Circuit& b);

// Mask form of the operator a Pauli circuit applies (gates applied first to
// last): P|j> = phase (-1)^popcount(j & z) |j ^ x>.  Returns false when a gate
// is not X, Y, Z or I.
bool pauli_masks(const Circuit& c, uint64_t& x, uint64_t& z, cplx& phase);

// ---------------------------------------------------------------------------
// QubitOperator (qubit_operator.h:12-72, qubit_operator.cc)
// ---------------------------------------------------------------------------
class QubitOperator {
  public:
    using Term = std::pair<cplx, Circuit>;
    void add_term(cplx c, const Circuit& circ) { terms_.emplace_back(c, circ); }
    void add_op(const QubitOperator& o) { terms_.insert(terms_.end(), o.terms_.begin(), o.terms_.end()); }
    void set_coeffs(const std::vector<cplx>& c);
    void mult_coeffs(cplx m);
    const std::vector<Term>& terms() const { return terms_; }
    std::vector<Term>& mutable_terms() { return terms_; }
    void order_terms();                                   // qubit_operator.cc:51-72
    void canonical_order();                               // qubit_operator.cc:74-78
    void simplify(bool combine_like_terms = true);        // qubit_operator.cc:80-103
    void operator_product(const QubitOperator& r, bool pre_simplify = true, bool post_simplify = true);
    bool check_op_equivalence(QubitOperator other, bool reorder);
    std::string str() const;
    size_t num_qubits() const;

  private:
    std::vector<Term> terms_;
};

// ---------------------------------------------------------------------------
// SQOperator (sq_operator.h, sq_operator.cc)
// ---------------------------------------------------------------------------
class SQOperator {
  public:
    using Term = std::tuple<cplx, std::vector<size_t>, std::vector<size_t>>;
    void add_term(cplx c, const std::vector<size_t>& cre, const std::vector<size_t>& ann) {
        terms_.emplace_back(c, cre, ann);
    }
    void add_op(const SQOperator& o) { terms_.insert(terms_.end(), o.terms_.begin(), o.terms_.end()); }
    void set_coeffs(const std::vector<cplx>& c);
    void mult_coeffs(cplx m);
    const std::vector<Term>& terms() const { return terms_; }
    void canonical_order();                                  // sq_operator.cc:49-65
    void simplify();                                         // sq_operator.cc:67-86
    QubitOperator jw_transform(bool qubit_excitation = false);  // sq_operator.cc:138-169
    std::string str() const;

  private:
    std::vector<Term> terms_;
};

// ---------------------------------------------------------------------------
// QubitOpPool (qubit_op_pool.h, qubit_op_pool.cc)
// ---------------------------------------------------------------------------
class QubitOpPool {
  public:
    using Term = std::pair<cplx, QubitOperator>;
    void add_term(cplx c, const QubitOperator& op) { terms_.emplace_back(c, op); }
    void set_coeffs(const std::vector<cplx>& c);
    void set_op_coeffs(const std::vector<cplx>& c);
    void set_terms(const std::vector<Term>& t) { terms_.insert(terms_.end(), t.begin(), t.end()); }
    const std::vector<Term>& terms() const { return terms_; }
    void square(bool upper_triangle_only);
    void join_op_from_right_lazy(const QubitOperator& q);
    void join_op_from_right(const QubitOperator& q);
    void join_op_from_left(const QubitOperator& q);
    void join_as_commutator(const QubitOperator& q);
    void fill_pool(const std::string& pool_type, size_t nqb);
    std::string str() const;

  private:
    std::vector<Term> terms_;
};

// ---------------------------------------------------------------------------
// SQOpPool (sq_op_pool.h, sq_op_pool.cc)
// ---------------------------------------------------------------------------
class SQOpPool {
  public:
    using Term = std::pair<cplx, SQOperator>;
    void add_term(cplx c, const SQOperator& op) { terms_.emplace_back(c, op); }
    void set_coeffs(const std::vector<cplx>& c);
    const std::vector<Term>& terms() const { return terms_; }
    std::vector<Term>& mutable_terms() { return terms_; }
    void set_orb_spaces(const std::vector<int>& ref, const std::vector<size_t>& orb_irreps_to_int = {});
    QubitOpPool get_qubit_op_pool();
    QubitOperator get_qubit_operator(const std::string& order_type, bool combine_like_terms = true,
                                     bool qubit_excitations = false);
    void fill_pool(const std::string& pool_type);
    std::string str() const;

  private:
    uint64_t ref_int_ = 0;
    int n_spinorb_ = 0;
    int n_occ_alpha_ = 0, n_occ_beta_ = 0, n_vir_alpha_ = 0, n_vir_beta_ = 0;
    std::vector<size_t> orb_irreps_to_int_;
    std::vector<Term> terms_;
    void fill_excitations(int max_rank);
    void fill_gsd();
    void fill_sa_sd();
};

size_t find_irrep(const std::vector<size_t>& orb_irreps_to_int, const std::vector<size_t>& spinorb_indices);

// ---------------------------------------------------------------------------
// exponentiate / trotterize (utils/exponentiate.py:9-87,
// utils/trotterization.py:10-57) -- native, with exponential segments
// ---------------------------------------------------------------------------
std::pair<Circuit, cplx> exponentiate_pauli_string(cplx coefficient, const Circuit& term, bool use_cRz = false,
                                                   long ancilla_idx = -1, bool use_open_cRz = false);
std::pair<Circuit, cplx> trotterize(const QubitOperator& op, cplx factor = 1.0, int trotter_number = 1,
                                    int trotter_order = 1);
std::pair<Circuit, cplx> trotterize_w_cRz(const QubitOperator& op, size_t ancilla, cplx factor, bool use_open_cRz,
                                          int trotter_number, int trotter_order);

// ---------------------------------------------------------------------------
// Computer (computer.h:22-185): the state lives in HBM behind a qfe_state.
// ---------------------------------------------------------------------------
class Computer {
  public:
    explicit Computer(int nqubit, double print_threshold = 1.0e-6);
    Computer(const Computer& other);
    Computer& operator=(const Computer&) = delete;
    ~Computer();

    void apply_circuit(const Circuit& c);
    void apply_gate(const Gate& g);
    void apply_circuit_safe(const Circuit& c);
    void apply_gate_safe(const Gate& g);
    void apply_operator(const QubitOperator& op);
    void apply_constant(cplx a);
    cplx direct_op_exp_val(const QubitOperator& op);
    std::vector<cplx> direct_oppl_exp_val(const QubitOpPool& pool);
    std::vector<cplx> direct_idxd_oppl_exp_val(const QubitOpPool& pool, const std::vector<int>& idxs);
    std::vector<cplx> direct_oppl_exp_val_w_mults(const QubitOpPool& pool, const std::vector<cplx>& mults);
    cplx direct_circ_exp_val(const Circuit& c);
    cplx direct_pauli_circ_exp_val(const Circuit& c);
    std::vector<double> measure_circuit(const Circuit& c, size_t n_measurements);
    std::vector<std::vector<int>> measure_z_readouts_fast(size_t na, size_t nb, size_t n_measurements);
    std::vector<std::vector<int>> measure_readouts(const Circuit& c, size_t n_measurements);
    double perfect_measure_circuit(const Circuit& c);

    std::vector<cplx> get_coeff_vec() const;
    void get_coeff_into(cplx* out) const;
    void set_coeff_vec(const cplx* data, size_t n);
    cplx coeff(const QubitBasis& b) const;
    void set_state(const std::vector<std::pair<QubitBasis, cplx>>& state);
    void null_state();
    void reset();
    size_t get_nqubit() const { return (size_t)n_; }
    size_t get_nbasis() const { return 1ull << n_; }
    std::string str() const;
    qfe_state handle() const { return s_; }
    double print_threshold() const { return thr_; }

  private:
    int n_;
    double thr_;
    qfe_state s_ = nullptr;
    // Lower a circuit to an op stream (gates, and exponential segments as one op).
    void lower(const Circuit& c, std::vector<qfe_op>& ops) const;
    void apply_ops(const std::vector<qfe_op>& ops);
    // copy of the state in the measurement basis of c (computer.cc:141-165)
    Computer basis_rotated(const Circuit& c) const;
    std::vector<uint64_t> sample(size_t n_measurements) const;
};

bool operator==(const Computer& a, const Computer& b);  // computer.cc:968-973
cplx dot(const Computer& a, const Computer& b);        // computer.cc:975-980

// Throws std::runtime_error with the library message when rc != QFE_OK.
void check(int rc, const char* what);

// Compile a QubitOperator whose terms are all Pauli circuits; false if not.
bool compile_pauli_sum(const QubitOperator& op, std::vector<qfe_pauli>& out, cplx mult = 1.0);

// The JW-mapped molecular Hamiltonian of norb spatial orbitals as a linear
// map from integrals to Pauli coefficients: term k is M(x_k, z_k) with
// M(x, z)|j> = (-1)^{popcount(j & z)} |j ^ x>, and its coefficient is
// sum over e in [offsets[k], offsets[k+1]) of value[e] * inputs[index[e]].
// Inputs (ninputs = 1 + norb(norb+1)/2 + P(P+1)/2, P = norb(norb+1)/2):
// [constant, h_pq for p >= q, (pq|rs) for pq >= rs], pair index p(p+1)/2 + q.
// The spin-orbital expansion is that of molecule_adapters.py:183-210.
struct HamiltonianMap {
    size_t norb = 0;
    size_t ninputs = 0;
    std::vector<uint64_t> x, z;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> index;
    std::vector<double> value;
};
HamiltonianMap molecular_hamiltonian_map(size_t norb);
// Pack (constant, h[norb*norb], g[norb^4], chemist notation) into map inputs.
std::vector<double> pack_integrals(size_t norb, double constant, const double* h, const double* g);
// The Hamiltonian of a map at given inputs, as canonical Pauli circuits.
QubitOperator map_to_operator(const HamiltonianMap& m, const std::vector<double>& inputs);

}  // namespace qfa
