// Python module qforte_amd._core: the qforte API surface of the reference's
// pybind11 module (bindings.cc:23-350) over the native host core, plus the
// compiled-program and serving entry points that sit beside it.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <memory>
#include <unordered_map>

#include "core.h"

namespace py = pybind11;
using namespace pybind11::literals;
using namespace qfa;

namespace {

const char* kOneQubit[] = {"X", "Y", "Z", "H", "R", "Rx", "Ry", "Rz", "V", "S", "T", "I"};
const char* kTwoQubitNoParam[] = {"SWAP", "cV", "CNOT", "cX", "aCNOT", "acX", "cY", "cZ"};
const char* kOneQubitNoParam[] = {"X", "Y", "Z", "H", "V", "S", "T", "I"};

template <size_t N>
bool in(const std::string& s, const char* (&list)[N]) {
    return std::any_of(std::begin(list), std::end(list), [&](const char* t) { return s == t; });
}

// Basis indices drawn from |psi|^2 on the device (qfe_sample), seeded.
py::array_t<uint64_t> sample_indices(const Computer& c, size_t n, uint64_t seed) {
    py::array_t<uint64_t> out(n);
    {
        py::gil_scoped_release rel;
        check(qfe_sample(c.handle(), n, seed, out.mutable_data()), "Computer.sample");
    }
    return out;
}

py::array_t<std::complex<double>> coeff_array(const Computer& c) {
    py::array_t<std::complex<double>> out(c.get_nbasis());
    {
        py::gil_scoped_release rel;
        c.get_coeff_into(out.mutable_data());
    }
    return out;
}

void set_coeffs_from(Computer& c, py::object obj) {
    auto arr = py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>::ensure(obj);
    if (!arr) throw std::invalid_argument("set_coeff_vec: expected a sequence of complex numbers");
    py::gil_scoped_release rel;
    c.set_coeff_vec(arr.data(), (size_t)arr.size());
}

// ---- compiled UCC program -----------------------------------------------------------
class UCCProgram {
  public:
    UCCProgram(int nqubit, uint64_t ref_index, const std::vector<QubitOperator>& generators,
               const std::vector<int>& param_index, const std::vector<double>& scale, size_t nparam,
               const QubitOperator& hamiltonian)
        : n_(nqubit), nparam_(nparam) {
        if (generators.size() != param_index.size() || generators.size() != scale.size())
            throw std::invalid_argument("UCCProgram: generators, param_index and scale must have equal length");
        std::vector<qfe_pauli> terms, one;
        std::vector<uint64_t> off{0};
        for (const auto& g : generators) {
            if (!compile_pauli_sum(g, one)) throw std::invalid_argument("UCCProgram: generators must be sums of Pauli strings");
            terms.insert(terms.end(), one.begin(), one.end());
            off.push_back(terms.size());
        }
        std::vector<qfe_pauli> ham;
        if (!compile_pauli_sum(hamiltonian, ham)) throw std::invalid_argument("UCCProgram: the Hamiltonian must be a sum of Pauli strings");
        std::vector<int32_t> pi(param_index.begin(), param_index.end());
        check(qfe_program_create(nqubit, ref_index, terms.data(), off.data(), generators.size(), pi.data(), scale.data(),
                                 nparam, ham.data(), ham.size(), &p_),
              "UCCProgram");
    }
    ~UCCProgram() {
        if (p_) qfe_program_destroy(p_);
    }
    UCCProgram(const UCCProgram&) = delete;
    UCCProgram& operator=(const UCCProgram&) = delete;

    double energy(const std::vector<double>& t) {
        need(t);
        double e = 0;
        py::gil_scoped_release rel;
        check(qfe_program_energy(p_, t.data(), &e), "UCCProgram.energy");
        return e;
    }
    double energy_graph(const std::vector<double>& t) {
        need(t);
        double e = 0;
        py::gil_scoped_release rel;
        check(qfe_program_energy_graph(p_, t.data(), &e), "UCCProgram.energy_graph");
        return e;
    }
    std::pair<double, std::vector<double>> gradient(const std::vector<double>& t) {
        need(t);
        std::vector<double> g(nparam_);
        double e = 0;
        py::gil_scoped_release rel;
        check(qfe_program_gradient(p_, t.data(), g.data(), &e), "UCCProgram.gradient");
        return {e, g};
    }
    void prepare(const std::vector<double>& t, Computer& out) {
        need(t);
        py::gil_scoped_release rel;
        check(qfe_program_prepare(p_, t.data(), out.handle()), "UCCProgram.prepare");
    }
    std::vector<cplx> residuals(const std::vector<double>& t, const std::vector<uint64_t>& idx) {
        need(t);
        std::vector<cplx> out(idx.size());
        py::gil_scoped_release rel;
        check(qfe_program_residuals(p_, t.data(), idx.data(), idx.size(), reinterpret_cast<qfe_c64*>(out.data())),
              "UCCProgram.residuals");
        return out;
    }
    std::vector<double> pool_gradients(const std::vector<double>& t, const std::vector<QubitOperator>& ops,
                                       size_t batch) {
        need(t);
        std::vector<qfe_pauli> terms, one;
        std::vector<uint64_t> off{0};
        for (const auto& g : ops) {
            if (!compile_pauli_sum(g, one)) throw std::invalid_argument("pool_gradients: operators must be Pauli sums");
            terms.insert(terms.end(), one.begin(), one.end());
            off.push_back(terms.size());
        }
        std::vector<double> out(ops.size());
        py::gil_scoped_release rel;
        check(qfe_program_pool_gradients(p_, t.data(), terms.data(), off.data(), ops.size(), batch, out.data()),
              "UCCProgram.pool_gradients");
        return out;
    }
    int nqubit() const { return n_; }
    size_t nparam() const { return nparam_; }
    qfe_program handle() const { return p_; }

  private:
    void need(const std::vector<double>& t) const {
        if (t.size() != nparam_) throw std::invalid_argument("UCCProgram: expected " + std::to_string(nparam_) + " amplitudes");
    }
    int n_;
    size_t nparam_;
    qfe_program p_ = nullptr;
};

// Determinant-sector template: one size class of UCC energy evaluations.
class SectorTemplate {
  public:
    SectorTemplate(int nqubit, uint64_t ref_index, const std::vector<QubitOperator>& generators,
                   const std::vector<int>& param_index, const std::vector<double>& scale, size_t nparam,
                   const QubitOperator& hamiltonian)
        : n_(nqubit), nparam_(nparam) {
        std::vector<qfe_pauli> gen, ham;
        std::vector<uint64_t> off;
        std::vector<int32_t> pi;
        compile_generators(generators, param_index, scale, gen, off, pi);
        if (!compile_pauli_sum(hamiltonian, ham))
            throw std::invalid_argument("SectorTemplate: the Hamiltonian must be a sum of Pauli strings");
        check(qfe_template_create(nqubit, ref_index, gen.data(), off.data(), generators.size(), pi.data(), scale.data(),
                                  nparam, ham.data(), ham.size(), nullptr, nullptr, nullptr, 0, &t_),
              "SectorTemplate");
    }
    SectorTemplate(int nqubit, uint64_t ref_index, const std::vector<QubitOperator>& generators,
                   const std::vector<int>& param_index, const std::vector<double>& scale, size_t nparam,
                   const HamiltonianMap& map)
        : n_(nqubit), nparam_(nparam), ninputs_(map.ninputs) {
        std::vector<qfe_pauli> gen, ham(map.x.size());
        std::vector<uint64_t> off;
        std::vector<int32_t> pi;
        compile_generators(generators, param_index, scale, gen, off, pi);
        for (size_t k = 0; k < ham.size(); ++k) ham[k] = qfe_pauli{map.x[k], map.z[k], {1.0, 0.0}, {0.0, 0.0}};
        std::vector<qfe_c64> val(map.value.size());
        for (size_t e = 0; e < val.size(); ++e) val[e] = qfe_c64{map.value[e], 0.0};
        check(qfe_template_create(nqubit, ref_index, gen.data(), off.data(), generators.size(), pi.data(), scale.data(),
                                  nparam, ham.data(), ham.size(), map.offsets.data(), map.index.data(), val.data(),
                                  map.ninputs, &t_),
              "SectorTemplate");
    }
    ~SectorTemplate() {
        if (t_) qfe_template_destroy(t_);
    }
    SectorTemplate(const SectorTemplate&) = delete;
    SectorTemplate& operator=(const SectorTemplate&) = delete;

    double energy(const std::vector<double>& t, const std::vector<double>& inputs) {
        need(t, inputs);
        double e = 0.0;
        py::gil_scoped_release rel;
        check(qfe_template_energy(t_, t.data(), inputs.data(), &e), "SectorTemplate.energy");
        return e;
    }
    void prepare(const std::vector<double>& t, Computer& out) {
        if (t.size() != nparam_) throw std::invalid_argument("SectorTemplate: wrong number of amplitudes");
        py::gil_scoped_release rel;
        check(qfe_template_prepare(t_, t.data(), out.handle()), "SectorTemplate.prepare");
    }
    std::pair<double, std::vector<double>> gradient(const std::vector<double>& t, const std::vector<double>& inputs) {
        need(t, inputs);
        std::vector<double> g(nparam_);
        double e = 0.0;
        py::gil_scoped_release rel;
        check(qfe_template_gradient(t_, t.data(), inputs.data(), g.data(), &e), "SectorTemplate.gradient");
        return {e, g};
    }
    std::vector<cplx> residuals(const std::vector<double>& t, const std::vector<uint64_t>& idx,
                                const std::vector<double>& inputs) {
        need(t, inputs);
        std::vector<cplx> out(idx.size());
        py::gil_scoped_release rel;
        check(qfe_template_residuals(t_, t.data(), inputs.data(), idx.data(), idx.size(),
                                     reinterpret_cast<qfe_c64*>(out.data())),
              "SectorTemplate.residuals");
        return out;
    }
    std::vector<double> pool_gradients(const std::vector<double>& t, const std::vector<QubitOperator>& ops, size_t batch,
                                       const std::vector<double>& inputs) {
        need(t, inputs);
        std::vector<qfe_pauli> terms, one;
        std::vector<uint64_t> off{0};
        for (const auto& op : ops) {
            if (!compile_pauli_sum(op, one)) throw std::invalid_argument("SectorTemplate.pool_gradients: operators must be Pauli sums");
            terms.insert(terms.end(), one.begin(), one.end());
            off.push_back(terms.size());
        }
        std::vector<double> out(ops.size());
        py::gil_scoped_release rel;
        check(qfe_template_pool_gradients(t_, t.data(), inputs.data(), terms.data(), off.data(), ops.size(), batch,
                                          out.data()),
              "SectorTemplate.pool_gradients");
        return out;
    }
    uint64_t dim() const {
        uint64_t d = 0;
        qfe_template_info(t_, &d, nullptr);
        return d;
    }
    uint64_t request_bytes() const {
        uint64_t b = 0;
        qfe_template_info(t_, nullptr, &b);
        return b;
    }
    // {nodes, kernel nodes, kernel nodes with checked arguments, memcpy nodes with checked parameters}
    std::vector<uint64_t> graph_audit(bool gradient) {
        std::vector<uint64_t> out(4);
        {
            py::gil_scoped_release rel;
            check(qfe_template_graph_audit(t_, gradient ? 1 : 0, out.data()), "SectorTemplate.graph_audit");
        }
        return out;
    }
    int nqubit() const { return n_; }
    size_t nparam() const { return nparam_; }
    size_t ninputs() const { return ninputs_; }
    qfe_template handle() const { return t_; }
    void need(const std::vector<double>& t, const std::vector<double>& inputs) const {
        if (t.size() != nparam_) throw std::invalid_argument("SectorTemplate: expected " + std::to_string(nparam_) + " amplitudes");
        if (inputs.size() != ninputs_)
            throw std::invalid_argument("SectorTemplate: expected " + std::to_string(ninputs_) + " Hamiltonian inputs");
    }

  private:
    static void compile_generators(const std::vector<QubitOperator>& generators, const std::vector<int>& param_index,
                                   const std::vector<double>& scale, std::vector<qfe_pauli>& terms,
                                   std::vector<uint64_t>& off, std::vector<int32_t>& pi) {
        if (generators.size() != param_index.size() || generators.size() != scale.size())
            throw std::invalid_argument("SectorTemplate: generators, param_index and scale must have equal length");
        std::vector<qfe_pauli> one;
        off.assign(1, 0);
        for (const auto& g : generators) {
            if (!compile_pauli_sum(g, one)) throw std::invalid_argument("SectorTemplate: generators must be sums of Pauli strings");
            terms.insert(terms.end(), one.begin(), one.end());
            off.push_back(terms.size());
        }
        pi.assign(param_index.begin(), param_index.end());
    }
    int n_;
    size_t nparam_;
    size_t ninputs_ = 0;
    qfe_template t_ = nullptr;
};

// Continuous-batching serving engine over sector templates.
class BatchEngine {
  public:
    BatchEngine(uint64_t budget, int max_active) { check(qfe_batch_create(budget, max_active, &b_), "BatchEngine"); }
    ~BatchEngine() {
        if (b_) qfe_batch_destroy(b_);
    }
    uint64_t submit(std::shared_ptr<SectorTemplate> t, const std::vector<double>& amps, const std::vector<double>& inputs,
                    bool gradient) {
        t->need(amps, inputs);
        uint64_t id = 0;
        if (gradient)
            check(qfe_batch_submit_gradient(b_, t->handle(), amps.data(), inputs.data(), &id), "BatchEngine.submit");
        else
            check(qfe_batch_submit(b_, t->handle(), amps.data(), inputs.data(), &id), "BatchEngine.submit");
        keep_.emplace(id, std::move(t));
        return id;
    }
    std::vector<double> take_gradient(uint64_t id, size_t nparam) {
        std::vector<double> g(nparam);
        check(qfe_batch_take_gradient(b_, id, g.data(), nparam), "BatchEngine.take_gradient");
        return g;
    }
    std::vector<std::pair<uint64_t, double>> step(size_t cap) { return run(cap, true); }
    std::vector<std::pair<uint64_t, double>> poll(size_t cap) { return run(cap, false); }
    std::pair<uint64_t, uint64_t> pending() {
        uint64_t q = 0, a = 0;
        check(qfe_batch_pending(b_, &q, &a), "BatchEngine.pending");
        return {q, a};
    }

  private:
    std::vector<std::pair<uint64_t, double>> run(size_t cap, bool block) {
        std::vector<uint64_t> ids(cap);
        std::vector<double> es(cap);
        uint64_t nd = 0;
        {
            py::gil_scoped_release rel;
            if (block)
                check(qfe_batch_step(b_, ids.data(), es.data(), cap, &nd), "BatchEngine.step");
            else
                check(qfe_batch_poll(b_, ids.data(), es.data(), cap, &nd), "BatchEngine.poll");
        }
        std::vector<std::pair<uint64_t, double>> out;
        for (uint64_t i = 0; i < nd; ++i) {
            out.emplace_back(ids[i], es[i]);
            keep_.erase(ids[i]);  // the engine no longer reads the request's template
        }
        return out;
    }
    qfe_batch b_ = nullptr;
    // the template of every queued or running request, released when its result is returned
    std::unordered_map<uint64_t, std::shared_ptr<SectorTemplate>> keep_;
};

// HIP-event timer on the engine stream.
class DeviceTimer {
  public:
    DeviceTimer() {
        check(qfe_event_create(&a_), "DeviceTimer");
        check(qfe_event_create(&b_), "DeviceTimer");
    }
    ~DeviceTimer() {
        qfe_event_destroy(a_);
        qfe_event_destroy(b_);
    }
    void start() { check(qfe_event_record(a_), "DeviceTimer.start"); }
    float stop() {
        check(qfe_event_record(b_), "DeviceTimer.stop");
        float ms = 0.0f;
        py::gil_scoped_release rel;
        check(qfe_event_elapsed_ms(a_, b_, &ms), "DeviceTimer.stop");
        return ms;
    }

  private:
    void* a_ = nullptr;
    void* b_ = nullptr;
};

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "qforte_amd native core: qforte data model over the MI355X state-vector engine";

    py::class_<Circuit, std::shared_ptr<Circuit>>(m, "Circuit")
        .def(py::init<>())
        .def(py::init<const Circuit&>())
        .def("add", &Circuit::add_gate)
        .def("add", &Circuit::add_circuit)
        .def("add_gate", &Circuit::add_gate)
        .def("add_circuit", &Circuit::add_circuit)
        .def("insert_gate", &Circuit::insert_gate)
        .def("remove_gate", &Circuit::remove_gate, "pos"_a)
        .def("swap_gates", &Circuit::swap_gates, "pos1"_a, "pos2"_a)
        .def("insert_circuit", &Circuit::insert_circuit, "pos"_a, "circ"_a)
        .def("remove_gates", &Circuit::remove_gates, "pos1"_a, "pos2"_a)
        .def("replace_gate", &Circuit::replace_gate, "pos"_a, "gate"_a)
        .def("gates", &Circuit::gates)
        .def("gate", [](const Circuit& c, size_t pos) { return c.gates().at(pos); })
        .def("size", &Circuit::size)
        .def("adjoint", &Circuit::adjoint)
        .def("canonicalize_pauli_circuit", &Circuit::canonicalize_pauli_circuit)
        .def("set_parameters", &Circuit::set_parameters)
        .def("set_parameter", &Circuit::set_parameter)
        .def("get_parameters", &Circuit::get_parameters)
        .def("get_num_cnots", &Circuit::get_num_cnots)
        .def("is_pauli", &Circuit::is_pauli)
        .def("num_qubits", &Circuit::num_qubits)
        .def("num_exponential_segments", [](const Circuit& c) { return c.segments().size(); })
        .def("str", &Circuit::str)
        .def("__str__", &Circuit::str)
        .def("__repr__", &Circuit::str)
        .def("__len__", &Circuit::size)
        .def("__eq__", [](const Circuit& a, const Circuit& b) { return a == b; });

    py::class_<SQOperator>(m, "SQOperator")
        .def(py::init<>())
        .def("add", &SQOperator::add_term)
        .def("add", &SQOperator::add_op)
        .def("add_term", &SQOperator::add_term)
        .def("add_op", &SQOperator::add_op)
        .def("set_coeffs", &SQOperator::set_coeffs)
        .def("mult_coeffs", &SQOperator::mult_coeffs)
        .def("terms", &SQOperator::terms)
        .def("canonical_order", &SQOperator::canonical_order)
        .def("simplify", &SQOperator::simplify)
        .def("jw_transform", &SQOperator::jw_transform, py::arg("qubit_excitation") = false)
        .def("str", &SQOperator::str)
        .def("__str__", &SQOperator::str)
        .def("__repr__", &SQOperator::str);

    py::class_<SQOpPool>(m, "SQOpPool")
        .def(py::init<>())
        .def("add", &SQOpPool::add_term)
        .def("add_term", &SQOpPool::add_term)
        .def("set_coeffs", &SQOpPool::set_coeffs)
        .def("terms", &SQOpPool::terms)
        .def("set_orb_spaces", &SQOpPool::set_orb_spaces, py::arg("ref"), py::arg("orb_irreps_to_int") = std::vector<size_t>{})
        .def("get_qubit_op_pool", &SQOpPool::get_qubit_op_pool)
        .def("get_qubit_operator", &SQOpPool::get_qubit_operator, py::arg("order_type"),
             py::arg("combine_like_terms") = true, py::arg("qubit_excitations") = false)
        .def("fill_pool", &SQOpPool::fill_pool)
        .def("str", &SQOpPool::str)
        .def("__getitem__", [](const SQOpPool& p, size_t i) { return p.terms().at(i); })
        .def("__iter__", [](const SQOpPool& p) { return py::make_iterator(p.terms()); }, py::keep_alive<0, 1>())
        .def("__len__", [](const SQOpPool& p) { return p.terms().size(); })
        .def("__str__", &SQOpPool::str)
        .def("__repr__", &SQOpPool::str);

    py::class_<QubitOperator>(m, "QubitOperator")
        .def(py::init<>())
        .def("add", &QubitOperator::add_term)
        .def("add", &QubitOperator::add_op)
        .def("add_term", &QubitOperator::add_term)
        .def("add_op", &QubitOperator::add_op)
        .def("set_coeffs", &QubitOperator::set_coeffs)
        .def("mult_coeffs", &QubitOperator::mult_coeffs)
        .def("terms", &QubitOperator::terms)
        .def("order_terms", &QubitOperator::order_terms)
        .def("canonical_order", &QubitOperator::canonical_order)
        .def("simplify", &QubitOperator::simplify, py::arg("combine_like_terms") = true)
        .def("operator_product", &QubitOperator::operator_product, py::arg("rqo"), py::arg("pre_simplify") = true,
             py::arg("post_simplify") = true)
        .def("check_op_equivalence", &QubitOperator::check_op_equivalence)
        .def("num_qubits", &QubitOperator::num_qubits)
        .def("str", &QubitOperator::str)
        .def("__iter__", [](const QubitOperator& o) { return py::make_iterator(o.terms()); }, py::keep_alive<0, 1>())
        .def("__str__", &QubitOperator::str)
        .def("__repr__", &QubitOperator::str);

    py::class_<QubitOpPool>(m, "QubitOpPool")
        .def(py::init<>())
        .def("add", &QubitOpPool::add_term)
        .def("add_term", &QubitOpPool::add_term)
        .def("set_coeffs", &QubitOpPool::set_coeffs)
        .def("set_op_coeffs", &QubitOpPool::set_op_coeffs)
        .def("set_terms", &QubitOpPool::set_terms)
        .def("terms", &QubitOpPool::terms)
        .def("join_op_from_right_lazy", &QubitOpPool::join_op_from_right_lazy)
        .def("join_op_from_right", &QubitOpPool::join_op_from_right)
        .def("join_op_from_left", &QubitOpPool::join_op_from_left)
        .def("join_as_commutator", &QubitOpPool::join_as_commutator)
        .def("square", &QubitOpPool::square)
        .def("fill_pool", &QubitOpPool::fill_pool)
        .def("str", &QubitOpPool::str)
        .def("__iter__", [](const QubitOpPool& p) { return py::make_iterator(p.terms()); }, py::keep_alive<0, 1>())
        .def("__str__", &QubitOpPool::str)
        .def("__repr__", &QubitOpPool::str);

    py::class_<QubitBasis>(m, "QubitBasis")
        .def(py::init<uint64_t>(), "n"_a = 0)
        .def("str", &QubitBasis::str)
        .def("__str__", [](const QubitBasis& b) { return b.str(64); })
        .def("__repr__", [](const QubitBasis& b) { return b.str(64); })
        .def("flip_bit", &QubitBasis::flip_bit)
        .def("set_bit", &QubitBasis::set_bit)
        .def("index", &QubitBasis::index)
        .def("get_bit", &QubitBasis::get_bit);

    py::class_<Computer, std::shared_ptr<Computer>>(m, "Computer")
        .def(py::init<int, double>(), "nqubits"_a, "print_threshold"_a = 1.0e-6)
        .def(py::init<const Computer&>())
        .def("apply_circuit_safe", &Computer::apply_circuit_safe)
        .def("apply_operator", &Computer::apply_operator, py::call_guard<py::gil_scoped_release>())
        .def("apply_circuit", &Computer::apply_circuit, py::call_guard<py::gil_scoped_release>())
        .def("apply_gate_safe", &Computer::apply_gate_safe)
        .def("apply_gate", &Computer::apply_gate, py::call_guard<py::gil_scoped_release>())
        .def("apply_constant", &Computer::apply_constant, py::call_guard<py::gil_scoped_release>())
        .def("measure_circuit", &Computer::measure_circuit)
        .def("measure_z_readouts_fast", &Computer::measure_z_readouts_fast)
        .def("measure_readouts", &Computer::measure_readouts)
        .def("perfect_measure_circuit", &Computer::perfect_measure_circuit)
        .def("direct_oppl_exp_val", &Computer::direct_oppl_exp_val, py::call_guard<py::gil_scoped_release>())
        .def("direct_idxd_oppl_exp_val", &Computer::direct_idxd_oppl_exp_val, py::call_guard<py::gil_scoped_release>())
        .def("direct_oppl_exp_val_w_mults", &Computer::direct_oppl_exp_val_w_mults)
        .def("direct_op_exp_val", &Computer::direct_op_exp_val, py::call_guard<py::gil_scoped_release>())
        .def("direct_circ_exp_val", &Computer::direct_circ_exp_val)
        .def("direct_pauli_circ_exp_val", &Computer::direct_pauli_circ_exp_val)
        .def("coeff", &Computer::coeff)
        .def("get_coeff_vec", &coeff_array)
        .def("sample", &sample_indices, "n_measurements"_a, "seed"_a)
        .def("get_nqubit", &Computer::get_nqubit)
        .def("set_coeff_vec", &set_coeffs_from)
        .def("set_coeff_vec_from_numpy", &set_coeffs_from)
        .def("set_state", &Computer::set_state)
        .def("null_state", &Computer::null_state)
        .def("reset", &Computer::reset, "Reset the quantum computer to the state |0>")
        .def("get_timings", [](const Computer&) { return std::vector<std::pair<std::string, double>>{}; })
        .def("clear_timings", [](Computer&) {})
        .def("str", &Computer::str)
        .def("__str__", &Computer::str)
        .def("__repr__", &Computer::str)
        .def("__eq__", [](const Computer& a, const Computer& b) { return a == b; });

    py::class_<Gate>(m, "Gate")
        .def("gate_id", &Gate::gate_id)
        .def("target", &Gate::target)
        .def("control", &Gate::control)
        .def("adjoint", &Gate::adjoint)
        .def("nqubits", &Gate::nqubits)
        .def("str", &Gate::str)
        .def("has_parameter", &Gate::has_parameter)
        .def("parameter", &Gate::parameter)
        .def("matrix", [](const Gate& g) {
            py::array_t<std::complex<double>> a({4, 4});
            auto r = a.mutable_unchecked<2>();
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) r(i, j) = g.matrix()[i][j];
            return a;
        })
        .def("__str__", &Gate::str)
        .def("__repr__", &Gate::repr)
        .def("__eq__", [](const Gate& a, const Gate& b) { return a == b; })
        .def("__neq__", [](const Gate& a, const Gate& b) { return a != b; })
        .def("update_parameter", [](const Gate& g, double p) {
            if (!g.has_parameter()) throw std::invalid_argument("Gate does not have a parameter.");
            return make_gate(g.gate_id(), g.target(), g.control(), p);
        });

    // gate(...) overloads in the reference's registration order (bindings.cc:238-290)
    m.def(
        "gate",
        [](std::string type, size_t target, std::complex<double> parameter) {
            if (parameter.imag() != 0.0) throw std::invalid_argument("Gate parameter must be real.");
            if (in(type, kOneQubit)) return make_gate(type, target, target, parameter.real());
            throw std::invalid_argument("make_gate()\ttarget = " + std::to_string(target) + ", parameter = " +
                                        std::to_string(parameter.real()) + " is not a valid quantum gate input for type = " + type);
        },
        "type"_a, "target"_a, "parameter"_a = 0.0);
    m.def(
        "gate",
        [](std::string type, size_t target, size_t control) {
            if (in(type, kTwoQubitNoParam)) return make_gate(type, target, control, 0.0);
            if (target == control && in(type, kOneQubitNoParam)) return make_gate(type, target, control, 0.0);
            throw std::invalid_argument("make_gate()\ttarget = " + std::to_string(target) + ", control = " +
                                        std::to_string(control) + ", is not a valid quantum gate input for type = " + type);
        },
        "type"_a, "target"_a, "control"_a);
    m.def(
        "gate",
        [](std::string type, size_t target, size_t control, std::complex<double> parameter) {
            if (parameter.imag() != 0.0) throw std::invalid_argument("Gate parameter must be real.");
            return make_gate(type, target, control, parameter.real());
        },
        "type"_a, "target"_a, "control"_a, "parameter"_a = 0.0);
    m.def("control_gate", &make_control_gate, "control"_a, "Gate"_a);
    m.def("prepare_computer_from_circuit", [](int n, const Circuit& c) {
        auto comp = std::make_shared<Computer>(n);
        comp->apply_circuit(c);
        return comp;
    }, "nqubit"_a, "gates"_a);
    m.def("prepare_computer_from_gates", [](int n, const std::vector<Gate>& gates) {
        auto comp = std::make_shared<Computer>(n);
        for (const auto& g : gates) comp->apply_gate(g);
        return comp;
    }, "nqubits"_a, "gates"_a);
    m.def("add_gate_to_computer", [](const Gate& g, const Computer& c) {
        auto comp = std::make_shared<Computer>(c);
        comp->apply_gate(g);
        return comp;
    }, "gate"_a, "computer"_a);
    m.def("inner_product", &qfa::dot, "a"_a, "b"_a);
    m.def("cross_expectations", [](const Computer& bra, const Computer& ket, const std::vector<QubitOperator>& ops) {
        // <bra|O_k|ket> for every Pauli-sum operator O_k, one device launch
        std::vector<qfe_pauli> terms, one;
        std::vector<uint64_t> off{0};
        for (const auto& o : ops) {
            if (!compile_pauli_sum(o, one)) throw std::invalid_argument("cross_expectations: operators must be Pauli sums");
            terms.insert(terms.end(), one.begin(), one.end());
            off.push_back(terms.size());
        }
        std::vector<qfe_c64> r(ops.size());
        {
            py::gil_scoped_release rel;
            check(qfe_cross_expval_many(bra.handle(), ket.handle(), terms.data(), off.data(), ops.size(), r.data()),
                  "cross_expectations");
        }
        std::vector<std::complex<double>> out;
        for (const auto& v : r) out.emplace_back(v.re, v.im);
        return out;
    }, "bra"_a, "ket"_a, "operators"_a);
    m.def("norm2", [](const Computer& c) {
        double v = 0.0;
        check(qfe_state_norm2(c.handle(), &v), "norm2");
        return v;
    }, "Squared norm of a Computer's state, reduced on the device");
    m.def("find_irrep", &find_irrep, py::arg("orb_irrep_to_int"), py::arg("spinorb_indices"));
    m.def("normalize_angle", &normalize_angle);

    // native exponentiation and Trotterization
    m.def("exponentiate_pauli_string", &exponentiate_pauli_string, "coefficient"_a, "term"_a, "Use_cRz"_a = false,
          "ancilla_idx"_a = -1, "Use_open_cRz"_a = false);
    m.def("trotterize", &trotterize, "operator"_a, "factor"_a = std::complex<double>(1.0, 0.0), "trotter_number"_a = 1,
          "trotter_order"_a = 1);
    m.def("trotterize_w_cRz", &trotterize_w_cRz, "operator"_a, "ancilla_qubit_idx"_a,
          "factor"_a = std::complex<double>(1.0, 0.0), "Use_open_cRz"_a = false, "trotter_number"_a = 1,
          "trotter_order"_a = 1);

    // mask form of a Pauli circuit (application order)
    m.def("pauli_masks", [](const Circuit& c) -> py::object {
        uint64_t x, z;
        std::complex<double> ph;
        if (!pauli_masks(c, x, z, ph)) return py::none();
        return py::make_tuple(x, z, ph);
    });

    py::class_<UCCProgram, std::shared_ptr<UCCProgram>>(m, "UCCProgram")
        .def(py::init<int, uint64_t, const std::vector<QubitOperator>&, const std::vector<int>&,
                      const std::vector<double>&, size_t, const QubitOperator&>(),
             "nqubit"_a, "ref_index"_a, "generators"_a, "param_index"_a, "scale"_a, "nparam"_a, "hamiltonian"_a)
        .def("energy", &UCCProgram::energy)
        .def("energy_graph", &UCCProgram::energy_graph)
        .def("gradient", &UCCProgram::gradient)
        .def("prepare", &UCCProgram::prepare)
        .def("residuals", &UCCProgram::residuals, "amplitudes"_a, "indices"_a)
        .def("pool_gradients", &UCCProgram::pool_gradients, "amplitudes"_a, "operators"_a, "batch"_a = 64)
        .def_property_readonly("nqubit", &UCCProgram::nqubit)
        .def_property_readonly("nparam", &UCCProgram::nparam);

    py::class_<HamiltonianMap, std::shared_ptr<HamiltonianMap>>(m, "HamiltonianMap")
        .def(py::init([](size_t norb) { return std::make_shared<HamiltonianMap>(molecular_hamiltonian_map(norb)); }),
             "norb"_a)
        .def_readonly("norb", &HamiltonianMap::norb)
        .def_readonly("ninputs", &HamiltonianMap::ninputs)
        .def_property_readonly("nterms", [](const HamiltonianMap& mp) { return mp.x.size(); })
        .def("pack",
             [](const HamiltonianMap& mp, double constant, py::array_t<double, py::array::c_style | py::array::forcecast> h,
                py::array_t<double, py::array::c_style | py::array::forcecast> g) {
                 const size_t n = mp.norb;
                 if ((size_t)h.size() != n * n || (size_t)g.size() != n * n * n * n)
                     throw std::invalid_argument("HamiltonianMap.pack: h must be norb^2 and g norb^4");
                 return pack_integrals(n, constant, h.data(), g.data());
             },
             "constant"_a, "h"_a, "g"_a)
        .def("operator", &map_to_operator, "inputs"_a);

    py::class_<SectorTemplate, std::shared_ptr<SectorTemplate>>(m, "SectorTemplate")
        .def(py::init<int, uint64_t, const std::vector<QubitOperator>&, const std::vector<int>&,
                      const std::vector<double>&, size_t, const QubitOperator&>(),
             "nqubit"_a, "ref_index"_a, "generators"_a, "param_index"_a, "scale"_a, "nparam"_a, "hamiltonian"_a)
        .def(py::init<int, uint64_t, const std::vector<QubitOperator>&, const std::vector<int>&,
                      const std::vector<double>&, size_t, const HamiltonianMap&>(),
             "nqubit"_a, "ref_index"_a, "generators"_a, "param_index"_a, "scale"_a, "nparam"_a, "hamiltonian_map"_a)
        .def("energy", &SectorTemplate::energy, "amplitudes"_a, "inputs"_a = std::vector<double>())
        .def("prepare", &SectorTemplate::prepare)
        .def("residuals", &SectorTemplate::residuals, "amplitudes"_a, "indices"_a,
             "inputs"_a = std::vector<double>())
        .def("gradient", &SectorTemplate::gradient, "amplitudes"_a, "inputs"_a = std::vector<double>())
        .def("pool_gradients", &SectorTemplate::pool_gradients, "amplitudes"_a, "operators"_a, "batch"_a = 64,
             "inputs"_a = std::vector<double>())
        .def_property_readonly("dim", &SectorTemplate::dim)
        .def_property_readonly("request_bytes", &SectorTemplate::request_bytes)
        .def("graph_audit", &SectorTemplate::graph_audit, "gradient"_a = false)
        .def_property_readonly("nqubit", &SectorTemplate::nqubit)
        .def_property_readonly("nparam", &SectorTemplate::nparam)
        .def_property_readonly("ninputs", &SectorTemplate::ninputs);

    py::class_<BatchEngine>(m, "BatchEngine")
        .def(py::init<uint64_t, int>(), "hbm_budget_bytes"_a, "max_active"_a)
        .def("submit", &BatchEngine::submit, "template"_a, "amplitudes"_a, "inputs"_a = std::vector<double>(),
             "gradient"_a = false)
        .def("take_gradient", &BatchEngine::take_gradient, "request_id"_a, "nparam"_a)
        .def("step", &BatchEngine::step, "cap"_a = 4096)
        .def("poll", &BatchEngine::poll, "cap"_a = 4096)
        .def("pending", &BatchEngine::pending);

    // engine controls
    m.def("device_count", []() {
        int n = 0;
        qfe_device_count(&n);
        return n;
    });
    m.def("set_tuning", [](int k, int l, int ops) { check(qfe_set_tuning(k, l, ops), "set_tuning"); }, "tile_qubits"_a = 0,
          "low_qubits"_a = 0, "max_ops_per_pass"_a = 0);
    m.def("set_pass_tuning", [](int o) { check(qfe_set_pass_tuning(o), "set_pass_tuning"); }, "orbit_max_dims"_a);
    m.def("set_sector_word_limit", [](uint64_t n) { check(qfe_set_sector_word_limit(n), "set_sector_word_limit"); },
          "max_pair_words"_a);
    m.def("get_pass_tuning", []() {
        int o = 0;
        qfe_get_pass_tuning(&o);
        return o;
    });
    m.def("get_tuning", []() {
        int k, l, o;
        qfe_get_tuning(&k, &l, &o);
        return py::make_tuple(k, l, o);
    });
    m.def("get_counters", []() {
        uint64_t f, s, r, v[7];
        double b;
        qfe_get_counters(&f, &s, &r, &b);
        qfe_get_pass_counts(v, 7);
        return py::dict("fused_passes"_a = f, "single_passes"_a = s, "reduce_passes"_a = r, "bytes"_a = b,
                        "tile_passes"_a = v[3], "orbit_passes"_a = v[4], "lane_orbit_passes"_a = v[5],
                        "nt_passes"_a = v[6]);
    });
    m.def("reset_counters", []() { qfe_reset_counters(); });
    m.def("synchronize", []() { check(qfe_synchronize(), "synchronize"); });
    py::class_<DeviceTimer>(m, "DeviceTimer")
        .def(py::init<>())
        .def("start", &DeviceTimer::start)
        .def("stop", &DeviceTimer::stop, "Milliseconds between start and stop on the engine stream");
    m.def("stream_handle", []() {
        void* s = nullptr;
        check(qfe_get_stream(&s), "stream_handle");
        return (uintptr_t)s;
    });
}
