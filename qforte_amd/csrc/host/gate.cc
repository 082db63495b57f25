// Gate catalogue and gate algebra (reference make_gate.cc, gate.cc).
#include <cmath>
#include <cstdio>

#include "core.h"

namespace qfa {

std::string fmt_complex(cplx v) {
    // helpers.cc:17-26: "{:+f}" real, "{:+f}j" imaginary, "{:+f} {:+f}i" general
    char buf[96];
    if (v.imag() == 0.0) {
        std::snprintf(buf, sizeof buf, "%+f", v.real());
    } else if (v.real() == 0.0) {
        std::snprintf(buf, sizeof buf, "%+fj", v.imag());
    } else {
        std::snprintf(buf, sizeof buf, "%+f %+fi", v.real(), v.imag());
    }
    return buf;
}

std::string join(const std::vector<std::string>& parts, const std::string& sep) {
    std::string out;
    for (size_t i = 0; i < parts.size(); ++i) {
        if (i) out += sep;
        out += parts[i];
    }
    return out;
}

std::string QubitBasis::str(size_t nqubit) const {
    std::string s = "|";
    for (size_t q = 0; q < nqubit; ++q) s += get_bit(q) ? '1' : '0';
    return s + ">";
}

Gate::Gate(std::string label, size_t target, size_t control, const Mat4& m,
           std::optional<std::pair<double, bool>> parameter)
    : label_(std::move(label)), target_(target), control_(control), parameter_(parameter), m_(m) {}

double normalize_angle(double angle, double period) {
    angle = std::fmod(angle, period);
    if (angle < -period / 2) {
        angle += period;
    } else if (angle > period / 2) {
        angle -= period;
    }
    return angle;
}

namespace {

const cplx I1(0.0, 1.0);

Mat4 zero4() {
    Mat4 m;
    for (auto& r : m) r.fill(cplx(0.0, 0.0));
    return m;
}

Mat4 one_qubit(cplx a, cplx b, cplx c, cplx d) {
    Mat4 m = zero4();
    m[0][0] = a;
    m[0][1] = b;
    m[1][0] = c;
    m[1][1] = d;
    return m;
}

// I on the control-0 block, u on the control-1 block (basis 2*c + t).
Mat4 controlled(cplx a, cplx b, cplx c, cplx d) {
    Mat4 m = zero4();
    m[0][0] = 1.0;
    m[1][1] = 1.0;
    m[2][2] = a;
    m[2][3] = b;
    m[3][2] = c;
    m[3][3] = d;
    return m;
}

std::string invalid_type_msg(const std::string& type) {
    return "make_gate()\ntype = " + type + " is not a valid quantum gate type";
}

}  // namespace

Gate make_gate(const std::string& type, size_t target, size_t control, double parameter) {
    using P = std::optional<std::pair<double, bool>>;
    if (target == control) {
        const cplx r2 = 1.0 / std::sqrt(2.0);  // complex, so -r2 carries a signed zero like the reference's
        if (type == "X") return Gate(type, target, control, one_qubit(0.0, 1.0, 1.0, 0.0));
        if (type == "Y") return Gate(type, target, control, one_qubit(0.0, -I1, I1, 0.0));
        if (type == "Z") return Gate(type, target, control, one_qubit(1.0, 0.0, 0.0, -1.0));
        if (type == "H") return Gate(type, target, control, one_qubit(r2, r2, r2, -r2));
        if (type == "I") return Gate(type, target, control, one_qubit(1.0, 0.0, 0.0, 1.0));
        if (type == "S") return Gate(type, target, control, one_qubit(1.0, 0.0, 0.0, I1));
        if (type == "T") return Gate(type, target, control, one_qubit(1.0, 0.0, 0.0, (1.0 + I1) / std::sqrt(2.0)));
        if (type == "V") {
            const cplx a = 0.5 + 0.5 * I1, b = 0.5 - 0.5 * I1;
            return Gate(type, target, control, one_qubit(a, b, b, a));
        }
        if (type == "R") {
            const double th = normalize_angle(parameter, 2 * M_PI);
            return Gate(type, target, control, one_qubit(1.0, 0.0, 0.0, std::exp(I1 * th)), P({th, true}));
        }
        if (type == "Rx") {
            const double th = normalize_angle(parameter, 4 * M_PI);
            const cplx a = std::cos(0.5 * th), b = I1 * std::sin(0.5 * th);
            return Gate(type, target, control, one_qubit(a, -b, -b, a), P({th, true}));
        }
        if (type == "Ry") {
            const double th = normalize_angle(parameter, 4 * M_PI);
            const cplx a = std::cos(0.5 * th), b = std::sin(0.5 * th);
            return Gate(type, target, control, one_qubit(a, -b, b, a), P({th, true}));
        }
        if (type == "Rz") {
            const double th = normalize_angle(parameter, 4 * M_PI);
            return Gate(type, target, control,
                        one_qubit(std::exp(-I1 * 0.5 * th), 0.0, 0.0, std::exp(I1 * 0.5 * th)), P({th, true}));
        }
    } else {
        if (type == "A") {
            // particle-number preserving A gate (make_gate.cc:153-168)
            const double th = normalize_angle(parameter, 2 * M_PI);
            const cplx c = std::cos(th), s = std::sin(th);
            Mat4 m = zero4();
            m[0][0] = 1.0;
            m[1][1] = c;
            m[1][2] = s;
            m[2][1] = s;
            m[2][2] = -c;
            m[3][3] = 1.0;
            return Gate(type, target, control, m, P({th, false}));
        }
        if (type == "cX" || type == "CNOT") return Gate(type, target, control, controlled(0.0, 1.0, 1.0, 0.0));
        if (type == "acX" || type == "aCNOT") {
            Mat4 m = zero4();
            m[0][1] = 1.0;
            m[1][0] = 1.0;
            m[2][2] = 1.0;
            m[3][3] = 1.0;
            return Gate(type, target, control, m);
        }
        if (type == "cY") return Gate(type, target, control, controlled(0.0, -I1, I1, 0.0));
        if (type == "cZ") return Gate(type, target, control, controlled(1.0, 0.0, 0.0, -1.0));
        if (type == "cR") {
            const double th = normalize_angle(parameter, 2 * M_PI);
            return Gate(type, target, control, controlled(1.0, 0.0, 0.0, std::exp(I1 * th)), P({th, true}));
        }
        if (type == "cV") {
            const cplx a = 0.5 + 0.5 * I1, b = 0.5 - 0.5 * I1;
            return Gate(type, target, control, controlled(a, b, b, a));
        }
        if (type == "cRz") {
            const double th = normalize_angle(parameter, 4 * M_PI);
            return Gate(type, target, control,
                        controlled(std::exp(-I1 * 0.5 * th), 0.0, 0.0, std::exp(I1 * 0.5 * th)), P({th, true}));
        }
        if (type == "SWAP") {
            Mat4 m = zero4();
            m[0][0] = 1.0;
            m[1][2] = 1.0;
            m[2][1] = 1.0;
            m[3][3] = 1.0;
            return Gate(type, target, control, m);
        }
    }
    throw std::invalid_argument(invalid_type_msg(type));
}

Gate make_control_gate(size_t control, const Gate& U) {
    if (U.target() == control) throw std::invalid_argument("Cannot create Control-U where targer == control !");
    const Mat4& u = U.matrix();
    std::optional<std::pair<double, bool>> p;
    if (U.has_parameter()) p = std::make_pair(*U.parameter(), U.minus_parameter_on_adjoint());
    return Gate("cU", U.target(), control, controlled(u[0][0], u[0][1], u[1][0], u[1][1]), p);
}

Gate Gate::adjoint() const {
    Mat4 adj;
    bool self_adjoint = true;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            adj[j][i] = std::conj(m_[i][j]);
        }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            if (std::norm(adj[i][j] - m_[i][j]) > 1.0e-12) self_adjoint = false;
    if (has_parameter()) {
        const double sign = minus_parameter_on_adjoint() ? -1.0 : 1.0;
        return Gate(label_, target_, control_, adj, std::make_pair(parameter_->first * sign, parameter_->second));
    }
    if (self_adjoint) return *this;
    // adjoints of S and T are phase gates R(-pi/2), R(-pi/4) (gate.cc:162-165)
    if (label_ == "T") return Gate("R", target_, control_, adj, std::make_pair(-M_PI / 4, true));
    if (label_ == "S") return Gate("R", target_, control_, adj, std::make_pair(-M_PI / 2, true));
    if (label_.size() > 4 && label_.compare(0, 4, "adj(") == 0 && label_.back() == ')')
        return Gate(label_.substr(4, label_.size() - 5), target_, control_, adj, parameter_);
    return Gate("adj(" + label_ + ")", target_, control_, adj, parameter_);
}

std::string Gate::str() const {
    if (target_ == control_) return label_ + std::to_string(target_);
    return label_ + std::to_string(target_) + "_" + std::to_string(control_);
}

std::string Gate::repr() const {
    std::string s = label_ + " gate, target qubit:" + std::to_string(target_) +
                    ", contol qubit:" + std::to_string(control_) + "\n";
    const int dim = nqubits() == 1 ? 2 : 4;
    char buf[96];
    for (int i = 0; i < dim; ++i) {
        for (int j = 0; j < dim; ++j) {
            std::snprintf(buf, sizeof buf, "  %+f %+f i", m_[i][j].real(), m_[i][j].imag());
            s += buf;
        }
        s += '\n';
    }
    return s;
}

qfe_gate Gate::to_abi() const {
    qfe_gate g;
    g.target = (int32_t)target_;
    g.control = (int32_t)control_;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) g.m[i * 4 + j] = qfe_c64{m_[i][j].real(), m_[i][j].imag()};
    return g;
}

char Gate::pauli_letter() const {
    if (target_ != control_) return 0;
    if (label_ == "X") return 'X';
    if (label_ == "Y") return 'Y';
    if (label_ == "Z") return 'Z';
    if (label_ == "I") return 'I';
    return 0;
}

bool operator==(const Gate& a, const Gate& b) {
    if (a.gate_id() != b.gate_id() || a.target() != b.target() || a.control() != b.control()) return false;
    if (a.has_parameter() && b.has_parameter()) return std::fabs(*a.parameter() - *b.parameter()) < 1.0e-12;
    return true;
}

bool operator!=(const Gate& a, const Gate& b) { return !(a == b); }

bool operator<(const Gate& a, const Gate& b) {
    if (a.gate_id() != b.gate_id()) return a.gate_id() < b.gate_id();
    if (a.target() != b.target()) return a.target() < b.target();
    if (a.control() != b.control()) return a.control() < b.control();
    if (a.has_parameter() && b.has_parameter()) return *a.parameter() < *b.parameter();
    return false;
}

size_t find_irrep(const std::vector<size_t>& orb_irreps_to_int, const std::vector<size_t>& spinorb_indices) {
    size_t irrep = 0;
    for (size_t idx : spinorb_indices) irrep ^= orb_irreps_to_int.at(idx / 2);
    return irrep;
}

}  // namespace qfa
