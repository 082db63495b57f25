"""Dense FP64 state-vector oracle for qforte's ``Computer`` hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``qforte_amd``) may import,
call or link this module; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

It is a CPU restatement, in numpy, of the arithmetic the reference performs on
the amplitude vector.  Conventions (reference ``src/qforte``):

* basis index bit ``q`` is qubit ``q`` (``qubit_basis.h:27-36``);
* a 2-qubit gate matrix is indexed by ``2*control_bit + target_bit``
  (``gate.h:134-150``, ``computer.cc:427-456``);
* a ``Circuit`` is applied first gate first (``computer.cc:93-97``);
* ``apply_operator`` replaces psi by sum_k c_k * (circuit_k applied to psi)
  (``computer.cc:80-91``);
* ``direct_op_exp_val`` returns ``<psi| sum_k c_k circ_k |psi>`` with the bra
  conjugated (``computer.cc:753-771``, default non-OpenMP build).

Every function names the reference lines it follows.  Each one is pinned in
``tests/test_oracle.py`` by closed forms, brute-force Kronecker products,
``scipy.linalg.expm``, the reference's own known-answer vectors and the
reference C++ core compiled from its sources (``oracle/Makefile`` ->
``oracle/_ref``), with planted mistakes that those checks must catch.
"""

from __future__ import annotations

import math
from typing import Iterable, Sequence

import numpy as np

# ---------------------------------------------------------------------------
# Gate catalogue (reference make_gate.cc:7-256, gate.cc:135-179)
# ---------------------------------------------------------------------------

ONE_QUBIT = ("X", "Y", "Z", "H", "R", "Rx", "Ry", "Rz", "V", "S", "T", "I")
TWO_QUBIT = ("A", "cX", "CNOT", "acX", "aCNOT", "cY", "cZ", "cR", "cV", "cRz", "SWAP")
PERIOD = {"R": 2 * math.pi, "Rx": 4 * math.pi, "Ry": 4 * math.pi, "Rz": 4 * math.pi,
          "A": 2 * math.pi, "cR": 2 * math.pi, "cRz": 4 * math.pi}


def normalize_angle(angle: float, period: float) -> float:
    """Fold ``angle`` into [-period/2, period/2] (make_gate.cc:7-16)."""
    angle = math.fmod(angle, period)
    if angle < -period / 2:
        angle += period
    elif angle > period / 2:
        angle -= period
    return angle


def gate_matrix(gtype: str, parameter: float = 0.0) -> np.ndarray:
    """Matrix of a named gate (make_gate.cc:18-235).

    Returns a 2x2 array for one-qubit gates and a 4x4 array (basis
    ``2*c + t``) for two-qubit gates.  The parameter is folded with
    :func:`normalize_angle` exactly as the reference does before use.
    """
    if gtype in PERIOD:
        parameter = normalize_angle(parameter, PERIOD[gtype])
    c2 = 1.0 / math.sqrt(2.0)
    if gtype == "X":
        return np.array([[0, 1], [1, 0]], dtype=complex)
    if gtype == "Y":
        return np.array([[0, -1j], [1j, 0]], dtype=complex)
    if gtype == "Z":
        return np.array([[1, 0], [0, -1]], dtype=complex)
    if gtype == "H":
        return np.array([[c2, c2], [c2, -c2]], dtype=complex)
    if gtype == "R":
        return np.array([[1, 0], [0, np.exp(1j * parameter)]], dtype=complex)
    if gtype == "Rx":
        a, b = math.cos(0.5 * parameter), 1j * math.sin(0.5 * parameter)
        return np.array([[a, -b], [-b, a]], dtype=complex)
    if gtype == "Ry":
        a, b = math.cos(0.5 * parameter), math.sin(0.5 * parameter)
        return np.array([[a, -b], [b, a]], dtype=complex)
    if gtype == "Rz":
        return np.array([[np.exp(-0.5j * parameter), 0], [0, np.exp(0.5j * parameter)]], dtype=complex)
    if gtype == "V":
        a, b = 0.5 + 0.5j, 0.5 - 0.5j
        return np.array([[a, b], [b, a]], dtype=complex)
    if gtype == "S":
        return np.array([[1, 0], [0, 1j]], dtype=complex)
    if gtype == "T":
        return np.array([[1, 0], [0, (1 + 1j) / math.sqrt(2.0)]], dtype=complex)
    if gtype == "I":
        return np.eye(2, dtype=complex)

    def controlled(u):
        m = np.eye(4, dtype=complex)
        m[2:, 2:] = u
        return m

    if gtype == "A":
        c, s = math.cos(parameter), math.sin(parameter)
        return np.array([[1, 0, 0, 0], [0, c, s, 0], [0, s, -c, 0], [0, 0, 0, 1]], dtype=complex)
    if gtype in ("cX", "CNOT"):
        return controlled(gate_matrix("X"))
    if gtype in ("acX", "aCNOT"):
        m = np.eye(4, dtype=complex)
        m[:2, :2] = gate_matrix("X")
        return m
    if gtype == "cY":
        return controlled(gate_matrix("Y"))
    if gtype == "cZ":
        return controlled(gate_matrix("Z"))
    if gtype == "cR":
        return controlled(np.array([[1, 0], [0, np.exp(1j * parameter)]], dtype=complex))
    if gtype == "cV":
        return controlled(gate_matrix("V"))
    if gtype == "cRz":
        return controlled(np.array([[np.exp(-0.5j * parameter), 0], [0, np.exp(0.5j * parameter)]], dtype=complex))
    if gtype == "SWAP":
        return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)
    raise ValueError(f"unknown gate type {gtype}")


class G:
    """A gate as plain data: (type, target, control, parameter, matrix)."""

    __slots__ = ("gtype", "target", "control", "parameter", "matrix")

    def __init__(self, gtype, target, control=None, parameter=0.0, matrix=None):
        self.gtype = gtype
        self.target = int(target)
        self.control = int(target if control is None else control)
        self.parameter = float(parameter)
        self.matrix = gate_matrix(gtype, parameter) if matrix is None else np.asarray(matrix, dtype=complex)

    @property
    def nqubits(self):
        return 1 if self.target == self.control else 2

    def adjoint(self):
        """Conjugate transpose (gate.cc:135-179).  Rotation-type gates keep
        their name with the parameter negated, as the reference does; the A
        gate is self-adjoint."""
        m = self.matrix.conj().T
        p = -self.parameter if self.gtype in ("R", "Rx", "Ry", "Rz", "cR", "cRz") else self.parameter
        return G(self.gtype, self.target, self.control, p, matrix=m)

    def __repr__(self):
        return f"G({self.gtype},{self.target},{self.control},{self.parameter})"


# ---------------------------------------------------------------------------
# Gate application (reference computer.cc:105-114, 319-751)
# ---------------------------------------------------------------------------

def _bits_index(n: int, bits: Sequence[int]) -> np.ndarray:
    """Indices with the listed bits cleared, in ascending order (enumerates
    the 2^(n-len(bits)) orbits a k-qubit operator acts on)."""
    rest = [q for q in range(n) if q not in bits]
    k = np.arange(1 << len(rest), dtype=np.int64)
    out = np.zeros_like(k)
    for i, q in enumerate(rest):
        out |= ((k >> i) & 1) << q
    return out


def apply_gate(psi: np.ndarray, n: int, g: G) -> np.ndarray:
    """Return the state after one gate (the arithmetic of computer.cc:340-751).

    Qubit indices at or beyond ``n`` read as bit 0, as the reference's bit
    tests do; a control outside the register therefore never fires.
    """
    psi = np.array(psi, dtype=complex, copy=True)
    if g.nqubits == 1:
        t = g.target
        if t >= n:
            raise ValueError("target out of range")
        base = _bits_index(n, [t])
        i0, i1 = base, base | (1 << t)
        a, b = psi[i0].copy(), psi[i1].copy()
        m = g.matrix
        psi[i0] = m[0, 0] * a + m[0, 1] * b
        psi[i1] = m[1, 0] * a + m[1, 1] * b
        return psi
    t, c = g.target, g.control
    m = g.matrix
    if c >= n:
        # control bit reads 0: only the control-0 block acts
        if not (np.allclose(m[:2, :2], np.eye(2)) and np.allclose(m[:2, 2:], 0) and np.allclose(m[2:, :2], 0)):
            raise ValueError("non-controlled 2-qubit gate with a qubit out of range")
        return psi
    if t >= n:
        raise ValueError("target out of range")
    base = _bits_index(n, [t, c])
    idx = [base | (cb << c) | (tb << t) for cb in (0, 1) for tb in (0, 1)]
    sub = np.stack([psi[i] for i in idx])
    new = m @ sub
    for a in range(4):
        psi[idx[a]] = new[a]
    return psi


def apply_circuit(psi: np.ndarray, n: int, gates: Iterable[G]) -> np.ndarray:
    """Apply gates first to last (computer.cc:93-97)."""
    for g in gates:
        psi = apply_gate(psi, n, g)
    return psi


def circuit_matrix(n: int, gates: Iterable[G]) -> np.ndarray:
    """Brute-force 2^n x 2^n matrix of a circuit from Kronecker products.

    Independent of :func:`apply_gate`: each gate is lifted with ``np.kron`` over
    qubits ordered most-significant first, the form the reference's sparse
    lifting reproduces (``gate.cc:29-91``)."""
    dim = 1 << n
    total = np.eye(dim, dtype=complex)
    for g in gates:
        if g.nqubits == 1:
            ops = [np.eye(2, dtype=complex)] * n
            ops = list(ops)
            ops[n - 1 - g.target] = g.matrix
            full = ops[0]
            for o in ops[1:]:
                full = np.kron(full, o)
        else:
            full = np.zeros((dim, dim), dtype=complex)
            for j in range(dim):
                cb, tb = (j >> g.control) & 1, (j >> g.target) & 1
                col = 2 * cb + tb
                for row in range(4):
                    v = g.matrix[row, col]
                    if v != 0:
                        i = j & ~(1 << g.control) & ~(1 << g.target)
                        i |= ((row >> 1) & 1) << g.control
                        i |= (row & 1) << g.target
                        full[i, j] += v
        total = full @ total
    return total


# ---------------------------------------------------------------------------
# Pauli strings in mask form
# ---------------------------------------------------------------------------

def pauli_masks(term: Sequence[tuple[str, int]]):
    """(x_mask, z_mask, phase) of the operator a Pauli circuit applies.

    ``term`` is the circuit as (letter, qubit) pairs, applied first to last
    (computer.cc:93-97).  The product is formed letter by letter with
    X|b> = |b^1>, Z|b> = (-1)^b |b>, Y|b> = i(-1)^b |b^1>, giving
    P|j> = phase * (-1)^popcount(j & z) |j ^ x>.
    """
    x = z = 0
    phase = 1 + 0j
    for letter, q in term:
        bit = 1 << q
        # left-multiply the current operator by the new letter
        if letter == "X":
            # X (phase (-1)^{j.z} |j^x>) : the sign depends on the pre-image;
            # X commutes through as a pure flip of the output index.
            x ^= bit
        elif letter == "Z":
            # Z acts on the output index j^x: (-1)^{(j^x)_q}
            #   = (-1)^{j_q} (-1)^{x_q}
            if x & bit:
                phase = -phase
            z ^= bit
        elif letter == "Y":
            # Y = i X Z applied after: i * (-1)^{(j^x)_q} then flip q
            phase *= 1j
            if x & bit:
                phase = -phase
            z ^= bit
            x ^= bit
        elif letter == "I":
            continue
        else:
            raise ValueError(f"not a Pauli letter: {letter}")
    return x, z, phase


def apply_pauli(psi: np.ndarray, x: int, z: int, phase: complex) -> np.ndarray:
    """(P psi)[k] = phase * (-1)^popcount((k^x) & z) * psi[k ^ x]."""
    k = np.arange(psi.shape[0], dtype=np.int64)
    src = k ^ x
    sign = 1 - 2 * (_popcount(src & z) & 1)
    return phase * sign * psi[src]


def _popcount(a: np.ndarray) -> np.ndarray:
    """Bit count of every element (SWAR reduction on uint64, wrapping multiply)."""
    x = a.astype(np.uint64)
    x = x - ((x >> np.uint64(1)) & np.uint64(0x5555555555555555))
    x = (x & np.uint64(0x3333333333333333)) + ((x >> np.uint64(2)) & np.uint64(0x3333333333333333))
    x = (x + (x >> np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
    with np.errstate(over="ignore"):
        x = (x * np.uint64(0x0101010101010101)) >> np.uint64(56)
    return x.astype(np.int64)


def apply_pauli_circuit(psi: np.ndarray, n: int, term: Sequence[tuple[str, int]]) -> np.ndarray:
    """Gate-by-gate application of a Pauli circuit (computer.cc:93-97)."""
    for letter, q in term:
        psi = apply_gate(psi, n, G(letter, q))
    return psi


def exp_pauli(psi: np.ndarray, coeff: complex, term, n: int | None = None) -> np.ndarray:
    """exp(coeff * P) psi for purely imaginary ``coeff`` and a canonical term.

    The reference builds this as basis change, CNOT ladder, Rz(-2 Im coeff),
    inverse ladder, inverse basis change (exponentiate.py:9-87), which equals
    cos(phi) I + i sin(phi) P with phi = Im(coeff) for a product of Paulis on
    distinct qubits.  A real part above 1e-14 is rejected as the reference
    does (exponentiate.py:25-28); the empty term is the phase exp(coeff)
    (exponentiate.py:31-32).
    """
    if abs(complex(coeff).real) > 1.0e-14:
        raise ValueError("exp_pauli called with a real coefficient")
    if len(term) == 0:
        return psi * np.exp(coeff)
    x, z, phase = pauli_masks(term)
    phi = complex(coeff).imag
    return math.cos(phi) * psi + 1j * math.sin(phi) * apply_pauli(psi, x, z, phase)


def exponentiate_gates(coeff: complex, term) -> list[G]:
    """The gate list exponentiate_pauli_string builds (exponentiate.py:34-85)."""
    to_z, cx, prev, last = [], [], None, None
    for letter, q in term:
        if letter == "X":
            to_z.append(G("H", q))
        elif letter == "Y":
            to_z.append(G("Rx", q, q, math.pi / 2))
        elif letter == "I":
            continue
        if prev is not None:
            cx.append(G("cX", q, prev))
        prev = last = q
    rz = G("Rz", last, last, -2.0 * complex(coeff).imag)
    adj = lambda gs: [g.adjoint() for g in reversed(gs)]
    return to_z + cx + [rz] + adj(cx) + adj(to_z)


# ---------------------------------------------------------------------------
# Operators (reference computer.cc:80-91, 753-806; qubit_operator.cc)
# ---------------------------------------------------------------------------

def apply_operator(psi: np.ndarray, n: int, op) -> np.ndarray:
    """sum_k c_k circuit_k(psi) (computer.cc:80-91).  ``op`` is a list of
    (coeff, [(letter, qubit), ...]) or (coeff, [G, ...])."""
    out = np.zeros_like(psi, dtype=complex)
    for c, term in op:
        if term and isinstance(term[0], G):
            out += c * apply_circuit(psi, n, term)
        else:
            out += c * apply_pauli_circuit(psi, n, term)
    return out


def expval(psi: np.ndarray, n: int, op) -> complex:
    """<psi| op |psi> (computer.cc:753-771, default build)."""
    return complex(np.vdot(psi, apply_operator(psi, n, op)))


def expval_masks(psi: np.ndarray, terms) -> complex:
    """Same value from mask form: sum_t c_t <psi|P_t|psi>, terms as
    (coeff, x, z, phase)."""
    tot = 0j
    for c, x, z, ph in terms:
        tot += c * np.vdot(psi, apply_pauli(psi, x, z, ph))
    return complex(tot)


def operator_matrix(n: int, op) -> np.ndarray:
    """Dense matrix of sum_k c_k P_k from Kronecker products (brute force)."""
    dim = 1 << n
    out = np.zeros((dim, dim), dtype=complex)
    for c, term in op:
        gates = [G(l, q) for l, q in term]
        out += c * circuit_matrix(n, gates)
    return out


# ---------------------------------------------------------------------------
# Second quantisation: Jordan-Wigner (sq_operator.cc:106-169)
# ---------------------------------------------------------------------------

def jw_dense(terms, n: int) -> np.ndarray:
    """Dense matrix of a fermionic operator via Jordan-Wigner, built from
    a_p = 1/2 (X_p + i Y_p) prod_{k<p} Z_k and a_p^dag = 1/2 (X_p - i Y_p)
    prod_{k<p} Z_k (sq_operator.cc:106-136).  ``terms`` holds
    (coeff, creators, annihilators); the product is cre[0]^ cre[1]^ ...
    ann[0] ann[1] ... in written order (sq_operator.cc:138-169)."""
    dim = 1 << n
    eye = np.eye(dim, dtype=complex)

    def single(p, dagger):
        z = [("Z", k) for k in range(p)]
        mx = operator_matrix(n, [(1.0, z + [("X", p)])])
        my = operator_matrix(n, [(1.0, z + [("Y", p)])])
        return 0.5 * (mx - 1j * my) if dagger else 0.5 * (mx + 1j * my)

    out = np.zeros((dim, dim), dtype=complex)
    for c, cre, ann in terms:
        m = eye.copy()
        for p in cre:
            m = m @ single(p, True)
        for p in ann:
            m = m @ single(p, False)
        out += c * m
    return out


# ---------------------------------------------------------------------------
# UCC energy / gradient on the dense vector (ansatz.py:25-82,
# algorithm.py:271-334, uccvqeabc.py:122-251)
# ---------------------------------------------------------------------------

def ucc_state(n: int, ref_occ: Sequence[int], exps) -> np.ndarray:
    """|ref> then the exponentials in list order.  ``exps`` is a list of
    (imag_coeff_times_amplitude, canonical_term)."""
    psi = np.zeros(1 << n, dtype=complex)
    idx = sum(1 << q for q, o in enumerate(ref_occ) if o)
    psi[idx] = 1.0
    for c, term in exps:
        psi = exp_pauli(psi, c, term)
    return psi
