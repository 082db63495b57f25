"""QITE driver (reference src/qforte/ite/qite.py:20-491).

Per imaginary-time step: b_I = -i / sqrt(1 - 2 db E) <psi| sigma_I H |psi> and
S_IJ = <sigma_I psi| sigma_J psi>, solve S x = b, apply the Trotterised
exp(-i db sum_I x_I sigma_I).  The overlaps run as batched device reductions
(one launch for all b_I, one for all S_IJ) instead of one state copy and
vdot per element.
"""

import numpy as np
from scipy.linalg import lstsq

from .. import _core
from ..abc.algorithm import Algorithm, Trotterizable
from ..maths.eigsolve import canonical_geig_solve
from ..utils.trotterization import trotterize


class QITE(Trotterizable, Algorithm):
    def run(self, beta=1.0, db=0.2, expansion_type="SD", sparseSb=True, b_thresh=1.0e-6, x_thresh=1.0e-10,
            do_lanczos=False, lanczos_gap=2):
        self._beta = beta
        self._db = db
        self._nbeta = int(beta / db) + 1
        self._expansion_type = expansion_type
        self._sparseSb = sparseSb
        self._total_phase = 1.0 + 0.0j
        self._Uqite = _core.Circuit()
        self._b_thresh = b_thresh
        self._x_thresh = x_thresh
        self._n_classical_params = 0
        self._n_cnot = self._Uprep.get_num_cnots()
        self._n_pauli_trm_measures = 0
        self._do_lanczos = do_lanczos
        self._lanczos_gap = lanczos_gap
        qc_ref = _core.Computer(self._nqb)
        qc_ref.apply_circuit(self._Uprep)
        self._Ekb = [np.real(qc_ref.direct_op_exp_val(self._qb_ham))]
        self.build_expansion_pool()
        self.evolve()
        if self._do_lanczos:
            self.do_qlanczos()
        self.verify_required_attributes()
        print(f"Final QITE Energy:                        {round(self._Egs, 10)}")

    def build_expansion_pool(self):
        """Odd-Y Pauli strings of the (G)SD..-pool JW images (qite.py:214-254)."""
        self._sig = _core.QubitOpPool()
        if self._expansion_type == "complete_qubit":
            if self._nqb > 6:
                raise ValueError("Using complete qubits expansion will result in a very large number of terms!")
            self._sig.fill_pool("complete_qubit", self._nqb)
        elif self._expansion_type == "cqoy":
            self._sig.fill_pool("cqoy", self._nqb)
        elif self._expansion_type in {"SD", "GSD", "SDT", "SDTQ", "SDTQP", "SDTQPH"}:
            P = _core.SQOpPool()
            P.set_orb_spaces(self._ref)
            P.fill_pool(self._expansion_type)
            sig_temp = P.get_qubit_operator("commuting_grp_lex", False)
            for _, rho in sig_temp.terms():
                ny = 0
                circ = _core.Circuit()
                for g in rho.gates():
                    circ.add(_core.gate(g.gate_id(), g.target(), g.control()))
                    ny += g.gate_id() == "Y"
                if ny % 2 == 1:
                    op = _core.QubitOperator()
                    op.add(1.0, circ)
                    self._sig.add(1.0, op)
        else:
            raise ValueError("Invalid expansion type specified.")
        self._NI = len(self._sig.terms())
        self._sig_ops = [op for _, op in self._sig.terms()]

    def _overlaps(self, idx):
        """S_IJ = <psi|sigma_I sigma_J|psi> for I, J in idx (qite.py:256-303)."""
        ops = []
        pairs = []
        for a, I in enumerate(idx):
            for b_, J in enumerate(idx[:a]):
                prod = _core.QubitOperator()
                circ = _core.Circuit()
                circ.add(self._sig_ops[J].terms()[0][1])
                circ.add(self._sig_ops[I].terms()[0][1])
                prod.add(self._sig_ops[I].terms()[0][0] * self._sig_ops[J].terms()[0][0], circ)
                ops.append(prod)
                pairs.append((a, b_))
        S = np.eye(len(idx))
        if ops:
            vals = _core.cross_expectations(self._qc, self._qc, ops)
            for (a, b_), v in zip(pairs, vals):
                S[a, b_] = S[b_, a] = np.real(v)
        return S

    def build_b(self):
        """b_I = -i / sqrt(1 - 2 db E) <psi|sigma_I H|psi> (qite.py:305-326)."""
        prefactor = -1.0j / np.sqrt(1 - 2 * self._db * self._Ekb[-1])
        self._n_pauli_trm_measures += self._Nl * self._NI
        Hpsi = _core.Computer(self._qc)
        Hpsi.apply_operator(self._qb_ham)
        vals = _core.cross_expectations(self._qc, Hpsi, self._sig_ops)
        return np.real(prefactor * np.asarray(vals))

    def build_S(self):
        self._n_pauli_trm_measures += int(self._NI * (self._NI + 1) * 0.5)
        return self._overlaps(list(range(self._NI)))

    def build_sparse_S_b(self, b):
        idx = [I for I, bI in enumerate(b) if abs(bI) > self._b_thresh]
        self._n_pauli_trm_measures += int(len(idx) * (len(idx) + 1) * 0.5)
        return idx, self._overlaps(idx), np.real([b[I] for I in idx])

    def do_qite_step(self):
        btot = self.build_b()
        A = _core.QubitOperator()
        if self._sparseSb:
            sp, S, btot = self.build_sparse_S_b(btot)
        else:
            S = self.build_S()
            sp = list(range(self._NI))
        x = np.real(lstsq(S, btot)[0]) if len(sp) else np.zeros(0)
        for I, spI in enumerate(sp):
            if abs(x[I]) > self._x_thresh:
                A.add(-1.0j * self._db * x[I], self._sig_ops[spI].terms()[0][1])
                self._n_classical_params += 1
        eiA, phase = trotterize(A, trotter_number=self._trotter_number)
        self._total_phase *= phase
        self._Uqite.add(eiA)
        self._qc.apply_circuit(eiA)
        self._Ekb.append(np.real(self._qc.direct_op_exp_val(self._qb_ham)))
        self._n_cnot += eiA.get_num_cnots()

    def evolve(self):
        self._Uqite.add(self._Uprep)
        self._qc = _core.Computer(self._nqb)
        self._qc.apply_circuit(self._Uqite)
        if self._do_lanczos:
            self._lanczos_vecs = [self._qc.get_coeff_vec()]
            h = _core.Computer(self._qc)
            h.apply_operator(self._qb_ham)
            self._Hlanczos_vecs = [h.get_coeff_vec()]
        for kb in range(1, self._nbeta):
            self.do_qite_step()
            if self._do_lanczos and kb % self._lanczos_gap == 0:
                self._lanczos_vecs.append(self._qc.get_coeff_vec())
                h = _core.Computer(self._qc)
                h.apply_operator(self._qb_ham)
                self._Hlanczos_vecs.append(h.get_coeff_vec())
            print(f" {kb * self._db:7.3f}    {self._Ekb[kb]:+15.9f}")
        self._Egs = self._Ekb[-1]

    def do_qlanczos(self):
        """Quantum Lanczos on the saved QITE states (qite.py:444-491)."""
        n = len(self._lanczos_vecs)
        h = np.zeros((n, n), dtype=complex)
        s = np.zeros((n, n), dtype=complex)
        evals = None
        for m in range(n):
            for k in range(m + 1):
                h[m, k] = np.vdot(self._lanczos_vecs[m], self._Hlanczos_vecs[k])
                h[k, m] = np.conj(h[m, k])
                s[m, k] = np.vdot(self._lanczos_vecs[m], self._lanczos_vecs[k])
                s[k, m] = np.conj(s[m, k])
            evals, _ = canonical_geig_solve(s[: m + 1, : m + 1], h[: m + 1, : m + 1], sort_ret_vals=True)
        self._Egs_lanczos = evals[0]
