"""UCC-VQE base class (reference src/qforte/abc/uccvqeabc.py:22-344)."""

import numpy as np

from .. import _core
from ..utils.trotterization import trotterize
from ..utils.compact_excitation_circuits import compact_excitation_circuit
from .algorithm import AnsatzAlgorithm, UCC


class UCCVQE(UCC, AnsatzAlgorithm):
    """Variational minimisation of E(t) = <ref|U(t)^dag H U(t)|ref> over a
    disentangled UCC ansatz; analytic gradients by the adjoint recursion."""

    def __init__(self, *args, **kwargs):
        self.computer_initializable = True
        super().__init__(*args, **kwargs)

    # --- gradients -------------------------------------------------------------------------
    def measure_gradient(self, params=None):
        """dE/dt_mu for every operator of the ansatz (uccvqeabc.py:122-251).

        With the compiled program: one forward sweep, sigma = H psi, and a
        reverse sweep on device.  Otherwise the reference's recursion on the
        Computer API."""
        if not self._fast:
            raise ValueError("self._fast must be True for gradient measurement.")
        amps = list(self._tamps if params is None else params)
        if self._can_use_program():
            _, g = self._get_program().gradient([float(x) for x in amps])
            return np.asarray(g)
        M = len(self._tamps)
        grads = np.zeros(M)
        Utot = self.build_Uvqc(amps)
        qc_psi = self.get_initial_computer()
        qc_psi.apply_circuit(Utot)
        qc_sig = _core.Computer(qc_psi)
        qc_sig.apply_operator(self._qb_ham)
        mu = M - 1
        K_prev = self._pool_obj[self._tops[mu]][1].jw_transform(self._qubit_excitations)
        K_prev.mult_coeffs(self._pool_obj[self._tops[mu]][0])
        qc_temp = _core.Computer(qc_psi)
        qc_temp.apply_operator(K_prev)
        grads[mu] = 2.0 * np.real(np.vdot(qc_sig.get_coeff_vec(), qc_temp.get_coeff_vec()))
        for mu in reversed(range(M - 1)):
            K = self._pool_obj[self._tops[mu]][1].jw_transform(self._qubit_excitations)
            K.mult_coeffs(self._pool_obj[self._tops[mu]][0])
            if self._compact_excitations:
                # exp(-t kappa) of the next operator; for sa_SD the de-excitation
                # form carries the minus sign (uccvqeabc.py:181-207)
                Umu = _core.Circuit()
                terms = self._pool_obj[self._tops[mu + 1]][1].terms()
                if self._pool_type == "sa_SD":
                    for coeff, cr, ann in terms[: len(terms) // 2]:
                        Umu.add(compact_excitation_circuit(amps[mu + 1] * coeff, ann, cr, self._qubit_excitations))
                else:
                    coeff, cr, ann = terms[1]
                    Umu.add(compact_excitation_circuit(-amps[mu + 1] * coeff, cr, ann, self._qubit_excitations))
            elif self._pool_type == "sa_SD":
                Umu = _core.Circuit()
                terms = self._pool_obj[self._tops[mu + 1]][1].terms()
                for coeff, cr, ann in terms[: len(terms) // 2]:
                    sq = _core.SQOperator()
                    sq.add_term(coeff, cr, ann)
                    sq.add_term(-coeff, ann, cr)
                    U, _ = trotterize(sq.jw_transform(self._qubit_excitations), factor=-amps[mu + 1],
                                      trotter_number=self._trotter_number)
                    Umu.add(U)
            else:
                Umu, pmu = trotterize(K_prev, factor=-amps[mu + 1], trotter_number=self._trotter_number)
                if pmu != 1.0 + 0.0j:
                    raise ValueError("Encountered phase change, phase not equal to (1.0 + 0.0i)")
            qc_sig.apply_circuit(Umu)
            qc_psi.apply_circuit(Umu)
            qc_temp = _core.Computer(qc_psi)
            qc_temp.apply_operator(K)
            grads[mu] = 2.0 * np.real(np.vdot(qc_sig.get_coeff_vec(), qc_temp.get_coeff_vec()))
            K_prev = K
        return grads

    def measure_gradient3(self):
        """2 Re <H psi | K_mu psi> for every pool operator (uccvqeabc.py:253-283),
        batched 64 operators per launch."""
        if not self._fast:
            raise ValueError("self._fast must be True for gradient measurement.")
        ops = []
        for coeff, operator in self._pool_obj:
            K = operator.jw_transform(self._qubit_excitations)
            K.mult_coeffs(coeff)
            ops.append(K)
        if self._can_use_program():
            return np.asarray(self._get_program().pool_gradients([float(x) for x in self._tamps], ops, 64))
        Utot = self.build_Uvqc()
        qc_psi = self.get_initial_computer()
        qc_psi.apply_circuit(Utot)
        qc_sig = _core.Computer(qc_psi)
        qc_sig.apply_operator(self._qb_ham)
        sig = qc_sig.get_coeff_vec()
        grads = np.zeros(len(ops))
        for mu, K in enumerate(ops):
            qc_temp = _core.Computer(qc_psi)
            qc_temp.apply_operator(K)
            grads[mu] = 2.0 * np.real(np.vdot(sig, qc_temp.get_coeff_vec()))
        return grads

    def gradient_ary_feval(self, params):
        grads = np.asarray(self.measure_gradient(params))
        self._curr_grad_norm = float(np.linalg.norm(grads))
        self._res_vec_evals += 1
        self._res_m_evals += len(self._tamps)
        return grads

    def report_iteration(self, x):
        self._k_counter += 1
        if self._k_counter == 1:
            print("\n    k iteration         Energy               dE           Ngvec ev      Ngm ev*         ||g||")
            print("-" * 100)
        dE = self._curr_energy - self._prev_energy
        print(f"     {self._k_counter:7}        {self._curr_energy:+12.10f}      {dE:+12.10f}      "
              f"{self._res_vec_evals:4}        {self._res_m_evals:6}       {self._curr_grad_norm:+12.10f}")
        self._prev_energy = self._curr_energy
