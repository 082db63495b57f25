"""UCCN-PQE driver (reference src/qforte/ucc/uccnpqe.py, abc/uccpqeabc.py,
abc/pqeabc.py): solve the projective conditions r_mu(t) = 0 with the Jacobi /
DIIS solver."""

import numpy as np

from .. import _core
from ..abc.algorithm import AnsatzAlgorithm, UCC
from ..maths import optimizer


class UCCNPQE(UCC, AnsatzAlgorithm):
    """r_mu = <Phi_mu| U(t)^dag H U(t) |Phi_0> * phase_mu for every pool operator."""

    def run(self, pool_type="SD", opt_thresh=1.0e-5, opt_maxiter=40, noise_factor=0.0, optimizer="jacobi"):
        if self._state_prep_type != "occupation_list":
            raise ValueError("PQE implementation can only handle occupation_list Hartree-Fock reference.")
        self._pool_type = pool_type
        self._optimizer = optimizer
        self._opt_thresh = opt_thresh
        self._opt_maxiter = opt_maxiter
        self._noise_factor = noise_factor
        self._tops, self._tamps = [], []
        self._converged = 0
        self._res_vec_evals = 0
        self._res_m_evals = 0
        self._excited_dets = []
        self._n_classical_params = 0
        self._n_cnot = 0
        self._n_pauli_trm_measures = 0
        self.fill_pool()
        self._tops = list(range(len(self._pool_obj)))
        self._tamps = [0.0] * len(self._pool_obj)
        self.fill_excited_dets()
        self.build_orb_energies()
        self.solve()
        self.verify_required_attributes()
        print(f"Final UCCN-PQE Energy:                       {round(self._Egs, 10)}")

    def solve(self):
        """Jacobi iterations, or scipy minimisation of sum_mu r_mu^2 (uccpqeabc.py:101-122)."""
        if self._optimizer.lower() == "jacobi":
            return optimizer.jacobi_solver(self)
        if self._optimizer.lower() in ("nelder-mead", "powell", "bfgs", "l-bfgs-b", "cg", "slsqp"):
            return optimizer.scipy_solver(self, self.get_sum_residual_square)
        raise NotImplementedError(
            "Currently only Jacobi, Nelder-Mead, Powell, BFGS, L-BFGS-B, CG, and SLSQP solvers are implemented")

    def get_sum_residual_square(self, tamps):
        r = np.asarray(self.get_residual_vector(tamps))
        return float(np.sum(r * r))

    def report_iteration(self, x):
        """scipy callback: one line per iteration (uccpqeabc.py:63-99)."""
        self._k_counter += 1
        if self._k_counter == 1:
            print("\n    k iteration         Energy               dE           Nrvec ev      Nrm ev*         ||r||")
        self._curr_energy = self.energy_feval(x)
        dE = self._curr_energy - self._prev_energy
        print(f"     {self._k_counter:7}        {self._curr_energy:+12.10f}      {dE:+12.10f}      "
              f"{self._res_vec_evals:4}        {self._res_m_evals:6}       {self._res_vec_norm:+12.10f}")
        self._prev_energy = self._curr_energy

    def fill_excited_dets(self):
        """Excited determinant and phase of every pool operator (uccnpqe.py:193-242)."""
        for _, sq_op in self._pool_obj:
            temp_idx = sq_op.terms()[0][2][-1]
            if self._ref[temp_idx]:
                cre, ann = sq_op.terms()[0][1], sq_op.terms()[0][2]
            else:
                cre, ann = sq_op.terms()[0][2], sq_op.terms()[0][1]
            det = _core.QubitBasis(0)
            for k, occ in enumerate(self._ref):
                det.set_bit(k, bool(occ))
            for p in reversed(ann):
                if not det.get_bit(p):
                    raise ValueError("no ops should destroy reference, something went wrong!!")
                det.set_bit(p, False)
            for p in reversed(cre):
                if det.get_bit(p):
                    raise ValueError("no ops should destroy reference, something went wrong!!")
                det.set_bit(p, True)
            I = det.index()
            qc = _core.Computer(self._nqb)
            qc.apply_circuit(self._refprep)
            qc.apply_operator(sq_op.jw_transform(self._qubit_excitations))
            self._excited_dets.append((I, qc.coeff(_core.QubitBasis(I))))

    def get_residual_vector(self, trial_amps):
        """U(t)|ref>, apply H, un-apply U, read the excited amplitudes (uccnpqe.py:249-292)."""
        if self._pool_type == "sa_SD":
            raise ValueError("Must use single term particle-hole nbody operators for residual calculation")
        if self._can_use_program():
            # U^dag H U |ref> on the compiled program, read at the excited determinants
            amps = self._get_program().residuals([float(x) for x in trial_amps], [I for I, _ in self._excited_dets])
        else:
            U = self.ansatz_circuit(trial_amps)
            qc = _core.Computer(self._nqb)
            qc.apply_circuit(self._Uprep)
            qc.apply_circuit(U)
            qc.apply_operator(self._qb_ham)
            qc.apply_circuit(U.adjoint())
            amps = [qc.coeff(_core.QubitBasis(I)) for I, _ in self._excited_dets]
        residuals = []
        for a, (I, phase) in zip(amps, self._excited_dets):
            r = a * phase
            if abs(np.imag(r)) > 1.0e-14:
                raise ValueError("residual has imaginary component, something went wrong!!")
            residuals.append(float(np.real(r)))
        self._res_vec_norm = float(np.linalg.norm(residuals))
        self._res_vec_evals += 1
        self._res_m_evals += len(self._tamps)
        return residuals

    def gradient_ary_feval(self, params):
        raise NotImplementedError
