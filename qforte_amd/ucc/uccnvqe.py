"""UCCN-VQE driver (reference src/qforte/ucc/uccnvqe.py:17-303)."""

import copy

import numpy as np
from scipy.optimize import minimize

from ..abc.uccvqeabc import UCCVQE
from ..maths import optimizer


class UCCNVQE(UCCVQE):
    """Optimise a fixed disentangled UCCN ansatz (all pool operators)."""

    def run(self, opt_thresh=1.0e-5, opt_ftol=1.0e-5, opt_maxiter=200, pool_type="SD", optimizer="BFGS",
            use_analytic_grad=True, noise_factor=0.0):
        self._opt_thresh = opt_thresh
        self._opt_ftol = opt_ftol
        self._opt_maxiter = opt_maxiter
        self._use_analytic_grad = use_analytic_grad
        self._optimizer = optimizer
        if use_analytic_grad and optimizer in {"nelder-mead", "powell", "cobyla"}:
            self._use_analytic_grad = False
        self._pool_type = pool_type
        self._noise_factor = noise_factor
        self._tops, self._tamps = [], []
        self._converged = 0
        self._n_classical_params = 0
        self._n_cnot = 0
        self._n_pauli_trm_measures = 0
        self._res_vec_evals = 0
        self._res_m_evals = 0
        self._k_counter = 0
        self._curr_grad_norm = 0.0
        self.fill_pool()
        self.initialize_ansatz()
        self.solve()
        self._n_nonzero_params = sum(1 for t in self._tamps if abs(t) > 1.0e-12)
        self.verify_required_attributes()
        print(f"Final UCCN-VQE Energy:                       {round(self._Egs, 10)}")

    def initialize_ansatz(self):
        self._tops = list(range(len(self._pool_obj)))
        self._tamps = [0.0] * len(self._pool_obj)

    def solve(self):
        if self._optimizer.lower() == "jacobi":
            self.build_orb_energies()
            return optimizer.jacobi_solver(self)
        return self.scipy_solve()

    def scipy_solve(self):
        """scipy.optimize.minimize with the analytic gradient (uccnvqe.py:196-270)."""
        opts = {"disp": False, "maxiter": self._opt_maxiter}
        if self._optimizer in ["BFGS", "CG", "L-BFGS-B", "TNC", "trust-constr"]:
            opts["gtol"] = self._opt_thresh
        if self._optimizer == "Nelder-Mead":
            opts["fatol"] = self._opt_ftol
        if self._optimizer in ["Powell", "L-BFGS-B", "TNC", "SLSQP"]:
            opts["ftol"] = self._opt_ftol
        if self._optimizer == "COBYLA":
            opts["tol"] = self._opt_ftol
        if self._optimizer in ["L-BFGS-B", "TNC"]:
            opts["maxfun"] = self._opt_maxiter
        x0 = copy.deepcopy(self._tamps)
        self._prev_energy = self.energy_feval(x0)
        jac = self.gradient_ary_feval if self._use_analytic_grad else None
        res = minimize(self.energy_feval, x0, method=self._optimizer, jac=jac, options=opts,
                       callback=self.report_iteration)
        if self._use_analytic_grad:
            self._n_pauli_trm_measures += sum(int(2 * self._Nl * res.njev) for t in res.x if abs(t) > 1.0e-12)
        self._n_pauli_trm_measures += int(self._Nl * res.nfev)
        self._Egs = float(res.fun)
        self._final_result = res
        self._tamps = list(res.x)
        self._n_classical_params = len(self._tamps)
        self._n_cnot = self.build_Uvqc().get_num_cnots()

    def get_num_ham_measurements(self):
        return self._final_result.nfev

    def get_num_commut_measurements(self):
        return 0
