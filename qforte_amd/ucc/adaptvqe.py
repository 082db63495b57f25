"""ADAPT-VQE driver (reference src/qforte/ucc/adaptvqe.py:21-512)."""

import copy

import numpy as np
from scipy.optimize import minimize

from ..abc.uccvqeabc import UCCVQE
from ..maths import optimizer


class ADAPTVQE(UCCVQE):
    """Grow the ansatz one pool operator at a time, picking the operator with
    the largest |<psi|[H, K_mu]|psi>| (batched pool gradients on device)."""

    def run(self, avqe_thresh=1.0e-2, pool_type="sa_SD", opt_thresh=1.0e-5, opt_maxiter=200, adapt_maxiter=20,
            optimizer="BFGS", use_analytic_grad=True, use_cumulative_thresh=False, add_equiv_ops=False):
        self._avqe_thresh = avqe_thresh
        self._opt_thresh = opt_thresh
        self._adapt_maxiter = adapt_maxiter
        self._opt_maxiter = opt_maxiter
        self._use_analytic_grad = use_analytic_grad
        self._optimizer = optimizer
        if use_analytic_grad and optimizer in {"nelder-mead", "powell", "cobyla"}:
            self._use_analytic_grad = False
        self._pool_type = pool_type
        self._use_cumulative_thresh = use_cumulative_thresh
        self._add_equiv_ops = add_equiv_ops
        self._results, self._energies, self._grad_norms = [], [], []
        self._tops, self._tamps = [], []
        self._converged = 0
        self._n_classical_params = 0
        self._n_cnot = 0
        self._n_cnot_lst = []
        self._n_classical_params_lst = []
        self._n_pauli_trm_measures = 0
        self._n_pauli_trm_measures_lst = []
        self._res_vec_evals = 0
        self._res_m_evals = 0
        self._k_counter = 0
        self._curr_grad_norm = 0.0
        self.fill_pool()
        self._prev_energy = self.energy_feval([])
        it = 0
        hit_maxiter = False
        while not self._converged:
            self.update_ansatz()
            if self._converged:
                break
            self.solve()
            it += 1
            if it > self._adapt_maxiter - 1:
                hit_maxiter = True
                break
        self._Egs = self.get_final_energy(hit_max_avqe_iter=hit_maxiter)
        self._n_classical_params = len(self._tamps)
        self._n_cnot = self._n_cnot_lst[-1] if self._n_cnot_lst else 0
        self._n_pauli_trm_measures = sum(self._n_pauli_trm_measures_lst)
        self.verify_required_attributes()
        print(f"Final ADAPT-VQE Energy:                      {round(self._Egs, 10)}")

    def solve(self):
        if self._optimizer.lower() == "jacobi":
            self.build_orb_energies()
            return optimizer.jacobi_solver(self)
        return self.scipy_solve()

    def scipy_solve(self):
        """Re-optimise all amplitudes (adaptvqe.py:285-378)."""
        self._k_counter = 0
        opts = {"gtol": self._opt_thresh, "disp": False, "maxiter": self._opt_maxiter}
        x0 = copy.deepcopy(self._tamps)
        self._prev_energy = self.energy_feval(x0)
        jac = self.gradient_ary_feval if self._use_analytic_grad else None
        res = minimize(self.energy_feval, x0, method=self._optimizer, jac=jac, options=opts,
                       callback=self.report_iteration)
        self._n_pauli_measures_k += self._Nl * res.nfev
        if self._use_analytic_grad:
            for m in self._tops:
                self._n_pauli_measures_k += self._Nm[m] * self._Nl * res.njev
        self._energies.append(float(res.fun))
        self._results.append(res)
        self._tamps = list(res.x)
        self._n_pauli_trm_measures_lst.append(self._n_pauli_measures_k)
        self._n_cnot_lst.append(self.build_Uvqc().get_num_cnots())

    def update_ansatz(self):
        """Append the operator with the largest gradient (adaptvqe.py:380-458)."""
        grads = self.measure_gradient3()
        self._n_pauli_measures_k = self._Nl * sum(self._Nm[m] for m in range(len(grads)))
        self._curr_grad_norm = float(np.sqrt(sum(g * g for g in grads)))
        self._grad_norms.append(self._curr_grad_norm)
        # operators whose |gradient| set a new strict maximum in pool order:
        # the last one is the choice; the one before it, if any, is the
        # candidate that add_equiv_ops admits when their magnitudes tie
        records, top = [], 0.0
        for m, g in enumerate(grads):
            if abs(g) > top:
                top = abs(g)
                records.append(m)
        self.conv_status()
        if self._converged:
            return
        if self._use_cumulative_thresh:
            # the operators left once the smallest squared gradients summing to
            # at most thresh^2 are dropped, largest first
            order = sorted(range(len(grads)), key=lambda m: (grads[m] * grads[m], m))
            running = np.cumsum([grads[m] * grads[m] for m in order])
            chosen = [m for m, r in zip(order, running) if r > self._avqe_thresh * self._avqe_thresh][::-1]
            self._tops.extend(chosen)
            self._tamps.extend([0.0] * len(chosen))
        else:
            best = records[-1] if records else None
            self._tops.append(best)
            self._tamps.append(0.0)
            if self._add_equiv_ops and len(records) > 1 and abs(grads[best]) - abs(grads[records[-2]]) < 1.0e-5:
                self._tops.append(records[-2])
                self._tamps.append(0.0)
        self._n_classical_params_lst.append(len(self._tops))

    def conv_status(self):
        if abs(self._curr_grad_norm) < abs(self._avqe_thresh):
            self._converged = True
            self._final_energy = self._energies[-1] if self._energies else self._prev_energy
        else:
            self._converged = False

    def get_final_energy(self, hit_max_avqe_iter=False):
        if hit_max_avqe_iter:
            print("\nADAPT-VQE at maximum number of iterations!")
            self._final_energy = self._energies[-1]
        return self._final_energy

    def get_num_ham_measurements(self):
        return sum(r.nfev for r in self._results)

    def get_num_commut_measurements(self):
        return len(self._tamps) * len(self._pool_obj)
