"""Quasi-Newton (Jacobi) amplitude solver with DIIS acceleration and the
scipy minimisation of projective residuals (reference
src/qforte/maths/optimizer.py:6-170)."""

import copy

import numpy as np
from scipy.optimize import minimize


def diis(diis_max_dim, t_diis, e_diis):
    """Direct inversion in the iterative subspace (optimizer.py:6-37)."""
    if len(t_diis) > diis_max_dim:
        del t_diis[0]
        del e_diis[0]
    dim = len(t_diis) - 1
    B = -np.ones((dim + 1, dim + 1))
    rhs = np.zeros(dim + 1)
    B[-1, -1] = 0.0
    rhs[-1] = -1.0
    for i, ei in enumerate(e_diis):
        for j, ej in enumerate(e_diis):
            B[i, j] = np.dot(np.real(ei), np.real(ej))
    B[:-1, :-1] /= np.abs(B[:-1, :-1]).max()
    x = np.linalg.lstsq(B, rhs, rcond=None)[0][:-1]
    t_new = np.zeros(len(t_diis[0]))
    for l in range(dim):
        t_new = t_new + x[l] * np.asarray(t_diis[l + 1])
    return list(np.real(t_new))


def jacobi_solver(self):
    """t <- t + r / D (MP denominators), DIIS-extrapolated, until ||r|| or
    ||g|| < opt_thresh (optimizer.py:40-116)."""
    pqe = self.__class__.__name__ in ("UCCNPQE", "SPQE")
    t_diis = [copy.deepcopy(self._tamps)]
    e_diis = []
    Ek0 = self.energy_feval(self._tamps)
    Ek = Ek0
    k = 0
    for k in range(1, self._opt_maxiter + 1):
        t_old = copy.deepcopy(self._tamps)
        r_k = self.get_residual_vector(self._tamps) if pqe else self.gradient_ary_feval(self._tamps)
        r_k = self.get_res_over_mpdenom(r_k)
        self._tamps = list(np.add(self._tamps, r_k))
        t_diis.append(copy.deepcopy(self._tamps))
        e_diis.append(np.subtract(self._tamps, t_old))
        if self._diis_max_dim >= 2:
            self._tamps = diis(self._diis_max_dim, t_diis, e_diis)
        Ek = self.energy_feval(self._tamps)
        dE = Ek - Ek0
        Ek0 = Ek
        norm = self._res_vec_norm if pqe else self._curr_grad_norm
        print(f"     {k:7}        {Ek:+12.10f}      {dE:+12.10f}      {norm:+12.10f}")
        if norm < self._opt_thresh:
            break
    self._Egs = Ek
    if k == self._opt_maxiter:
        print("\nMaximum number of Jacobi iterations reached!")
    if hasattr(self, "_energies"):
        self._energies.append(Ek)
    self._n_classical_params = len(self._tamps)
    if hasattr(self, "_n_pauli_trm_measures"):
        self._n_pauli_trm_measures = (self._n_pauli_trm_measures or 0) + 2 * self._Nl * k * len(self._tamps) + self._Nl * k
    self._n_cnot = self.build_Uvqc().get_num_cnots()
    if hasattr(self, "_n_cnot_lst"):
        self._n_cnot_lst.append(self._n_cnot)
    if hasattr(self, "_n_pauli_trm_measures_lst"):
        self._n_pauli_trm_measures_lst.append(self._Nl * k * (2 * len(self._tamps) + 1))


def scipy_solver(self, function_to_minimize):
    """scipy.optimize.minimize of ``function_to_minimize`` (the squared
    residual norm for PQE) with the reference's per-method tolerances and
    bookkeeping (optimizer.py:119-170)."""
    method = self._optimizer.lower()
    opts = {"disp": False, "maxiter": self._opt_maxiter}
    if method in ("bfgs", "cg", "l-bfgs-b"):
        opts["gtol"] = self._opt_thresh
    if method == "nelder-mead":
        opts["fatol"] = self._opt_thresh
        opts["adaptive"] = True
    if method in ("powell", "l-bfgs-b", "slsqp"):
        opts["ftol"] = self._opt_thresh
    x0 = copy.deepcopy(self._tamps)
    self._prev_energy = self.energy_feval(x0)
    self._k_counter = 0
    res = minimize(function_to_minimize, x0, method=self._optimizer, options=opts, callback=self.report_iteration)
    if not res.success:
        print("  => WARNING: minimization result may not be tightly converged.")
    self._tamps = list(res.x)
    self._Egs = self.energy_feval(self._tamps)
    if hasattr(self, "_energies"):
        self._energies.append(self._Egs)
    self._n_classical_params = len(self._tamps)
    if hasattr(self, "_n_pauli_trm_measures"):
        self._n_pauli_trm_measures = ((self._n_pauli_trm_measures or 0) + 2 * self._Nl * self._k_counter *
                                      len(self._tamps) + self._Nl * self._k_counter)
    self._n_cnot = self.build_Uvqc().get_num_cnots()
    if hasattr(self, "_n_cnot_lst"):
        self._n_cnot_lst.append(self._n_cnot)
