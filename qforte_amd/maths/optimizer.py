"""Amplitude solvers for the UCC drivers.

Two drivers of the reference's interface (src/qforte/maths/optimizer.py:40,
:119): ``jacobi_solver(driver)`` and ``scipy_solver(driver, objective)``.

The Jacobi solver is the quasi-Newton fixed-point iteration of projective and
variational coupled-cluster solvers, t <- t + r(t) / D, where r is the
residual vector (PQE) or the energy gradient (VQE) and D the Moller-Plesset
orbital-energy denominators, so that a diagonal approximation of the
Jacobian scales each step.  Each step is accelerated by Pulay's direct
inversion in the iterative subspace (P. Pulay, Chem. Phys. Lett. 73, 393
(1980)): over the last m trial vectors t_i with error vectors e_i (here the
step that produced t_i), find weights c minimising |sum_i c_i e_i|^2 subject
to sum_i c_i = 1, i.e. solve the bordered system

    [ B   1 ] [ c ]   [ 0 ]
    [ 1^T 0 ] [ l ] = [ 1 ],     B_ij = <e_i, e_j>,

and continue from sum_i c_i t_i.  B is divided by its largest entry first,
which leaves c unchanged and keeps the system well scaled as the errors
shrink.  The window holds ``diis_max_dim - 1`` pairs, the same subspace the
reference keeps, so iterations follow the same path.

The scipy solver hands the objective (the energy for VQE, the squared
residual norm for PQE) to ``scipy.optimize.minimize`` with a per-method
convergence option set to the driver's threshold.

Both record the driver's resource counters (classical parameters, Pauli-term
measurements, CNOT counts) that the reference drivers report.
"""

from collections import deque

import numpy as np
from scipy.optimize import minimize


class PulayDIIS:
    """Bounded history of (trial vector, error vector) pairs and the Pulay
    extrapolation over it."""

    def __init__(self, max_pairs):
        self.pairs = deque(maxlen=max(1, int(max_pairs)))

    def push(self, trial, error):
        self.pairs.append((np.asarray(trial, dtype=float).copy(), np.real(np.asarray(error)).astype(float)))

    def extrapolate(self):
        m = len(self.pairs)
        errs = np.array([e for _, e in self.pairs])
        gram = errs @ errs.T
        scale = np.abs(gram).max()
        if scale > 0.0:
            gram = gram / scale
        system = np.zeros((m + 1, m + 1))
        system[:m, :m] = gram
        system[:m, m] = 1.0
        system[m, :m] = 1.0
        rhs = np.zeros(m + 1)
        rhs[m] = 1.0
        # least squares: an exhausted subspace makes B singular
        weights = np.linalg.lstsq(system, rhs, rcond=None)[0][:m]
        trials = np.array([t for t, _ in self.pairs])
        return weights @ trials


def _is_projective(driver):
    return type(driver).__name__ in ("UCCNPQE", "SPQE")


def _record_resources(driver, iterations, per_solve_list=True):
    """Counters the reference drivers report after a solve: the number of
    classical parameters, Pauli-term measurements (Nl terms per energy, per
    gradient or residual component, per iteration) and the ansatz CNOT count.
    ADAPT keeps the measurements of each Jacobi solve in a list as well."""
    nparam = len(driver._tamps)
    driver._n_classical_params = nparam
    per_solve = driver._Nl * iterations * (2 * nparam + 1)
    if hasattr(driver, "_n_pauli_trm_measures"):
        driver._n_pauli_trm_measures = (driver._n_pauli_trm_measures or 0) + per_solve
    driver._n_cnot = driver.build_Uvqc().get_num_cnots()
    if hasattr(driver, "_n_cnot_lst"):
        driver._n_cnot_lst.append(driver._n_cnot)
    if per_solve_list and hasattr(driver, "_n_pauli_trm_measures_lst"):
        driver._n_pauli_trm_measures_lst.append(per_solve)


def jacobi_solver(driver):
    """Jacobi iteration with DIIS until the residual (PQE) or gradient (VQE)
    norm falls below ``driver._opt_thresh`` or ``driver._opt_maxiter`` steps."""
    projective = _is_projective(driver)

    def scaled_step(t):
        r = driver.get_residual_vector(t) if projective else driver.gradient_ary_feval(t)
        return np.real(np.asarray(driver.get_res_over_mpdenom(r), dtype=complex)).astype(float)

    use_diis = driver._diis_max_dim >= 2
    history = PulayDIIS(driver._diis_max_dim - 1)
    t = np.asarray(driver._tamps, dtype=float)
    energy = driver.energy_feval(list(t))
    step_count = 0
    converged = False
    while step_count < driver._opt_maxiter:
        step_count += 1
        step = scaled_step(list(t))
        trial = t + step
        history.push(trial, step)
        t = history.extrapolate() if use_diis else trial
        previous, energy = energy, driver.energy_feval(list(t))
        norm = driver._res_vec_norm if projective else driver._curr_grad_norm
        print(f"     {step_count:7}        {energy:+12.10f}      {energy - previous:+12.10f}      {norm:+12.10f}")
        if norm < driver._opt_thresh:
            converged = True
            break
    driver._tamps = [float(v) for v in t]
    driver._Egs = energy
    if not converged and step_count == driver._opt_maxiter:
        print("\nMaximum number of Jacobi iterations reached!")
    if hasattr(driver, "_energies"):
        driver._energies.append(energy)
    _record_resources(driver, step_count)


# convergence option each scipy method reads, set to the driver's threshold
_TOLERANCE_OPTIONS = {
    "bfgs": ("gtol",),
    "cg": ("gtol",),
    "l-bfgs-b": ("gtol", "ftol"),
    "nelder-mead": ("fatol",),
    "powell": ("ftol",),
    "slsqp": ("ftol",),
}


def scipy_solver(driver, objective):
    """scipy.optimize.minimize of ``objective`` from the driver's amplitudes;
    the driver's ``report_iteration`` is the per-iteration callback."""
    method = driver._optimizer.lower()
    options = {"disp": False, "maxiter": driver._opt_maxiter}
    for key in _TOLERANCE_OPTIONS.get(method, ()):
        options[key] = driver._opt_thresh
    if method == "nelder-mead":
        options["adaptive"] = True
    start = [float(v) for v in driver._tamps]
    driver._prev_energy = driver.energy_feval(start)
    driver._k_counter = 0
    result = minimize(objective, start, method=driver._optimizer, options=options, callback=driver.report_iteration)
    if not result.success:
        print("  => WARNING: minimization result may not be tightly converged.")
    driver._tamps = [float(v) for v in result.x]
    driver._Egs = driver.energy_feval(driver._tamps)
    if hasattr(driver, "_energies"):
        driver._energies.append(driver._Egs)
    _record_resources(driver, driver._k_counter, per_solve_list=False)
