"""Generalised eigenproblem H C = S C e by canonical orthogonalisation
(Szabo & Ostlund eq. 3.169ff; reference src/qforte/maths/eigsolve.py)."""

import numpy as np
from scipy import linalg


def canonical_geig_solve(S, H, print_mats=False, sort_ret_vals=False, threshold=1e-15):
    s, U = linalg.eigh(S)
    keep = s > threshold
    X = U[:, keep] / np.sqrt(s[keep])
    Hp = X.conj().T @ H @ X
    e, Cp = linalg.eig(Hp)
    C = X @ Cp
    if sort_ret_vals:
        order = np.argsort(np.real(e))
        e, C = e[order], C[:, order]
    return e, C
