"""Trotter products of Pauli-sum exponentials (reference
src/qforte/utils/trotterization.py:10-128)."""

from .. import _core


def trotterize(operator, factor=1.0, trotter_number=1, trotter_order=1):
    """Circuit for prod_k exp(factor * c_k P_k), term 0 applied first.

    With ``trotter_number`` N > 1 all terms repeat N times with c_k / N
    (trotterization.py:37-56).  Returns ``(Circuit, phase)``.
    """
    if trotter_order > 1:
        raise NotImplementedError("Higher order trotterization is not yet implemented.")
    return _core.trotterize(operator, complex(factor), int(trotter_number), int(trotter_order))


def trotterize_w_cRz(operator, ancilla_qubit_idx, factor=1.0, Use_open_cRz=False, trotter_number=1, trotter_order=1):
    """Controlled version with cRz on ``ancilla_qubit_idx`` (trotterization.py:60-128)."""
    if trotter_order > 1:
        raise NotImplementedError("Higher order trotterization is not yet implemented.")
    return _core.trotterize_w_cRz(
        operator, int(ancilla_qubit_idx), complex(factor), bool(Use_open_cRz), int(trotter_number), int(trotter_order)
    )
