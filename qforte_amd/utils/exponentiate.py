"""exp(c P) for a Pauli string P and purely imaginary c.

Same gate list as the reference's ``exponentiate_pauli_string``
(src/qforte/utils/exponentiate.py:9-87): basis change (H for X, Rx(pi/2) for
Y), a CNOT ladder, Rz(-2 Im c) on the last qubit, then the ladder and basis
change undone.  The construction runs natively (``_core``) and tags the gate
run as one exponential, which ``Computer.apply_circuit`` executes as a single
bit-mask permutation-and-phase pass.
"""

from .. import _core


def exponentiate_pauli_string(coefficient, term, Use_cRz=False, ancilla_idx=None, Use_open_cRz=False):
    """Return ``(Circuit, phase)`` with ``Circuit`` = exp(coefficient * term).

    Raises ``ValueError`` for a coefficient with a real part above 1e-14, as
    the reference does (exponentiate.py:25-28).  An empty term returns
    ``(Circuit(), exp(coefficient))`` (exponentiate.py:31-32).
    """
    anc = -1 if ancilla_idx is None else int(ancilla_idx)
    return _core.exponentiate_pauli_string(complex(coefficient), term, bool(Use_cRz), anc, bool(Use_open_cRz))
