"""Reference-determinant helpers (reference src/qforte/utils/state_prep.py)."""

from .. import _core


def build_refprep(ref):
    """X on every occupied spin orbital (state_prep.py:5-11)."""
    circ = _core.Circuit()
    for j, occupied in enumerate(ref):
        if occupied:
            circ.add(_core.gate("X", j, j))
    return circ


def ref_to_basis_idx(ref):
    """Basis index of an occupation list; entry q is bit q."""
    idx = 0
    for q, occ in enumerate(ref):
        if occ:
            idx |= 1 << q
    return idx


def ref_string(ref, nqb):
    """``|q0 q1 ...>`` string of an occupation list (state_prep.py:14-18)."""
    return _core.QubitBasis(ref_to_basis_idx(ref)).str(nqb)


def integer_to_ref(n, nqubits):
    """Occupation list of basis index ``n`` (state_prep.py:21-48)."""
    return [1 if (n >> i) & 1 else 0 for i in range(nqubits)]
