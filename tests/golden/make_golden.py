"""Writes golden vectors from the reference C++ core compiled into oracle/_ref.

c1_reference.npz: the 12-qubit plain workload (random_circuit(12, 500, seed=7),
random_pauli_sum(12, 200, seed=8)) run through the reference Computer:
final amplitudes and <H>.  Run from the repository root after
``make -C oracle ref``:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import load_reference  # noqa: E402
from qforte_amd.engine import workloads as W  # noqa: E402


def main():
    ref = load_reference()
    if ref is None:
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    n = 12
    gates = W.random_circuit(n, 500, seed=7)
    terms = W.random_pauli_sum(n, 200, seed=8)
    qc = ref.Computer(n)
    qc.apply_circuit(W.to_circuit(ref, gates))
    e = qc.direct_op_exp_val(W.to_operator(ref, terms))
    amps = np.array(qc.get_coeff_vec(), dtype=np.complex128)
    out = os.path.join(ROOT, "tests", "golden", "c1_reference.npz")
    np.savez_compressed(out, amplitudes=amps, energy=np.array([e.real, e.imag]))
    print("wrote", out, e)


if __name__ == "__main__":
    main()
