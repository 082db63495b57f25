"""Writes golden vectors from the reference C++ core compiled into oracle/_ref.

c1_reference.npz: the 12-qubit plain workload (random_circuit(12, 500, seed=7),
random_pauli_sum(12, 200, seed=8)) run through the reference Computer:
final amplitudes and <H>.

nt26_reference.npz: a 26-qubit sample of the 30-qubit roofline workload (2
brickwork layers, 36 Pauli exponentials of weight 4-8 as the reference's own
gate ladders, a 48-term Pauli sum) run through the reference Computer: 4096
amplitudes at seeded indices, <H> and the squared norm.  At 26 qubits the
engine takes its non-temporal kernel variants (states beyond the caches).
About a minute and ~8 GB of host memory.

Run from the repository root after ``make -C oracle ref``:
python tests/golden/make_golden.py [c1] [nt26]
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import _load_oracle, load_reference  # noqa: E402
from qforte_amd.engine import workloads as W  # noqa: E402

# the 26-qubit sample: sizes and seeds shared with tests/test_gpu_nt26.py
NT26 = {"n": 26, "layers": 2, "nexp": 36, "nterms": 48, "seeds": (301, 302, 303), "sample_seed": 304, "nsample": 4096}


def nt26_workload():
    """(layers, strings, terms) of the 26-qubit sample as workload lists."""
    c = NT26
    layers = W.brickwork_layers(c["n"], c["layers"], seed=c["seeds"][0])
    strings = W.random_pauli_strings(c["n"], c["nexp"], seed=c["seeds"][1])
    terms = W.random_pauli_sum(c["n"], c["nterms"], seed=c["seeds"][2])
    return layers, strings, terms


def nt26_indices():
    c = NT26
    return np.sort(np.random.default_rng(c["sample_seed"]).choice(1 << c["n"], c["nsample"], replace=False))


def write_nt26(ref):
    oracle = _load_oracle()
    n = NT26["n"]
    layers, strings, terms = nt26_workload()
    qc = ref.Computer(n)
    qc.apply_circuit(W.to_circuit(ref, layers))
    ladders = ref.Circuit()
    for theta, term in strings:
        for g in oracle.exponentiate_gates(-1j * theta, term):
            ladders.add(ref.gate(g.gtype, g.target, g.control, g.parameter) if g.gtype in ("Rx", "Ry", "Rz")
                        else ref.gate(g.gtype, g.target, g.control))
    qc.apply_circuit(ladders)
    e = qc.direct_op_exp_val(W.to_operator(ref, terms))
    amps = np.array(qc.get_coeff_vec(), dtype=np.complex128)
    idx = nt26_indices()
    out = os.path.join(ROOT, "tests", "golden", "nt26_reference.npz")
    np.savez_compressed(out, indices=idx, amplitudes=amps[idx], energy=np.array([e.real, e.imag]),
                        norm2=np.array([float(np.vdot(amps, amps).real)]))
    print("wrote", out, e)


def main():
    ref = load_reference()
    if ref is None:
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    which = sys.argv[1:] or ["c1", "nt26"]
    if "nt26" in which:
        write_nt26(ref)
    if "c1" not in which:
        return
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
