"""The non-temporal kernel variants against the compiled reference.

From 24 qubits on the state no longer fits the caches and the engine switches
its streaming kernels to non-temporal loads and stores (the LDS tile pass,
the register and lane orbits, the expectation pass): the code the 30-qubit
roofline workload runs.  A norm check cannot see a wrong phase or sign, so
this test runs a 26-qubit sample of that workload (2 brickwork layers, 36
Pauli exponentials of weight 4-8, a 48-term Pauli sum) and compares 4096
amplitudes and <H> with ``tests/golden/nt26_reference.npz``, which the
reference Computer compiled from its sources wrote by applying the
reference's own gate ladders (``tests/golden/make_golden.py``).  Each tuning
routes the exponentials through a different orbit kernel.

Tolerances: amplitudes 1e-12 absolute, energy 1e-10 Ha, norm 1e-12.
"""

import importlib.util
import os

import numpy as np
import pytest

from qforte_amd.engine import workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _generator():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture
def restore_tuning(qf, gpu):
    saved, dims = qf.get_tuning(), qf.get_pass_tuning()
    yield
    qf.set_tuning(*saved)
    qf.set_pass_tuning(dims)


# (orbit dimensions, (tile qubits, low qubits, ops per pass), counter that must be non-zero)
CASES = [
    (7, (11, 4, 24), "lane_orbit_passes"),  # k_orbit_lanes<4, 3, NT>, k_fused<11>
    (6, (11, 4, 24), "lane_orbit_passes"),  # k_orbit_lanes<4, 2, NT>, k_fused<11>
    (5, (13, 4, 64), "lane_orbit_passes"),  # k_orbit_lanes<4, 1, NT>, k_fused<13>
    (4, (11, 4, 24), "orbit_passes"),       # k_orbit<M <= 4, NT>
]


@pytest.mark.parametrize("orbit_dims,tuning,kind", CASES)
def test_nt26_sample_matches_reference(qf, restore_tuning, orbit_dims, tuning, kind):
    gen = _generator()
    gold = np.load(os.path.join(GOLDEN, "nt26_reference.npz"))
    n = gen.NT26["n"]
    idx = gen.nt26_indices()
    assert np.array_equal(idx, gold["indices"])
    layers, strings, terms = gen.nt26_workload()
    qf.set_tuning(*tuning)
    qf.set_pass_tuning(orbit_dims)
    print(f"nt26: orbit_dims={orbit_dims} tuning={tuning}", flush=True)
    qc = qf.Computer(n)
    qf.reset_counters()
    qc.apply_circuit(W.to_circuit(qf, layers))
    c = qf.get_counters()
    assert c["tile_passes"] > 0 and c["nt_passes"] == c["tile_passes"], c  # every tile pass streams non-temporally
    U, _ = qf.trotterize(W.exp_operator(qf, strings))
    qf.reset_counters()
    qc.apply_circuit(U)
    c = qf.get_counters()
    assert c[kind] > 0 and c["nt_passes"] > 0, c
    qf.reset_counters()
    e = qc.direct_op_exp_val(W.to_operator(qf, terms))
    c = qf.get_counters()
    assert c["reduce_passes"] > 0 and c["nt_passes"] > 0, c
    amps = np.asarray(qc.get_coeff_vec())
    diff = np.abs(amps[idx] - gold["amplitudes"])
    scale = np.abs(gold["amplitudes"]).max()
    print(f"nt26: max amplitude diff {diff.max():.3e} (largest sampled amplitude {scale:.3e}), "
          f"energy diff {abs(e.real - gold['energy'][0]):.3e}", flush=True)
    assert diff.max() < 1e-12
    assert diff.max() < 1e-9 * scale  # relative to the amplitudes' own size (about 2^-13)
    assert abs(e.real - gold["energy"][0]) < 1e-10
    assert abs(e.imag - gold["energy"][1]) < 1e-10
    assert abs(float(np.vdot(amps, amps).real) - gold["norm2"][0]) < 1e-12
