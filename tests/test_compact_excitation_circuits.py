"""Compact excitation circuits (reference utils/compact_excitation_circuits.py)
against the mathematics they implement: every circuit equals exp(theta kappa)
with kappa the Jordan-Wigner (or qubit) image of the pool operator, built here
from a_p = 1/2 (X_p + i Y_p) Z_{<p} independently of the package's JW; and the
CNOT counts of qubit excitations follow the closed form of the reference's
tests/test_feb_qeb_circuits.py:83-90."""

import numpy as np
import pytest
import scipy.linalg as sla

from qforte_amd.utils.compact_excitation_circuits import compact_excitation_circuit


def circuit_matrix(oracle, n, circ):
    gates = [oracle.G(g.gate_id(), g.target(), g.control(), g.parameter() or 0.0) for g in circ.gates()]
    return oracle.circuit_matrix(n, gates)


def kappa_matrix(oracle, n, terms, qubit):
    """sum_k c_k prod(a^dag cre) prod(a ann) with or without the Z strings."""
    if not qubit:
        return oracle.jw_dense(terms, n)
    dim = 1 << n
    out = np.zeros((dim, dim), dtype=complex)

    def single(p, dagger):
        mx = oracle.operator_matrix(n, [(1.0, [("X", p)])])
        my = oracle.operator_matrix(n, [(1.0, [("Y", p)])])
        return 0.5 * (mx - 1j * my) if dagger else 0.5 * (mx + 1j * my)

    for c, cre, ann in terms:
        m = np.eye(dim, dtype=complex)
        for p in cre:
            m = m @ single(p, True)
        for p in ann:
            m = m @ single(p, False)
        out += c * m
    return out


POOLS = [("SD", 2, 6), ("GSD", 2, 6), ("SDT", 3, 6), ("SDTQ", 4, 8)]


@pytest.mark.parametrize("pool_type,nocc,nqb", POOLS)
@pytest.mark.parametrize("qubit", [False, True])
def test_compact_circuit_is_the_exponential(qf, oracle, pool_type, nocc, nqb, qubit):
    pool = qf.SQOpPool()
    pool.set_orb_spaces([1] * nocc + [0] * (nqb - nocc))
    pool.fill_pool(pool_type)
    rng = np.random.default_rng(11)
    ops = [op for _, op in pool]
    stride = max(1, len(ops) // 24)  # every rank is kept: the highest ranks come last
    ops = ops[::stride] + ops[-3:]
    checked = 0
    for op in ops:
        terms = op.terms()
        t = float(rng.uniform(-1.2, 1.2))
        coeff, cre, ann = terms[1]
        U = circuit_matrix(oracle, nqb, compact_excitation_circuit(t * coeff, cre, ann, qubit))
        E = sla.expm(t * kappa_matrix(oracle, nqb, terms, qubit))
        assert np.abs(U - E).max() < 1e-12, (terms, qubit)
        checked += 1
    assert checked > 0


def test_sa_sd_halves_and_gradient_forms(qf, oracle):
    # the ansatz splits spin-adapted operators into (c, cre, ann) - (c, ann, cre)
    # (ansatz.py:43-58); the gradient applies the de-excitation form
    # compact(t c, ann, cre) = exp(-t kappa) (uccvqeabc.py:181-196)
    nqb = 6
    pool = qf.SQOpPool()
    pool.set_orb_spaces([1, 1, 0, 0, 0, 0])
    pool.fill_pool("sa_SD")
    for _, op in pool:
        terms = op.terms()
        for coeff, cre, ann in terms[: len(terms) // 2]:
            for qubit in (False, True):
                half = [(coeff, cre, ann), (-coeff, ann, cre)]
                K = kappa_matrix(oracle, nqb, half, qubit)
                fwd = circuit_matrix(oracle, nqb, compact_excitation_circuit(0.3 * -coeff, ann, cre, qubit))
                assert np.abs(fwd - sla.expm(0.3 * K)).max() < 1e-12
                rev = circuit_matrix(oracle, nqb, compact_excitation_circuit(0.3 * coeff, ann, cre, qubit))
                assert np.abs(rev - sla.expm(-0.3 * K)).max() < 1e-12


@pytest.mark.parametrize("rank", [1, 2, 3])
def test_qubit_excitation_cnot_count(rank):
    # tests/test_feb_qeb_circuits.py:88-90: 2^(2r-1) + 4r - 2 CNOTs per rank-r qubit excitation
    cre = list(range(2 * rank, 3 * rank))
    ann = list(range(rank))
    circ = compact_excitation_circuit(0.1, cre, ann, True)
    assert circ.get_num_cnots() == (1 << (2 * rank - 1)) + 4 * rank - 2


def test_rejects_non_conserving_and_double_repeats():
    with pytest.raises(ValueError, match="particle-number-conserving"):
        compact_excitation_circuit(0.1, [3, 2], [0], False)
    with pytest.raises(ValueError, match="p\\^ q\\^ p q"):
        compact_excitation_circuit(0.1, [1, 0], [1, 0], False)
