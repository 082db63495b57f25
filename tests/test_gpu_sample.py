"""Computational-basis sampling on the device (qfe_sample) and the Computer
measurement methods built on it (reference computer.cc:135-315).

The reference seeds std::mt19937 from std::random_device, so its draws are
not reproducible; parity is the distribution: point masses at chunk and
sub-block edges, a chi-square test of 2^20 draws against |psi|^2 at 10 qubits,
per-qubit marginals of a 26-qubit product state, and the readout / eigenvalue
bookkeeping against states whose outcomes are fixed."""

import ctypes
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qforte_amd", "libqfe.so")


@pytest.fixture(scope="module")
def lib(qf):
    lib = ctypes.CDLL(LIB)
    lib.qfe_last_error.restype = ctypes.c_char_p
    assert lib.qfe_init(0) == 0
    return lib


class State:
    def __init__(self, lib, n):
        self.lib, self.n = lib, n
        self.h = ctypes.c_void_p()
        assert lib.qfe_state_create(n, ctypes.byref(self.h)) == 0, lib.qfe_last_error()

    def upload(self, psi):
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        assert self.lib.qfe_state_upload(self.h, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(a.size)) == 0

    def set_point(self, idx):
        i = (ctypes.c_uint64 * 1)(idx)
        v = (ctypes.c_double * 2)(1.0, 0.0)
        assert self.lib.qfe_state_set_amps(self.h, i, v, ctypes.c_uint64(1)) == 0

    def sample(self, count, seed):
        out = np.zeros(count, dtype=np.uint64)
        rc = self.lib.qfe_sample(self.h, ctypes.c_uint64(count), ctypes.c_uint64(seed),
                                 out.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0, self.lib.qfe_last_error()
        return out

    def __del__(self):
        self.lib.qfe_state_destroy(self.h)


@pytest.mark.parametrize("n", [1, 3, 6, 7, 14, 15, 20])
def test_point_masses_at_block_edges(lib, n):
    N = 1 << n
    s = State(lib, n)
    for idx in sorted({0, 63, 64, N // 2 - 1, N // 2, (1 << 14) - 1, 1 << 14, N - 1}):
        if idx >= N:
            continue
        print(f"point mass n={n} idx={idx}", flush=True)
        s.set_point(idx)
        out = s.sample(1000, seed=idx + 1)
        assert (out == idx).all(), (n, idx)


def test_chi_square_against_probabilities(lib):
    n = 10
    rng = np.random.default_rng(3)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi[rng.random(1 << n) < 0.2] = 0.0  # zero-mass amplitudes are never drawn
    psi /= np.linalg.norm(psi)
    p = np.abs(psi) ** 2
    s = State(lib, n)
    s.upload(psi)
    S = 1 << 20
    out = s.sample(S, seed=12345)
    assert out.max() < (1 << n)
    counts = np.bincount(out.astype(np.int64), minlength=1 << n)
    assert counts[p == 0].sum() == 0
    m = S * p > 5
    chi2 = float((((counts - S * p) ** 2)[m] / (S * p)[m]).sum())
    dof = int(m.sum()) - 1
    assert chi2 < dof + 6.0 * math.sqrt(2.0 * dof), (chi2, dof)
    # the seed fixes the draw; another seed gives another draw
    assert (s.sample(S, seed=12345) == out).all()
    assert (s.sample(S, seed=54321) != out).any()


def test_marginals_of_a_26_qubit_product_state(qf, lib):
    n = 26
    comp = qf.Computer(n)
    angles = [0.3 + 0.1 * q for q in range(n)]
    for q, a in enumerate(angles):
        comp.apply_gate(qf.gate("Ry", q, q, a))
    p1 = np.array([math.sin(a / 2) ** 2 for a in angles])
    S = 1 << 16
    reads = np.asarray(comp.measure_z_readouts_fast(0, n - 1, S), dtype=np.float64)
    assert reads.shape == (S, n)
    freq = reads.mean(axis=0)
    sigma = np.sqrt(p1 * (1 - p1) / S)
    assert (np.abs(freq - p1) < 6 * sigma).all(), np.abs(freq - p1) / sigma


def test_measure_circuit_eigenvalues_and_perfect_measure(qf, oracle):
    n = 3
    # |+> (x) |-i> (x) |1>: X_0 -> +1, Y_1 -> -1, Z_2 -> -1 with certainty
    comp = qf.Computer(n)
    comp.apply_gate(qf.gate("H", 0, 0))
    comp.apply_gate(qf.gate("X", 1, 1))
    comp.apply_gate(qf.gate("H", 1, 1))
    comp.apply_gate(qf.gate("S", 1, 1))
    comp.apply_gate(qf.gate("X", 2, 2))
    before = np.asarray(comp.get_coeff_vec())
    for text, value in [("X_0", 1.0), ("Y_1", -1.0), ("Z_2", -1.0), ("X_0 Y_1 Z_2", 1.0), ("X_0 Z_2", -1.0)]:
        circ = qf.build_circuit(text)
        vals = comp.measure_circuit(circ, 200)
        assert set(vals) == {value}, text
        assert comp.perfect_measure_circuit(circ) == pytest.approx(value, abs=1e-14)
        reads = comp.measure_readouts(circ, 50)
        assert all(len(r) == len(circ.gates()) for r in reads)
        assert all(r == reads[0] for r in reads)
    # measuring leaves the state as it was
    assert np.array_equal(np.asarray(comp.get_coeff_vec()), before)


def test_perfect_measure_matches_the_dense_rotation(qf, oracle):
    # sum_k prod_g (1 - 2 bit_g(k)) |a'_k|^2 after Z -> I, X -> H, Y -> Rx(pi/2)
    # (computer.cc:275-315); every gate of the circuit contributes its sign
    n = 5
    rng = np.random.default_rng(21)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    comp = qf.Computer(n)
    comp.set_coeff_vec(psi)
    for text in ["Z_0", "X_1", "Y_2", "X_0 Y_3 Z_4", "Y_1 Y_2 X_4", "Z_0 Z_0"]:
        circ = qf.build_circuit(text)
        rot = []
        for g in circ.gates():
            if g.gate_id() == "X":
                rot.append(oracle.G("H", g.target()))
            elif g.gate_id() == "Y":
                rot.append(oracle.G("Rx", g.target(), parameter=np.pi / 2))
        a = oracle.apply_circuit(psi.copy(), n, rot)
        k = np.arange(1 << n)
        v = np.ones(1 << n)
        for g in circ.gates():
            v *= 1.0 - 2.0 * ((k >> g.target()) & 1)
        expect = float(np.sum(v * np.abs(a) ** 2))
        assert comp.perfect_measure_circuit(circ) == pytest.approx(expect, abs=1e-13), text
        vals = np.asarray(comp.measure_circuit(circ, 1 << 16))
        assert abs(vals.mean() - expect) < 6.0 * math.sqrt(max(1e-12, 1.0 - expect**2) / (1 << 16)) + 1e-12, text


def test_zero_norm_state_is_an_error(lib):
    s = State(lib, 4)
    assert lib.qfe_state_zero(s.h) == 0
    out = np.zeros(4, dtype=np.uint64)
    assert lib.qfe_sample(s.h, ctypes.c_uint64(4), ctypes.c_uint64(1), out.ctypes.data_as(ctypes.c_void_p)) != 0
    assert b"zero norm" in lib.qfe_last_error()


def test_computer_sample_binding_is_seeded_and_in_range(qf):
    comp = qf.Computer(12)
    for q in range(12):
        comp.apply_gate(qf.gate("H", q, q))
    a = comp.sample(4096, 99)
    b = comp.sample(4096, 99)
    assert a.dtype == np.uint64 and a.shape == (4096,)
    assert (a == b).all() and int(a.max()) < (1 << 12)
    # uniform superposition: every outcome equally likely
    counts = np.bincount(a.astype(np.int64), minlength=1 << 12)
    assert counts.max() < 12
