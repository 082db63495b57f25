"""Seeded synthetic workloads of the hot path, as plain data.

Gates are ``(type, target, control, parameter)`` tuples and Pauli terms are
``(coeff, [(letter, qubit), ...])``; ``to_circuit`` / ``to_operator`` turn them
into objects of any module with the qforte API (this package, or a compiled
reference build for comparison).

* ``random_circuit`` / ``random_pauli_sum``: a seeded random circuit over
  {H, X, Y, Z, S, T, Rx, Ry, Rz, cX, cZ, SWAP} on uniformly random distinct
  qubits with angles U[0, 2pi), and a random Pauli sum with weight U{1..w},
  uniform letters and N(0, 1) real coefficients (12 qubits / 500 gates /
  200 terms is the plain CPU-runnable workload).
* ``brickwork_layers``, ``random_pauli_strings``: the 30-qubit roofline
  workload -- brickwork layers of per-qubit Ry(U[0, 2pi)) Rz(U[0, 2pi)) then
  cX on alternating neighbour pairs (control = lower index), followed by
  exp(-i theta P) for random strings of weight 4-8 with theta ~ N(0, 0.1).
"""

import math

import numpy as np

GATES_1Q = ("H", "X", "Y", "Z", "S", "T", "Rx", "Ry", "Rz")
GATES_2Q = ("cX", "cZ", "SWAP")
PARAM = {"Rx", "Ry", "Rz"}
LETTERS = ("X", "Y", "Z")


def random_circuit(nqubit, ngates, seed):
    rng = np.random.default_rng(seed)
    kinds = GATES_1Q + GATES_2Q
    out = []
    for _ in range(ngates):
        g = kinds[int(rng.integers(len(kinds)))]
        if g in GATES_2Q:
            t, c = (int(v) for v in rng.choice(nqubit, size=2, replace=False))
            out.append((g, t, c, 0.0))
        else:
            q = int(rng.integers(nqubit))
            out.append((g, q, q, float(rng.uniform(0.0, 2 * math.pi)) if g in PARAM else 0.0))
    return out


def random_pauli_sum(nqubit, nterms, seed, wmin=1, wmax=6, real_coeffs=True):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nterms):
        w = int(rng.integers(wmin, wmax + 1))
        qs = sorted(int(v) for v in rng.choice(nqubit, size=w, replace=False))
        letters = [LETTERS[int(rng.integers(3))] for _ in qs]
        c = float(rng.normal(0.0, 1.0))
        out.append((c if real_coeffs else complex(c), list(zip(letters, qs))))
    return out


def brickwork_layers(nqubit, nlayers, seed):
    rng = np.random.default_rng(seed)
    out = []
    for layer in range(nlayers):
        for q in range(nqubit):
            out.append(("Ry", q, q, float(rng.uniform(0.0, 2 * math.pi))))
            out.append(("Rz", q, q, float(rng.uniform(0.0, 2 * math.pi))))
        start = 0 if layer % 2 == 0 else 1
        for lo in range(start, nqubit - 1, 2):
            out.append(("cX", lo + 1, lo, 0.0))
    return out


def random_pauli_strings(nqubit, count, seed, wmin=4, wmax=8):
    """(theta, term) pairs for exp(-i theta P) with theta ~ N(0, 0.1)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        w = int(rng.integers(wmin, wmax + 1))
        qs = sorted(int(v) for v in rng.choice(nqubit, size=w, replace=False))
        letters = [LETTERS[int(rng.integers(3))] for _ in qs]
        out.append((float(rng.normal(0.0, 0.1)), list(zip(letters, qs))))
    return out


def to_circuit(mod, gates):
    c = mod.Circuit()
    for g, t, ctl, p in gates:
        if t == ctl:
            if g in PARAM:
                c.add(mod.gate(g, t, t, p))
            else:
                c.add(mod.gate(g, t, t))
        else:
            c.add(mod.gate(g, t, ctl))
    return c


def term_circuit(mod, term):
    c = mod.Circuit()
    for letter, q in term:
        c.add(mod.gate(letter, q, q))
    return c


def to_operator(mod, terms):
    op = mod.QubitOperator()
    for c, term in terms:
        op.add(c, term_circuit(mod, term))
    return op


def exp_operator(mod, strings):
    """QubitOperator whose Trotter product is prod_k exp(-i theta_k P_k)."""
    op = mod.QubitOperator()
    for theta, term in strings:
        op.add(-1j * theta, term_circuit(mod, term))
    return op


# ---- UCC serving workload ---------------------------------------------------------------


def half_filled_reference(nqubit):
    """Occupation list of the half-filled determinant of ``nqubit`` spin
    orbitals: nqubit/2 electrons, N_alpha = ceil(N/2), N_beta = floor(N/2),
    alpha on even and beta on odd spin orbitals (state_prep.py:5-11)."""
    norb = nqubit // 2
    ne = nqubit // 2
    na, nb = (ne + 1) // 2, ne // 2
    ref = [0] * nqubit
    for a in range(na):
        ref[2 * a] = 1
    for b in range(nb):
        ref[2 * b + 1] = 1
    return norb, na, nb, ref


def ucc_size_classes(mod, sizes=(8, 10, 12, 14, 16, 18, 20)):
    """One compiled UCCSD size class per qubit count: the SD pool of the
    half-filled reference, the integral map, and the sector template."""
    out = {}
    for n in sizes:
        norb, na, nb, ref = half_filled_reference(n)
        pool = mod.SQOpPool()
        pool.set_orb_spaces(ref)
        pool.fill_pool("SD")
        tops = list(range(len(pool)))
        hmap = mod.HamiltonianMap(norb)
        tpl = mod.compile_ucc_template(n, ref, hmap, pool, tops)
        out[n] = {"norb": norb, "ref": ref, "pool": pool, "tops": tops, "map": hmap, "template": tpl}
    return out


def ucc_requests(mod, classes, count, seed):
    """``count`` requests with qubit counts uniform over the classes, each
    with its own seeded synthetic integrals (h symmetric N(0, 0.1) - 2p/norb on
    the diagonal, g 8-fold symmetric N(0, 0.01) + 0.5 on g_pppp) and amplitudes
    N(0, 0.05).  Returns (n, inputs, amplitudes) tuples."""
    rng = np.random.default_rng(seed)
    sizes = sorted(classes)
    out = []
    for k in range(count):
        n = sizes[int(rng.integers(len(sizes)))]
        c = classes[n]
        h, g = mod.synthetic_integrals(c["norb"], seed * 1000003 + k)
        inputs = c["map"].pack(0.0, h, g)
        t = rng.normal(0.0, 0.05, c["template"].nparam)
        out.append((n, inputs, [float(v) for v in t]))
    return out
