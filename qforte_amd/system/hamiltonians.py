"""Molecular Hamiltonians from integrals.

``build_sq_hamiltonian`` restates the spin-orbital expansion of
src/qforte/adapters/molecule_adapters.py:183-210 (chemist-notation two-electron
integrals ``g[p, q, r, s] = (pq|rs)``, alpha = even and beta = odd spin
orbitals).  ``synthetic_integrals`` draws the seeded synthetic integrals the
benchmarks use: h symmetric N(0, 0.1) plus -2 p / norb on the diagonal, g
8-fold symmetric N(0, 0.01) plus 0.5 on g[p, p, p, p].
"""

import numpy as np

from .. import _core
from .molecular_info import Molecule


def synthetic_integrals(norb, seed):
    """Seeded (h, g) for ``norb`` spatial orbitals (see module docstring)."""
    rng = np.random.default_rng(seed)
    h = rng.normal(0.0, 0.1, (norb, norb))
    h = 0.5 * (h + h.T)
    h[np.diag_indices(norb)] += -2.0 * np.arange(norb) / norb
    g = rng.normal(0.0, 0.01, (norb,) * 4)
    # average over the 8 index permutations of real chemist-notation integrals
    g = (
        g
        + g.transpose(1, 0, 2, 3)
        + g.transpose(0, 1, 3, 2)
        + g.transpose(1, 0, 3, 2)
        + g.transpose(2, 3, 0, 1)
        + g.transpose(3, 2, 0, 1)
        + g.transpose(2, 3, 1, 0)
        + g.transpose(3, 2, 1, 0)
    ) / 8.0
    for p in range(norb):
        g[p, p, p, p] += 0.5
    return h, g


def build_sq_hamiltonian(h, g, constant=0.0):
    """SQOperator of H = const + sum h_ij a_i^ a_j + 1/2 sum (il|kj) a_i^ a_j^ a_k a_l
    over spin orbitals, term by term in the order of molecule_adapters.py:183-210."""
    norb = h.shape[0]
    H = _core.SQOperator()
    H.add(constant, [], [])
    for i in range(norb):
        ia, ib = 2 * i, 2 * i + 1
        for j in range(norb):
            ja, jb = 2 * j, 2 * j + 1
            H.add(h[i, j], [ia], [ja])
            H.add(h[i, j], [ib], [jb])
            for k in range(norb):
                ka, kb = 2 * k, 2 * k + 1
                for l in range(norb):
                    la, lb = 2 * l, 2 * l + 1
                    # one term per spin block, mixed blocks first, each
                    # skipped when a creator or annihilator pair repeats
                    v = g[i, l, k, j] / 2
                    if ia != jb and kb != la:
                        H.add(v, [ia, jb], [kb, la])
                    if ib != ja and ka != lb:
                        H.add(v, [ib, ja], [ka, lb])
                    if ia != ja and ka != la:
                        H.add(v, [ia, ja], [ka, la])
                    if ib != jb and kb != lb:
                        H.add(v, [ib, jb], [kb, lb])
    return H


def hf_occupation(norb, nalpha, nbeta):
    ref = [0] * (2 * norb)
    for a in range(nalpha):
        ref[2 * a] = 1
    for b in range(nbeta):
        ref[2 * b + 1] = 1
    return ref


def synthetic_molecule(norb, nalpha, nbeta, seed):
    """A ``Molecule`` with seeded synthetic integrals (JW-mapped Hamiltonian)."""
    h, g = synthetic_integrals(norb, seed)
    mol = Molecule(description=f"synthetic norb={norb} seed={seed}")
    mol.sq_hamiltonian = build_sq_hamiltonian(h, g)
    mol.hamiltonian = mol.sq_hamiltonian.jw_transform()
    mol.hf_reference = hf_occupation(norb, nalpha, nbeta)
    mol.point_group = ["C1", ["A"]]
    mol.orb_irreps = ["A"] * norb
    mol.orb_irreps_to_int = [0] * norb
    mol.integrals = (h, g)
    return mol
