"""Compact (few-CNOT) circuits for fermionic and qubit excitations
(reference src/qforte/utils/compact_excitation_circuits.py:9-91).

``compact_excitation_circuit(theta, creation, annihilation, qubit_excitations)``
returns a ``Circuit`` equal to exp(theta * kappa) with kappa = tau - tau^dag,
tau = a^dag_{c0} a^dag_{c1} ... a_{a0} a_{a1} ... (fermionic, Jordan-Wigner) or
its qubit analogue Q - Q^dag (``qubit_excitations=True``).  The gate sequence
is the one the reference emits, so ``size()``, ``get_num_cnots()`` and ``str``
of the circuits agree:

* fermionic sign: a CNOT staircase over the qubits whose Z parity survives
  the product of the Jordan-Wigner strings, closed by a cZ onto the first
  creation qubit, and its adjoint after the excitation
  (compact_excitation_circuits.py:94-149);
* the excitation proper: CNOTs from the first creation (annihilation) qubit
  onto the other creation (annihilation) qubits and one onto the first
  annihilation qubit, a multi-controlled Ry on the first creation qubit, and
  the CNOTs undone (152-210);
* the multi-controlled Ry as 2^k single-qubit Ry rotations of angle
  +-theta / 2^(k-1) interleaved with CNOT / anti-controlled CNOT gates in
  Gray-code order (213-290), with the rank-dependent sign of fermionic
  excitations and the sign flip of generalized excitations p^ q^ p s.

On the device the ansatz drivers apply exp(theta * kappa) directly (one pass
per excitation); these circuits fix the gate counts the drivers report.
"""

from __future__ import annotations

from .. import _core


def compact_excitation_circuit(theta, creation, annihilation, qubit_excitations):
    """exp(theta * kappa) as a compact circuit (compact_excitation_circuits.py:9-91)."""
    creation = list(creation)
    annihilation = list(annihilation)
    if len(creation) != len(annihilation):
        raise ValueError(
            "Compact fermionic/qubit excitations are implemented for particle-number-conserving operators only.")
    cre_only = [p for p in creation if p not in annihilation]
    ann_only = [p for p in annihilation if p not in creation]
    shared = sorted(set(creation) & set(annihilation))
    if len(shared) > 1:
        raise ValueError("A GSD excitation of the form p^ q^ p q was encountered!")

    shared_sign = 1
    same_slot = False
    if shared:
        shared_sign = -1
        # the repeated index sits in the same slot of both lists: the sign
        # staircase then runs over the distinct indices only
        slots = []
        for i in range(2):
            if creation[i] == shared[0]:
                slots.append(i)
            if annihilation[i] == shared[0]:
                slots.append(i)
        same_slot = slots in ([0, 0], [1, 1])

    circ = _core.Circuit()
    stair = None
    if not qubit_excitations:
        stair = fermion_sign_circuit(cre_only, ann_only) if same_slot else fermion_sign_circuit(creation, annihilation)
        circ.add(stair)
    circ.add(qubit_excitation(theta, cre_only, ann_only, shared, shared_sign, qubit_excitations))
    if stair is not None:
        circ.add(stair.adjoint())
    return circ


def fermion_sign_circuit(creation, annihilation):
    """CNOT staircase and cZ that carry the Jordan-Wigner sign of an
    excitation (compact_excitation_circuits.py:94-149)."""
    ops = list(creation) + list(annihilation)
    n = max(ops) + 1
    # qubit j keeps a Z when an odd number of the operators' strings cover it
    parity = [sum(1 for p in ops if j < p) & 1 for j in range(n)]
    for p in ops:
        parity[p] = 0
    chain = [j for j in range(n - 1, -1, -1) if parity[j]]
    circ = _core.Circuit()
    if not chain:
        return circ
    for hi, lo in zip(chain[:-1], chain[1:]):
        circ.add(_core.gate("CNOT", lo, hi))
    circ.add(_core.gate("cZ", creation[0], chain[-1]))
    return circ


def qubit_excitation(theta, creation, annihilation, gsd_control, gsd_sign, qubit_excitations):
    """The excitation proper on the distinct indices (compact_excitation_circuits.py:152-210)."""
    ladder = _core.Circuit()
    for q in creation[1:]:
        ladder.add(_core.gate("CNOT", q, creation[0]))
    for q in annihilation[1:]:
        ladder.add(_core.gate("CNOT", q, annihilation[0]))
    ladder.add(_core.gate("CNOT", annihilation[0], creation[0]))
    circ = _core.Circuit()
    circ.add(ladder)
    circ.add(multi_qubit_controlled_Ry(theta, creation[0], list(creation[1:]), list(annihilation), list(gsd_control),
                                       gsd_sign, qubit_excitations))
    circ.add(ladder.adjoint())
    return circ


def multi_qubit_controlled_Ry(theta, target, control_creation, control_annihilation, gsd_control, gsd_sign,
                              qubit_excitations):
    """Ry(2 theta) on ``target`` controlled by the other excitation qubits,
    as 2^k rotations and k-qubit Gray-code CNOT steps
    (compact_excitation_circuits.py:213-290)."""
    controls = list(control_annihilation) + list(reversed(control_creation)) + list(gsd_control)
    k = len(controls)
    nrot = 1 << k
    angle = theta / (1 << (k - 1))
    factor = 1
    if not qubit_excitations:
        # fermionic excitations of rank r carry (-1) unless r = 0 or 1 (mod 4)
        rank = len(control_annihilation)
        factor = (1 if rank % 4 in (0, 1) else -1) * gsd_sign
    circ = _core.Circuit()
    for i in range(nrot):
        sign = (1 if i % 2 == 0 else -1) * factor
        circ.add(_core.gate("Ry", target, target, sign * angle))
        for j, control in enumerate(controls):
            if (i + 1) % (nrot >> (j + 1)) == 0:
                if (i + 1) % (nrot >> 1) == 0 or i == nrot - 1 or gsd_control:
                    circ.add(_core.gate("CNOT", target, control))
                else:
                    circ.add(_core.gate("aCNOT", target, control))
                break
    return circ
