"""``system_factory`` (reference src/qforte/system/system_factory.py:5-57).

Supported build types: ``external`` (JSON integrals, the reference's
create_external_mol, molecule_adapters.py:232-297) and ``synthetic`` (seeded
synthetic integrals).  ``psi4`` is not available in this package.
"""

import json

from .. import _core
from .molecular_info import Molecule
from .hamiltonians import synthetic_molecule

_IRREPS = {
    "C1": ["A"],
    "Ci": ["Ag", "Au"],
    "C2": ["A", "B"],
    "Cs": ["Ap", "App"],
    "D2": ["A", "B1", "B2", "B3"],
    "C2v": ["A1", "A2", "B1", "B2"],
    "C2h": ["Ag", "Bg", "Au", "Bu"],
    "D2h": ["Ag", "B1g", "B2g", "B3g", "Au", "B1u", "B2u", "B3u"],
}


def irreps_of_point_group(point_group):
    for k, v in _IRREPS.items():
        if k.lower() == str(point_group).lower():
            return v
    raise ValueError(f"unknown point group {point_group}")


def create_external_mol(**kwargs):
    """Molecule from a JSON file with scalar_energy, oei [p, q, h_pq],
    tei [p, q, r, s, <pq||rs>], nso, na, nb (molecule_adapters.py:232-297)."""
    mol = Molecule(multiplicity=kwargs.get("multiplicity", 1), charge=kwargs.get("charge", 0),
                   filename=kwargs["filename"])
    with open(kwargs["filename"]) as f:
        data = json.load(f)
    try:
        point_group = data["point_group"]["data"]
    except KeyError:
        point_group = "C1"
    irreps = irreps_of_point_group(point_group)
    mol.point_group = [point_group, irreps]
    nso = data["nso"]["data"]
    if point_group == "C1":
        mol.orb_irreps = ["A"] * (nso // 2)
        mol.orb_irreps_to_int = [0] * (nso // 2)
    else:
        mol.orb_irreps_to_int = list(data["symmetry"]["data"][::2])
        mol.orb_irreps = [irreps[i] for i in mol.orb_irreps_to_int]
    H = _core.SQOperator()
    H.add(data["scalar_energy"]["data"], [], [])
    for p, q, h_pq in data["oei"]["data"]:
        H.add(h_pq, [p], [q])
    for p, q, r, s, h_pqrs in data["tei"]["data"]:
        H.add(h_pqrs / 4.0, [p, q], [s, r])
    ref = [0] * nso
    for a in range(data["na"]["data"]):
        ref[2 * a] = 1
    for b in range(data["nb"]["data"]):
        ref[2 * b + 1] = 1
    mol.hf_reference = ref
    mol.sq_hamiltonian = H
    mol.hamiltonian = H.jw_transform()
    return mol


def system_factory(system_type="molecule", build_type="external", **kwargs):
    """Build a system.  ``build_type="external"`` reads JSON integrals
    (``filename=``); ``build_type="synthetic"`` takes ``norb``, ``nalpha``,
    ``nbeta`` and ``seed``."""
    if system_type != "molecule":
        raise TypeError("system type not supported, supported types are 'molecule'.")
    if build_type == "external":
        return create_external_mol(**kwargs)
    if build_type == "synthetic":
        return synthetic_molecule(kwargs["norb"], kwargs["nalpha"], kwargs["nbeta"], kwargs.get("seed", 0))
    raise TypeError(f"build type {build_type} not supported, supported types are: external, synthetic")
