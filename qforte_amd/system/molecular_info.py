"""System containers (reference src/qforte/system/molecular_info.py)."""


class System:
    """A quantum many-body system: Hamiltonian plus reference determinant."""

    def __init__(self):
        self.hamiltonian = None
        self.sq_hamiltonian = None
        self.hf_reference = None


class Molecule(System):
    """Molecular data as the drivers read it: ``hamiltonian`` (QubitOperator),
    ``sq_hamiltonian`` (SQOperator), ``hf_reference`` (occupation list with
    alpha = even, beta = odd spin orbitals), optional ``hf_energy``,
    ``point_group``, ``orb_irreps`` and ``orb_irreps_to_int``."""

    def __init__(self, multiplicity=1, charge=0, filename="", basis=None, description=""):
        super().__init__()
        self.multiplicity = multiplicity
        self.charge = charge
        self.filename = filename
        self.basis = basis
        self.description = description
