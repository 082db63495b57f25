"""qforte_amd -- qforte's ``Computer`` hot path on AMD Instinct MI355X.

A drop-in for the reference ``qforte`` flat namespace on the state-vector
path: ``Computer``, ``Circuit``, ``Gate``/``gate``, ``QubitOperator``,
``SQOperator``, ``SQOpPool``, ``QubitOpPool``, ``QubitBasis`` and the helpers
the UCC-VQE/PQE, ADAPT and QITE drivers call (reference
``src/qforte/__init__.py:5``).  The data model is native C++
(``qforte_amd/csrc/host``); amplitudes live in HBM and every sweep runs in
hand-written HIP kernels for gfx950 behind the C ABI in ``include/qfe.h``.

There is no CPU fallback: importing works on any host (the data model is
CPU code), but creating a ``Computer`` without a GPU raises ``RuntimeError``.
"""

__version__ = "0.1.0"

from ._core import (  # noqa: F401
    BatchEngine,
    Circuit,
    Computer,
    DeviceTimer,
    Gate,
    HamiltonianMap,
    QubitBasis,
    QubitOperator,
    QubitOpPool,
    SQOperator,
    SQOpPool,
    SectorTemplate,
    UCCProgram,
    add_gate_to_computer,
    control_gate,
    cross_expectations,
    device_count,
    find_irrep,
    gate,
    get_counters,
    get_pass_tuning,
    get_tuning,
    inner_product,
    norm2,
    normalize_angle,
    pauli_masks,
    prepare_computer_from_circuit,
    prepare_computer_from_gates,
    reset_counters,
    set_pass_tuning,
    set_sector_word_limit,
    set_tuning,
    stream_handle,
    synchronize,
)

from .utils.exponentiate import exponentiate_pauli_string  # noqa: E402,F401
from .utils import trotterization  # noqa: E402,F401
from .utils.trotterization import trotterize, trotterize_w_cRz  # noqa: E402,F401
from .utils.state_prep import build_refprep, ref_string, integer_to_ref, ref_to_basis_idx  # noqa: E402,F401
from .helper.io_helper import build_circuit, build_operator, smart_print  # noqa: E402,F401
from .system.molecular_info import Molecule, System  # noqa: E402,F401
from .system.system_factory import system_factory  # noqa: E402,F401
from .system.hamiltonians import (  # noqa: E402,F401
    build_sq_hamiltonian,
    synthetic_integrals,
    synthetic_molecule,
)
from .engine.program import compile_ucc_program, compile_ucc_template, ucc_generators  # noqa: E402,F401
from .engine.serving import ServingEngine  # noqa: E402,F401
from .abc.algorithm import Algorithm, AnsatzAlgorithm  # noqa: E402,F401
from .abc.uccvqeabc import UCCVQE  # noqa: E402,F401
from .ucc.uccnvqe import UCCNVQE  # noqa: E402,F401
from .ucc.adaptvqe import ADAPTVQE  # noqa: E402,F401
from .ucc.uccnpqe import UCCNPQE  # noqa: E402,F401
from .ite.qite import QITE  # noqa: E402,F401
from .maths.eigsolve import canonical_geig_solve  # noqa: E402,F401
