"""Algorithm base classes (reference src/qforte/abc/algorithm.py,
abc/ansatz.py, abc/mixin.py, abc/vqeabc.py).

The drivers keep the reference's constructor options and attribute names.
Energies and gradients of UCC ansatze go through a compiled program -- the
same exponentials, term order and Hamiltonian as the reference's circuit
path, resident in HBM.  When every generator keeps the alpha and beta
particle numbers the program runs on the reference's (N_alpha, N_beta)
determinant sector (``SectorTemplate``), otherwise, or with
``use_sector=False``, on the full state (``UCCProgram``); ``use_program=False``
selects the gate-level path through ``Computer``.  All three agree within
1e-10 and are tested.  With ``compact_excitations=True`` the circuits are the
reference's compact excitation circuits, each equal to exp(t kappa), so the
compiled program serves them too; the circuits fix the reported CNOT counts.
"""

import numpy as np

from .. import _core
from ..utils.state_prep import build_refprep, ref_string
from ..utils.trotterization import trotterize
from ..utils.compact_excitation_circuits import compact_excitation_circuit
from ..engine.program import compile_ucc_program, compile_ucc_template


class Trotterizable:
    """trotter_order / trotter_number options (mixin.py:10-32)."""

    def __init__(self, *args, trotter_order=1, trotter_number=1, **kwargs):
        super().__init__(*args, **kwargs)
        self._trotter_order = trotter_order
        self._trotter_number = trotter_number

    def print_trotter_options(self):
        print("Trotter order (rho):                     ", self._trotter_order)
        print("Trotter number (m):                      ", self._trotter_number)


class Algorithm:
    """Common state of every algorithm (algorithm.py:13-224)."""

    def __init__(self, system, reference=None, state_prep_type="occupation_list", fast=True, verbose=False,
                 print_summary_file=False, **kwargs):
        self._sys = system
        self._state_prep_type = state_prep_type
        if state_prep_type == "occupation_list":
            if reference is None:
                self._ref = system.hf_reference
            else:
                if not isinstance(reference, list):
                    raise ValueError("occupation_list reference must be list of 1s and 0s.")
                self._ref = reference
            self._refprep = build_refprep(self._ref)
            self._Uprep = _core.Circuit(self._refprep)
        elif state_prep_type == "unitary_circ":
            if not isinstance(reference, _core.Circuit):
                raise ValueError("unitary_circ reference must be a Circuit.")
            self._ref = system.hf_reference
            self._refprep = build_refprep(self._ref)
            self._Uprep = reference
        elif state_prep_type == "computer":
            if not isinstance(reference, _core.Computer):
                raise ValueError("computer reference must be a Computer.")
            if not fast:
                raise ValueError("`self._fast = False` specifies not to skip steps, but "
                                 "`self._state_prep_type = computer` specifies to skip state initialization. "
                                 "That's inconsistent.")
            if reference.get_nqubit() != len(system.hf_reference):
                raise ValueError(f"Computer needs {len(system.hf_reference)} qubits, found {reference.get_nqubit()}.")
            if not getattr(self, "computer_initializable", False):
                raise ValueError("Class cannot be initialized with a computer.")
            self._ref = system.hf_reference
            self._refprep = build_refprep(self._ref)
            self._Uprep = _core.Circuit()
            self.computer = reference
        else:
            raise ValueError("QForte only supports references as occupation lists, Circuits, or Computers.")
        self._nqb = len(self._ref)
        self._qb_ham = system.hamiltonian
        if self._qb_ham.num_qubits() > self._nqb:
            raise ValueError(f"The reference has {self._nqb} qubits, but the Hamiltonian has "
                             f"{self._qb_ham.num_qubits()}. This is inconsistent.")
        self._hf_energy = getattr(system, "hf_energy", 0.0)
        self._Nl = len(self._qb_ham.terms())
        self._fast = fast
        self._verbose = verbose
        self._print_summary_file = print_summary_file
        self._noise_factor = 0.0
        self._Egs = None
        self._Umaxdepth = None
        self._n_classical_params = None
        self._n_cnot = None
        self._n_pauli_trm_measures = None

    def get_gs_energy(self):
        return self._Egs

    def verify_required_attributes(self):
        for attr in ("_Egs", "_n_classical_params", "_n_cnot", "_n_pauli_trm_measures"):
            if getattr(self, attr) is None:
                raise NotImplementedError(f"Concrete Algorithm class must define self.{attr} attribute.")

    def print_generic_options(self):
        print("Trial reference state:                   ", ref_string(self._ref, self._nqb))
        print("Number of Hamiltonian Pauli terms:       ", self._Nl)
        print("Trial state preparation method:          ", self._state_prep_type)
        if isinstance(self, Trotterizable):
            self.print_trotter_options()
        print("Use fast version of algorithm:           ", str(self._fast))


class AnsatzAlgorithm(Algorithm):
    """Operator-pool ansatz machinery (algorithm.py:227-480)."""

    def __init__(self, *args, qubit_excitations=False, compact_excitations=False, diis_max_dim=8,
                 max_moment_rank=0, moment_dt=None, penalty=None, use_program=True, use_sector=True, **kwargs):
        super().__init__(*args, **kwargs)
        if max_moment_rank:
            raise NotImplementedError("moment energy corrections are not part of this package")
        if penalty is not None:
            raise NotImplementedError("Hamiltonian penalty terms are not part of this package")
        self._curr_energy = 0
        self._Nm = []
        self._tamps = []
        self._tops = []
        self._pool_obj = _core.SQOpPool()
        self._qubit_excitations = qubit_excitations
        self._compact_excitations = compact_excitations
        self._diis_max_dim = diis_max_dim
        self._max_moment_rank = max_moment_rank
        self._use_program = use_program
        self._use_sector = use_sector
        self._program = None
        self._program_key = None

    # --- circuits (algorithm.py:271-288) -------------------------------------------
    def build_Uvqc(self, amplitudes=None):
        U = self.ansatz_circuit(amplitudes)
        Uvqc = _core.Circuit()
        Uvqc.add(self._Uprep)
        Uvqc.add(U)
        return Uvqc

    def fill_pool(self):
        """Populate the SQOpPool (algorithm.py:290-307)."""
        if self._pool_type in {"sa_SD", "GSD", "SD", "SDT", "SDTQ", "SDTQP", "SDTQPH"}:
            self._pool_obj = _core.SQOpPool()
            if hasattr(self._sys, "orb_irreps_to_int"):
                self._pool_obj.set_orb_spaces(self._ref, self._sys.orb_irreps_to_int)
            else:
                self._pool_obj.set_orb_spaces(self._ref)
            self._pool_obj.fill_pool(self._pool_type)
        elif isinstance(self._pool_type, _core.SQOpPool):
            self._pool_obj = self._pool_type
        else:
            raise ValueError("Invalid operator pool type specified.")
        self._Nm = [len(op.jw_transform().terms()) for _, op in self._pool_obj]

    # --- compiled program -------------------------------------------------------------
    def _can_use_program(self):
        # A compact excitation circuit equals exp(t kappa) exactly, as the
        # program applies it, whatever the Trotter number.
        return (self._use_program and self._fast and self._state_prep_type == "occupation_list"
                and (self._trotter_number == 1 or self._compact_excitations))

    def _get_program(self):
        key = (tuple(self._tops), self._pool_type if isinstance(self._pool_type, str) else id(self._pool_type))
        if self._program is None or self._program_key != key:
            pool_type = self._pool_type if isinstance(self._pool_type, str) else None
            args = (self._nqb, self._ref, self._qb_ham, self._pool_obj, list(self._tops), pool_type,
                    self._qubit_excitations)
            self._program = None
            if self._use_sector and self._nqb % 2 == 0:
                try:
                    self._program = compile_ucc_template(*args)
                except ValueError:  # a generator leaves the sector: full state instead
                    self._program = None
            if self._program is None:
                self._program = compile_ucc_program(*args)
            self._program_key = key
        return self._program

    # --- energies (algorithm.py:309-334, 458-480) ---------------------------------------------
    def measure_energy(self, Ucirc, computer=None):
        if not self._fast:
            raise NotImplementedError("only the fast (state-vector) path is available")
        if computer is None:
            computer = _core.Computer(self._nqb)
        computer.apply_circuit(Ucirc)
        val = computer.direct_op_exp_val(self._qb_ham)
        assert np.isclose(np.imag(val), 0.0)
        return np.real(val)

    def energy_feval(self, params):
        if self._can_use_program():
            energy = self._get_program().energy([float(x) for x in params])
        else:
            Ucirc = self.build_Uvqc(amplitudes=params)
            energy = self.measure_energy(Ucirc, self.get_initial_computer())
        self._curr_energy = energy
        return energy

    def get_initial_computer(self):
        if hasattr(self, "computer"):
            return _core.Computer(self.computer)
        return _core.Computer(self._nqb)


class UCC(Trotterizable):
    """Disentangled UCC ansatz (ansatz.py:17-154)."""

    def ansatz_circuit(self, amplitudes=None):
        temp_pool = _core.SQOpPool()
        tamps = self._tamps if amplitudes is None else amplitudes
        if self._pool_type == "sa_SD":
            for tamp, top in zip(tamps, self._tops):
                terms = self._pool_obj[top][1].terms()
                half = len(terms) // 2
                for coeff, cr, ann in terms[:half]:
                    sq = _core.SQOperator()
                    sq.add_term(coeff, cr, ann)
                    sq.add_term(-coeff, ann, cr)
                    temp_pool.add(tamp, sq)
        else:
            for tamp, top in zip(tamps, self._tops):
                temp_pool.add(tamp, self._pool_obj[top][1])
        if self._compact_excitations:  # ansatz.py:60-71
            U = _core.Circuit()
            for tamp, sq_op in temp_pool:
                coeff, cre, ann = sq_op.terms()[1]
                U.add(compact_excitation_circuit(tamp * coeff, cre, ann, self._qubit_excitations))
            return U
        A = temp_pool.get_qubit_operator("commuting_grp_lex", qubit_excitations=self._qubit_excitations)
        U, phase1 = trotterize(A, trotter_number=self._trotter_number)
        if phase1 != 1.0 + 0.0j:
            raise ValueError("Encountered phase change, phase not equal to (1.0 + 0.0i)")
        return U

    def build_orb_energies(self):
        """Orbital energies from Koopmans-like differences (ansatz.py:84-118)."""
        self._orb_e = []
        if hasattr(self._sys, "hf_orbital_energies"):
            for e in self._sys.hf_orbital_energies:
                self._orb_e += [e] * 2
            return
        qc = _core.Computer(self._nqb)
        qc.apply_circuit(self._refprep)
        E0 = qc.direct_op_exp_val(self._qb_ham)
        nocc = sum(self._ref)
        for i in range(self._nqb):
            qc = _core.Computer(self._nqb)
            qc.apply_circuit(self._refprep)
            qc.apply_gate(_core.gate("X", i, i))
            Ei = qc.direct_op_exp_val(self._qb_ham)
            self._orb_e.append(np.real(E0 - Ei if i < nocc else Ei - E0))

    def get_res_over_mpdenom(self, residuals):
        """Residuals over Moller-Plesset denominators (ansatz.py:120-154)."""
        out = []
        for mu, m in enumerate(self._tops):
            sq_op = self._pool_obj[m][1]
            temp_idx = sq_op.terms()[0][2][-1]
            if self._ref[temp_idx]:
                cre, ann = sq_op.terms()[0][1], sq_op.terms()[0][2]
            else:
                cre, ann = sq_op.terms()[0][2], sq_op.terms()[0][1]
            denom = sum(self._orb_e[x] for x in ann) - sum(self._orb_e[x] for x in cre)
            out.append(residuals[mu] / denom)
        return out
