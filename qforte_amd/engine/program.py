"""Compile a disentangled UCC ansatz and a Hamiltonian into a device-resident
program (``qfe_program``).

The generators and their term order are exactly what the reference's
``UCC.ansatz_circuit`` feeds to ``trotterize`` (src/qforte/abc/ansatz.py:25-82):
for each selected pool operator, JW(operator) simplified and ordered as
``SQOpPool.get_qubit_operator("commuting_grp_lex")`` does
(sq_op_pool.cc:97-104); for ``sa_SD`` pools every spin-orbital
anti-Hermitian piece becomes its own generator sharing the amplitude
(ansatz.py:38-56).  The program applies exp(t_k * sum_j c_j P_j) term by term
in that order, so energies match the reference's gate-level circuit.
"""

from .. import _core
from ..utils.state_prep import ref_to_basis_idx


def _generator(sq_op, qubit_excitations):
    q = sq_op.jw_transform(qubit_excitations)
    q.simplify()
    q.order_terms()
    return q


def ucc_generators(pool, tops, pool_type=None, qubit_excitations=False):
    """Generators (QubitOperators), their amplitude indices and scales for
    the operators ``tops`` of ``pool`` (an SQOpPool)."""
    gens, index, scale = [], [], []
    for mu, top in enumerate(tops):
        sq_op = pool[top][1]
        if pool_type == "sa_SD":
            terms = sq_op.terms()
            half = len(terms) // 2
            for coeff, cr, ann in terms[:half]:
                piece = _core.SQOperator()
                piece.add_term(coeff, cr, ann)
                piece.add_term(-coeff, ann, cr)
                gens.append(_generator(piece, qubit_excitations))
                index.append(mu)
                scale.append(1.0)
        else:
            gens.append(_generator(sq_op, qubit_excitations))
            index.append(mu)
            scale.append(1.0)
    return gens, index, scale


def compile_ucc_program(nqubit, reference, hamiltonian, pool, tops, pool_type=None, qubit_excitations=False):
    """A ``UCCProgram`` for the ansatz U(t) = prod_mu exp(t_mu K_mu) (operator
    tops[0] first) on the determinant ``reference`` (occupation list)."""
    gens, index, scale = ucc_generators(pool, tops, pool_type, qubit_excitations)
    return _core.UCCProgram(nqubit, ref_to_basis_idx(reference), gens, index, scale, len(tops), hamiltonian)


def compile_ucc_template(nqubit, reference, hamiltonian, pool, tops, pool_type=None, qubit_excitations=False):
    """A ``SectorTemplate``: the same ansatz as ``compile_ucc_program`` kept on
    the determinants of the reference's (N_alpha, N_beta) sector.
    ``hamiltonian`` is a ``QubitOperator`` (fixed) or a ``HamiltonianMap``
    (coefficients from per-request integrals)."""
    gens, index, scale = ucc_generators(pool, tops, pool_type, qubit_excitations)
    return _core.SectorTemplate(nqubit, ref_to_basis_idx(reference), gens, index, scale, len(tops), hamiltonian)
