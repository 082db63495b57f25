"""The restated drivers on the engine reproduce the reference's integration
tests: FCI energies of He/cc-pVDZ (UCCSD-VQE, BFGS and Jacobi; UCCSD-PQE),
H4/STO-6G (ADAPT-VQE), H2 active space (PQE) and H2 QITE
(reference tests/test_uccsd.py, test_adapt_vqe.py, test_external_builder.py,
test_qite.py, test_computer_init.py)."""

import os

import numpy as np
import pytest
from pytest import approx

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def he(qf):
    return qf.system_factory(build_type="external", filename=os.path.join(GOLDEN, "He-ccpvdz.json"))


def test_he_uccsd_vqe(qf):
    # tests/test_uccsd.py:14-29
    alg = qf.UCCNVQE(he(qf))
    alg.run(pool_type="SD", use_analytic_grad=True)
    assert alg.get_gs_energy() == approx(-2.887594831090938, abs=1.0e-10)


def test_he_uccsd_vqe_gate_path(qf):
    # the same optimisation through the gate-level circuit path of Computer
    alg = qf.UCCNVQE(he(qf), use_program=False)
    alg.run(pool_type="SD", use_analytic_grad=True)
    assert alg.get_gs_energy() == approx(-2.887594831090938, abs=1.0e-10)


def test_he_uccsd_vqe_full_space_program(qf):
    # the compiled program on the full state instead of the determinant sector
    alg = qf.UCCNVQE(he(qf), use_sector=False)
    alg.run(pool_type="SD", use_analytic_grad=True)
    assert alg.get_gs_energy() == approx(-2.887594831090938, abs=1.0e-10)
    assert type(alg._get_program()).__name__ == "UCCProgram"


def test_default_drivers_run_on_the_sector(qf):
    alg = qf.UCCNVQE(he(qf))
    alg.run(pool_type="SD", use_analytic_grad=True)
    assert type(alg._get_program()).__name__ == "SectorTemplate"


def test_he_uccsd_vqe_jacobi(qf):
    # tests/test_uccsd.py:31-47
    alg = qf.UCCNVQE(he(qf))
    alg.run(pool_type="SD", use_analytic_grad=True, optimizer="jacobi")
    assert alg.get_gs_energy() == approx(-2.887594831090938, abs=1.0e-11)


def test_he_uccsd_pqe(qf):
    # tests/test_uccsd.py:69-84
    alg = qf.UCCNPQE(he(qf))
    alg.run(pool_type="SD", opt_thresh=1.0e-7)
    assert alg.get_gs_energy() == approx(-2.887594831090938, abs=1.0e-11)


def test_h2_active_space_pqe(qf):
    # tests/test_external_builder.py:9-29
    mol = qf.system_factory(build_type="external", filename=os.path.join(GOLDEN, "as_ints.json"))
    alg = qf.UCCNPQE(mol)
    alg.run(pool_type="SD", opt_thresh=1.0e-7)
    assert alg.get_gs_energy() == approx(-1.129181390402433, abs=1.0e-12)


def test_h4_adapt_vqe(qf):
    # tests/test_adapt_vqe.py:12-40
    mol = qf.system_factory(build_type="external", filename=os.path.join(GOLDEN, "H4-sto6g-075a.json"))
    alg = qf.ADAPTVQE(mol, print_summary_file=False)
    alg.run(adapt_maxiter=20, avqe_thresh=1.0e-4, opt_thresh=1.0e-5, use_analytic_grad=True, pool_type="SDTQ")
    assert alg.get_gs_energy() == approx(-2.1628978832666865, abs=5.0e-11)


def test_h4_vqe_from_computer_reference(qf):
    # tests/test_computer_init.py:13-30: a Computer as the reference state,
    # analytic and finite-difference gradients reach the same minimum
    mol = qf.system_factory(build_type="external", filename=os.path.join(GOLDEN, "H4-sto6g-075a.json"))
    n = len(mol.hf_reference)
    comp = qf.Computer(n)
    v = np.zeros(2**n)
    v[int("00001111", 2)] = 1
    v[int("00110011", 2)] = 0.2
    v[int("00111100", 2)] = 0.1
    v[int("11001100", 2)] = 0.04
    comp.set_coeff_vec(v / np.linalg.norm(v))
    e = []
    for analytic in (True, False):
        alg = qf.UCCNVQE(mol, reference=comp, state_prep_type="computer")
        alg.run(pool_type="SD", use_analytic_grad=analytic)
        e.append(alg.get_gs_energy())
    assert e[0] == approx(e[1], abs=1.0e-8)
    # ADAPT-GSD from the HF determinant and from the Computer reach FCI
    hf = qf.ADAPTVQE(mol)
    hf.run(use_analytic_grad=True, pool_type="GSD", avqe_thresh=1e-5)
    cp = qf.ADAPTVQE(mol, reference=comp, state_prep_type="computer")
    cp.run(use_analytic_grad=True, pool_type="GSD", avqe_thresh=1e-5)
    assert hf.get_gs_energy() == approx(cp.get_gs_energy(), abs=1.0e-8)
    assert hf.get_gs_energy() == approx(-2.162897881184882, abs=1.0e-8)


def test_h2_qite(qf):
    # tests/test_qite.py:8-81: H2 at 1.5 A in STO-3G, Hamiltonian data given there
    coef = [-0.4917857774144603, 0.09345649662931771, 0.09345649662931771, -0.0356448161226769,
            -0.0356448161226769, 0.1381758457453024, 0.05738398402634884, -0.0573839840263488,
            -0.0573839840263488, 0.05738398402634884, 0.08253705485911705, 0.13992103888546592,
            0.13992103888546592, 0.08253705485911705, 0.1458551902800438]
    circs = ["", "Z_0", "Z_1", "Z_2", "Z_3", "Z_0 Z_1", "Y_0 X_1 X_2 Y_3", "X_0 X_1 Y_2 Y_3", "Y_0 Y_1 X_2 X_3",
             "X_0 Y_1 Y_2 X_3", "Z_0 Z_2", "Z_0 Z_3", "Z_1 Z_2", "Z_1 Z_3", "Z_2 Z_3"]
    H = qf.QubitOperator()
    for c, s in zip(coef, circs):
        H.add(c, qf.build_circuit(s) if s else qf.Circuit())
    mol = qf.Molecule()
    mol.hamiltonian = H
    alg = qf.QITE(mol, reference=[1, 1, 0, 0])
    alg.run(beta=18.0, do_lanczos=True, lanczos_gap=49)
    assert alg.get_gs_energy() == approx(-0.9981493534, abs=1.0e-10)


def h4(qf):
    return qf.system_factory(build_type="external", filename=os.path.join(GOLDEN, "H4-sto6g-075a.json"))


@pytest.mark.parametrize("qubit", [False, True])
def test_compact_excitations_match_standard_circuits(qf, qubit):
    # tests/test_feb_qeb_circuits.py:6-62 on H4/STO-6G: the same energies with
    # fewer CNOTs.  The drivers run both forms through the compiled program
    # (each compact circuit is exp(t kappa)); the gate circuits themselves are
    # exercised below with use_program=False.
    mol = h4(qf)
    cases = [(qf.UCCNVQE, {"pool_type": "SD"}), (qf.UCCNVQE, {"pool_type": "sa_SD"}),
             (qf.ADAPTVQE, {"pool_type": "SD", "avqe_thresh": 1.0e-3}), (qf.UCCNPQE, {"pool_type": "SD"})]
    for method, options in cases:
        kw = {"print_summary_file": False} if method is qf.ADAPTVQE else {}
        std = method(mol, compact_excitations=False, qubit_excitations=qubit, **kw)
        std.run(**options)
        eff = method(mol, compact_excitations=True, qubit_excitations=qubit, **kw)
        eff.run(**options)
        assert std.get_gs_energy() == approx(eff.get_gs_energy(), abs=1.0e-10), (method.__name__, options)
        assert std._n_cnot > eff._n_cnot


def test_compact_excitations_on_the_gate_path(qf):
    # the compact circuits and the compact reverse sweep of the analytic
    # gradient (uccvqeabc.py:181-207) through Computer
    mol = h4(qf)
    for pool_type in ("SD", "sa_SD"):
        std = qf.UCCNVQE(mol)
        std.run(pool_type=pool_type, use_analytic_grad=True)
        eff = qf.UCCNVQE(mol, compact_excitations=True, use_program=False)
        eff.run(pool_type=pool_type, use_analytic_grad=True)
        assert eff.get_gs_energy() == approx(std.get_gs_energy(), abs=1.0e-9)
    # one gradient of each pool: the compact reverse sweep equals the standard
    # gate-level recursion; for SD both are the exact derivative the program
    # returns.  (For sa_SD the reference's recursion is not the exact
    # derivative: its spin-adapted halves do not commute.)
    for pool_type in ("SD", "sa_SD"):
        std = qf.UCCNVQE(mol, use_program=False)
        eff = qf.UCCNVQE(mol, compact_excitations=True, use_program=False)
        for alg in (std, eff):
            alg._pool_type = pool_type
            alg.fill_pool()
            alg._tops = list(range(len(alg._pool_obj)))
            alg._tamps = [0.05 * (k % 5 - 2) for k in range(len(alg._tops))]
        g_std = np.asarray(std.measure_gradient())
        g_eff = np.asarray(eff.measure_gradient())
        assert np.abs(g_std - g_eff).max() < 1.0e-12, pool_type
        if pool_type == "SD":
            g_prog = np.asarray(eff._get_program().gradient(list(eff._tamps))[1])
            assert np.abs(g_prog - g_eff).max() < 1.0e-12


def test_compact_pqe_jacobi_and_bfgs_agree(qf):
    # tests/test_uccsd.py:86-122: compact-excitation PQE by Jacobi and by BFGS
    # on the squared residuals reach the same energy
    mol = h4(qf)
    jacobi = qf.UCCNPQE(mol, compact_excitations=True, qubit_excitations=False, diis_max_dim=8)
    jacobi.run(optimizer="jacobi", pool_type="SD")
    bfgs = qf.UCCNPQE(mol, compact_excitations=True, qubit_excitations=False)
    bfgs.run(optimizer="BFGS", pool_type="SD")
    assert jacobi.get_gs_energy() == approx(bfgs.get_gs_energy(), abs=1.0e-8)


def test_pqe_residuals_sector_full_space_and_gate_path(qf):
    # r_mu = <Phi_mu| U^dag H U |ref> phase_mu (uccnpqe.py:249-292) on the
    # sector template, on the full-space program and through Computer with the
    # standard and the compact circuits
    mol = h4(qf)
    t = None
    results = []
    for compact, use_program, use_sector in [(False, True, True), (False, True, False), (False, False, True),
                                             (True, False, True)]:
        alg = qf.UCCNPQE(mol, compact_excitations=compact, use_program=use_program, use_sector=use_sector)
        alg._pool_type = "SD"
        alg.fill_pool()
        alg._tops = list(range(len(alg._pool_obj)))
        if t is None:
            t = [0.07 * ((3 * k) % 7 - 3) for k in range(len(alg._tops))]
        alg._tamps = list(t)
        alg._excited_dets, alg._res_vec_evals, alg._res_m_evals = [], 0, 0
        alg.fill_excited_dets()
        results.append(np.asarray(alg.get_residual_vector(t)))
        if use_program:
            kind = "SectorTemplate" if use_sector else "UCCProgram"
            assert type(alg._get_program()).__name__ == kind
    assert np.abs(results[0]).max() > 1e-3
    for r in results[1:]:
        assert np.abs(r - results[0]).max() < 1e-12


def test_compact_adapt_and_pqe_jacobi_match_bfgs(qf):
    # tests/test_adapt_vqe.py:42-90 (compact qubit excitations, GSD pool) and
    # tests/test_uccsd.py:86-122 (compact fermionic PQE): Jacobi with DIIS and
    # BFGS reach the same energy
    mol = h4(qf)
    jacobi = qf.ADAPTVQE(mol, compact_excitations=True, qubit_excitations=True, diis_max_dim=8,
                         print_summary_file=False)
    jacobi.run(optimizer="jacobi", pool_type="GSD", avqe_thresh=0.001)
    bfgs = qf.ADAPTVQE(mol, compact_excitations=True, qubit_excitations=True, print_summary_file=False)
    bfgs.run(optimizer="BFGS", pool_type="GSD", avqe_thresh=0.001)
    assert jacobi.get_gs_energy() == approx(bfgs.get_gs_energy(), abs=1.0e-8)
    assert jacobi._n_cnot > 0 and bfgs._n_cnot > 0
