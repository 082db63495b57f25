"""Determinant-sector templates and the serving engine.

The sector path must reproduce the full-space compiled program (itself
pinned against the dense oracle and the reference's gate-level path): the
prepared state amplitude by amplitude (1e-12) and the energy (1e-10 Ha), for
fixed Hamiltonians and for Hamiltonians built from per-request integrals by
``HamiltonianMap``.  The serving engine must return, for every request, the
same energy in a mixed batch as alone."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
E_TOL = 1e-10


def sd_pool(qf, ref):
    pool = qf.SQOpPool()
    pool.set_orb_spaces(ref)
    pool.fill_pool("SD")
    return pool


def synthetic_case(qf, norb, nalpha, nbeta, seed):
    h, g = qf.synthetic_integrals(norb, seed)
    ref = qf.system.hamiltonians.hf_occupation(norb, nalpha, nbeta)
    return h, g, ref


def test_h4_template_matches_program(qf):
    mol = qf.system_factory(build_type="external", filename=os.path.join(GOLDEN, "H4-sto6g-075a.json"))
    n = len(mol.hf_reference)
    for pool_type in ("SD", "GSD", "SDTQ"):
        pool = qf.SQOpPool()
        pool.set_orb_spaces(mol.hf_reference)
        pool.fill_pool(pool_type)
        tops = list(range(len(pool)))
        prog = qf.compile_ucc_program(n, mol.hf_reference, mol.hamiltonian, pool, tops)
        tpl = qf.compile_ucc_template(n, mol.hf_reference, mol.hamiltonian, pool, tops)
        assert tpl.dim == 36  # C(4,2)^2 determinants of 2 alpha + 2 beta electrons in 4 orbitals
        for seed in (1, 2):
            t = list(np.random.default_rng(seed).normal(0.0, 0.2, len(tops)))
            assert abs(tpl.energy(t) - prog.energy(t)) < E_TOL, pool_type
            a, b = qf.Computer(n), qf.Computer(n)
            prog.prepare(t, a)
            tpl.prepare(t, b)
            diff = np.abs(np.asarray(a.get_coeff_vec()) - np.asarray(b.get_coeff_vec())).max()
            assert diff < 1e-12, (pool_type, diff)


def test_template_energy_against_dense_oracle(qf, oracle):
    from test_gpu_ucc import oracle_ucc_energy, terms_list

    h, g, ref = synthetic_case(qf, 4, 2, 2, 11)
    H = qf.build_sq_hamiltonian(h, g, 0.25).jw_transform()
    pool = sd_pool(qf, ref)
    tops = list(range(len(pool)))
    tpl = qf.compile_ucc_template(8, ref, H, pool, tops)
    gens, _, _ = qf.ucc_generators(pool, tops)
    t = list(np.random.default_rng(3).normal(0.0, 0.05, len(tops)))
    e_or, _ = oracle_ucc_energy(oracle, 8, ref, gens, t, terms_list(H))
    assert abs(tpl.energy(t) - e_or) < E_TOL


@pytest.mark.parametrize("norb,na,nb", [(4, 2, 2), (5, 3, 2), (6, 3, 3), (7, 4, 3)])
def test_integral_map_template_matches_program(qf, norb, na, nb):
    n = 2 * norb
    h, g, ref = synthetic_case(qf, norb, na, nb, 100 + norb)
    H = qf.build_sq_hamiltonian(h, g, 0.5).jw_transform()
    pool = sd_pool(qf, ref)
    tops = list(range(len(pool)))
    prog = qf.compile_ucc_program(n, ref, H, pool, tops)
    hmap = qf.HamiltonianMap(norb)
    tpl = qf.compile_ucc_template(n, ref, hmap, pool, tops)
    inputs = hmap.pack(0.5, h, g)
    t = list(np.random.default_rng(norb).normal(0.0, 0.05, len(tops)))
    assert abs(tpl.energy(t, inputs) - prog.energy(t)) < E_TOL


def test_non_conserving_generator_is_rejected(qf):
    # a single Pauli string changes the particle number: no sector form exists
    gen = qf.QubitOperator()
    gen.add(0.5j, qf.build_circuit("X_0 Y_1"))
    H = qf.QubitOperator()
    H.add(1.0, qf.build_circuit("Z_0"))
    with pytest.raises(Exception, match="conserve|anti-Hermitian"):
        qf.SectorTemplate(4, 0b0011, [gen], [0], [1.0], 1, H)


def test_serving_results_are_the_same_alone_and_in_the_mix(qf):
    cases = []
    for norb, na, nb in [(4, 2, 2), (5, 3, 2), (6, 3, 3), (7, 4, 3), (8, 4, 4)]:
        _, _, ref = synthetic_case(qf, norb, na, nb, 0)
        pool = sd_pool(qf, ref)
        hmap = qf.HamiltonianMap(norb)
        tpl = qf.compile_ucc_template(2 * norb, ref, hmap, pool, list(range(len(pool))))
        cases.append((norb, tpl, hmap))
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(60):
        norb, tpl, hmap = cases[k % len(cases)]
        h, g = qf.synthetic_integrals(norb, 1000 + k)
        inputs = hmap.pack(0.0, h, g)
        t = list(rng.normal(0.0, 0.05, tpl.nparam))
        reqs.append((tpl, t, inputs))
    alone = [tpl.energy(t, inp) for tpl, t, inp in reqs]
    eng = qf.ServingEngine(hbm_budget_bytes=1 << 30, max_active=4096)
    ids = []
    # arrivals in three waves: later waves are admitted while earlier launches run
    for wave in range(3):
        for tpl, t, inp in reqs[wave * 20:(wave + 1) * 20]:
            ids.append(eng.submit(tpl, t, inp))
        eng.poll()
    eng.drain()
    mixed = [eng.results[i] for i in ids]
    assert np.max(np.abs(np.array(mixed) - np.array(alone))) < 1e-12
    p = eng.latency_percentiles()
    assert 0 < p["p50"] <= p["p99"]


def test_serving_gradients_match_the_template_alone(qf):
    """Gradient requests interleaved with energy requests of every size class:
    each returns the energy and gradient of ``SectorTemplate.gradient`` alone."""
    cases = []
    for norb, na, nb in [(4, 2, 2), (6, 3, 3), (8, 4, 4), (10, 5, 5)]:
        _, _, ref = synthetic_case(qf, norb, na, nb, 0)
        pool = sd_pool(qf, ref)
        hmap = qf.HamiltonianMap(norb)
        tpl = qf.compile_ucc_template(2 * norb, ref, hmap, pool, list(range(len(pool))))
        cases.append((norb, tpl, hmap))
    rng = np.random.default_rng(9)
    reqs = []
    for k in range(40):
        norb, tpl, hmap = cases[k % len(cases)]
        h, g = qf.synthetic_integrals(norb, 2000 + k)
        reqs.append((tpl, list(rng.normal(0.0, 0.05, tpl.nparam)), hmap.pack(0.0, h, g), k % 3 != 0))
    eng = qf.ServingEngine(hbm_budget_bytes=1 << 30, max_active=4096)
    ids = []
    for wave in range(2):
        for tpl, t, inp, want_grad in reqs[wave * 20:(wave + 1) * 20]:
            ids.append(eng.submit(tpl, t, inp, gradient=want_grad))
        eng.poll()
    eng.drain()
    assert set(eng.gradients) == {i for i, r in zip(ids, reqs) if r[3]}
    for rid, (tpl, t, inp, want_grad) in zip(ids, reqs):
        if want_grad:
            e, gr = tpl.gradient(t, inp)
            assert abs(eng.results[rid] - e) < 1e-12
            assert np.max(np.abs(np.array(eng.gradients[rid]) - np.array(gr))) < 1e-12
        else:
            assert abs(eng.results[rid] - tpl.energy(t, inp)) < 1e-12


def test_serving_respects_the_budget(qf):
    _, _, ref = synthetic_case(qf, 6, 3, 3, 0)
    pool = sd_pool(qf, ref)
    hmap = qf.HamiltonianMap(6)
    tpl = qf.compile_ucc_template(12, ref, hmap, pool, list(range(len(pool))))
    h, g = qf.synthetic_integrals(6, 1)
    inputs = hmap.pack(0.0, h, g)
    t = [0.01] * tpl.nparam
    # room for about two requests per launch: every request still completes
    eng = qf.ServingEngine(hbm_budget_bytes=3 * 2 * tpl.request_bytes + 1024, max_active=8)
    ids = [eng.submit(tpl, t, inputs) for _ in range(10)]
    eng.drain()
    e = tpl.energy(t, inputs)
    assert all(abs(eng.results[i] - e) < 1e-12 for i in ids)
    with pytest.raises(Exception, match="budget"):
        small = qf.ServingEngine(hbm_budget_bytes=tpl.request_bytes, max_active=8)
        small.submit(tpl, t, inputs)


def test_sector_gradient_matches_program_and_finite_differences(qf):
    mol = qf.system_factory(build_type="external", filename=os.path.join(GOLDEN, "H4-sto6g-075a.json"))
    n = len(mol.hf_reference)
    pool = sd_pool(qf, mol.hf_reference)
    tops = list(range(len(pool)))
    prog = qf.compile_ucc_program(n, mol.hf_reference, mol.hamiltonian, pool, tops)
    tpl = qf.compile_ucc_template(n, mol.hf_reference, mol.hamiltonian, pool, tops)
    t = np.random.default_rng(4).normal(0.0, 0.1, len(tops))
    e, g = tpl.gradient(list(t))
    e_p, g_p = prog.gradient(list(t))
    assert abs(e - e_p) < E_TOL
    assert np.max(np.abs(np.array(g) - np.array(g_p))) < 1e-10
    h = 1e-5
    for k in range(len(tops)):
        tp, tm = t.copy(), t.copy()
        tp[k] += h
        tm[k] -= h
        fd = (tpl.energy(list(tp)) - tpl.energy(list(tm))) / (2 * h)
        assert abs(fd - g[k]) < 1e-8, (k, fd, g[k])


def test_sector_gradient_with_integral_map(qf):
    norb, na, nb = 5, 3, 2
    h, g, ref = synthetic_case(qf, norb, na, nb, 21)
    H = qf.build_sq_hamiltonian(h, g, 0.0).jw_transform()
    pool = sd_pool(qf, ref)
    tops = list(range(len(pool)))
    prog = qf.compile_ucc_program(2 * norb, ref, H, pool, tops)
    hmap = qf.HamiltonianMap(norb)
    tpl = qf.compile_ucc_template(2 * norb, ref, hmap, pool, tops)
    t = list(np.random.default_rng(6).normal(0.0, 0.05, len(tops)))
    e, gr = tpl.gradient(t, hmap.pack(0.0, h, g))
    e_p, g_p = prog.gradient(t)
    assert abs(e - e_p) < E_TOL
    assert np.max(np.abs(np.array(gr) - np.array(g_p))) < 1e-10


def test_sector_pool_gradients_match_program(qf):
    mol = qf.system_factory(build_type="external", filename=os.path.join(GOLDEN, "H4-sto6g-075a.json"))
    n = len(mol.hf_reference)
    pool = sd_pool(qf, mol.hf_reference)
    tops = list(range(len(pool)))
    prog = qf.compile_ucc_program(n, mol.hf_reference, mol.hamiltonian, pool, tops)
    tpl = qf.compile_ucc_template(n, mol.hf_reference, mol.hamiltonian, pool, tops)
    gsd = qf.SQOpPool()
    gsd.set_orb_spaces(mol.hf_reference)
    gsd.fill_pool("GSD")
    ops, _, _ = qf.ucc_generators(gsd, list(range(len(gsd))))
    t = list(np.random.default_rng(8).normal(0.0, 0.1, len(tops)))
    a = np.array(tpl.pool_gradients(t, ops, 7))
    b = np.array(prog.pool_gradients(t, ops, 64))
    assert a.shape == b.shape
    assert np.max(np.abs(a - b)) < 1e-10


def test_large_sector_grid_wide_path_matches_program(qf):
    """An 18-qubit template (sector too large for one workgroup's LDS: the
    grid-wide single-request path, one launch per step and per generator)
    against the full-space program: energy, amplitudes and gradient.  The
    last generator is two excitations in one operator (two runs: the
    multi-step branch of the reverse sweep)."""
    from qforte_amd.engine import workloads as W

    n = 18
    norb, na, nb, ref = W.half_filled_reference(n)
    h, g = qf.synthetic_integrals(norb, 31)
    hmap = qf.HamiltonianMap(norb)
    inputs = hmap.pack(0.0, h, g)
    H = hmap.operator(inputs)
    pool = sd_pool(qf, ref)
    tops = list(range(0, len(pool), 7))  # a sample of singles and doubles
    gens, idx, scale = qf.ucc_generators(pool, tops)
    combo = qf.QubitOperator()
    combo.add_op(gens[1])
    combo.add_op(gens[-1])
    gens = list(gens) + [combo]
    idx = list(idx) + [len(tops)]
    scale = list(scale) + [0.5]
    nparam = len(tops) + 1
    ridx = qf.ref_to_basis_idx(ref)
    prog = qf.UCCProgram(n, ridx, gens, idx, scale, nparam, H)
    tpl = qf.SectorTemplate(n, ridx, gens, idx, scale, nparam, hmap)
    assert tpl.dim == 126 * 126
    t = list(np.random.default_rng(3).normal(0.0, 0.1, nparam))
    assert abs(tpl.energy(t, inputs) - prog.energy(t)) < E_TOL
    a, b = qf.Computer(n), qf.Computer(n)
    prog.prepare(t, a)
    tpl.prepare(t, b)
    assert np.max(np.abs(np.asarray(a.get_coeff_vec()) - np.asarray(b.get_coeff_vec()))) < 1e-12
    e, gr = tpl.gradient(t, inputs)
    e_p, g_p = prog.gradient(t)
    assert abs(e - e_p) < E_TOL
    assert np.max(np.abs(np.array(gr) - np.array(g_p))) < 1e-10
    # repeated calls give the same numbers (fixed-order sums)
    e2, gr2 = tpl.gradient(t, inputs)
    assert e2 == e and list(gr2) == list(gr)
    # PQE residual amplitudes <I| U^dag H U |ref> at sector determinants:
    # the reference and the largest amplitudes of U|ref>
    amps = np.asarray(a.get_coeff_vec())
    sel = [int(ridx)] + [int(i) for i in np.argsort(-np.abs(amps))[:40]]
    r_t = np.asarray(tpl.residuals(t, sel, inputs))
    r_p = np.asarray(prog.residuals(t, sel))
    assert np.abs(r_p).max() > 1e-3
    assert np.max(np.abs(r_t - r_p)) < 1e-11
    assert abs(r_p[0].real - e_p) < E_TOL  # <ref|U^dag H U|ref> is the energy


def test_pair_words_and_list_form_agree_bit_for_bit(qf):
    """Templates above the pair-word limit rotate from the string lists; both
    forms apply the same arithmetic to every pair, so energies, gradients
    and served results are identical (tiers 0, 1 and 2: 8, 16 and 18 qubits)."""
    rng = np.random.default_rng(17)
    for norb, na, nb in [(4, 2, 2), (8, 4, 4), (9, 5, 4)]:
        _, _, ref = synthetic_case(qf, norb, na, nb, 0)
        pool = sd_pool(qf, ref)
        hmap = qf.HamiltonianMap(norb)
        tops = list(range(0, len(pool), 3))
        qf.set_sector_word_limit(0)
        try:
            listed = qf.compile_ucc_template(2 * norb, ref, hmap, pool, tops)
        finally:
            qf.set_sector_word_limit(1 << 25)
        worded = qf.compile_ucc_template(2 * norb, ref, hmap, pool, tops)
        h, g = qf.synthetic_integrals(norb, 50 + norb)
        inputs = hmap.pack(0.0, h, g)
        t = list(rng.normal(0.0, 0.1, len(tops)))
        assert listed.energy(t, inputs) == worded.energy(t, inputs), norb
        e1, g1 = listed.gradient(t, inputs)
        e2, g2 = worded.gradient(t, inputs)
        assert e1 == e2 and list(g1) == list(g2), norb
        eng = qf.ServingEngine(hbm_budget_bytes=1 << 30, max_active=64)
        ids = [eng.submit(tpl, t, inputs) for tpl in (listed, worded, listed, worded)]
        gids = [eng.submit(tpl, t, inputs, gradient=True) for tpl in (listed, worded)]
        eng.drain()
        vals = [eng.results[i] for i in ids]
        assert vals[0] == vals[1] == vals[2] == vals[3], norb
        assert list(eng.gradients[gids[0]]) == list(eng.gradients[gids[1]]), norb


def test_captured_sequences_pass_the_host_audit(qf):
    """Every node of a captured single-request sequence is checked on the host
    before the graph is kept (one chain; the expected memcpy or kernel per
    node; non-zero grids, blocks of 1..1024 threads; argument values and copy
    extents equal to the addresses the template owns).  A full 18-qubit SD
    template (the smallest size class on the captured grid-wide path) must
    pass for the energy and the gradient sequence with every kernel's
    arguments compared, and the replayed results must equal the full-space
    program's."""
    from qforte_amd.engine import workloads as W

    n = 18
    norb, na, nb, ref = W.half_filled_reference(n)
    h, g = qf.synthetic_integrals(norb, 41)
    hmap = qf.HamiltonianMap(norb)
    inputs = hmap.pack(0.0, h, g)
    pool = sd_pool(qf, ref)
    tops = list(range(len(pool)))
    tpl = qf.compile_ucc_template(n, ref, hmap, pool, tops)
    for grad in (False, True):
        nodes, kernels, kernels_checked, copies_checked = tpl.graph_audit(grad)
        print(f"audit n={n} gradient={grad}: {nodes} nodes, {kernels} kernels, {kernels_checked} with arguments "
              f"compared, {copies_checked} copies compared", flush=True)
        copies = 3 if grad else 2  # request head up; energy (and generator overlaps) down
        assert nodes == kernels + copies
        assert kernels > len(tops)  # at least one launch per step
        assert kernels_checked == kernels
    t = list(np.random.default_rng(5).normal(0.0, 0.05, len(tops)))
    prog = qf.compile_ucc_program(n, ref, hmap.operator(inputs), pool, tops)
    e, gr = tpl.gradient(t, inputs)
    e_p, g_p = prog.gradient(t)
    assert abs(tpl.energy(t, inputs) - e_p) < E_TOL
    assert abs(e - e_p) < E_TOL
    assert np.max(np.abs(np.array(gr) - np.array(g_p))) < 1e-10
    # a template whose sequence runs in one workgroup is never captured
    _, _, ref8 = synthetic_case(qf, 4, 2, 2, 0)
    small = qf.compile_ucc_template(8, ref8, qf.HamiltonianMap(4), sd_pool(qf, ref8), [0, 1])
    assert small.graph_audit(True) == [0, 0, 0, 0]
