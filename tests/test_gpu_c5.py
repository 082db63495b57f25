"""The differentiation workload at full size: UCCSD at 20 qubits.

10 spatial orbitals, 5 alpha + 5 beta electrons, synthetic integrals, 875
amplitudes drawn N(0, 0.05).  The energy gradient comes from the hand-written
adjoint backward behind ``torch.autograd`` (``UCCEnergy``; the recursion of
uccvqeabc.py:122-251).  It is checked three ways: against central finite
differences (h = 1e-5, tolerance 1e-8) on sampled components; sector path
against full-space path (1e-10), each run through autograd; and ADAPT
commutator gradients 2 Re<H psi|A_k psi> of the generalized singles and
doubles pool (uccvqeabc.py:253-283), 64 operators per launch, against the
full-space program on whole batches.  The full-space program is itself pinned
at small sizes against the reference's gate-level recursion and the dense
oracle (test_gpu_ucc.py).
"""

import numpy as np
import pytest

from qforte_amd.engine import workloads as W

pytestmark = pytest.mark.gpu
N = 20


@pytest.fixture(scope="module")
def c5(qf, gpu):
    norb, na, nb, ref = W.half_filled_reference(N)
    assert (norb, na, nb) == (10, 5, 5)
    h, g = qf.synthetic_integrals(norb, 20)
    hmap = qf.HamiltonianMap(norb)
    inputs = hmap.pack(0.0, h, g)
    pool = qf.SQOpPool()
    pool.set_orb_spaces(ref)
    pool.fill_pool("SD")
    tops = list(range(len(pool)))
    assert len(tops) == 875
    t = np.random.default_rng(20).normal(0.0, 0.05, len(tops))
    prog = qf.compile_ucc_program(N, ref, hmap.operator(inputs), pool, tops)
    tpl = qf.compile_ucc_template(N, ref, hmap, pool, tops)
    return {"ref": ref, "inputs": inputs, "t": t, "prog": prog, "tpl": tpl}


def autograd_gradient(program, t):
    import torch

    from qforte_amd.engine.autograd import UCCEnergy

    x = torch.tensor(t, dtype=torch.float64, requires_grad=True)
    e = UCCEnergy.apply(x, program)
    e.backward()
    return float(e), x.grad.numpy().copy()


def test_c5_gradient_against_finite_differences(qf, c5):
    from qforte_amd.engine.autograd import SectorProgram

    sp = SectorProgram(c5["tpl"], c5["inputs"])
    t = c5["t"]
    e_sec, g_sec = autograd_gradient(sp, t)
    e_full, g_full = autograd_gradient(c5["prog"], t)
    print(f"c5: E = {e_sec:.12f}, |g| = {np.linalg.norm(g_sec):.3e}", flush=True)
    assert abs(e_sec - e_full) < 1e-10
    assert np.max(np.abs(g_sec - g_full)) < 1e-10
    h = 1e-5
    comps = list(np.random.default_rng(2).choice(len(t), 8, replace=False)) + [int(np.argmax(np.abs(g_sec)))]
    worst = 0.0
    for k in comps:
        tp, tm = t.copy(), t.copy()
        tp[k] += h
        tm[k] -= h
        fd = (sp.energy(tp) - sp.energy(tm)) / (2 * h)
        worst = max(worst, abs(fd - g_sec[k]))
        assert abs(fd - g_sec[k]) < 1e-8, (k, fd, g_sec[k])
    print(f"c5: worst central-difference error {worst:.2e} over {len(comps)} components", flush=True)
    assert np.abs(g_sec).max() > 1e-3  # a non-trivial gradient


def test_c5_gsd_pool_gradients_in_batches_of_64(qf, c5):
    gsd = qf.SQOpPool()
    gsd.set_orb_spaces(c5["ref"])
    gsd.fill_pool("GSD")
    ops, _, _ = qf.ucc_generators(gsd, list(range(len(gsd))))
    print(f"c5: GSD pool of {len(ops)} operators", flush=True)
    assert len(ops) > 6000
    tl = [float(v) for v in c5["t"]]
    pg = np.array(c5["tpl"].pool_gradients(tl, ops, 64, c5["inputs"]))
    assert pg.shape == (len(ops),)
    # whole batches of 64 at the start, the middle and the end of the pool
    for start in (0, (len(ops) // 2) // 64 * 64, len(ops) - 64):
        batch = ops[start:start + 64]
        ref_vals = np.array(c5["prog"].pool_gradients(tl, batch, 64))
        assert np.max(np.abs(pg[start:start + 64] - ref_vals)) < 1e-10, start
    assert np.abs(pg).max() > 1e-3
