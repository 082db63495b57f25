#!/usr/bin/env python3
"""Benchmark of the Computer hot path on one MI355X.

Workload (one "step"): a 30-qubit complex128 state (16 GiB) reset to |0...0>,
then 40 brickwork layers (per-qubit Ry(U[0,2pi)) Rz(U[0,2pi)), then cX on
alternating neighbour pairs, control = lower index), then exp(-i theta P) for
2000 random Pauli strings of weight 4-8 (theta ~ N(0, 0.1)), then <psi|H|psi>
for a random 1000-term Pauli sum (weight 1-6, N(0,1) coefficients).  Inputs
are resident in HBM when the timed region starts.

``value`` is the achieved HBM bandwidth of the whole step: the algorithmic
bytes of every pass the engine launched (2 * 16 * 2^n per read-write sweep,
16 * 2^n per read-only expectation pass over one X-mask group) divided by the
step time.  ``roofline`` reports the dominant kernel (the register-orbit
passes of the exponential stage) timed with HIP events on the engine stream,
against the 8 TB/s HBM peak.  ``serving`` and ``gradient`` report the UCC
energy serving leg and the 20-qubit differentiation leg.  ``cpu_baseline`` times the reference's own C++ Computer
(compiled from its sources into oracle/_ref) single-threaded on a bounded
26-qubit sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
# HBM bytes per launch of the exponential-stage kernel, by (qubits, orbit
# width), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch),
# each scaled by the same counter's reading on k_scale, a plain sweep that
# reads and writes exactly 16 GiB (FETCH 8388621-8388622 KiB: factor 2.0000;
# WRITE 16778250-16778282 KiB: factor 0.99994).
#   k_orbit_lanes<4, 3>: FETCH 8388762, WRITE 16782743 -> 33559207 KiB
#   k_orbit_lanes<4, 2>: FETCH 8388759, WRITE 16786363 -> 33562824 KiB
TRAFFIC_PER_LAUNCH = {(30, 7): 33559207 * 1024, (30, 6): 33562824 * 1024}
METRIC = "HBM GB/s (% peak) at 30 qubits; batched UCC energy evals/s + p99 vs plain"


def build_workload(qf, W, n, nlayers, nexp, nterms):
    layers = W.brickwork_layers(n, nlayers, seed=101)
    strings = W.random_pauli_strings(n, nexp, seed=102)
    terms = W.random_pauli_sum(n, nterms, seed=103)
    layer_circ = W.to_circuit(qf, layers)
    U, phase = qf.trotterize(W.exp_operator(qf, strings))
    assert phase == 1.0
    H = W.to_operator(qf, terms)
    return layers, strings, terms, layer_circ, U, H


def time_exp_stage(qf, qc, U):
    """Average launch duration of the passes of the exponential stage (the
    dominant kernel: register-orbit passes, one read-write sweep of the state
    each), HIP events on the engine stream around one application of U."""
    qf.synchronize()
    qf.reset_counters()
    t = qf.DeviceTimer()
    t.start()
    qc.apply_circuit(U)
    ms = t.stop()
    c = qf.get_counters()
    passes = c["fused_passes"] + c["single_passes"]
    return ms / passes, passes


def time_single_sweep(qf, W, qc, n, reps=40):
    """Average duration of the single Pauli-exponential sweep (one op per
    pass) over `reps` back-to-back launches, HIP events on the engine stream."""
    strings = W.random_pauli_strings(n, reps, seed=104)
    U, _ = qf.trotterize(W.exp_operator(qf, strings))
    saved = qf.get_tuning()
    qf.set_tuning(0, 0, 1)  # one op per pass: every exponential is its own sweep
    try:
        qc.apply_circuit(U)  # warm
        qf.synchronize()
        t = qf.DeviceTimer()
        t.start()
        qc.apply_circuit(U)
        ms = t.stop()
    finally:
        qf.set_tuning(*saved)
    return ms / reps


def time_sampling(qf, qc, S, draws=1 << 20):
    """Computational-basis draws from the resident state (measure_circuit's
    sampling step, computer.cc:166-188): one probability scan of the state
    (S bytes read) plus one wave per draw; wall clock per call, indices
    included in host memory."""
    qc.sample(draws, 1)
    qf.synchronize()
    t0 = time.perf_counter()
    for k in range(3):
        idx = qc.sample(draws, 2 + k)
    dt = (time.perf_counter() - t0) / 3
    assert int(idx.max()) < (1 << qc.get_nqubit())
    return {"kernels": "k_sample_scan + k_sample_chunks + k_sample_guide + k_sample_draw", "draws": draws,
            "ms_per_call": round(dt * 1e3, 3), "scan_bytes": S,
            "effective_GBps": round(S / dt / 1e9, 1), "draws_per_s": round(draws / dt, 1)}


def time_single_gate_sweep(qf, qc, n, reps=40, seed=105):
    """Average duration of a dense one-qubit gate sweep (Ry then Rz merged
    into one 2x2 on a random qubit, one gate per pass), HIP events."""
    import math

    import numpy as np

    rng = np.random.default_rng(seed)
    c = qf.Circuit()
    for _ in range(reps):
        q = int(rng.integers(n))
        c.add(qf.gate("Ry", q, q, float(rng.uniform(0.0, 2 * math.pi))))
        c.add(qf.gate("Rz", q, q, float(rng.uniform(0.0, 2 * math.pi))))
    saved = qf.get_tuning()
    qf.set_tuning(0, 0, 1)
    try:
        qc.apply_circuit(c)  # warm
        qf.synchronize()
        t = qf.DeviceTimer()
        t.start()
        qc.apply_circuit(c)
        ms = t.stop()
    finally:
        qf.set_tuning(*saved)
    return ms / reps


def time_attainable_sweep(qf, qc, reps=10):
    """Average duration of the plain in-place sweep (apply_constant: read and
    write every amplitude once, no pairing), the attainable streaming rate."""
    qc.apply_constant(1.0)
    qf.synchronize()
    t = qf.DeviceTimer()
    t.start()
    for _ in range(reps):
        qc.apply_constant(1.0)
    return t.stop() / reps


def run_open_loop(qf, classes, reqs, rate, seed):
    """Poisson arrivals at `rate` requests/s into a fresh serving engine;
    returns (makespan seconds, latencies in arrival order, results)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    eng = qf.ServingEngine(hbm_budget_bytes=64 << 30, max_active=8192)
    ids = []
    i = 0
    t0 = time.perf_counter()
    while len(eng.results) < len(reqs):
        now = time.perf_counter() - t0
        while i < len(reqs) and arrivals[i] <= now:
            n, inputs, t = reqs[i]
            ids.append(eng.submit(classes[n]["template"], t, inputs, arrival_time=t0 + arrivals[i]))
            i += 1
        if not eng.poll() and i < len(reqs):
            wait = arrivals[i] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(min(wait, 2e-4))
    makespan = time.perf_counter() - t0
    lat = np.array([eng.latencies[r] for r in ids])
    return makespan, lat, [eng.results[r] for r in ids]


def latency_by_class(reqs, lat):
    """p50 / p99 latency (ms) of each size class of an open-loop run."""
    import numpy as np

    out = {}
    for n in sorted({r[0] for r in reqs}):
        v = np.array([lat[i] for i, r in enumerate(reqs) if r[0] == n])
        out[n] = {"p50_ms": round(1e3 * float(np.percentile(v, 50)), 3), "p99_ms": round(1e3 * float(np.percentile(v, 99)), 3)}
    return out


def serving_leg(qf, W, nreq, plain_per_class):
    """Continuous-batching UCC energy serving: capacity (all requests queued
    at once), three Poisson rates relative to it, and the plain
    one-request-per-launch paths on a per-class sample."""
    import numpy as np

    t_setup = time.perf_counter()
    classes = W.ucc_size_classes(qf)
    reqs = W.ucc_requests(qf, classes, nreq, seed=7)
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] serving: {len(classes)} size classes compiled, {nreq} requests built in {setup_s:.1f} s",
          file=sys.stderr, flush=True)

    # warm every class once (first-launch costs), then capacity: all queued at t=0;
    # plain_per_class == 0 leaves every single-request call out (batched lanes only)
    if plain_per_class > 0:
        for n, c in classes.items():
            c["template"].energy([0.0] * c["template"].nparam, [0.0] * c["template"].ninputs)
    eng = qf.ServingEngine(hbm_budget_bytes=64 << 30, max_active=8192)
    t0 = time.perf_counter()
    ids = [eng.submit(classes[n]["template"], t, inputs, arrival_time=t0) for n, inputs, t in reqs]
    eng.drain()
    cap_s = time.perf_counter() - t0
    batched = [eng.results[r] for r in ids]
    capacity = nreq / cap_s
    print(f"[bench] serving capacity {capacity:.0f} evals/s ({cap_s:.3f} s for {nreq})", file=sys.stderr, flush=True)

    rates = {}
    for f in (0.5, 0.9, 1.5):
        span, lat, res = run_open_loop(qf, classes, reqs, f * capacity, seed=int(f * 10))
        same = float(np.max(np.abs(np.array(res) - np.array(batched))))
        rates[f"{f}x"] = {
            "offered_per_s": round(f * capacity, 1),
            "evals_per_s": round(nreq / span, 1),
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "max_abs_diff_vs_capacity_run": same,
            "by_class": latency_by_class(reqs, lat),
        }
        print(f"[bench] serving at {f}x: {rates[f'{f}x']}", file=sys.stderr, flush=True)

    # one request at a time through the serving engine: a class's service time
    # with nothing else queued, the reference for its latency under load
    alone, alone_diff = {}, 0.0
    eng1 = qf.ServingEngine(hbm_budget_bytes=64 << 30, max_active=8192)
    for n in sorted(classes):
        lat1 = []
        for i in [i for i, r in enumerate(reqs) if r[0] == n][:8]:
            _, inputs, t = reqs[i]
            rid = eng1.submit(classes[n]["template"], t, inputs)
            eng1.drain()
            lat1.append(eng1.latencies[rid])
            alone_diff = max(alone_diff, abs(eng1.results[rid] - batched[i]))
        alone[n] = float(np.median(lat1))
    print(f"[bench] served alone, median ms by class: { {n: round(1e3 * v, 3) for n, v in alone.items()} }",
          file=sys.stderr, flush=True)

    # plain paths on a sample: per class, the first plain_per_class requests
    by_class = {n: [i for i, r in enumerate(reqs) if r[0] == n][:plain_per_class] for n in classes}
    plain_full, plain_sector, create_full, plain_graph, capture, max_err = {}, {}, {}, {}, {}, 0.0
    for n, idx in by_class.items():
        if not idx:
            continue
        c = classes[n]
        gens, pidx, scale = qf.ucc_generators(c["pool"], c["tops"])
        tf, ts, tc, tg, tcap = [], [], [], [], []
        for i in idx:
            _, inputs, t = reqs[i]
            a = time.perf_counter()
            prog = qf.UCCProgram(n, qf.ref_to_basis_idx(c["ref"]), gens, pidx, scale, len(c["tops"]),
                                 c["map"].operator(inputs))
            b = time.perf_counter()
            e_full = prog.energy(t)
            d = time.perf_counter()
            e_sec = c["template"].energy(t, inputs)
            f = time.perf_counter()
            prog.energy_graph(t)  # plan with device-side angles and capture
            g = time.perf_counter()
            e_graph = prog.energy_graph(t)  # replay
            h = time.perf_counter()
            tc.append(b - a)
            tf.append(d - b)
            ts.append(f - d)
            tcap.append(g - f)
            tg.append(h - g)
            max_err = max(max_err, abs(e_full - batched[i]), abs(e_sec - batched[i]), abs(e_graph - batched[i]))
            del prog
        plain_full[n], plain_sector[n], create_full[n] = float(np.mean(tf)), float(np.mean(ts)), float(np.mean(tc))
        plain_graph[n], capture[n] = float(np.mean(tg)), float(np.mean(tcap))
        print(f"[bench] plain n={n}: full-space {1e3 * plain_full[n]:.3f} ms/eval (+{1e3 * create_full[n]:.1f} ms "
              f"program build), graph replay {1e3 * plain_graph[n]:.3f} ms (capture {1e3 * capture[n]:.1f} ms), "
              f"sector one-launch {1e3 * plain_sector[n]:.3f} ms/eval", file=sys.stderr, flush=True)
    # the request mix is uniform over the classes
    plain_rate = 1.0 / float(np.mean(list(plain_full.values()))) if plain_full else None
    sector_rate = 1.0 / float(np.mean(list(plain_sector.values()))) if plain_sector else None
    grad = serving_gradient_leg(qf, classes, reqs, by_class)
    return {
        "workload": (f"{nreq} UCCSD energy requests, n uniform over {sorted(classes)}, half filled, own seeded "
                     "synthetic integrals per request, amplitudes N(0,0.05)"),
        "capacity_evals_per_s": round(capacity, 1),
        "poisson": rates,
        "served_alone_ms_by_n": {n: round(1e3 * v, 3) for n, v in alone.items()},
        "served_alone_max_abs_diff_vs_capacity_run": alone_diff,
        "plain_full_space_evals_per_s": round(plain_rate, 2) if plain_rate else None,
        "plain_sector_one_launch_evals_per_s": round(sector_rate, 2) if sector_rate else None,
        "speedup_vs_plain": round(capacity / plain_rate, 1) if plain_rate else None,
        "speedup_vs_plain_sector": round(capacity / sector_rate, 1) if sector_rate else None,
        "plain_ms_per_eval_by_n": {n: round(1e3 * v, 3) for n, v in plain_full.items()},
        "plain_sector_ms_per_eval_by_n": {n: round(1e3 * v, 3) for n, v in plain_sector.items()},
        "plain_graph_replay_ms_by_n": {n: round(1e3 * v, 3) for n, v in plain_graph.items()},
        "plain_graph_capture_ms_by_n": {n: round(1e3 * v, 2) for n, v in capture.items()},
        "plain_program_build_ms_by_n": {n: round(1e3 * v, 2) for n, v in create_full.items()},
        "plain_sample_per_class": plain_per_class,
        "max_abs_energy_diff_vs_plain": max_err,
        "sector_dims": {n: c["template"].dim for n, c in classes.items()},
        "gradient": grad,
    }


def serving_gradient_leg(qf, classes, reqs, by_class):
    """The same requests as energy + gradient requests: capacity (all queued
    at once) against the plain per-request gradient paths on the sample."""
    import numpy as np

    eng = qf.ServingEngine(hbm_budget_bytes=64 << 30, max_active=8192)
    t0 = time.perf_counter()
    ids = [eng.submit(classes[n]["template"], t, inputs, arrival_time=t0, gradient=True) for n, inputs, t in reqs]
    eng.drain()
    cap_s = time.perf_counter() - t0
    capacity = len(reqs) / cap_s
    print(f"[bench] gradient serving capacity {capacity:.0f} evals/s ({cap_s:.3f} s for {len(reqs)})",
          file=sys.stderr, flush=True)
    plain_full, plain_sector, max_err = {}, {}, 0.0
    for n, idx in by_class.items():
        if not idx:
            continue
        c = classes[n]
        gens, pidx, scale = qf.ucc_generators(c["pool"], c["tops"])
        tf, ts = [], []
        for i in idx:
            _, inputs, t = reqs[i]
            prog = qf.UCCProgram(n, qf.ref_to_basis_idx(c["ref"]), gens, pidx, scale, len(c["tops"]),
                                 c["map"].operator(inputs))
            a = time.perf_counter()
            e_full, g_full = prog.gradient(t)
            b = time.perf_counter()
            e_sec, g_sec = c["template"].gradient(t, inputs)
            d = time.perf_counter()
            tf.append(b - a)
            ts.append(d - b)
            g_srv = np.array(eng.gradients[ids[i]])
            max_err = max(max_err, float(np.max(np.abs(g_srv - np.array(g_full)))),
                          float(np.max(np.abs(g_srv - np.array(g_sec)))), abs(eng.results[ids[i]] - e_full))
            del prog
        plain_full[n], plain_sector[n] = float(np.mean(tf)), float(np.mean(ts))
        print(f"[bench] plain gradient n={n}: full-space {1e3 * plain_full[n]:.3f} ms, sector one-request "
              f"{1e3 * plain_sector[n]:.3f} ms", file=sys.stderr, flush=True)
    plain_rate = 1.0 / float(np.mean(list(plain_full.values()))) if plain_full else None
    sector_rate = 1.0 / float(np.mean(list(plain_sector.values()))) if plain_sector else None
    return {
        "capacity_evals_per_s": round(capacity, 1),
        "plain_full_space_evals_per_s": round(plain_rate, 2) if plain_rate else None,
        "plain_sector_one_request_evals_per_s": round(sector_rate, 2) if sector_rate else None,
        "speedup_vs_plain": round(capacity / plain_rate, 1) if plain_rate else None,
        "speedup_vs_plain_sector": round(capacity / sector_rate, 1) if sector_rate else None,
        "plain_ms_per_gradient_by_n": {n: round(1e3 * v, 3) for n, v in plain_full.items()},
        "plain_sector_ms_per_gradient_by_n": {n: round(1e3 * v, 3) for n, v in plain_sector.items()},
        "max_abs_diff_vs_plain": max_err,
    }


def timed(fn, reps=3):
    """Median wall time (s) of fn() over reps calls (each call returns to the host)."""
    ts, out = [], None
    for _ in range(reps):
        a = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - a)
    return sorted(ts)[len(ts) // 2], out


def gradient_leg(qf, W, nqubit=20, seed=20, fd_components=6, full_pool_sample=128):
    """Full UCCSD gradient at 20 qubits through torch.autograd (adjoint
    backward), checked against central differences and the full-space
    program, plus ADAPT GSD pool gradients in batches of 64."""
    import numpy as np
    import torch

    from qforte_amd.engine.autograd import SectorProgram, UCCEnergy

    norb, na, nb, ref = W.half_filled_reference(nqubit)
    h, g = qf.synthetic_integrals(norb, seed)
    hmap = qf.HamiltonianMap(norb)
    inputs = hmap.pack(0.0, h, g)
    H = hmap.operator(inputs)
    pool = qf.SQOpPool()
    pool.set_orb_spaces(ref)
    pool.fill_pool("SD")
    tops = list(range(len(pool)))
    t = np.random.default_rng(seed).normal(0.0, 0.05, len(tops))
    tl = [float(v) for v in t]

    a = time.perf_counter()
    prog = qf.compile_ucc_program(nqubit, ref, H, pool, tops)
    build_full = time.perf_counter() - a
    a = time.perf_counter()
    tpl = qf.compile_ucc_template(nqubit, ref, hmap, pool, tops)
    build_sector = time.perf_counter() - a
    sp = SectorProgram(tpl, inputs)

    prog.energy(tl)
    sp.gradient(tl)  # warm
    t_e_full, e_full = timed(lambda: prog.energy(tl))
    t_g_full, (_, g_full) = timed(lambda: prog.gradient(tl))
    t_e_sec, e_sec = timed(lambda: sp.energy(tl))

    def autograd_step():
        x = torch.tensor(t, dtype=torch.float64, requires_grad=True)
        UCCEnergy.apply(x, sp).backward()
        return x.grad.numpy().copy()

    t_g_sec, g_auto = timed(autograd_step)
    g_full = np.array(g_full)
    fd_err = 0.0
    hstep = 1e-5
    for k in np.random.default_rng(1).choice(len(tops), fd_components, replace=False):
        tp, tm = t.copy(), t.copy()
        tp[k] += hstep
        tm[k] -= hstep
        fd = (sp.energy(list(tp)) - sp.energy(list(tm))) / (2 * hstep)
        fd_err = max(fd_err, abs(fd - g_auto[k]))
    print(f"[bench] gradient n={nqubit}: full-space {1e3 * t_g_full:.1f} ms, sector autograd {1e3 * t_g_sec:.1f} ms",
          file=sys.stderr, flush=True)

    gsd = qf.SQOpPool()
    gsd.set_orb_spaces(ref)
    gsd.fill_pool("GSD")
    a = time.perf_counter()
    ops, _, _ = qf.ucc_generators(gsd, list(range(len(gsd))))
    jw_s = time.perf_counter() - a
    t_pool_sec, pg = timed(lambda: tpl.pool_gradients(tl, ops, 64, inputs), reps=2)
    sample = ops[:full_pool_sample]
    t_pool_full, pg_full = timed(lambda: prog.pool_gradients(tl, sample, 64), reps=1)
    pool_err = float(np.max(np.abs(np.array(pg[: len(sample)]) - np.array(pg_full))))
    print(f"[bench] ADAPT GSD pool: {len(ops)} ops, sector {1e3 * t_pool_sec:.1f} ms, full-space "
          f"{1e3 * t_pool_full / len(sample):.3f} ms/op", file=sys.stderr, flush=True)
    return {
        "workload": (f"UCCSD gradient at {nqubit} qubits ({norb} orbitals, {na}a+{nb}b, {len(tops)} amplitudes, "
                     f"{sum(len(q.terms()) for q in qf.ucc_generators(pool, tops)[0])} Pauli exponentials, "
                     f"{hmap.nterms} Hamiltonian terms), synthetic integrals seed {seed}, amplitudes N(0,0.05)"),
        "sector_dim": tpl.dim,
        "energy_ms": {"full_space": round(1e3 * t_e_full, 3), "sector": round(1e3 * t_e_sec, 3)},
        "gradient_ms": {"full_space_adjoint": round(1e3 * t_g_full, 3), "sector_adjoint_autograd": round(1e3 * t_g_sec, 3)},
        "build_ms": {"full_space": round(1e3 * build_full, 1), "sector": round(1e3 * build_sector, 1)},
        "max_abs_grad_diff_sector_vs_full": float(np.max(np.abs(g_auto - g_full))),
        "energy_diff_sector_vs_full": abs(e_sec - e_full),
        "max_abs_fd_error": fd_err,
        "fd_step": hstep,
        "adapt_gsd_pool": {
            "ops": len(ops),
            "batch": 64,
            "sector_ms": round(1e3 * t_pool_sec, 3),
            "full_space_ms_per_op": round(1e3 * t_pool_full / len(sample), 4),
            "full_space_sample": len(sample),
            "max_abs_diff_vs_full_space": pool_err,
            "jw_compile_s": round(jw_s, 2),
        },
    }


def host_description():
    """The host the CPU baseline ran on: cores this process may use, CPU model
    and RAM (the baseline itself runs on one of those cores)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    ram = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    ram = round(int(line.split()[1]) / 2**20, 1)
                    break
    except OSError:
        pass
    return {"affinity_cores": len(os.sched_getaffinity(0)), "logical_cpus": os.cpu_count(), "cpu_model": model,
            "ram_gib": ram, "threads_used": 1}


def cpu_baseline(W, n_cpu, oracle_ref_loader, qf=None):
    """Reference Computer (oracle/_ref), one thread, on a bounded sample:
    1 brickwork layer + 5 exponentials + 5 expectation terms at n_cpu qubits.
    With ``qf`` the same sample also runs on the GPU (the exponentials as
    one-pass Pauli exponentials rather than gate ladders) and is compared with
    the reference: energy and 4096 sampled amplitudes."""
    ref = oracle_ref_loader()
    if ref is None:
        return None
    import importlib.util

    spec = importlib.util.spec_from_file_location("sv_oracle", os.path.join(ROOT, "oracle", "sv_oracle.py"))
    oracle = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(oracle)
    layers = W.brickwork_layers(n_cpu, 1, seed=101)
    strings = W.random_pauli_strings(n_cpu, 5, seed=102)
    terms = W.random_pauli_sum(n_cpu, 5, seed=103)
    circ = W.to_circuit(ref, layers)
    exp_circ = ref.Circuit()
    for theta, term in strings:
        for g in oracle.exponentiate_gates(-1j * theta, term):
            exp_circ.add(ref.gate(g.gtype, g.target, g.control, g.parameter) if g.gtype in ("Rx", "Ry", "Rz")
                         else ref.gate(g.gtype, g.target, g.control))
    H = W.to_operator(ref, terms)
    qc = ref.Computer(n_cpu)
    t0 = time.perf_counter()
    qc.apply_circuit(circ)
    qc.apply_circuit(exp_circ)
    e = qc.direct_op_exp_val(H)
    dt = time.perf_counter() - t0
    S = 16.0 * (1 << n_cpu)
    # same accounting as the GPU's effective rate: 2S per gate or exponential, S per term
    eff_bytes = 2 * S * (len(layers) + len(strings)) + S * len(terms)
    parity = None
    if qf is not None:
        import numpy as np

        g = qf.Computer(n_cpu)
        g.apply_circuit(W.to_circuit(qf, layers))
        U, _ = qf.trotterize(W.exp_operator(qf, strings))
        g.apply_circuit(U)
        e_gpu = g.direct_op_exp_val(W.to_operator(qf, terms))
        amps = np.asarray(g.get_coeff_vec())
        idx = np.random.default_rng(7).choice(1 << n_cpu, 4096, replace=False)
        ref_amps = np.array([complex(qc.coeff(ref.QubitBasis(int(i)))) for i in idx])
        parity = {"energy_abs_diff": float(abs(complex(e) - complex(e_gpu))),
                  "amplitude_max_abs_diff": float(np.max(np.abs(amps[idx] - ref_amps))),
                  "amplitudes_compared": len(idx)}
        del g, amps
    return {
        "seconds": dt,
        "eff_bytes": eff_bytes,
        "gates": len(layers),
        "exps": len(strings),
        "terms": len(terms),
        "energy": complex(e),
        "reference_gate_sweeps": len(layers) + exp_circ.size() + sum(len(t[1]) + 4 for t in terms),
        "parity": parity,
    }


def _median_ms(fn, reps, warm=2):
    import numpy as np

    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        a = time.perf_counter()
        fn()
        ts.append(1e3 * (time.perf_counter() - a))
    return float(np.median(ts))


def small_configs_leg(qf, W, oracle_ref_loader):
    """The two small configurations, end to end through the Python API (host
    synchronisation included), next to the reference C++ Computer
    (oracle/_ref, one thread) on the same inputs:
    C1: 12 qubits from |0..0>, random_circuit(12, 500, seed=7), then <H> of
        random_pauli_sum(12, 200, seed=8) (the inputs of
        tests/golden/c1_reference.npz);
    C2: the 16-qubit UCCSD energy (8 orbitals, 4a+4b, synthetic integrals seed
        2024, SD pool: 360 amplitudes, 2688 Pauli exponentials, amplitudes
        N(0, 0.05) seed 16) through the full-space program, its captured graph
        and the sector template; the reference applies the same exponentials
        as its own gate ladders and takes <H> of the JW Hamiltonian.
    Both run with the library's default tile shape (chosen by state size),
    not the shape tuned for the 30-qubit workload."""
    import importlib.util

    import numpy as np

    saved_tuning = qf.get_tuning()
    qf.set_tuning(-1, -1, -1)
    try:
        return _small_configs(qf, W, oracle_ref_loader, importlib, np)
    finally:
        qf.set_tuning(*saved_tuning)


def _small_configs(qf, W, oracle_ref_loader, importlib, np):
    ref = oracle_ref_loader()
    spec = importlib.util.spec_from_file_location("sv_oracle", os.path.join(ROOT, "oracle", "sv_oracle.py"))
    oracle = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(oracle)

    def terms_of(op):
        return [(c, [(g.gate_id(), g.target()) for g in circ.gates()]) for c, circ in op.terms()]

    def ladders(mod, exps):
        c = mod.Circuit()
        for coef, term in exps:
            for g in oracle.exponentiate_gates(coef, term):
                c.add(mod.gate(g.gtype, g.target, g.control, g.parameter) if g.gtype in ("Rx", "Ry", "Rz")
                      else mod.gate(g.gtype, g.target, g.control))
        return c

    out = {}
    # C1
    n = 12
    gates = W.random_circuit(n, 500, seed=7)
    terms = W.random_pauli_sum(n, 200, seed=8)
    circ, H = W.to_circuit(qf, gates), W.to_operator(qf, terms)
    qc = qf.Computer(n)

    def c1_gpu():
        qc.reset()
        qc.apply_circuit(circ)
        return qc.direct_op_exp_val(H)

    c1 = {"gpu_ms": round(_median_ms(c1_gpu, 20), 3)}
    e_gpu = c1_gpu()
    amps = np.asarray(qc.get_coeff_vec())
    if ref is not None:
        rc, rH = W.to_circuit(ref, gates), W.to_operator(ref, terms)
        rq = ref.Computer(n)
        a = time.perf_counter()
        rq.apply_circuit(rc)
        e_ref = rq.direct_op_exp_val(rH)
        c1["reference_cpu_ms"] = round(1e3 * (time.perf_counter() - a), 3)
        c1["speedup"] = round(c1["reference_cpu_ms"] / c1["gpu_ms"], 1)
        ref_amps = np.array(rq.get_coeff_vec(), dtype=np.complex128)
        c1["energy_abs_diff"] = float(abs(complex(e_gpu) - complex(e_ref)))
        c1["amplitude_max_abs_diff"] = float(np.max(np.abs(amps - ref_amps)))
    out["c1_12q_500_gates_200_terms"] = c1
    del qc
    # C2
    norb, nq = 8, 16
    h, g = qf.synthetic_integrals(norb, 2024)
    Hq = qf.build_sq_hamiltonian(h, g, 0.0).jw_transform()
    occ = qf.system.hamiltonians.hf_occupation(norb, 4, 4)
    pool = qf.SQOpPool()
    pool.set_orb_spaces(occ)
    pool.fill_pool("SD")
    tops = list(range(len(pool)))
    gens, _, _ = qf.ucc_generators(pool, tops)
    t = list(np.random.default_rng(16).normal(0.0, 0.05, len(tops)))
    exps = [(t[k] * c, term) for k, g_ in enumerate(gens) for c, term in terms_of(g_)]
    a = time.perf_counter()
    prog = qf.compile_ucc_program(nq, occ, Hq, pool, tops)
    build_prog_ms = 1e3 * (time.perf_counter() - a)
    hmap = qf.HamiltonianMap(norb)
    inputs = hmap.pack(0.0, h, g)
    a = time.perf_counter()
    tpl = qf.compile_ucc_template(nq, occ, hmap, pool, tops)
    build_tpl_ms = 1e3 * (time.perf_counter() - a)
    c2 = {"exponentials": len(exps), "hamiltonian_terms": len(Hq.terms()),
          "gpu_full_space_ms": round(_median_ms(lambda: prog.energy(t), 20), 3),
          "gpu_full_space_graph_ms": round(_median_ms(lambda: prog.energy_graph(t), 20), 3),
          "gpu_sector_ms": round(_median_ms(lambda: tpl.energy(t, inputs), 20), 3),
          "build_ms": {"full_space": round(build_prog_ms, 1), "sector": round(build_tpl_ms, 1)}}
    e_prog, e_graph, e_tpl = prog.energy(t), prog.energy_graph(t), tpl.energy(t, inputs)
    c2["energy"] = e_prog
    c2["graph_minus_eager"] = float(abs(e_graph - e_prog))
    c2["sector_minus_full_space"] = float(abs(e_tpl - e_prog))
    if ref is not None:
        rc = ref.Circuit()
        for q in range(nq):
            if occ[q]:
                rc.add(ref.gate("X", q, q))
        rc.add(ladders(ref, exps))
        rH = ref.QubitOperator()
        for c, term in terms_of(Hq):
            cc = ref.Circuit()
            for letter, q in term:
                cc.add(ref.gate(letter, q, q))
            rH.add(c, cc)
        rq = ref.Computer(nq)
        a = time.perf_counter()
        rq.apply_circuit(rc)
        e_ref = rq.direct_op_exp_val(rH)
        c2["reference_cpu_ms"] = round(1e3 * (time.perf_counter() - a), 1)
        c2["reference_gates"] = rc.size()
        c2["speedup_full_space"] = round(c2["reference_cpu_ms"] / c2["gpu_full_space_ms"], 1)
        c2["speedup_sector"] = round(c2["reference_cpu_ms"] / c2["gpu_sector_ms"], 1)
        c2["energy_abs_diff_vs_reference"] = float(abs(complex(e_ref).real - e_prog))
    out["c2_16q_uccsd_energy"] = c2
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nqubit", type=int, default=30)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--nexp", type=int, default=2000)
    ap.add_argument("--nterms", type=int, default=1000)
    ap.add_argument("--cpu-nqubit", type=int, default=26)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-roofline", action="store_true", help="skip the dominant-kernel timing")
    ap.add_argument("--tuning", type=str, default="", help="tile_qubits,low_qubits,max_ops (skips the autotuner)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default fusion width and tile shape")
    ap.add_argument("--no-serving", action="store_true", help="skip the UCC serving leg")
    ap.add_argument("--no-small", action="store_true", help="skip the 12- and 16-qubit configurations")
    ap.add_argument("--no-gradient", action="store_true", help="skip the 20-qubit gradient leg")
    ap.add_argument("--gradient-only", action="store_true", help="only the 20-qubit gradient leg (printed as JSON)")
    ap.add_argument("--serving-only", action="store_true", help="only the UCC serving leg (printed as JSON)")
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--plain-per-class", type=int, default=8)
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the exponential stage and a calibration sweep (for hardware-counter runs)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported by this build "
              "(the state-vector path runs on one GPU).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import qforte_amd as qf
    from qforte_amd.engine import workloads as W

    if qf.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.tuning:
        qf.set_tuning(*[int(v) for v in args.tuning.split(",")])
    tune_report = None

    if args.serving_only:
        print(json.dumps(serving_leg(qf, W, args.requests, args.plain_per_class)))
        return
    if args.gradient_only:
        print(json.dumps(gradient_leg(qf, W)))
        return

    n = args.nqubit
    if not args.no_autotune and not args.tuning and n >= 20 and not args.roofline_only:
        # fusion width and tile shape timed on a 2-qubit-smaller replica of the workload
        from qforte_amd.engine.autotune import autotune

        m = n - 2
        rl, _, _, rlc, rU, _ = build_workload(qf, W, m, args.layers, args.nexp, 1)
        tune_report = autotune(qf, m, [rlc], [rU],
                               log=lambda msg: print("[bench] " + msg, file=sys.stderr, flush=True))
        del rl, rlc, rU
    print(f"[bench] building workload n={n} layers={args.layers} exps={args.nexp} terms={args.nterms}",
          file=sys.stderr, flush=True)
    layers, strings, terms, layer_circ, U, H = build_workload(qf, W, n, args.layers, args.nexp, args.nterms)
    qc = qf.Computer(n)

    if args.roofline_only:
        # short run for hardware-counter passes: the exponential stage once,
        # then two plain read-write sweeps (k_scale) as the calibration kernel
        qc.apply_circuit(U)
        avg_ms, passes = time_exp_stage(qf, qc, U)
        qc.apply_constant(1.0)
        qc.apply_constant(1.0)
        qf.synchronize()
        print(json.dumps({"exp_stage_passes": passes, "avg_launch_ms": round(avg_ms, 4),
                          "state_bytes": 16 * (1 << n)}))
        return

    def step():
        qc.reset()
        qc.apply_circuit(layer_circ)
        qc.apply_circuit(U)
        return qc.direct_op_exp_val(H)

    for i in range(args.warmup):
        step()
        qf.synchronize()
        print(f"[bench] warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    qf.reset_counters()
    qf.synchronize()
    t0 = time.perf_counter()
    energy = None
    for i in range(args.steps):
        energy = step()
        qf.synchronize()
        print(f"[bench] step {i + 1}/{args.steps} done", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    counters = qf.get_counters()
    norm_err = abs(qf.norm2(qc) - 1.0)
    S = 16.0 * (1 << n)
    ms_per_step = 1000.0 * dt / args.steps
    gbps = counters["bytes"] / dt / 1e9
    eff_bytes = 2 * S * (len(layers) + len(strings)) + S * len(terms)
    eff_gbps = eff_bytes * args.steps / dt / 1e9

    roofline = None
    if not args.no_roofline:
        avg_ms, passes = time_exp_stage(qf, qc, U)
        achieved = 2 * S / (avg_ms * 1e-3) / 1e9
        single_ms = time_single_sweep(qf, W, qc, n)
        gate_ms = time_single_gate_sweep(qf, qc, n)
        copy_ms = time_attainable_sweep(qf, qc)
        dims = qf.get_pass_tuning()
        if dims > 4:
            kname = (f"k_orbit_lanes<4, {dims - 4}> ({dims}-dimension register orbits over {1 << (dims - 4)} lanes "
                     f"of 16 amplitudes")
        else:
            kname = f"k_orbit<{dims}> ({dims}-dimension register orbits in one lane"
        roofline = {
            "kernel": (kname + "; the exponential stage: %d passes for %d exponentials, 2S = %.2f GB per launch)"
                       % (passes, len(strings), 2 * S / 1e9)),
            "stage_ms": round(avg_ms * passes, 1),
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": TRAFFIC_PER_LAUNCH.get((n, dims)),
            "avg_launch_ms": round(avg_ms, 4),
            "single_exponential_sweep": {"kernel": "k_pexp", "avg_launch_ms": round(single_ms, 4),
                                         "achieved": round(2 * S / (single_ms * 1e-3) / 1e9, 1),
                                         "frac": round(2 * S / (single_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
            "single_gate_sweep": {"kernel": "k_u1 (dense 2x2 on one qubit)", "avg_launch_ms": round(gate_ms, 4),
                                  "achieved": round(2 * S / (gate_ms * 1e-3) / 1e9, 1),
                                  "frac": round(2 * S / (gate_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
            "attainable_in_place_sweep": {"kernel": "k_scale (reads and writes every amplitude once)",
                                          "avg_launch_ms": round(copy_ms, 4),
                                          "achieved": round(2 * S / (copy_ms * 1e-3) / 1e9, 1)},
            "sampling": time_sampling(qf, qc, S),
        }

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from conftest import load_reference

        print(f"[bench] cpu baseline: reference Computer at {args.cpu_nqubit} qubits", file=sys.stderr, flush=True)
        res = cpu_baseline(W, args.cpu_nqubit, load_reference, qf)
        if res is not None:
            cpu = {
                "value": round(res["eff_bytes"] / res["seconds"] / 1e9, 3),
                "unit": "GB/s effective (2S per gate or exponential + S per expectation term)",
                "cores": 1,
                "kind": "reference",
                "sample": (f"reference C++ Computer (oracle/_ref, default non-OpenMP build) at {args.cpu_nqubit} "
                           f"qubits: 1 brickwork layer ({res['gates']} gates) + {res['exps']} Pauli exponentials "
                           f"(gate ladders) + {res['terms']}-term expectation, {res['seconds']:.2f} s"),
                "seconds": round(res["seconds"], 3),
                "host": host_description(),
                "why_one_thread": ("the reference's gate and exponential sweeps are serial loops; its one OpenMP "
                                   "loop (expectation terms, computer.cc:895) compiles serially in its default "
                                   "build (CMakeLists.txt:10 USE_OpenMP OFF, the OpenMP block commented out)"),
                "gpu_effective_GBps_same_unit": round(eff_gbps, 1),
                "parity_vs_gpu": res["parity"],
            }
    serving = None
    if not args.no_serving:
        del qc
        serving = serving_leg(qf, W, args.requests, args.plain_per_class)
    gradient = None
    if not args.no_gradient:
        gradient = gradient_leg(qf, W)
    small = None
    if not args.no_small:
        print("[bench] 12- and 16-qubit configurations", file=sys.stderr, flush=True)
        loader = lambda: None  # noqa: E731
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            from conftest import load_reference as loader
        small = small_configs_leg(qf, W, loader)
    line = {
        "metric": METRIC,
        "value": round(gbps, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128 (f64)",
        "data": "synthetic (seeded brickwork layers, random Pauli strings and Pauli sum)",
        "config": {
            "workload": f"{n}-qubit roofline: {args.layers} brickwork layers + {args.nexp} Pauli exponentials "
                        f"(weight 4-8) + {args.nterms}-term expectation",
            "nqubit": n,
            "state_bytes": int(S),
            "tuning": list(qf.get_tuning()),
            "orbit_dims": qf.get_pass_tuning(),
            "autotune": tune_report,
        },
        "pct_of_peak": round(100.0 * gbps / HBM_PEAK_GBPS, 2),
        "effective_GBps": round(eff_gbps, 1),
        "passes_per_step": {
            "fused": counters["fused_passes"] // args.steps,
            "single": counters["single_passes"] // args.steps,
            "expectation_groups": counters["reduce_passes"] // args.steps,
        },
        "energy": [energy.real, energy.imag],
        "norm_error": norm_err,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "serving": serving,
        "gradient": gradient,
        "small_configs": small,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
