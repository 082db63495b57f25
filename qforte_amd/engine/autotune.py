"""Autotuner for the pass planner: fusion width and tile shape.

The planner has three knobs (``set_tuning``: LDS tile qubits k, contiguous low
qubits L, ops per fused pass; ``set_pass_tuning``: the widest register orbit,
4 to 7 dimensions).  Which setting is fastest depends on the op mix and on
the state size, so the tuner times candidates on the caller's own circuits,
applied to a replica state of ``probe_qubits`` qubits (the circuits must act
only on those qubits), with HIP events on the engine stream, and keeps the
fastest.  Orbit width is chosen on the exponential circuits, the tile shape on
the gate circuits, each with the other knob at its current value.
"""

import json

DEFAULT_TILES = ((10, 4, 16), (11, 4, 16), (11, 4, 24), (11, 4, 32), (12, 4, 32))
DEFAULT_ORBITS = (4, 5, 6, 7)


def _time_circuits(qf, qc, circuits, reps):
    best = None
    for _ in range(reps):
        qf.synchronize()
        t = qf.DeviceTimer()
        t.start()
        for c in circuits:
            qc.apply_circuit(c)
        ms = t.stop()
        best = ms if best is None else min(best, ms)
    return best


def autotune(qf, probe_qubits, gate_circuits=(), exp_circuits=(), tiles=DEFAULT_TILES, orbits=DEFAULT_ORBITS,
             reps=1, log=None):
    """Pick (tile_qubits, low_qubits, max_ops) and the orbit width that run
    ``gate_circuits`` and ``exp_circuits`` fastest on a ``probe_qubits``
    replica; applies the choice and returns a report dict."""
    qc = qf.Computer(probe_qubits)
    report = {"probe_qubits": probe_qubits, "orbit_ms": {}, "tile_ms": {}}
    start_tiles = qf.get_tuning()
    if exp_circuits:
        for c in exp_circuits:  # warm: first launches, staging buffers
            qc.apply_circuit(c)
        for o in orbits:
            qf.set_pass_tuning(o)
            ms = _time_circuits(qf, qc, exp_circuits, reps)
            report["orbit_ms"][o] = round(ms, 3)
            if log:
                log(f"autotune orbit_dims={o}: {ms:.1f} ms")
        best_o = min(report["orbit_ms"], key=report["orbit_ms"].get)
        qf.set_pass_tuning(best_o)
    report["orbit_dims"] = qf.get_pass_tuning()
    if gate_circuits:
        for c in gate_circuits:
            qc.apply_circuit(c)
        for tun in tiles:
            qf.set_tuning(*tun)
            ms = _time_circuits(qf, qc, gate_circuits, reps)
            report["tile_ms"][",".join(map(str, tun))] = round(ms, 3)
            if log:
                log(f"autotune tile={tun}: {ms:.1f} ms")
        best_t = min(report["tile_ms"], key=report["tile_ms"].get)
        qf.set_tuning(*[int(v) for v in best_t.split(",")])
    else:
        qf.set_tuning(*start_tiles)
    report["tuning"] = list(qf.get_tuning())
    del qc
    return report


if __name__ == "__main__":  # pragma: no cover - manual use
    import qforte_amd as qf
    from qforte_amd.engine import workloads as W

    n = 24
    L = W.to_circuit(qf, W.brickwork_layers(n, 20, seed=101))
    U, _ = qf.trotterize(W.exp_operator(qf, W.random_pauli_strings(n, 400, seed=102)))
    print(json.dumps(autotune(qf, n, [L], [U], log=print)))
