"""Continuous-batching serving loop for UCC energy requests.

Requests carry a compiled program and its amplitudes.  The native engine
(``qfe_batch``) admits them into active slots while the HBM budget
(16 * 2^n bytes per state) allows, advances all active requests together and
retires them as they finish; this module adds arrival-time bookkeeping and
latency statistics.
"""

import time

from .. import _core


class ServingEngine:
    def __init__(self, hbm_budget_bytes=64 << 30, max_active=256):
        self._eng = _core.BatchEngine(int(hbm_budget_bytes), int(max_active))
        self._arrival = {}
        self.latencies = {}
        self.results = {}

    def submit(self, program, amplitudes, arrival_time=None):
        rid = self._eng.submit(program, list(amplitudes))
        self._arrival[rid] = time.perf_counter() if arrival_time is None else arrival_time
        return rid

    def step(self):
        done = self._eng.step()
        now = time.perf_counter()
        for rid, e in done:
            self.results[rid] = e
            self.latencies[rid] = now - self._arrival.pop(rid)
        return done

    def pending(self):
        return self._eng.pending()

    def drain(self):
        while True:
            q, a = self.pending()
            if q == 0 and a == 0:
                return
            self.step()
