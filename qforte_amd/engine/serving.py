"""Continuous-batching serving loop for UCC energy and gradient requests.

A request is a ``SectorTemplate`` (one size class: ansatz and Hamiltonian
structure), its amplitudes and its Hamiltonian inputs (integrals packed by
``HamiltonianMap.pack``).  The native engine (``qfe_batch``) queues requests
on the host; every ``step``/``poll`` launches what can be admitted onto one of
a few streams (one workgroup per request, requests of any size class in the
same launch, bounded by the HBM budget), so work that arrives while other
launches run is admitted into the next launch and launches overlap on the
device.  A gradient request (``gradient=True``) returns the energy and the
adjoint gradient of the amplitudes (``SectorTemplate.gradient``); it is
launched with other gradient requests of its size class.  This module adds
arrival-time bookkeeping and latency statistics.
"""

import time

import numpy as np

from .. import _core


class ServingEngine:
    def __init__(self, hbm_budget_bytes=64 << 30, max_active=4096):
        self._eng = _core.BatchEngine(int(hbm_budget_bytes), int(max_active))
        self._arrival = {}
        self.latencies = {}
        self.results = {}
        self.gradients = {}
        self._grad_len = {}

    def submit(self, template, amplitudes, inputs=(), arrival_time=None, gradient=False):
        rid = self._eng.submit(template, list(amplitudes), list(inputs), bool(gradient))
        self._arrival[rid] = time.perf_counter() if arrival_time is None else arrival_time
        if gradient:
            self._grad_len[rid] = template.nparam
        return rid

    def _record(self, done):
        now = time.perf_counter()
        for rid, e in done:
            self.results[rid] = e
            n = self._grad_len.pop(rid, None)
            if n is not None:
                self.gradients[rid] = self._eng.take_gradient(rid, n)
            self.latencies[rid] = now - self._arrival.pop(rid)
        return done

    def step(self):
        """Launch what can be admitted; wait for the oldest launch only when
        nothing has completed yet."""
        return self._record(self._eng.step())

    def poll(self):
        """Launch what can be admitted and return what has completed, without waiting."""
        return self._record(self._eng.poll())

    def pending(self):
        return self._eng.pending()

    def drain(self):
        while True:
            q, a = self.pending()
            if q == 0 and a == 0:
                return
            self.step()

    def latency_percentiles(self, ids=None, q=(50, 99)):
        lat = np.array([self.latencies[i] for i in (ids if ids is not None else self.latencies)])
        return {f"p{p}": float(np.percentile(lat, p)) for p in q}
