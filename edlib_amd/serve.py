"""Serving: continuous batching of impurity solves on one GPU.

Requests (an impurity model plus solve parameters) arrive over time.  A
request is admitted when its estimated device footprint fits in the HBM
budget next to the requests already running; admitted requests share one
Batch, so their eigensolver and Lanczos runs advance in the same grouped
launches while each request moves through its own phases (eigenpairs of
every sector, then the Green's-function start vectors).  Requests that
arrive while others run join at the next sweep.  Latency is measured from
arrival to result.  Each request's arithmetic is independent of the others
(see batch.py), so its result is the same alone and in the mix.
"""
import collections
import math
import time

import numpy as np

from . import binom
from .batch import Batch
from .solver import BatchedSolve, sector_list


def footprint_bytes(Ns, p):
    """Upper estimate of a solve's device memory: Krylov bases (two of
    ncv+1 vectors), eigenvectors and tables over every sector, plus three
    vectors per concurrent Green's-function run of the largest sector."""
    m = min(p.ncv_eff, 47)
    tot = 0
    big = 0
    for s in sector_list(Ns, p):
        n = binom(Ns, s[0]) * binom(Ns, s[1])
        tot += n * 8 * (2 * (m + 1) + p.nev + 1)
        big = max(big, n)
    ngf = 4 * len(p.orbitals) * p.nev * 4
    return tot + 3 * 8 * big * ngf + (64 << 20)


class Server:
    def __init__(self, device=0, hbm_budget=200 << 30, sweeps_per_tick=4, max_active=64):
        self.batch = Batch(device)
        self.budget = hbm_budget
        self.sweeps = sweeps_per_tick
        self.max_active = max_active
        self.queue = collections.deque()
        self.active = []
        self.done = {}
        self.reserved = 0

    def submit(self, rid, model, params, t_arrival=None):
        self.queue.append((rid, model, params, time.perf_counter() if t_arrival is None else t_arrival))

    def _admit(self):
        while self.queue and len(self.active) < self.max_active:
            rid, model, p, ta = self.queue[0]
            need = footprint_bytes(model.Ns, p)
            if self.active and self.reserved + need > self.budget:
                break  # wait for memory; the running requests free theirs when they finish
            self.queue.popleft()
            s = BatchedSolve(self.batch, model, p)
            s.start()
            self.active.append((rid, s, ta, time.perf_counter(), need))
            self.reserved += need

    def tick(self):
        self._admit()
        if self.active:
            self.batch.step(self.sweeps)
        still = []
        for rid, s, ta, tadm, need in self.active:
            if s.advance():
                now = time.perf_counter()
                self.done[rid] = dict(s.result, latency=now - ta, queued=tadm - ta, t_done=now)
                self.reserved -= need
            else:
                still.append((rid, s, ta, tadm, need))
        self.active = still
        return bool(self.queue or self.active)

    def run_stream(self, requests, rate, seed):
        """Poisson arrivals at `rate` requests/s (exponential gaps from
        `seed`); requests = [(rid, model, params)].  Returns the results and
        the wall time from the first arrival to the last completion."""
        rng = np.random.default_rng(seed)
        gaps = rng.exponential(1.0 / rate, len(requests)) if rate > 0 else np.zeros(len(requests))
        t0 = time.perf_counter()
        arrivals = t0 + np.cumsum(gaps) - gaps[0]
        i = 0
        while i < len(requests) or self.queue or self.active:
            now = time.perf_counter()
            while i < len(requests) and arrivals[i] <= now:
                rid, model, p = requests[i]
                self.submit(rid, model, p, t_arrival=arrivals[i])
                i += 1
            if not self.tick() and i < len(requests):
                time.sleep(max(0.0, min(arrivals[i] - time.perf_counter(), 0.01)))
        return self.done, time.perf_counter() - t0

    def close(self):
        self.batch.close()


def latency_stats(done):
    lat = np.array([r["latency"] for r in done.values()])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "n": int(len(lat))}
