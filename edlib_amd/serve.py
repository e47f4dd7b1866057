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


def table_bytes_estimate(Ns, n_up, n_dn, nterms):
    """Upper estimate of one sector's device tables: per spin the basis
    patterns, single-spin diagonal, CSR (row pointers, columns, values,
    packed entries) and sliced ELL with at most `nterms` entries per row,
    the rank tables (2^Ns per spin) and the impurity interaction table."""
    d_up, d_dn = binom(Ns, n_up), binom(Ns, n_dn)
    per_state = 4 + 8 + 4 + nterms * (4 + 8 + 4 + 4 + 4)
    return (d_up + d_dn) * per_state + 2 * 4 * (1 << Ns) + (1 << 20)


def footprint_bytes(Ns, p, nterms=None):
    """Upper estimate of a solve's device memory: Krylov bases (two of
    ncv+1 vectors), eigenvectors and tables over every sector, plus three
    vectors per concurrent Green's-function run of the largest sector
    (4 start vectors per orbital and pair, for 4 pairs per sector)."""
    m = min(p.ncv_eff, 47)
    nterms = 2 * Ns if nterms is None else nterms
    tot = 0
    big = 0
    for s in sector_list(Ns, p):
        n = binom(Ns, s[0]) * binom(Ns, s[1])
        tot += n * 8 * (2 * (m + 1) + p.nev + 1) + table_bytes_estimate(Ns, s[0], s[1], nterms)
        big = max(big, n)
    ngf = 4 * (len(p.gf_orbitals) + len(p.pairs)) * p.nev * 4
    return tot + 3 * 8 * big * ngf


def device_budget(device=0, reserve=0.08):
    """HBM budget for admission: the device's free memory (hipMemGetInfo)
    less a `reserve` fraction of its total (the batch's own descriptors,
    reduction scratch and the allocator's slack)."""
    import ctypes
    from . import _lib
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(_lib.lib().edh_device_memory(device, ctypes.byref(free), ctypes.byref(total)))
    return max(0, int(free.value - reserve * total.value))


class Server:
    """Continuous batching with admission by footprint.

    A request is admitted (FIFO) when its footprint fits next to the running
    requests' in the HBM budget; one whose footprint exceeds the whole
    budget is rejected with an error result instead of being run.  A request
    whose solve raises is ended with an error result (its processes and
    sectors are freed) while the others go on.  ``batch``, ``solve_cls`` and
    ``footprint`` can be replaced (tests drive the admission logic on the
    host with stand-ins).
    """

    def __init__(self, device=0, hbm_budget=None, sweeps_per_tick=4, max_active=64, batch=None,
                 solve_cls=BatchedSolve, footprint=None):
        self.batch = Batch(device) if batch is None else batch
        self.budget = device_budget(device) if hbm_budget is None else int(hbm_budget)
        self.sweeps = sweeps_per_tick
        self.max_active = max_active
        self.solve_cls = solve_cls
        self.footprint = footprint or (lambda model, p: footprint_bytes(model.Ns, p))
        self.queue = collections.deque()
        self.active = []
        self.done = {}
        self.reserved = 0
        self.peak_reserved = 0
        self.blocked_by = collections.Counter()  # what held the queue head back: "memory" / "slots"

    def submit(self, rid, model, params, t_arrival=None):
        self.queue.append((rid, model, params, time.perf_counter() if t_arrival is None else t_arrival))

    def _fail(self, rid, ta, msg, solve=None, tadm=None):
        if solve is not None:
            try:
                solve.abort()
            except Exception:  # noqa: BLE001 -- best effort: the request has already failed
                pass
        now = time.perf_counter()
        self.done[rid] = {"error": msg, "latency": now - ta, "queued": (tadm or now) - ta, "t_done": now}

    def _admit(self):
        while self.queue:
            rid, model, p, ta = self.queue[0]
            need = self.footprint(model, p)
            if need > self.budget:
                self.queue.popleft()
                self._fail(rid, ta, "footprint %d bytes exceeds the HBM budget of %d bytes" % (need, self.budget))
                continue
            if len(self.active) >= self.max_active:
                self.blocked_by["slots"] += 1
                break
            if self.reserved + need > self.budget:
                self.blocked_by["memory"] += 1
                break  # wait for memory; the running requests free theirs when they finish
            self.queue.popleft()
            tadm = time.perf_counter()
            s = self.solve_cls(self.batch, model, p)
            try:
                s.start()
            except Exception as e:  # noqa: BLE001 -- isolate the failing request
                self._fail(rid, ta, "start failed: %s" % e, s, tadm)
                continue
            self.active.append((rid, s, ta, tadm, need))
            self.reserved += need
            self.peak_reserved = max(self.peak_reserved, self.reserved)

    def tick(self):
        self._admit()
        if self.active:
            self.batch.step(self.sweeps)
        still = []
        for rid, s, ta, tadm, need in self.active:
            try:
                finished = s.advance()
            except Exception as e:  # noqa: BLE001 -- isolate the failing request
                self._fail(rid, ta, "solve failed: %s" % e, s, tadm)
                self.reserved -= need
                continue
            if finished:
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
    lat = np.array([r["latency"] for r in done.values() if "error" not in r])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "n": int(len(lat))}
