"""Server admission by HBM footprint, on the host with stand-ins for the
batch and the solves (no GPU): a request waits while the running ones hold
the memory it needs, is admitted as soon as they free it, is rejected with
an error result when it can never fit, and a request whose solve fails ends
with an error result while the others finish."""
import math

import pytest

from edlib_amd import serve
from edlib_amd.solver import SolveParams


class FakeBatch:
    def __init__(self):
        self.sweeps = 0

    def step(self, n):
        self.sweeps += n

    def close(self):
        pass


class FakeModel:
    def __init__(self, name, fp, ticks, fail=None):
        self.name, self.fp, self.ticks, self.fail = name, fp, ticks, fail


class FakeSolve:
    log = []

    def __init__(self, batch, model, p):
        self.b, self.m, self.left, self.result, self.aborted = batch, model, model.ticks, None, False

    def start(self):
        FakeSolve.log.append(("start", self.m.name, self.b.sweeps))
        if self.m.fail == "start":
            raise RuntimeError("no memory for the tables")

    def advance(self):
        self.left -= 1
        if self.m.fail == "advance" and self.left == 1:
            raise RuntimeError("hipMallocAsync failed")
        if self.left <= 0:
            self.result = {"name": self.m.name}
            FakeSolve.log.append(("done", self.m.name, self.b.sweeps))
            return True
        return False

    def abort(self):
        self.aborted = True
        FakeSolve.log.append(("abort", self.m.name, self.b.sweeps))


def _server(budget, max_active=64):
    FakeSolve.log = []
    return serve.Server(hbm_budget=budget, sweeps_per_tick=1, max_active=max_active, batch=FakeBatch(),
                        solve_cls=FakeSolve, footprint=lambda m, p: m.fp)


def _drain(srv):
    n = 0
    while srv.tick():
        n += 1
        assert n < 1000
    return srv.done


def test_request_waits_for_memory_and_is_admitted_when_it_frees():
    srv = _server(100)
    srv.submit("A", FakeModel("A", 60, 3), None, t_arrival=0.0)
    srv.submit("B", FakeModel("B", 60, 2), None, t_arrival=0.0)
    srv.submit("C", FakeModel("C", 30, 2), None, t_arrival=0.0)
    done = _drain(srv)
    assert set(done) == {"A", "B", "C"} and all("error" not in r for r in done.values())
    starts = {n: s for kind, n, s in FakeSolve.log if kind == "start"}
    ends = {n: s for kind, n, s in FakeSolve.log if kind == "done"}
    assert starts["A"] == 0
    assert starts["B"] == ends["A"]          # admitted at the first tick after A freed its memory
    assert starts["C"] == starts["B"]        # FIFO: C fits but waits behind B
    assert srv.blocked_by["memory"] > 0 and srv.blocked_by["slots"] == 0
    assert srv.peak_reserved <= 100


def test_slot_limit_is_separate_from_memory():
    srv = _server(1000, max_active=1)
    for n in "AB":
        srv.submit(n, FakeModel(n, 10, 2), None, t_arrival=0.0)
    _drain(srv)
    assert srv.blocked_by["slots"] > 0 and srv.blocked_by["memory"] == 0


def test_request_larger_than_budget_is_rejected():
    srv = _server(100)
    srv.submit("big", FakeModel("big", 101, 1), None, t_arrival=0.0)
    srv.submit("ok", FakeModel("ok", 50, 1), None, t_arrival=0.0)
    done = _drain(srv)
    assert "exceeds the HBM budget" in done["big"]["error"]
    assert "error" not in done["ok"]
    assert not any(n == "big" for kind, n, _ in FakeSolve.log if kind == "start")


@pytest.mark.parametrize("where", ["start", "advance"])
def test_failing_request_is_isolated(where):
    srv = _server(100)
    srv.submit("bad", FakeModel("bad", 40, 4, fail=where), None, t_arrival=0.0)
    srv.submit("good", FakeModel("good", 40, 3), None, t_arrival=0.0)
    done = _drain(srv)
    assert "failed" in done["bad"]["error"]
    assert done["good"]["name"] == "good"
    assert ("abort", "bad") in [(k, n) for k, n, _ in FakeSolve.log]
    assert srv.reserved == 0


def test_footprint_estimate_counts_bases_vectors_and_tables():
    p = SolveParams(nev=4, orbitals=(0, 1, 2))
    f14 = serve.footprint_bytes(14, p)
    # two bases of ncv+1 = 12 vectors and nev+1 eigenvector slots over 4^14
    # states, then three vectors of the largest sector per Green's-function run
    bases = 4 ** 14 * 8 * (2 * 12 + 5)
    gf = 3 * 8 * 3432 ** 2 * (4 * 3 * 4 * 4)
    assert bases + gf < f14 < (bases + gf) * 1.1
    tables = sum(serve.table_bytes_estimate(14, a, b, 28) for a in range(15) for b in range(15))
    assert tables > 15 * 15 * 8 * (1 << 14)
    assert serve.footprint_bytes(9, p) < serve.footprint_bytes(11, p) < f14
    assert math.isfinite(f14)
