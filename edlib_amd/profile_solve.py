"""Profiling workload for rocprofv3: one batched impurity solve.

Anderson impurity, 1 orbital + 11 bath sites (seeded bath), U=4, all 169
(N_up, N_dn) sectors, NEV=4, NCV=11, tol 1e-14, then the impurity Green's
function (4 start vectors per thermally relevant pair, beta=10, 256
Matsubara points, 60 Lanczos steps), everything through the batched engine.
A first solve warms up the code paths; the second is the one reported, with
the wall time of its eigensolver and Green's-function phases.
With --kanamori14: the served stream's 14-site request instead (Kanamori
3 orbitals + 11 bath sites, U=4, J=0.7, half filling, all 225 sectors, G of
the three impurity orbitals), one solve.
Usage: python -m edlib_amd.profile_solve [--plain] [--no-graphs] [--kanamori14]
(--no-graphs launches every sweep's kernels one by one, for counter
collection, whose per-dispatch sampling does not follow replayed graphs)
"""
import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

SEED = 20261202


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--plain", action="store_true", help="also time the plain per-sector solve")
    ap.add_argument("--sweeps", type=int, default=16, help="sweeps per Batch.step call")
    ap.add_argument("--no-graphs", action="store_true", help="no hipGraph replay of steady sweeps")
    ap.add_argument("--kanamori14", action="store_true", help="the 14-site Kanamori request, one solve")
    a = ap.parse_args()
    from edlib_amd.batch import Batch
    from edlib_amd.solver import BatchedSolve, SolveParams, solve_plain
    from edlib_amd.workloads import siam_model
    if a.kanamori14:
        sp = SolveParams(nev=4, beta=10.0, nomega=256, nlanc=60, orbitals=(0, 1, 2), seed=SEED)
        m = siam_model(14, 3, SEED + 1000 + 1)
    else:
        sp = SolveParams(nev=4, beta=10.0, nomega=256, nlanc=60, orbitals=(0,), seed=SEED)
        m = siam_model(12, 1, SEED + 4)
    b = Batch(0)
    if a.no_graphs:
        b.set_graphs(False)
    for rep in range(1 if a.kanamori14 else 2):
        t0 = time.perf_counter()
        s = BatchedSolve(b, m, sp)
        s.start()
        if rep == 0:
            b.autotune(list(s.sid.values()))
        t1 = time.perf_counter()
        while s.phase in ("new", "eig"):
            b.step(a.sweeps)
            s.advance()
        t2 = time.perf_counter()
        while not s.advance():
            b.step(a.sweeps)
        t3 = time.perf_counter()
        print("solve %d: setup %.3f s, eigensolver %.3f s, Green's function %.3f s, total %.3f s (%d pairs)"
              % (rep, t1 - t0, t2 - t1, t3 - t2, t3 - t0, s.result["n_gf"]), flush=True)
    b.close()
    if a.plain:
        t0 = time.perf_counter()
        solve_plain(m, sp)
        print("plain: %.3f s" % (time.perf_counter() - t0), flush=True)


if __name__ == "__main__":
    main()
