#!/usr/bin/env python3
"""Benchmark: spin-resolved H*v on one MI355X (the north-star workload).

Workload: 16-site 4x4 periodic Hubbard cluster, t=1, U=6, mu=3, half-filled
Sz=0 sector (N_up, N_dn) = (8, 8), dimension 12870^2 = 165,636,900, FP64.
One step = one Lanczos-form H*v over the whole sector, w = H v + w
(Lanczos.h:55 calls av(v, w, n, false)), with v and w resident in HBM.

Algorithmic bytes per step: B = 24 N + T (read v, read w, write w, plus the
per-spin hopping tables T); the diagonal and interaction terms are generated
on the fly.  value = B / (wall time per step); roofline.achieved = B / (mean
kernel time from HIP events on the launch stream).

Also reported (same JSON line):
* pcie_inclusive_gbs: the same H*v called with host vectors (staged over
  PCIe both ways, as the reference's host-pointer av() contract does);
* solves: impurity solves/s of the batched engine and of the continuous-
  batching server against the plain per-sector GPU solve on the same
  problems, with p50/p99 latency (see edlib_amd/solver.py, serve.py), and
  the roofline of the batched engine's grouped H*v launch over all 169
  sectors of the Anderson 1 + 11 model (solves.batched.grouped_hv).

Prints ONE JSON line.  --gpus N>1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from edlib_amd.workloads import siam_model, torus  # noqa: E402

METRIC = "H·v GB/s (% HBM peak); impurity solves/s + p99 latency, batched vs plain"


def progress(msg):
    """One line on stderr per finished part (stdout carries the JSON line only)."""
    print("bench: " + msg, file=sys.stderr, flush=True)
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def load_traffic(variant, kernel_ms):
    """Memory traffic per launch of the given H*v variant, measured by the
    rocprofv3 FETCH_SIZE/WRITE_SIZE counters and kept under profiles/ with
    the kernel time of the profiled build; None when absent or when this
    run's kernel time differs from that by more than 5 % (the kernel has
    changed since the counters were taken)."""
    p = os.path.join(ROOT, "profiles", "hv_16site_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    t = d.get("traffic_bytes_per_launch_by_variant", {}).get(str(variant))
    ms = d.get("kernel_ms_by_variant", {}).get(str(variant))
    if t is None or ms is None or abs(kernel_ms - ms) > 0.05 * ms:
        return None
    return t


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(t, U, xmu, v, w, B):
    """SpinResolvedStorage::av restatement (oracle), OpenMP over up-rows, on
    this host: 3 accumulate-mode calls over the full sector."""
    from oracle import edoracle as eo
    m = eo.hubbard(16, t, U, xmu)
    srs = eo.SRS(m, 8, 8)
    srs.av(v, w, clear=False, omp=True)  # warm-up
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        srs.av(v, w, clear=False, omp=True)
    dt = (time.perf_counter() - t0) / reps
    return {"value": B / dt / 1e9, "unit": "GB/s", "cores": int(eo.lib().orc_num_threads()), "kind": "port",
            "sample": "3 accumulate-mode av calls over the full 165,636,900-state sector (oracle restatement of "
                      "SpinResolvedStorage::av, OpenMP over up-rows; SOCRSStorage needs ~106 GB here)",
            "ms_per_call": dt * 1e3, "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


def cpu_socrs_configs():
    """SOCRSStorage restatement (oracle; SOCRSStorage.h:46-87 av with the
    :114-160 fill, OpenMP over static row chunks) on the largest sector of
    each smaller workload: accumulate-mode av GB/s at 24 N bytes per call,
    median of 5 calls after one warm-up (fill time reported apart)."""
    from oracle import edoracle as eo
    from edlib_amd.workloads import siam_arrays
    ring = np.zeros((8, 8))
    for i in range(8):
        ring[i, (i + 1) % 8] = ring[(i + 1) % 8, i] = 1.0
    tt = torus(3, 4)
    for r in range(3):
        for c in range(4):
            i = r * 4 + c
            for (dr, dc) in ((1, 1), (1, -1)):
                j = ((r + dr) % 3) * 4 + (c + dc) % 4
                tt[i, j] = tt[j, i] = -0.3
    cfgs = [("8-site ring, sector (4,4)", lambda: eo.hubbard(8, ring, np.full(8, 4.0), np.full(8, 2.0)), (4, 4)),
            ("12-site 3x4 torus t,t', sector (6,6)", lambda: eo.hubbard(12, tt, np.full(12, 8.0), np.full(12, 4.0)),
             (6, 6)),
            ("AIM 1+11, sector (6,6)", lambda: eo.siam(*siam_arrays(12, 1, SEED + 4)), (6, 6)),
            ("Kanamori 3+11, sector (7,7)", lambda: eo.siam(*siam_arrays(14, 3, SEED + 1000)), (7, 7))]
    out = []
    for name, mk, sec in cfgs:
        m = mk()
        t0 = time.perf_counter()
        s = eo.SOCRS(m, *sec)
        tf = time.perf_counter() - t0
        v = np.random.default_rng(SEED + 1).uniform(-1, 1, s.n)
        w = np.random.default_rng(SEED + 2).uniform(-1, 1, s.n)
        w = s.av(v, w, clear=False)
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            w = s.av(v, w, clear=False)
            ts.append(time.perf_counter() - t0)
        dt = float(np.median(ts))
        out.append({"workload": name, "dim": s.n, "av_ms": round(dt * 1e3, 3), "gbs": round(24 * s.n / dt / 1e9, 3),
                    "fill_s": round(tf, 2), "threads": int(eo.lib().orc_num_threads())})
        progress("CPU SOCRS av, %s: %.1f ms" % (name, dt * 1e3))
        del s
    return out


def cpu_solve_estimate(hv_by_sector, nlanc):
    """CPU time of the AIM 1+11 batched solve's H*v work: for every sector,
    the H*v count the device solve needed there (eigensolver and
    Green's-function steps) times one SOCRSStorage av of that sector on this
    host (oracle restatement, OpenMP), summed.  A lower bound of a CPU
    solve: ARPACK's re-orthogonalisation and the host loops are not counted."""
    from oracle import edoracle as eo
    from edlib_amd.workloads import siam_arrays
    m = eo.siam(*siam_arrays(12, 1, SEED + 4))
    tot, nav = 0.0, 0
    for sec, cnt in sorted(hv_by_sector.items()):
        s = eo.SOCRS(m, *sec)
        v = np.random.default_rng(SEED + 3).uniform(-1, 1, s.n)
        w = s.av(v)
        t0 = time.perf_counter()
        s.av(v, w, clear=False)
        tot += (time.perf_counter() - t0) * cnt
        nav += cnt
        del s
    progress("CPU H*v time of the AIM 1+11 solve's H*v count: %.1f s" % tot)
    return {"cpu_s_per_solve_lower_bound": round(tot, 3), "solves_per_s_upper_bound": round(1.0 / tot, 5),
            "hv_count": nav,
            "what": "sum over sectors of (H*v count of the device solve in that sector) x (one SOCRS av of that "
                    "sector on this host); excludes ARPACK's orthogonalisation, so a CPU solve is slower"}



def grouped_hv_roofline(model, Ns, reps=50):
    """Roofline of the batched engine's grouped H*v: one launch over one
    operand per sector of every (N_up, N_dn) sector (an eigensolver sweep
    with all sectors active, w = H v), timed with HIP events on the batch's
    stream; algorithmic bytes per launch = sum over the sectors of 16 n + T."""
    from edlib_amd.batch import Batch
    b = Batch(0)
    try:
        secs = [(u, d) for u in range(Ns + 1) for d in range(Ns + 1)]
        ids, _ = b.add_sectors(model, secs)
        b.autotune(ids)
        ms, nbytes, nel = b.time_hv(ids, reps, SEED + 5)
    finally:
        b.close()
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "edh::hv_batch_wave_kernel", "sectors": len(secs), "elements": nel,
            "algorithmic_bytes": nbytes, "kernel_ms": round(ms, 4), "achieved_gbs": round(gbs, 1),
            "peak_gbs": HBM_PEAK_GBS, "frac": round(gbs / HBM_PEAK_GBS, 4),
            "form": "w = H v, one seeded operand per sector, the steady sweep's launch (graph replay), "
                    "mean of %d launches" % reps}


def solves_section(args):
    """Batched and served impurity solves against the plain per-sector solve."""
    from edlib_amd.serve import Server, latency_stats
    from edlib_amd.solver import SolveParams, solve_batched, solve_plain
    out = {}
    # batched: one Anderson impurity + 11 bath sites, all 169 sectors
    sp = SolveParams(nev=4, beta=10.0, nomega=256, nlanc=60, orbitals=(0,), seed=SEED)
    m = siam_model(12, 1, SEED + 4)
    solve_batched(m, sp)  # warm-up (code paths, allocator)
    t = time.perf_counter()
    rb = solve_batched(m, sp)
    tb = time.perf_counter() - t
    progress("AIM 1+11 batched solve %.2f s" % tb)
    t = time.perf_counter()
    rp = solve_plain(m, sp)
    tp = time.perf_counter() - t
    progress("AIM 1+11 plain solve %.2f s" % tp)
    out["batched"] = {
        "workload": "AIM 1 orbital + 11 bath sites, U=4, all 169 (N_up, N_dn) sectors, NEV=4, NCV=11, tol 1e-14; "
                    "G of the impurity orbital from 4 start vectors per pair (spin x particle/hole), beta=10, "
                    "256 Matsubara points, NLANC=60",
        "plain_solves_per_s": round(1.0 / tp, 4), "batched_solves_per_s": round(1.0 / tb, 4),
        "speedup": round(tp / tb, 3), "plain_s": round(tp, 3), "batched_s": round(tb, 3),
        "max_abs_dG": float(np.abs(rb["G"] - rp["G"]).max()),
        "max_rel_dE": float(np.max(np.abs(rb["energies"] - rp["energies"]) /
                                   np.maximum(np.abs(rp["energies"]), 1.0))),
        "n_gf": rb["n_gf"], "hv_by_sector": rb["hv_by_sector"]}
    out["batched"]["grouped_hv"] = grouped_hv_roofline(m, 12)
    progress("grouped H*v of all 169 sectors: %.1f us" % (1e3 * out["batched"]["grouped_hv"]["kernel_ms"]))
    # served: Kanamori 3 orbitals + 6..11 bath sites (9..14 sites), Poisson arrivals
    sp4 = SolveParams(nev=4, beta=10.0, nomega=256, nlanc=60, orbitals=(0, 1, 2), seed=SEED)
    rng = np.random.default_rng(SEED + 7)
    baths = [int(b) for b in rng.integers(6, 12, args.requests)]
    if 11 not in baths:  # the 14-site case is always part of the stream
        baths[int(np.argmax(baths))] = 11
    reqs = [(r, siam_model(3 + b, 3, SEED + 1000 + r), sp4) for r, b in enumerate(baths)]
    srv = Server(0, sweeps_per_tick=8, max_active=args.max_active)
    done, wall = srv.run_stream(reqs, rate=args.rate, seed=SEED + 7)
    progress("served stream of %d requests (bath sizes %s): %.2f s" % (len(reqs), baths, wall))
    blocked = dict(srv.blocked_by)
    budget, peak = srv.budget, srv.peak_reserved
    srv.close()
    st = latency_stats(done)
    # the plain per-sector solve on the same mix, one request after another
    t_plain = []
    dg = de = 0.0
    for r, m, p in reqs:
        t = time.perf_counter()
        rp = solve_plain(m, p)
        t_plain.append(time.perf_counter() - t)
        progress("plain solve of request %d (%d sites): %.2f s" % (r, m.Ns, t_plain[-1]))
        if "error" in done[r]:
            continue
        dg = max(dg, float(np.abs(done[r]["G"] - rp["G"]).max()))
        de = max(de, float(np.max(np.abs(done[r]["energies"] - rp["energies"]) /
                                  np.maximum(np.abs(rp["energies"]), 1.0))))
    tps = float(np.sum(t_plain))
    errors = [r for r, d in done.items() if "error" in d]
    out["served"] = {
        "workload": "AIM 3 orbitals (Kanamori U=4, J=0.7, U'=2.6, half filling) + 6..11 bath sites (9..14 sites, "
                    "sector dims up to 11,778,624), %d requests with bath sizes %s, Poisson arrivals at %.2f/s, "
                    "NEV=4, G of the 3 impurity orbitals, beta=10, 256 Matsubara points, NLANC=60; admitted by "
                    "HBM footprint (budget %.1f GB = free device memory less 8%%)" %
                    (len(reqs), baths, args.rate, budget / 1e9),
        "solves_per_s": round(len(done) / wall, 4), "p50_s": round(st["p50_s"], 3), "p99_s": round(st["p99_s"], 3),
        "plain_solves_per_s": round(len(reqs) / tps, 4), "plain_s_per_request": [round(x, 2) for x in t_plain],
        "speedup": round(tps / wall, 3), "wall_s": round(wall, 3),
        "plain_sample": "every request of the stream run alone by the plain per-sector solve, one after another",
        "max_abs_dG_vs_plain": dg, "max_rel_dE_vs_plain": de, "errors": errors,
        "admission": {"blocked_by_memory": blocked.get("memory", 0), "blocked_by_slots": blocked.get("slots", 0),
                      "peak_reserved_gb": round(peak / 1e9, 1)}}
    return out


def large_sector_section(model, v0buf_ctx, vb, args):
    """The 16-site sector (8,8) end to end on the device: 300 Lanczos steps
    from a random start (ground-state energy from the tridiagonal matrix),
    then the ground state from the batched eigensolver and 4 batched
    Green's-function runs (c+ and c of orbital 0, both spins) with their
    continued fractions, every vector resident in HBM."""
    from scipy.linalg import eigh_tridiagonal
    from edlib_amd.batch import Batch
    from edlib_amd.solver import matsubara
    ctx = v0buf_ctx
    nl = 300
    nlanc, a, b, ms = ctx.lanczos_time(vb, nl)
    e_lz = float(eigh_tridiagonal(a[:nlanc], b[1:nlanc])[0][0])
    out = {"workload": "16-site 4x4 Hubbard U=6 mu=3, sector (8,8), N=165636900: %d Lanczos steps; ground state "
                       "(NEV=1, NCV=20, tol 1e-10) and 4 Green's-function runs of NLANC=%d (c+, c on orbital 0, "
                       "both spins; sectors (9,8), (8,9), (7,8), (8,7)) through the batched engine, 256 Matsubara "
                       "points at beta=10" % (nl, args.gf_nlanc),
           "lanczos_steps": nlanc, "lanczos_s": round(ms * nlanc * 1e-3, 3), "lanczos_ms_per_step": round(ms, 3),
           "E0_lanczos": e_lz}
    bt = Batch(0)
    t0 = time.perf_counter()
    sid, _ = bt.add_sector(model, 8, 8)
    jid = bt.submit_eig(sid, 1, 20, 1e-10, 1000, SEED + 8)
    bt.run()
    ev, _ = bt.eig_result(jid)
    t1 = time.perf_counter()
    tg = {}
    for sec in ((9, 8), (8, 9), (7, 8), (8, 7)):
        tg[sec], _ = bt.add_sector(model, *sec)
    reqs = [(jid, 0, 0, 0, tg[(9, 8)]), (jid, 0, 16, 0, tg[(8, 9)]), (jid, 0, 0, 1, tg[(7, 8)]),
            (jid, 0, 16, 1, tg[(8, 7)])]
    ids = bt.submit_gf(reqs, args.gf_nlanc)
    bt.run()
    n2 = bt.lanczos_norm2(ids)
    wn = matsubara(10.0, 256)
    g = bt.cf_many(ids, list(n2), [1, 1, -1, -1], [1.0] * 4, [float(ev[0])] * 4, wn)
    t2 = time.perf_counter()
    info = bt.job_info(jid)
    gup, gdn = g[0] + g[2], g[1] + g[3]
    # sum rule: i w G(i w) -> 1 (the spectral weight of each spin)
    tail = max(abs(1j * wn[-1] * gup[-1] - 1.0), abs(1j * wn[-1] * gdn[-1] - 1.0))
    bt.close()
    out.update({"E0_eigensolver": float(ev[0]), "dE0_rel": abs(float(ev[0]) - e_lz) / abs(e_lz),
                "eigensolver_s": round(t1 - t0, 3), "eigensolver_hv": info["n_av"],
                "gf_s": round(t2 - t1, 3), "gf_weights": [round(float(x), 12) for x in n2],
                "sum_rule_dev_at_w_max": float(tail), "spin_symmetry_dG": float(np.abs(gup - gdn).max())})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solves", action="store_true")
    ap.add_argument("--requests", type=int, default=5)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--max-active", type=int, default=64)
    ap.add_argument("--no-large", action="store_true", help="skip the 16-site Lanczos/eigensolver/GF section")
    ap.add_argument("--gf-nlanc", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md on sharding)",
              file=sys.stderr)
        sys.exit(2)

    import edlib_amd as ea

    Ns = 16
    t = torus(4, 4)
    U = np.full(Ns, 6.0)
    xmu = np.full(Ns, 3.0)
    ctx = ea.Context(0)
    ctx.set_model(ea.HubbardModel(Ns, t, U, xmu))
    n = ctx.set_sector(8, 8)
    dup, ddn, tbytes = ctx.sector_info()
    B = 24 * n + tbytes
    rng = np.random.default_rng(SEED + 8)
    v = rng.uniform(-1.0, 1.0, n)
    v /= np.linalg.norm(v)
    w = np.random.default_rng(SEED + 9).uniform(-1.0, 1.0, n)
    vb, wb = ctx.buffer(n), ctx.buffer(n)
    vb.upload(v)
    wb.upload(w)

    for _ in range(args.warmup):
        ctx.av_device(vb, wb, clear=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.av_device(vb, wb, clear=False)
    ctx.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    kern_ms = ctx.av_time(vb, wb, clear=False, warmup=1, reps=max(args.steps, 5))
    chosen = ctx.hv_variant
    per_variant = {}
    for var, name in ((0, "single_pass_ms"), (1, "two_pass_ms")):
        ctx.hv_variant = var
        per_variant[name] = round(ctx.av_time(vb, wb, clear=False, warmup=1, reps=5), 4)
    ctx.hv_variant = chosen
    achieved = B / (kern_ms * 1e-3) / 1e9
    value = B / wall / 1e9
    progress("H*v %.3f ms per step (kernel %.3f ms)" % (wall * 1e3, kern_ms))
    # host vectors: staged over PCIe both ways (the reference's av() contract)
    import ctypes
    from edlib_amd import _lib
    wh = w.copy()
    t0 = time.perf_counter()
    _lib.check(_lib.lib().edh_av(ctx.handle, v.ctypes.data_as(ctypes.c_void_p), wh.ctypes.data_as(ctypes.c_void_p),
                                 n, 0, ea.EDH_MEM_HOST))
    pcie_gbs = B / (time.perf_counter() - t0) / 1e9
    del wh
    # static observables of the (normalised) vector on the device: one read
    # of v by rows and one by column tiles (StaticObservables.h:356-419)
    ctx.static_observables(vb, Ns)
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.static_observables(vb, Ns)
    obs_ms = (time.perf_counter() - t0) / 3 * 1e3

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform v, w; model parameters fixed)",
        "config": {"workload": "H*v, 16-site 4x4 periodic Hubbard t=1 U=6 mu=3, sector (8,8), N=165636900, "
                               "Lanczos form w=Hv+w", "sector_dim": n, "dim_up": dup, "dim_dn": ddn,
                   "algorithmic_bytes_per_step": B, "parallelism": "single GPU"},
        "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(chosen, kern_ms),
                     "kernel": {0: "edh::hv_rows_kernel (single pass)",
                                1: "edh::hv_rows_kernel (no up term) + edh::hv_upslab_kernel (two passes)"}[chosen],
                     "kernel_ms": round(kern_ms, 4), "variants": per_variant},
    }
    out["pcie_inclusive_gbs"] = round(pcie_gbs, 2)
    out["static_observables"] = {"ms": round(obs_ms, 3), "sites": Ns,
                                 "gbs": round(2 * 8 * n / (obs_ms * 1e-3) / 1e9, 1),
                                 "what": "N, N_up, N_dn, M, D_occ, M_i M_j, N_eff of one sector vector, host call "
                                         "including the small host reduction (2 reads of v)"}
    if not args.no_large:
        vb.upload(v)
        out["large_sector"] = large_sector_section(ea.HubbardModel(Ns, t, U, xmu), ctx, vb, args)
        progress("16-site sector: Lanczos, eigensolver and Green's functions done")
    ctx.close()
    if not args.no_solves:
        out["solves"] = solves_section(args)
    if not args.no_cpu_baseline:
        wcpu = w.copy()
        out["cpu_baseline"] = cpu_baseline(t, U, xmu, v, wcpu, B)
        progress("CPU av of the 16-site sector: %.1f ms" % out["cpu_baseline"]["ms_per_call"])
        out["cpu_baseline"]["per_config"] = cpu_socrs_configs()
        hv = out.get("solves", {}).get("batched", {}).pop("hv_by_sector", None)
        if hv:
            est = cpu_solve_estimate(hv, 60)
            est["vs_batched_gpu"] = round(out["solves"]["batched"]["batched_solves_per_s"] /
                                          est["solves_per_s_upper_bound"], 1)
            out["cpu_baseline"]["aim12_solve"] = est
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
