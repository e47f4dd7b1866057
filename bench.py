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

METRIC = "H·v GB/s (% HBM peak); impurity solves/s + p99 latency, batched vs plain"
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def torus(L1, L2, t=1.0):
    Ns = L1 * L2
    T = np.zeros((Ns, Ns))
    for r in range(L1):
        for c in range(L2):
            i = r * L2 + c
            for j in (((r + 1) % L1) * L2 + c, r * L2 + (c + 1) % L2):
                T[i, j] = T[j, i] = t
    return T


def load_traffic(variant):
    """Memory traffic per launch of the given H*v variant, measured by the
    rocprofv3 FETCH_SIZE/WRITE_SIZE counters and kept under profiles/, or None."""
    p = os.path.join(ROOT, "profiles", "hv_16site_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("traffic_bytes_per_launch_by_variant", {}).get(str(variant))
    return None


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
            "ms_per_call": dt * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(chosen),
                     "kernel": ("edh::hv_rows_kernel (single pass)" if chosen == 0 else
                                "edh::hv_rows_kernel (no up term) + edh::hv_upslab_kernel (two passes)"),
                     "kernel_ms": round(kern_ms, 4), "variants": per_variant},
    }
    if not args.no_cpu_baseline:
        wcpu = w.copy()
        out["cpu_baseline"] = cpu_baseline(t, U, xmu, v, wcpu, B)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
