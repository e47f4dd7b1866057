"""Profiling workload for rocprofv3 (kernel trace and HBM counters).

Runs `--reps` Lanczos-form H*v launches on the 16-site 4x4 Hubbard sector
(8,8) (`--sites 12`: the 3x4 t,t' torus sector (6,6); `--sites 14`: the
three-orbital Kanamori impurity with 11 bath sites of the served stream's
14-site request, sector (7,7)) and one device-to-device copy of
`--copy-mib` MiB (the blit kernel whose FETCH_SIZE / WRITE_SIZE calibrate
the counters against a known byte count).
Usage: python -m edlib_amd.profile_hv [--reps N] [--copy-mib M] [--sites S]
"""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def torus(L1, L2, t=1.0):
    Ns = L1 * L2
    T = np.zeros((Ns, Ns))
    for r in range(L1):
        for c in range(L2):
            i = r * L2 + c
            for j in (((r + 1) % L1) * L2 + c, r * L2 + (c + 1) % L2):
                T[i, j] = T[j, i] = t
    return T


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--sites", type=int, default=16, choices=(12, 14, 16))
    ap.add_argument("--variant", type=int, default=-1, choices=(-1, 0, 1),
                    help="H*v variant: -1 = autotuned, 0 = single pass, 1 = two passes")
    a = ap.parse_args()
    import edlib_amd as ea
    from edlib_amd import _lib
    if a.sites == 16:
        Ns, t, U = 16, torus(4, 4), 6.0
    elif a.sites == 14:
        Ns, t, U = 14, None, 4.0
    else:
        Ns, t, U = 12, torus(3, 4), 8.0
        for r in range(3):
            for c in range(4):
                i = r * 4 + c
                for (dr, dc) in ((1, 1), (1, -1)):
                    j = ((r + dr) % 3) * 4 + (c + dc) % 4
                    t[i, j] = t[j, i] = -0.3
    ctx = ea.Context(0)
    if a.sites == 14:
        from edlib_amd.workloads import siam_model
        ctx.set_model(siam_model(14, 3, 20261202 + 1001))
    else:
        ctx.set_model(ea.HubbardModel(Ns, t, np.full(Ns, U), np.full(Ns, U / 2)))
    n = ctx.set_sector(Ns // 2, Ns // 2)
    if a.variant >= 0:
        ctx.hv_variant = a.variant
    rng = np.random.default_rng(20261202 + 8)
    vb, wb = ctx.buffer(n), ctx.buffer(n)
    vb.upload(rng.uniform(-1, 1, n))
    wb.upload(rng.uniform(-1, 1, n))
    for _ in range(a.reps):
        ctx.av_device(vb, wb, clear=False)
    ctx.synchronize()
    nbytes = a.copy_mib << 20
    src, dst = ctx.buffer(nbytes // 8), ctx.buffer(nbytes // 8)
    _lib.check(_lib.lib().edh_memcpy(ctx.handle, dst.ptr, src.ptr, nbytes, 2))
    ctx.synchronize()
    _, _, tb = ctx.sector_info()
    print("profile_hv: sites=%d n=%d variant=%d reps=%d algorithmic_bytes=%d copy_bytes=%d" %
          (Ns, n, ctx.hv_variant, a.reps, 24 * n + tb, nbytes), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
