"""Oracle vector for the served-stream test's largest request: three-orbital
Kanamori impurity (U = 4, J = 0.7, U' = U - 2J, mu = (5U - 10J)/2) with 11
bath sites (14 sites; bath energies uniform [-2, 2], hybridisations uniform
[0.1, 0.8], impurity levels uniform [-0.3, 0.3] from seed 20261202 + 2011).
The lowest eigenvalue of sector (8,9) (6,012,006 states), which holds the
ground state, by scipy's ARPACK (NEV = 1, NCV = 11, tol 1e-14) on the
oracle's SpinResolvedStorage::av restatement with its H_loc interaction
rows (SpinResolvedStorage.h:58-155).  Written once on the CPU; writes
kanamori14_e0.json.
Usage: python tests/golden/make_kanamori14.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import edoracle as eo  # noqa: E402
from tests.models import oracle_model, siam_params  # noqa: E402

SEED = 20261202


def main():
    eo.AV_OMP = True
    p = siam_params(14, 3, np.random.default_rng(SEED + 2011))
    om = oracle_model("siam", p)
    srs = eo.SRS(om, 8, 9)
    e, _ = eo.storage_diag(srs, 1, 11)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "kanamori14_e0.json"), "w") as f:
        json.dump({"sector": [8, 9], "E0": float(e[0])}, f)
    print(e[0])


if __name__ == "__main__":
    main()
