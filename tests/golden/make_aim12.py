"""Oracle vectors for the batched impurity solve: single-orbital Anderson
impurity with 11 bath sites (bath energies uniform [-2, 2], hybridisations
uniform [0.1, 0.8] from seed 20261202 + 4, U = 4, mu = 2), every one of the
169 (N_up, N_dn) sectors, NEV = 4 (NCV = 11) per sector by scipy's ARPACK on
the oracle's SpinResolvedStorage::av restatement (Storage.h:43-146), the
ordered eigenpairs (Hamiltonian.h:54-101), and G of the impurity orbital on
256 Matsubara points at beta = 10 from 60 Lanczos steps per start vector
(the oracle's GreensFunction restatement, GreensFunction.h:109-159).
Written once on the CPU (several minutes); writes aim12_solve.json.
Usage: python tests/golden/make_aim12.py
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
    p = siam_params(12, 1, np.random.default_rng(SEED + 4))
    om = oracle_model("siam", p)
    params = {"NSITES": 12, "arpack.NEV": 4, "arpack.NCV": 11, "lanc.BETA": 10.0, "lanc.NOMEGA": 256,
              "lanc.NLANC": 60, "lanc.BOLTZMANN_CUTOFF": 1e-12}
    pairs = eo.hamiltonian_diag(om, params)
    G = eo.greens_function(om, pairs, params, gf_orbs=[0])["G"]
    out = {"energies": [q.eigenvalue for q in pairs], "sectors": [list(q.sector) for q in pairs],
           "G_re": [[float(x) for x in G[:, 0, s].real] for s in range(2)],
           "G_im": [[float(x) for x in G[:, 0, s].imag] for s in range(2)]}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "aim12_solve.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
