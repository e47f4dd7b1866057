"""Oracle vectors for the 12-site plain solve (written once on the CPU; the
oracle needs about three minutes for it, too long for a test run).

Model: 3x4 periodic Hubbard cluster, t=1, t'=-0.3 (stored as in
test/input/GF_Chi/generate_input.py), U=8, mu=4; sector (6,6) of 853,776
states; 6 eigenpairs (scipy's ARPACK on the oracle's SpinResolvedStorage::av
restatement); G_00 of both spins on 512 Matsubara points at beta=20 from
150 Lanczos steps per start vector (the oracle's GreensFunction restatement,
GreensFunction.h:109-159) over the pairs with Boltzmann factor >= 1e-4: the
ground state and the first excited state.  The next four of the six pairs
are one member each of three two-fold multiplets (a single-vector Krylov
method finds one vector of each from its start vector, ARPACK included), so
their terms depend on the start vector and are left out.  Writes
plain12_solve.json.
Usage: python tests/golden/make_plain12.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import edoracle as eo  # noqa: E402
from tests import models  # noqa: E402


def main():
    p = dict(Ns=12, t=models.torus_hopping(3, 4, 1.0, -0.3), U=np.full(12, 8.0), xmu=np.full(12, 4.0))
    om = models.oracle_model("hubbard", p)
    params = {"NSITES": 12, "arpack.NEV": 6, "lanc.BETA": 20.0, "lanc.NOMEGA": 512, "lanc.NLANC": 150,
              "lanc.BOLTZMANN_CUTOFF": 1e-4}
    pairs = eo.hamiltonian_diag(om, params, sectors=[(6, 6)])
    G = eo.greens_function(om, pairs, params, gf_orbs=[0])["G"]
    out = {"energies": [q.eigenvalue for q in pairs],
           "G_re": [[float(x) for x in G[:, 0, s].real] for s in range(2)],
           "G_im": [[float(x) for x in G[:, 0, s].imag] for s in range(2)]}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "plain12_solve.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
