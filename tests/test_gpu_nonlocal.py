"""Non-local Green's function G_ij and the self-energy on the plain and the
batched solve (GreensFunction.h:202-223, 290-359; FermionicModel.h:84-108):

* G_ij of listed orbital pairs from the batched solve == the plain solve ==
  the oracle's GreensFunction restatement (1e-8 absolute), on both meshes
  for the batched/plain comparison;
* at U = 0 the Dyson equation closes: Sigma = 0 for the Hubbard model and
  Sigma = -mu for the Anderson model (the reference's bare G0 leaves mu
  out, SingleImpurityAndersonModel.h:378-391), to 1e-8;
* at U > 0 the self-energy tends to its Hartree value U_i <n_{i,-s}> at the
  highest Matsubara frequency (the 1/(i w) correction bounds the gap);
* requests with orbital pairs through the server give the batched solve's
  G_ij and Sigma bit for bit.
"""
import math

import numpy as np
import pytest

from edlib_amd.solver import SolveParams, solve_batched, solve_plain
from oracle import edoracle as eo
from tests import models

pytestmark = pytest.mark.gpu

SEED = 20261203


def test_nonlocal_gf_batched_plain_oracle():
    p = models.random_hubbard_params(5, np.random.default_rng(SEED), with_J=True)
    m = models.product_model("hubbard", p)
    pairs = ((0, 1), (1, 2), (2, 0), (3, 1))
    sp = SolveParams(nev=3, beta=10.0, nomega=32, nlanc=80, orbitals=(0, 1, 2), pairs=pairs, seed=SEED)
    rb = solve_batched(m, sp)
    rp = solve_plain(m, sp)
    assert rb["G"].shape == (32, 4, 2) and rb["G_ij"].shape == (32, 25, 2)  # orbital 3 joins the local list
    assert np.abs(rb["G_ij"] - rp["G_ij"]).max() < 1e-8
    assert np.abs(rb["G"] - rp["G"]).max() < 1e-8
    om = models.oracle_model("hubbard", p)
    prm = {"arpack.NEV": 3, "lanc.BETA": 10.0, "lanc.NOMEGA": 32, "lanc.NLANC": 80, "lanc.BOLTZMANN_CUTOFF": 1e-12}
    go = eo.greens_function(om, eo.hamiltonian_diag(om, prm), prm, gf_orbs=[0, 1, 2], gf_pairs=list(pairs))
    listed = [i * 5 + j for i, j in pairs] + [o * 5 + o for o in range(4)]
    assert np.abs(rb["G_ij"][:, listed, :] - go["G_ij"][:, listed, :]).max() < 1e-8
    assert np.abs(go["G_ij"][:, listed[:4], :]).max() > 1e-3  # the pairs' entries are not small
    # the real axis: batched == plain
    sr = SolveParams(nev=3, beta=10.0, nomega=24, nlanc=80, orbitals=(0, 1, 2), pairs=pairs, seed=SEED,
                     mesh="real", emin=-3.0, emax=3.0)
    assert np.abs(solve_batched(m, sr)["G_ij"] - solve_plain(m, sr)["G_ij"]).max() < 1e-8


@pytest.mark.parametrize("mesh", ["matsubara", "real"])
def test_selfenergy_vanishes_at_U0(mesh):
    rng = np.random.default_rng(SEED + 1)
    Ns = 4
    t = rng.normal(size=(Ns, Ns))
    t = 0.5 * (t + t.T)
    np.fill_diagonal(t, 0.0)
    lev = np.linalg.eigvalsh(-t)
    p = dict(Ns=Ns, t=t, U=np.zeros(Ns), xmu=np.full(Ns, 0.5 * (lev[2] + lev[3])))
    m = models.product_model("hubbard", p)
    allp = tuple((i, j) for i in range(Ns) for j in range(Ns) if i != j)
    sp = SolveParams(nev=2, beta=200.0, nomega=16, nlanc=40, orbitals=tuple(range(Ns)), pairs=allp,
                     selfenergy=True, seed=SEED, mesh=mesh, emin=-2.0, emax=2.0)
    for r in (solve_batched(m, sp), solve_plain(m, sp)):
        assert np.abs(r["Sigma"]).max() < 1e-8
        assert np.abs(r["G_ij"] - r["G0"]).max() < 1e-8


def test_anderson_selfenergy_at_U0_is_minus_mu():
    p = models.siam_params(6, 1, np.random.default_rng(SEED + 2), U=0.0, mu=0.45)
    p["U"] = np.zeros_like(p["U"])
    m = models.product_model("siam", p)
    sp = SolveParams(nev=2, beta=200.0, nomega=16, nlanc=40, orbitals=(0,), selfenergy=True, seed=SEED)
    r = solve_batched(m, sp)
    assert r["Sigma"].shape == (16, 1, 2)
    assert np.abs(r["Sigma"] + 0.45).max() < 1e-8


def test_selfenergy_hartree_tail():
    Ns = 4
    U = 4.0
    p = dict(Ns=Ns, t=models.ring_hopping(Ns), U=np.full(Ns, U), xmu=np.full(Ns, 1.7))
    m = models.product_model("hubbard", p)
    allp = tuple((i, j) for i in range(Ns) for j in range(Ns) if i != j)
    sp = SolveParams(nev=4, beta=10.0, nomega=512, nlanc=100, orbitals=tuple(range(Ns)), pairs=allp,
                     selfenergy=True, seed=SEED)
    r = solve_batched(m, sp)
    w_max = (2 * 511 + 1) * math.pi / 10.0
    n = {0: r["observables"]["N_up"], 1: r["observables"]["N_dn"]}
    for s in range(2):
        for i in range(Ns):
            hartree = U * n[1 - s][i]
            # Sigma(i w) = U <n> + O(U^2 / w)
            assert abs(r["Sigma"][-1, i * Ns + i, s] - hartree) < 2 * U * U / w_max, (s, i)


def test_served_requests_with_orbital_pairs():
    """Non-local requests through the server: two Anderson impurities with
    two orbitals (pairs (0,1), (1,0), self-energy on) served together; each
    result's G_ij and Sigma are bit-identical to the same request solved
    alone by the batched solve."""
    from edlib_amd.serve import Server
    ps = [models.siam_params(n, 2, np.random.default_rng(SEED + 10 + n), U=3.0, J=0.5) for n in (5, 6)]
    ms = [models.product_model("siam", p) for p in ps]
    sp = SolveParams(nev=3, beta=10.0, nomega=32, nlanc=60, orbitals=(0, 1), pairs=((0, 1), (1, 0)),
                     selfenergy=True, seed=SEED)
    srv = Server(0, sweeps_per_tick=4)
    for r, m in enumerate(ms):
        srv.submit(r, m, sp)
    while srv.tick():
        pass
    done = srv.done
    srv.close()
    assert sorted(done) == [0, 1] and not any("error" in d for d in done.values())
    for r, m in enumerate(ms):
        alone = solve_batched(m, sp)
        assert done[r]["G_ij"].shape == (32, 4, 2)
        assert np.array_equal(alone["G_ij"], done[r]["G_ij"])
        assert np.array_equal(alone["Sigma"], done[r]["Sigma"])
        assert np.abs(done[r]["G_ij"][:, 1, :]).max() > 1e-3
