"""Pins of the oracle's non-local Green's function (GreensFunction.h:290-359:
start vectors (c+_i + c+_j)|psi>, (c_i + c_j)|psi>, then
G_ij = ((G_{i+j} - G_ii) - G_jj) / 2 per part) by closed forms.

At U = 0 the Green's function is the resolvent of the one-body matrix, so
the oracle's G_ij over every ordered pair must equal
* Hubbard: (z + mu + t)^-1 (the hopping values enter H as -t_ij c+_j c_i,
  the form HubbardModel::bare_greens_function inverts, HubbardModel.h:181-201);
* Anderson impurity, two orbitals with off-diagonal H0 and a shared bath:
  ((z + mu) - H0 - Delta(z))^-1 with Delta_ij = sum_k V_ik V_jk / (z - eps_k)
  (mu on the impurity only, SingleImpurityAndersonModel.h:282-306).
A dropped half, a wrong sign of the subtracted local parts or a pair vector
built from one orbital only fails these at O(1).  beta = 200 keeps every
excited pair's Boltzmann factor below the cutoff, so the thermal sum is the
ground state's alone (its sector's lowest pair is non-degenerate).
"""
import numpy as np

from oracle import edoracle as eo
from tests import models

BETA = 200.0
NW = 16
PRM = {"arpack.NEV": 2, "lanc.BETA": BETA, "lanc.NOMEGA": NW, "lanc.NLANC": 40, "lanc.BOLTZMANN_CUTOFF": 1e-12}


def _z():
    return 1j * eo.matsubara_points(BETA, NW)


def test_hubbard_nonlocal_gf_is_the_one_body_resolvent():
    rng = np.random.default_rng(11)
    Ns = 4
    t = rng.normal(size=(Ns, Ns))
    t = 0.5 * (t + t.T)
    np.fill_diagonal(t, 0.0)
    lev = np.linalg.eigvalsh(-t)
    mu = 0.5 * (lev[2] + lev[3])  # three one-body levels per spin below mu
    om = eo.hubbard(Ns, t, np.zeros(Ns), np.full(Ns, mu))
    pairs = eo.hamiltonian_diag(om, PRM)
    assert pairs[0].sector == (3, 3)
    allp = [(i, j) for i in range(Ns) for j in range(Ns) if i != j]
    go = eo.greens_function(om, pairs, PRM, gf_orbs=list(range(Ns)), gf_pairs=allp)
    G0 = np.array([np.linalg.inv((z + mu) * np.eye(Ns) + t) for z in _z()]).reshape(NW, Ns * Ns)
    for s in range(2):
        assert np.abs(go["G_ij"][:, :, s] - G0).max() < 1e-10
    # the local part alone: G_ii on the diagonal
    for i in range(Ns):
        assert np.abs(go["G"][:, i, 0] - G0[:, i * Ns + i]).max() < 1e-10


def test_anderson_two_orbital_nonlocal_gf_is_the_impurity_resolvent():
    p = models.siam_params(6, 2, np.random.default_rng(3), U=0.0, J=0.0, mu=0.37, h0_offdiag=0.15)
    p["U"] = np.zeros_like(p["U"])
    om = models.oracle_model("siam", p)
    pairs = eo.hamiltonian_diag(om, PRM)
    go = eo.greens_function(om, pairs, PRM, gf_orbs=[0, 1], gf_pairs=[(0, 1), (1, 0)])
    ml = 2
    for s in range(2):
        ref = []
        for z in _z():
            D = np.array([[np.sum(p["Vk"][i, :, s] * p["Vk"][j, :, s] / (z - p["Epsk"][:, s])) for j in range(ml)]
                          for i in range(ml)])
            ref.append(np.linalg.inv((z + p["mu"]) * np.eye(ml) - p["H0"][:, :, s] - D))
        ref = np.array(ref).reshape(NW, ml * ml)
        assert np.abs(go["G_ij"][:, :, s] - ref).max() < 1e-10
        # the off-diagonal entry is not small: the pin sees the pair's part
        assert np.abs(ref[:, 1]).max() > 1e-2
