"""Oracle pinning against the reference's own golden vectors and brute force.

Follows test/Lanczos_Test.cpp:29-121 (G, chi_Sz, chi_N vs gom1/xiats/xiatd),
test/HubbardModel_Test.cpp:27-71 (4-ring E0) and
test/StaticObservables_Test.cpp:27-72 (paramagnetic observables).
"""
import numpy as np
import pytest

from oracle import edoracle as eo
from tests import models
from tests.fixtures import hubbard_arrays, load_dat, load_input

GFCHI_PARAMS = {"NSITES": 4, "arpack.NEV": 100, "lanc.BETA": 20.0, "lanc.NOMEGA": 1000, "lanc.NLANC": 100,
                "lanc.BOLTZMANN_CUTOFF": 1e-12}


@pytest.fixture(scope="module")
def gfchi():
    d = load_input("GF_Chi_input.json")
    Ns, t, U, xmu, H, J = hubbard_arrays(d)
    m = eo.hubbard(Ns, t, U, xmu, H, J)
    pairs = eo.hamiltonian_diag(m, GFCHI_PARAMS)
    return m, pairs


def test_gf_golden(gfchi):
    m, pairs = gfchi
    res = eo.greens_function(m, pairs, GFCHI_PARAMS, gf_orbs=[0])
    ref = load_dat("gom1.dat")
    G = res["G"]
    assert np.allclose(eo.matsubara_points(20.0, 200), ref[:, 0], rtol=0, atol=1e-15)
    for s in range(2):
        diff = np.max(np.abs(G[:200, 0, s] - (ref[:, 1] + 1j * ref[:, 2])))
        assert diff < 1e-10, diff  # Lanczos_Test.cpp:115


def test_chi_golden(gfchi):
    m, pairs = gfchi
    obs = eo.static_observables(m, pairs, GFCHI_PARAMS)
    avg_m = np.sum(obs["M"]) / (2.0 * len(obs["M"]))
    chi = eo.chi_loc(m, pairs, GFCHI_PARAMS, "Sz", avg=avg_m, orbs=[0])
    ref = load_dat("xiats.dat")
    diff = np.max(np.abs(chi["G"][:200, 0] - (-0.25 * ref[:200, 1])))
    assert diff < 1e-9, diff  # Lanczos_Test.cpp:117
    avg_n = np.sum(obs["N"]) / len(obs["N"])
    chi = eo.chi_loc(m, pairs, GFCHI_PARAMS, "N", avg=avg_n, orbs=[0])
    ref = load_dat("xiatd.dat")
    diff = np.max(np.abs(chi["G"][:200, 0] - (-ref[:200, 1])))
    assert diff < 1e-9, diff  # Lanczos_Test.cpp:119


def test_gf_matches_exact_lehmann(gfchi):
    # independent of storages and Lanczos: dense Fock-space Lehmann sum
    m, pairs = gfchi
    wn = eo.matsubara_points(20.0, 40)
    exact = eo.lehmann_gf(m, 20.0, wn, 0, 0)
    res = eo.greens_function(m, pairs, dict(GFCHI_PARAMS, **{"lanc.NOMEGA": 40}), gf_orbs=[0])
    assert np.max(np.abs(res["G"][:, 0, 0] - exact)) < 1e-10


def test_4ring_ground_state():
    d = load_input("4ring_input.json")
    Ns, t, U, xmu, H, J = hubbard_arrays(d)
    m = eo.hubbard(Ns, t, U, xmu, H, J)
    params = {"arpack.NEV": 1, "lanc.BETA": 10.0}
    pairs = eo.hamiltonian_diag(m, params)
    assert abs(pairs[0].eigenvalue - (-11.8443)) < 1e-4  # HubbardModel_Test.cpp:54
    assert pairs[0].sector == (2, 2)
    obs = eo.static_observables(m, pairs, {"lanc.BETA": 10.0})
    for o in range(Ns):
        assert abs(obs["N"][o] - 1.0) < 1e-8
        assert obs["N_up"][o] > obs["N_dn"][o]
        assert obs["M"][o] > 0.0


def test_4ring_paramagnetic_observables():
    d = load_input("4ring_nomagfield_input.json")
    Ns, t, U, xmu, H, J = hubbard_arrays(d)
    m = eo.hubbard(Ns, t, U, xmu, H, J)
    pairs = eo.hamiltonian_diag(m, {"arpack.NEV": 1})
    obs = eo.static_observables(m, pairs, {"lanc.BETA": 10.0})
    assert len(obs["M_i M_j"]) == 16 and len(obs["N_eff"]) == 1
    for o in range(Ns):
        assert abs(obs["N"][o] - 1.0) < 1e-8
        assert abs(obs["N_up"][o] - 0.5) < 1e-8
        assert abs(obs["N_dn"][o] - 0.5) < 1e-8
        assert abs(obs["M"][o]) < 1e-7
        assert abs(obs["D_occ"][o] - obs["D_occ"][0]) < 1e-8
    assert obs["N_eff"][0] > 0.0


def _random_hubbard(Ns, rng):
    t = rng.normal(size=(Ns, Ns))
    t = 0.5 * (t + t.T)
    t[rng.random((Ns, Ns)) < 0.3] = 0.0
    t = 0.5 * (t + t.T)
    np.fill_diagonal(t, 0.0)
    return eo.hubbard(Ns, t, rng.uniform(1, 5, Ns), rng.uniform(0, 2, Ns), rng.uniform(-0.3, 0.3, Ns),
                      0.2 * rng.normal(size=(Ns, Ns)))


def _kanamori_tensor(ml, U, J):
    # scripts/generate_anderson_model_input.py:7-17 (all four spin blocks equal)
    Ut = np.zeros((ml, ml, ml, ml))
    for i in range(ml):
        Ut[i, i, i, i] = U
        for j in range(ml):
            if i != j:
                Ut[i, j, i, j] = U - 2 * J
                Ut[i, j, j, i] = J
                Ut[i, i, j, j] = J
    return np.broadcast_to(Ut, (2, 2) + Ut.shape).copy()


def _random_siam(Ns, ml, rng):
    nb = Ns - ml
    eps = np.repeat(rng.uniform(-2, 2, nb)[:, None], 2, axis=1)
    Vk = np.zeros((ml, nb, 2))
    for im in range(ml):
        Vk[im, :, :] = rng.uniform(0.1, 0.8, nb)[:, None]
    H0 = np.zeros((ml, ml, 2))
    for im in range(ml):
        H0[im, im, :] = rng.uniform(-0.5, 0.5)
    if ml > 1:
        H0[1, 0, :] = 0.2
        H0[0, 1, :] = 0.2
    return eo.siam(Ns, ml, eps, Vk, H0, _kanamori_tensor(ml, 4.0, 0.7), 6.5)


@pytest.mark.parametrize("kind", ["hubbard", "siam"])
def test_srs_av_matches_fock_space_brute_force(kind):
    rng = np.random.default_rng(11)
    m = _random_hubbard(4, rng) if kind == "hubbard" else _random_siam(4, 2, rng)
    H = eo.fock_hamiltonian(m)
    Ns = m.Ns
    for nu in range(Ns + 1):
        for nd in range(Ns + 1):
            srs = eo.SRS(m, nu, nd)
            idx = [(int(a) << Ns) | int(b) for a in eo.basis(Ns, nu) for b in eo.basis(Ns, nd)]
            Hs = H[np.ix_(idx, idx)]
            v = rng.normal(size=srs.n)
            w0 = rng.normal(size=srs.n)
            assert np.allclose(srs.av(v), Hs @ v, atol=1e-12)
            assert np.allclose(srs.av(v, w0, clear=False), Hs @ v + w0, atol=1e-12)
            soc = eo.SOCRS(m, nu, nd)
            assert np.allclose(soc.av(v), Hs @ v, atol=1e-12)


def test_siam_kanamori_term_count():
    # 3-orbital Kanamori: 36 off-diagonal Coulomb transitions (SingleImpurityAndersonModel.h:256-274)
    rng = np.random.default_rng(3)
    m = _random_siam(6, 3, rng)
    ints, vals = m.transitions(1)
    assert len(vals) == 36


def _literal_observables(m, sector, v):
    """StaticObservables.h:369-388 over full state bits (checkState at bit
    Ip - 1 - o, FermionicModel.h:35-40), independent of the occupation tables."""
    Ns, norb, Ip = m.Ns, m.interacting_orbitals(), 2 * m.Ns
    states = [(int(a) << Ns) | int(b) for a in eo.basis(Ns, sector[0]) for b in eo.basis(Ns, sector[1])]
    cs = lambda st, o: (st >> (Ip - 1 - o)) & 1
    mm = np.zeros((norb, norb))
    dd = np.zeros(norb)
    for st, x in zip(states, v):
        w = x * x
        for o in range(norb):
            up, dn = cs(st, o), cs(st, o + Ns)
            dd[o] += up * dn * w
            for o2 in range(norb):
                mm[o, o2] += (up - dn) * (cs(st, o2) - cs(st, o2 + norb)) * w
    return mm.reshape(-1), dd


@pytest.mark.parametrize("kind", ["hubbard", "siam"])
def test_static_observables_follow_reference_orbital_rule(kind):
    rng = np.random.default_rng(29)
    m = _random_hubbard(5, rng) if kind == "hubbard" else _random_siam(5, 2, rng)
    sector = (2, 3)
    v = rng.normal(size=eo.cnk(5, 2) * eo.cnk(5, 3))
    v /= np.linalg.norm(v)
    obs = eo.static_observables(m, [eo.EigenPair(0.0, v, 0, sector)], {"lanc.BETA": 1.0})
    mm, dd = _literal_observables(m, sector, v)
    assert np.allclose(obs["M_i M_j"], mm, atol=1e-14)
    assert np.allclose(obs["D_occ"], dd, atol=1e-14)
    if kind == "siam":
        # the rule differs from <m_i m_j> when bath sites exist
        nu, nd = eo.occupations(m, sector, m.Ns)
        mi = (nu[:, :2] - nd[:, :2]).astype(float)
        assert not np.allclose(obs["M_i M_j"], np.einsum("s,si,sj->ij", v * v, mi, mi).reshape(-1))


def test_openmp_lanczos_matches_serial():
    """The oracle's OpenMP Lanczos (for sectors too large for the serial
    loop) runs the serial recurrence of Lanczos.h:36-77, which the gom1.dat
    golden pins: coefficients equal to rounding on a random Hubbard model."""
    p = models.random_hubbard_params(8, np.random.default_rng(1))
    srs = eo.SRS(models.oracle_model("hubbard", p), 4, 4)
    v = np.random.default_rng(2).uniform(-1, 1, srs.n)
    v /= np.linalg.norm(v)
    n1, a1, b1 = srs.lanczos(v, 30)
    n2, a2, b2 = srs.lanczos(v, 30, omp=True)
    assert n1 == n2 == 30
    assert np.abs(a1 - a2).max() < 1e-12 and np.abs(b1 - b2).max() < 1e-12
