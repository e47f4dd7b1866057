"""Pins of the oracle's interaction terms by mathematics alone.

The reference holds no fixture for the SIAM/Kanamori interaction
(SingleImpurityAndersonModel.h:129-168 V_states with the density-density skip
rule of :256-274 and the 1/2 U value of :166; the diagonal of :282-306) nor
for the Hubbard exchange term J (HubbardModel.h:124-129).  These tests pin
the oracle's reading of them against facts fixed independently of that
reading:

(i)   the atomic-limit multiplets of the three-orbital Kanamori atom with
      U' = U - 2J (the t2g shell): E(N, S, L) = (U - 3J) N(N-1)/2
      - 2J S(S+1) - (J/2) L(L+1) + (5/2) J N, with their degeneracies;
(ii)  spin-rotation invariance: [H, S+] = [H, S^2] = 0 on the full Fock space
      of a three-orbital impurity with a spin-symmetric bath;
(iii) the textbook Hamiltonians written with Jordan-Wigner matrices of
      c/c^dagger built here from the sign rule (FermionicModel.h:35-68):
      Kanamori U n n + U' n n + (U' - J) n n - J spin flip + J pair hopping
      plus the hybridization, and the Hubbard model with the exchange
      J_ij m_i m_j (m = n_up - n_dn, i.e. 4 J_ij S^z_i S^z_j, i = j included);
(iv)  planted mistakes in the oracle's input tensor (a sign-flipped spin-flip
      pair, a sign-flipped pair-hopping pair, the off-diagonal terms scaled
      by 2 instead of 1/2) are caught by (i)-(iii).
"""
import numpy as np
import pytest

from oracle import edoracle as eo
from tests.models import kanamori_tensor


# ---------------------------------------------------------------------------
# Jordan-Wigner matrices in the reference's bit convention: orbital
# o = site + spin * Ns sits at bit 2Ns-1-o (spin up is the high half) and
# c^dagger_o picks up (-1)^(number of set bits above bit(o)).
# ---------------------------------------------------------------------------
def cdag(Ns, o):
    dim = 1 << (2 * Ns)
    pos = 2 * Ns - 1 - o
    C = np.zeros((dim, dim))
    for s in range(dim):
        if not (s >> pos) & 1:
            C[s | (1 << pos), s] = -1.0 if bin(s >> (pos + 1)).count("1") % 2 else 1.0
    return C


def ops(Ns):
    cd = [cdag(Ns, o) for o in range(2 * Ns)]
    c = [m.T.copy() for m in cd]
    return cd, c


def operator_matrix(model):
    """The oracle's Fock-space H as an operator matrix M[target, source]."""
    return eo.fock_hamiltonian(model).T


def spin_ops(Ns, cd, c):
    up = lambda i: i
    dn = lambda i: i + Ns
    Sp = sum(cd[up(i)] @ c[dn(i)] for i in range(Ns))
    Sm = Sp.T
    Sz = 0.5 * sum(cd[up(i)] @ c[up(i)] - cd[dn(i)] @ c[dn(i)] for i in range(Ns))
    S2 = 0.5 * (Sp @ Sm + Sm @ Sp) + Sz @ Sz
    return Sp, Sz, S2


def kanamori_atom(U, J, Ut=None):
    ml = 3
    Ut = kanamori_tensor(ml, U, J) if Ut is None else Ut
    return eo.siam(ml, ml, np.zeros((0, 2)), np.zeros((ml, 0, 2)), np.zeros((ml, ml, 2)), Ut, 0.0)


def multiplets(U, J):
    """E(N, S, L) of the t2g Kanamori atom and (2S+1)(2L+1) degeneracies."""
    table = {0: [(0, 0)], 1: [(0.5, 1)], 2: [(1, 1), (0, 2), (0, 0)], 3: [(1.5, 0), (0.5, 2), (0.5, 1)],
             4: [(1, 1), (0, 2), (0, 0)], 5: [(0.5, 1)], 6: [(0, 0)]}
    out = {}
    for N, lst in table.items():
        levels = []
        for S, L in lst:
            E = (U - 3 * J) * N * (N - 1) / 2 - 2 * J * S * (S + 1) - 0.5 * J * L * (L + 1) + 2.5 * J * N
            levels += [E] * int(round((2 * S + 1) * (2 * L + 1)))
        out[N] = np.sort(levels)
    return out


def spectrum_by_particle_number(M, Ns):
    Ntot = np.array([bin(s).count("1") for s in range(M.shape[0])])
    out = {}
    for N in range(2 * Ns + 1):
        idx = np.where(Ntot == N)[0]
        out[N] = np.sort(np.linalg.eigvalsh(0.5 * (M[np.ix_(idx, idx)] + M[np.ix_(idx, idx)].T)))
    return out


def multiplets_match(M, U, J):
    spec = spectrum_by_particle_number(M, 3)
    ref = multiplets(U, J)
    return all(np.allclose(spec[N], ref[N], atol=1e-10) for N in ref)


# ---------------------------------------------------------------------------
# (i) atomic-limit multiplets
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("U,J", [(4.0, 0.7), (2.0, 0.3), (5.0, 1.1)])
def test_kanamori_atom_multiplets(U, J):
    M = operator_matrix(kanamori_atom(U, J))
    spec = spectrum_by_particle_number(M, 3)
    ref = multiplets(U, J)
    # N = 2: U-3J (x9), U-J (x5), U+2J (x1); N = 3: 3U-9J (x4), 3U-6J (x10), 3U-4J (x6)
    assert np.allclose(ref[2], np.sort([U - 3 * J] * 9 + [U - J] * 5 + [U + 2 * J]))
    assert np.allclose(ref[3], np.sort([3 * U - 9 * J] * 4 + [3 * U - 6 * J] * 10 + [3 * U - 4 * J] * 6))
    for N in range(7):
        assert np.allclose(spec[N], ref[N], atol=1e-10), (N, spec[N], ref[N])


def test_kanamori_atom_spin_sector_spectra_agree():
    # SU(2): the spectrum of sector (nup, ndn) is contained in that of
    # (nup-1, ndn+1) multiplet by multiplet, when nup > ndn
    m = kanamori_atom(4.0, 0.7)
    H = eo.fock_hamiltonian(m)
    Ns = 3

    def sector_spec(nu, nd):
        idx = [(int(a) << Ns) | int(b) for a in eo.basis(Ns, nu) for b in eo.basis(Ns, nd)]
        return np.sort(np.linalg.eigvalsh(H[np.ix_(idx, idx)]))

    for N in range(1, 6):
        for nu in range(N + 1):
            nd = N - nu
            if nu <= nd or nu > Ns or nd < 0 or nd + 1 > Ns:
                continue
            hi, lo = sector_spec(nu, nd), sector_spec(nu - 1, nd + 1)
            for e in hi:  # every level of the larger |Sz| appears at the smaller one
                assert np.min(np.abs(lo - e)) < 1e-10


# ---------------------------------------------------------------------------
# (ii) spin-rotation invariance with a bath
# ---------------------------------------------------------------------------
def kanamori_with_bath(Ut, seed=5, nb=1):
    ml = 3
    rng = np.random.default_rng(seed)
    eps = np.repeat(rng.uniform(-2, 2, nb)[:, None], 2, axis=1)
    Vk = np.zeros((ml, nb, 2))
    for im in range(ml):
        Vk[im] = rng.uniform(0.1, 0.8, nb)[:, None]
    H0 = np.zeros((ml, ml, 2))
    for im in range(ml):
        H0[im, im, :] = rng.uniform(-0.3, 0.3)
    H0[1, 0, :] = H0[0, 1, :] = 0.25
    return eo.siam(ml + nb, ml, eps, Vk, H0, Ut, 6.5), dict(eps=eps, Vk=Vk, H0=H0, mu=6.5)


def su2_commutators(M, Ns):
    cd, c = ops(Ns)
    Sp, Sz, S2 = spin_ops(Ns, cd, c)
    return (np.max(np.abs(M @ Sp - Sp @ M)), np.max(np.abs(M @ S2 - S2 @ M)), np.max(np.abs(M @ Sz - Sz @ M)))


def test_kanamori_bath_su2_invariance():
    m, _ = kanamori_with_bath(kanamori_tensor(3, 4.0, 0.7))
    M = operator_matrix(m)
    assert np.max(np.abs(M - M.T)) < 1e-13
    cp, c2, cz = su2_commutators(M, 4)
    assert cp < 1e-12 and c2 < 1e-12 and cz < 1e-12, (cp, c2, cz)


# ---------------------------------------------------------------------------
# (iii) textbook Hamiltonians from Jordan-Wigner matrices
# ---------------------------------------------------------------------------
def textbook_kanamori(Ns, ml, U, J, eps, Vk, H0, mu):
    cd, c = ops(Ns)
    n = [cd[o] @ c[o] for o in range(2 * Ns)]
    up = lambda i: i
    dn = lambda i: i + Ns
    Up = U - 2 * J
    dim = 1 << (2 * Ns)
    H = np.zeros((dim, dim))
    for i in range(ml):
        H += U * n[up(i)] @ n[dn(i)]
        for j in range(ml):
            if i == j:
                continue
            H += Up * n[up(i)] @ n[dn(j)]
            if i < j:
                for s in (up, dn):
                    H += (Up - J) * n[s(i)] @ n[s(j)]
            H += -J * cd[up(i)] @ c[dn(i)] @ cd[dn(j)] @ c[up(j)]
            H += J * cd[up(i)] @ cd[dn(i)] @ c[dn(j)] @ c[up(j)]
    for sp, s in ((0, up), (1, dn)):
        for i in range(ml):
            H += (H0[i, i, sp] - mu) * n[s(i)]
            for j in range(ml):
                if i != j and H0[i, j, sp] != 0.0:
                    H += H0[i, j, sp] * cd[s(i)] @ c[s(j)]
        for k in range(Ns - ml):
            H += eps[k, sp] * n[s(ml + k)]
            for i in range(ml):
                v = Vk[i, k, sp]
                H += v * (cd[s(i)] @ c[s(ml + k)] + cd[s(ml + k)] @ c[s(i)])
    return H


def test_kanamori_matches_textbook_hamiltonian():
    m, p = kanamori_with_bath(kanamori_tensor(3, 4.0, 0.7))
    ref = textbook_kanamori(4, 3, 4.0, 0.7, p["eps"], p["Vk"], p["H0"], p["mu"])
    assert np.max(np.abs(operator_matrix(m) - ref)) < 1e-13


def textbook_hubbard(Ns, t, U, xmu, Hmag, J):
    cd, c = ops(Ns)
    n = [cd[o] @ c[o] for o in range(2 * Ns)]
    dim = 1 << (2 * Ns)
    H = np.zeros((dim, dim))
    for s in range(2):
        for i in range(Ns):
            for j in range(Ns):
                if t[i, j] != 0.0:
                    H += -t[i, j] * cd[j + s * Ns] @ c[i + s * Ns]
    for i in range(Ns):
        # chemical potential indexed by the spin (HubbardModel.h:120 quirk)
        H += -xmu[0] * n[i] - xmu[1] * n[i + Ns]
        H += U[i] * n[i] @ n[i + Ns]
        H += Hmag[i] * (n[i + Ns] - n[i])
    m_op = [n[i] - n[i + Ns] for i in range(Ns)]  # 2 S^z_i
    for i in range(Ns):
        for j in range(Ns):
            H += J[i, j] * m_op[i] @ m_op[j]  # 4 J_ij S^z_i S^z_j
    return H


def _hubbard_params(Ns, seed, t_scale=1.0):
    rng = np.random.default_rng(seed)
    t = rng.normal(size=(Ns, Ns)) * t_scale
    t = 0.5 * (t + t.T)
    np.fill_diagonal(t, 0.0)
    return dict(t=t, U=rng.uniform(1, 5, Ns), xmu=rng.uniform(0, 2, Ns), Hmag=rng.uniform(-0.3, 0.3, Ns),
                J=0.3 * rng.normal(size=(Ns, Ns)))


def test_hubbard_exchange_spin_picture_at_t0():
    Ns = 4
    p = _hubbard_params(Ns, 31, t_scale=0.0)
    m = eo.hubbard(Ns, p["t"], p["U"], p["xmu"], p["Hmag"], p["J"])
    Ip = 2 * Ns
    for s in range(1 << Ip):
        occ = lambda o: (s >> (Ip - 1 - o)) & 1
        Sz = [0.5 * (occ(i) - occ(i + Ns)) for i in range(Ns)]
        e = sum(-p["xmu"][0] * occ(i) - p["xmu"][1] * occ(i + Ns) + p["U"][i] * occ(i) * occ(i + Ns)
                + p["Hmag"][i] * (occ(i + Ns) - occ(i)) for i in range(Ns))
        e += sum(4.0 * p["J"][i, j] * Sz[i] * Sz[j] for i in range(Ns) for j in range(Ns))
        assert abs(m.diagonal(s) - e) < 1e-13, s


def test_hubbard_matches_textbook_hamiltonian():
    Ns = 4
    p = _hubbard_params(Ns, 37)
    m = eo.hubbard(Ns, p["t"], p["U"], p["xmu"], p["Hmag"], p["J"])
    ref = textbook_hubbard(Ns, p["t"], p["U"], p["xmu"], p["Hmag"], p["J"])
    assert np.max(np.abs(operator_matrix(m) - ref)) < 1e-13


# ---------------------------------------------------------------------------
# (iv) planted mistakes in the oracle's input are caught
# ---------------------------------------------------------------------------
def _flip_spin_flip(Ut, i, j):
    Ut = Ut.copy()
    for a, b in ((0, 1), (1, 0)):
        Ut[a, b, i, j, j, i] *= -1.0  # U_{ijji} opposite spins: only the spin-flip V terms read it
        Ut[a, b, j, i, i, j] *= -1.0
    return Ut


def _flip_pair_hop(Ut, i, j):
    Ut = Ut.copy()
    Ut[:, :, i, i, j, j] *= -1.0
    Ut[:, :, j, j, i, i] *= -1.0
    return Ut


def _double_offdiagonal(Ut):
    # the off-diagonal V terms at U instead of U/2: scale the tensor entries
    # that only V_states read (spin flip and pair hopping)
    Ut = Ut.copy()
    ml = Ut.shape[2]
    for i in range(ml):
        for j in range(ml):
            if i != j:
                Ut[:, :, i, i, j, j] *= 2.0
                for a, b in ((0, 1), (1, 0)):
                    Ut[a, b, i, j, j, i] *= 2.0
    return Ut


@pytest.mark.parametrize("plant", ["spin_flip", "pair_hop", "half_factor"])
def test_planted_interaction_mistakes_are_caught(plant):
    U, J = 4.0, 0.7
    base = kanamori_tensor(3, U, J)
    bad = {"spin_flip": _flip_spin_flip(base, 0, 2), "pair_hop": _flip_pair_hop(base, 1, 2),
           "half_factor": _double_offdiagonal(base)}[plant]
    # (i) atomic multiplets
    assert multiplets_match(operator_matrix(kanamori_atom(U, J)), U, J)
    assert not multiplets_match(operator_matrix(kanamori_atom(U, J, bad)), U, J)
    # (ii) spin rotation: a spin-flip sign or the factor breaks it; pair
    # hopping is a spin singlet, so its sign is caught by (i) and (iii) only
    m, p = kanamori_with_bath(bad)
    M = operator_matrix(m)
    cp, c2, _ = su2_commutators(M, 4)
    if plant in ("spin_flip", "half_factor"):
        assert cp > 1e-3 and c2 > 1e-3
    # (iii) textbook Hamiltonian
    ref = textbook_kanamori(4, 3, U, J, p["eps"], p["Vk"], p["H0"], p["mu"])
    assert np.max(np.abs(M - ref)) > 1e-3
