"""Oracle pinning: basis enumeration, ranks, sector order, signs.

Mirrors test/SzSymmetry_Test.cpp:10-39 and test/NSymmetry_Test.cpp:9-21, and
adds closed-form checks that a dropped term or transposed index would fail.
"""
import itertools
from math import comb

import numpy as np
import pytest

from oracle import edoracle as eo


def lex_rank_closed_form(bits, n, k):
    # rank = C(n,k) - 1 - sum_i C(n-1-p_i, k-i), p_0 < p_1 < ... set positions
    pos = [p for p in range(n) if (bits >> p) & 1]
    return comb(n, k) - 1 - sum(comb(n - 1 - p, k - i) for i, p in enumerate(pos))


def test_cnk():
    assert eo.cnk(3, 2) == 3  # SzSymmetry_Test.cpp:16
    for n in range(0, 17):
        for k in range(0, n + 1):
            assert eo.cnk(n, k) == comb(n, k)


@pytest.mark.parametrize("n", [1, 2, 4, 8, 12])
def test_basis_is_lexicographic_and_ranked(n):
    for k in range(0, n + 1):
        b = eo.basis(n, k)
        # lexicographic order of the sorted position tuples (Combination.h:33-45)
        tuples = [tuple(p for p in range(n) if (x >> p) & 1) for x in b]
        assert tuples == sorted(itertools.combinations(range(n), k))
        for r, x in enumerate(b):
            assert lex_rank_closed_form(int(x), n, k) == r
        # NSymmetry::index round trip (NSymmetry_Test.cpp:9-21)
        if n <= 8:
            for r, x in enumerate(b):
                assert eo.nsym_index(n, k, x) == r


def test_sector_order_sizes_and_members():
    for Ns in (4, 8, 12, 16):
        order = eo.sector_order(Ns)
        assert sorted(order) == [(i, j) for i in range(Ns + 1) for j in range(Ns + 1)]
        sizes = [comb(Ns, a) * comb(Ns, b) for a, b in order]
        assert sizes == sorted(sizes)


def test_sector_order_known_prefix():
    # 4 sites: the size-1 sectors come first; within equal sizes libstdc++'s
    # introsort with the non-strict comparator decides (SzSymmetry.h:51-54,94).
    order = eo.sector_order(4)
    assert set(order[:4]) == {(0, 0), (0, 4), (4, 0), (4, 4)}
    assert order[-1] == (2, 2)


def test_eigenpair_order_is_stable_sort():
    rng = np.random.default_rng(5)
    for _ in range(50):
        e = np.round(rng.normal(size=30), 1)  # many ties
        assert eo.eigenpair_order(e) == list(np.argsort(e, kind="stable"))


def _fock_cdag(model, orb):
    """c^+_orb on the full Fock space (index = 2Ns-bit state), from a_adag blocks."""
    Ns = model.Ns
    dim = 1 << (2 * Ns)
    C = np.zeros((dim, dim))
    for nu in range(Ns + 1):
        for nd in range(Ns + 1):
            bu, bd = eo.basis(Ns, nu), eo.basis(Ns, nd)
            src = [(int(a) << Ns) | int(b) for a in bu for b in bd]
            for j, sj in enumerate(src):
                e = np.zeros(len(src))
                e[j] = 1.0
                nxt, out = eo.a_adag(model, (nu, nd), orb, False, e)
                if out is None:
                    continue
                tu, td = eo.basis(Ns, nxt[0]), eo.basis(Ns, nxt[1])
                tgt = [(int(a) << Ns) | int(b) for a in tu for b in td]
                for q in np.nonzero(out)[0]:
                    C[tgt[q], sj] = out[q]
    return C


def test_canonical_anticommutation():
    # {c_i, c^+_j} = delta_ij and {c^+_i, c^+_j} = 0 on the whole Fock space of
    # 3 sites: a wrong sign rule or orbital-to-bit mapping breaks one of them.
    Ns = 3
    m = eo.hubbard(Ns, np.zeros((Ns, Ns)), np.zeros(Ns), np.zeros(Ns))
    Cd = [_fock_cdag(m, o) for o in range(2 * Ns)]
    dim = 1 << (2 * Ns)
    for i in range(2 * Ns):
        for j in range(2 * Ns):
            ac = Cd[i].T @ Cd[j] + Cd[j] @ Cd[i].T
            assert np.allclose(ac, np.eye(dim) if i == j else 0.0)
            assert np.allclose(Cd[i] @ Cd[j] + Cd[j] @ Cd[i], 0.0)


@pytest.mark.parametrize("Ns,ml", [(7, 2), (9, 3), (12, 1), (14, 3)])
def test_impurity_moves_shift_the_rank_by_a_constant(Ns, ml):
    """The device applies an impurity interaction term (a, b) -> (a', b')
    (impurity orbitals 0..ml-1, the highest bits of each spin half; the
    bath bits unchanged, the occupation count of the impurity kept) as one
    element offset: in the reference basis order (SzSymmetry.h:191-202) the
    states sharing their bath bits form a contiguous run ordered by the
    impurity pattern alone, so rank(bath | a') - rank(bath | a) depends on
    (a, a') only.  Checked for every pattern pair of every particle number."""
    nb = Ns - ml
    low = (1 << nb) - 1
    for k in range(Ns + 1):
        B = eo.basis(Ns, k)
        rank = {int(b): i for i, b in enumerate(B)}
        shift = {}
        for i, b in enumerate(B):
            a = int(b) >> nb
            for a2 in range(1 << ml):
                if bin(a2).count("1") != bin(a).count("1"):
                    continue
                j = rank[(int(b) & low) | (a2 << nb)]
                assert shift.setdefault((a, a2), j - i) == j - i, (k, a, a2)
