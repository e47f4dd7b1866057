"""GPU parity of the HIP path against the CPU oracle (calls go through the C ABI).

Tolerances: H*v and Lanczos coefficients are FP64 with a different summation
order than the reference, so they are compared to 1e-12 relative to the
vector's largest entry; integer work (a_adag index maps and signs) is
bit-exact; eigenvalues to 1e-10 relative; G(iw) / chi to 1e-8 absolute
(the reference's own test pins 1e-10 / 1e-9, which is also checked).
"""
import math

import numpy as np
import pytest

import edlib_amd as ea
from oracle import edoracle as eo
from tests import models
from tests.fixtures import hubbard_arrays, load_dat, load_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = ea.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b)))


_rng = np.random.default_rng(7)
CASES = [
    ("hubbard_random_J", "hubbard", models.random_hubbard_params(5, _rng)),
    ("hubbard_ring8", "hubbard", dict(Ns=8, t=models.ring_hopping(8), U=np.full(8, 4.0), xmu=np.full(8, 2.0))),
    ("hubbard_nonuniformU", "hubbard", models.random_hubbard_params(7, np.random.default_rng(1), with_J=False)),
    ("siam_kanamori_3orb", "siam", models.siam_params(7, 3, np.random.default_rng(3), h0_offdiag=0.2)),
    ("siam_1orb", "siam", models.siam_params(8, 1, np.random.default_rng(4))),
]


@pytest.mark.parametrize("name,kind,p", CASES, ids=[c[0] for c in CASES])
def test_av_parity(ctx, name, kind, p):
    om, pm = models.oracle_model(kind, p), models.product_model(kind, p)
    ctx.set_model(pm)
    Ns = p["Ns"]
    rng = np.random.default_rng(42)
    for sec in [(Ns // 2, Ns // 2), (Ns // 2 + 1, Ns // 2 - 1), (1, Ns - 1), (0, Ns), (Ns, Ns), (2, 1)]:
        n = ctx.set_sector(*sec)
        srs = eo.SRS(om, *sec)
        assert n == srs.n
        v = rng.uniform(-1, 1, n)
        w0 = rng.uniform(-1, 1, n)
        for var in (0, 1):  # every H*v variant, whatever the autotuner would pick
            ctx.hv_variant = var
            assert _rel(ctx.av(v), srs.av(v)) < 1e-12, var
            assert _rel(ctx.av(v, w0, clear=False), srs.av(v, w0, clear=False)) < 1e-12, var


@pytest.fixture(scope="module")
def torus12():
    # 3x4 torus with t, t' (the 853,776-dimensional half-filled sector)
    p = dict(Ns=12, t=models.torus_hopping(3, 4, 1.0, -0.3), U=np.full(12, 8.0), xmu=np.full(12, 4.0))
    om = models.oracle_model("hubbard", p)
    return p, om, eo.SRS(om, 6, 6)


VARIANTS = dict(argvalues=[0, 1], ids=["single_pass", "two_pass"])


@pytest.mark.parametrize("variant", **VARIANTS)
def test_av_parity_12site_torus(ctx, torus12, variant):
    p, om, srs = torus12
    ctx.set_model(models.product_model("hubbard", p))
    n = ctx.set_sector(6, 6)
    assert n == 853776 and ctx.hv_variant in (0, 1)
    ctx.hv_variant = variant
    rng = np.random.default_rng(20261202 + 2)
    v = rng.uniform(-1, 1, n)
    w0 = rng.uniform(-1, 1, n)
    assert _rel(ctx.av(v, w0, clear=False), srs.av(v, w0, clear=False, omp=True)) < 1e-12
    assert _rel(ctx.av(v), srs.av(v, omp=True)) < 1e-12


@pytest.mark.parametrize("variant", **VARIANTS)
def test_lanczos_12site_torus(ctx, torus12, variant):
    p, om, srs = torus12
    ctx.set_model(models.product_model("hubbard", p))
    n = ctx.set_sector(6, 6)
    ctx.hv_variant = variant
    v = np.random.default_rng(8).uniform(-1, 1, n)
    v /= np.linalg.norm(v)
    nl_o, a_o, b_o = srs.lanczos(v, 25)
    nl_g, a_g, b_g = ctx.lanczos(v, 25)
    assert nl_o == nl_g == 25
    assert np.max(np.abs(a_o - a_g)) < 1e-10
    assert np.max(np.abs(b_o[1:25] - b_g[1:25])) < 1e-10


@pytest.mark.parametrize("name,kind,p", CASES[1:2] + CASES[3:4], ids=["hubbard_ring8", "siam_kanamori"])
def test_lanczos_coefficients(ctx, name, kind, p):
    om, pm = models.oracle_model(kind, p), models.product_model(kind, p)
    ctx.set_model(pm)
    Ns = p["Ns"]
    sec = (Ns // 2, Ns // 2)
    n = ctx.set_sector(*sec)
    srs = eo.SRS(om, *sec)
    v = np.random.default_rng(5).uniform(-1, 1, n)
    v /= np.linalg.norm(v)
    nl_o, a_o, b_o = srs.lanczos(v, 40)
    nl_g, a_g, b_g = ctx.lanczos(v, 40)
    assert nl_o == nl_g
    k = min(20, nl_o)
    assert np.max(np.abs(a_o[:k] - a_g[:k])) < 1e-10
    assert np.max(np.abs(b_o[1:k] - b_g[1:k])) < 1e-10
    # the continued fraction built from the full recurrence agrees in G
    z = 1j * eo.matsubara_points(10.0, 64)
    g_o = eo.frac_points(z, 1.0, nl_o, 1, 1.0, a_o, b_o)
    g_g = eo.frac_points(z, 1.0, nl_g, 1, 1.0, a_g, b_g)
    assert np.max(np.abs(g_o - g_g)) < 1e-8


def test_lanczos_breakdown_small_sector(ctx):
    p = dict(Ns=4, t=models.ring_hopping(4), U=np.full(4, 4.0), xmu=np.full(4, 2.0))
    om, pm = models.oracle_model("hubbard", p), models.product_model("hubbard", p)
    ctx.set_model(pm)
    n = ctx.set_sector(1, 1)
    srs = eo.SRS(om, 1, 1)
    v = np.ones(n) / math.sqrt(n)
    nl_o, a_o, b_o = srs.lanczos(v, 100)
    nl_g, a_g, b_g = ctx.lanczos(v, 100)
    assert nl_o == nl_g < 100  # |beta| < 1e-10 stops both at the same step
    assert np.allclose(a_o[:nl_o], a_g[:nl_g], atol=1e-10)


@pytest.mark.parametrize("name,kind,p", CASES[1:2] + CASES[3:4], ids=["hubbard_ring8", "siam_kanamori"])
def test_adag_bit_exact(ctx, name, kind, p):
    om, pm = models.oracle_model(kind, p), models.product_model(kind, p)
    ctx.set_model(pm)
    Ns = p["Ns"]
    rng = np.random.default_rng(9)
    for sec in [(Ns // 2, Ns // 2), (Ns // 2 - 1, Ns // 2 + 1), (Ns, 0), (0, 0)]:
        n = ea.binom(Ns, sec[0]) * ea.binom(Ns, sec[1])
        v = rng.uniform(-1, 1, n)
        for orb in range(2 * Ns):
            for ann in (False, True):
                nxt_o, out_o = eo.a_adag(om, sec, orb, ann, v)
                nxt_g, out_g = ctx.adag(sec, orb, ann, v)
                assert nxt_o == nxt_g
                if out_o is None:
                    assert out_g is None
                else:
                    assert np.array_equal(out_o, out_g)


def test_local_operator(ctx):
    p = dict(Ns=6, t=models.ring_hopping(6), U=np.full(6, 4.0), xmu=np.full(6, 2.0))
    om, pm = models.oracle_model("hubbard", p), models.product_model("hubbard", p)
    ctx.set_model(pm)
    sec = (3, 2)
    n = ea.binom(6, 3) * ea.binom(6, 2)
    v = np.random.default_rng(1).uniform(-1, 1, n)
    nu, nd = eo.occupations(om, sec, 6)
    for orbs in ([0], [2], [1, 4], [3, 3]):
        sz = sum(0.5 * (nu[:, o].astype(float) - nd[:, o]) for o in orbs) * v
        nn = sum((nu[:, o].astype(float) + nd[:, o]) for o in orbs) * v
        assert np.array_equal(ctx.local_op(sec, ea.EDH_OP_SZ, orbs, v), sz)
        assert np.array_equal(ctx.local_op(sec, ea.EDH_OP_N, orbs, v), nn)


def test_continued_fraction(ctx):
    rng = np.random.default_rng(3)
    for nl in (1, 2, 3, 17, 150):
        a = rng.normal(size=nl)
        b = np.concatenate([[0.0], np.abs(rng.normal(size=nl))])
        for isign in (1, -1):
            z = 1j * eo.matsubara_points(20.0, 512) + rng.normal() * isign
            ref = eo.frac_points(z, 0.7, nl, isign, 0.3, a, b)
            got = ctx.cf_eval(z, 0.7, nl, isign, 0.3, a, b)
            assert np.max(np.abs(ref - got)) < 1e-12
    # overflow rescaling path: huge determinants
    a = np.full(300, 1e3)
    b = np.concatenate([[0.0], np.full(300, 1e-3)])
    z = 1j * eo.matsubara_points(20.0, 16)
    ref = eo.frac_points(z, 1.0, 300, 1, 1.0, a, b)
    got = ctx.cf_eval(z, 1.0, 300, 1, 1.0, a, b)
    assert np.max(np.abs(ref - got) / np.abs(ref)) < 1e-12


def test_eigsh_matches_arpack(ctx):
    """Storage::diag parity: the reference calls ARPACK (dsaupd/dseupd), which the
    oracle runs through scipy with the same which/tol/ncv.  A single start vector
    can resolve only part of a degenerate multiplet; both solvers then return the
    same truncated list, so the comparison is with ARPACK, and every value is also
    checked to be a true eigenvalue (dense spectrum) with a small residual."""
    p = dict(Ns=8, t=models.ring_hopping(8), U=np.full(8, 4.0), xmu=np.full(8, 2.0))
    om, pm = models.oracle_model("hubbard", p), models.product_model("hubbard", p)
    ctx.set_model(pm)
    for sec in [(4, 4), (3, 4), (2, 5)]:
        n = ctx.set_sector(*sec)
        srs = eo.SRS(om, *sec)
        H = srs.dense()
        dense = np.linalg.eigvalsh(0.5 * (H + H.T))
        ref, _ = eo.storage_diag(srs, 4, 11)
        ev, X, nav = ctx.eigsh(4, 11)
        assert len(ev) == 4
        assert np.max(np.abs(ev - ref) / np.abs(ref)) < 1e-10
        assert abs(ev[0] - dense[0]) < 1e-10 * abs(dense[0])
        for i in range(4):
            assert np.min(np.abs(dense - ev[i])) < 1e-10 * abs(ev[i])
            r = srs.av(X[:, i]) - ev[i] * X[:, i]
            assert np.linalg.norm(r) < 1e-8
        assert abs(np.linalg.norm(X[:, 0]) - 1.0) < 1e-12


def test_eigsh_small_and_trivial_sectors(ctx):
    p = dict(Ns=4, t=models.ring_hopping(4), U=np.full(4, 5.0), xmu=np.full(4, 2.5))
    om, pm = models.oracle_model("hubbard", p), models.product_model("hubbard", p)
    ctx.set_model(pm)
    ctx.set_sector(0, 0)
    ev, X, _ = ctx.eigsh(2)
    assert ev.shape == (1,) and ev[0] == 0.0 and X[0, 0] == 1.0
    n = ctx.set_sector(2, 2)
    srs = eo.SRS(om, 2, 2)
    ref = np.linalg.eigvalsh(srs.dense())
    ev, X, _ = ctx.eigsh(100)  # ncv = ntot: full spectrum except the top state (Storage.h:53-54)
    assert len(ev) == n - 1
    assert np.allclose(ev, ref[:-1], rtol=0, atol=1e-12)


def test_gf_chi_golden_via_device(ctx):
    """GreensFunction / ChiLoc flow on the device primitives vs gom1/xiats/xiatd."""
    d = load_input("GF_Chi_input.json")
    Ns, t, U, xmu, H, J = hubbard_arrays(d)
    pm = ea.HubbardModel(Ns, t, U, xmu, H, J)
    ctx.set_model(pm)
    beta, nomega, nlanc = 20.0, 200, 100
    pairs = []
    k = 0
    for sec in eo.sector_order(Ns):
        ctx.set_sector(*sec)
        ev, X, _ = ctx.eigsh(100)
        for i in range(len(ev)):
            pairs.append((ev[i], X[:, i].copy(), k, sec))
            k += 1
    order = eo.eigenpair_order([p[0] for p in pairs])
    pairs = [pairs[i] for i in order]
    e0 = pairs[0][0]
    Z = sum(math.exp(-(p[0] - e0) * beta) for p in pairs)
    z = 1j * eo.matsubara_points(beta, nomega)
    G = np.zeros((nomega, 2), complex)
    for (E, vec, _, sec) in pairs:
        bf = math.exp(-(E - e0) * beta)
        if bf < 1e-12:
            continue
        expb = 0.0 if beta * (E - e0) > 25 else math.exp(-beta * (E - e0))
        for spin in range(2):
            for ann, isign in ((False, 1), (True, -1)):
                nxt, out = ctx.adag(sec, 0 + spin * Ns, ann, vec)
                if out is None:
                    continue
                norm = ctx.vv(out, out)
                if not abs(norm) > 1e-10:
                    continue
                ctx.set_sector(*nxt)
                nl, a, b = ctx.lanczos(out / math.sqrt(norm), nlanc)
                G[:, spin] += ctx.cf_eval(z + E * isign, norm, nl, isign, expb, a, b)
    G /= Z
    ref = load_dat("gom1.dat")
    for s in range(2):
        assert np.max(np.abs(G[:, s] - (ref[:, 1] + 1j * ref[:, 2]))) < 1e-10


@pytest.mark.parametrize("name,kind,p", CASES, ids=[c[0] for c in CASES])
def test_static_observables_match_oracle(ctx, name, kind, p):
    """Per-eigenvector sums of StaticObservables.h:356-419 (including its
    orbital rule for the second down occupation) from device vectors."""
    om, pm = models.oracle_model(kind, p), models.product_model(kind, p)
    ctx.set_model(pm)
    Ns, norb = p["Ns"], om.interacting_orbitals()
    rng = np.random.default_rng(13)
    for sec in [(Ns // 2, Ns // 2), (Ns // 2 + 1, Ns // 2 - 1), (0, Ns), (2, 1)]:
        n = ctx.set_sector(*sec)
        v = rng.uniform(-1, 1, n)
        v /= np.linalg.norm(v)
        ref = eo.static_observables(om, [eo.EigenPair(0.0, v, 0, sec)], {"lanc.BETA": 1.0})
        buf = ctx.buffer(n)
        buf.upload(v)
        for got in (ctx.static_observables(v, norb), ctx.static_observables(buf, norb)):
            for key in ("N", "N_up", "N_dn", "M", "D_occ", "M_i M_j"):
                assert np.abs(got[key] - ref[key]).max() < 1e-12, (key, sec)
            assert abs(got["N_eff"][0] - ref["N_eff"][0]) < 1e-9 * ref["N_eff"][0]
        buf.free()


def test_static_observables_4ring_paramagnetic(ctx):
    """StaticObservables_Test.cpp:63-66 from the device ground state."""
    d = load_input("4ring_nomagfield_input.json")
    Ns, t, U, xmu, H, J = hubbard_arrays(d)
    ctx.set_model(ea.HubbardModel(Ns, t, U, xmu, H, J))
    n = ctx.set_sector(2, 2)
    evals, evecs, _ = ctx.eigsh(1)
    obs = ctx.static_observables(evecs[:, 0], Ns)
    for o in range(Ns):
        assert abs(obs["N"][o] - 1.0) < 1e-8
        assert abs(obs["N_up"][o] - 0.5) < 1e-8
        assert abs(obs["N_dn"][o] - 0.5) < 1e-8
        assert abs(obs["M"][o]) < 1e-7


def test_lanczos_graph_replay_matches_eager():
    """The Lanczos step replayed as hipGraphs gives the directly launched
    recurrence's coefficients, call after call with eager work in between."""
    import os
    Ns = 12
    ctx = ea.Context(0)
    ctx.set_model(ea.HubbardModel(Ns, models.torus_hopping(3, 4), np.full(Ns, 8.0), np.full(Ns, 4.0)))
    n = ctx.set_sector(6, 6)
    v0 = np.random.default_rng(20261202 + 11).uniform(-1, 1, n)
    v0 /= np.linalg.norm(v0)
    os.environ["EDH_LANCZOS_GRAPH"] = "0"
    try:
        ref = ctx.lanczos(v0, 40)
    finally:
        os.environ.pop("EDH_LANCZOS_GRAPH", None)
    for rep in range(3):
        ctx.av(v0)  # eager work between the replays
        got = ctx.lanczos(v0, 40)
        assert got[0] == ref[0]
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2]), rep
    ctx.close()


def test_av_parity_16site_sector():
    """The 16-site 4x4 Hubbard sector (8,8) (N = 165,636,900) against the
    oracle's SpinResolvedStorage::av restatement, every H*v variant."""
    Ns = 16
    t = models.torus_hopping(4, 4)
    om = eo.hubbard(Ns, t, np.full(Ns, 6.0), np.full(Ns, 3.0))
    srs = eo.SRS(om, 8, 8)
    n = srs.n
    rng = np.random.default_rng(20261202 + 8)
    v = rng.uniform(-1.0, 1.0, n)
    w0 = np.random.default_rng(20261202 + 9).uniform(-1.0, 1.0, n)
    ref = srs.av(v, w0.copy(), clear=False, omp=True)
    ctx = ea.Context(0)
    ctx.set_model(ea.HubbardModel(Ns, t, np.full(Ns, 6.0), np.full(Ns, 3.0)))
    assert ctx.set_sector(8, 8) == n
    for var in (0, 1):
        ctx.hv_variant = var
        got = ctx.av(v, w0, clear=False)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (var, err)
    ctx.close()


@pytest.mark.parametrize("nup", [1, 2, 3])
def test_av_parity_16site_long_rows(ctx, nup):
    """Rows of 12,870 columns (four column chunks) with few rows: sectors
    (nup, 8) of the 4x4 torus, every H*v variant against the oracle,
    accumulate and clear modes."""
    Ns = 16
    t = models.torus_hopping(4, 4)
    om = eo.hubbard(Ns, t, np.full(Ns, 6.0), np.full(Ns, 3.0))
    srs = eo.SRS(om, nup, 8)
    ctx.set_model(ea.HubbardModel(Ns, t, np.full(Ns, 6.0), np.full(Ns, 3.0)))
    n = ctx.set_sector(nup, 8)
    assert n == srs.n
    rng = np.random.default_rng(20261202 + nup)
    v = rng.uniform(-1, 1, n)
    w0 = rng.uniform(-1, 1, n)
    ref_acc = srs.av(v, w0.copy(), clear=False)
    ref = srs.av(v)
    for var in (0, 1):
        ctx.hv_variant = var
        assert _rel(ctx.av(v, w0, clear=False), ref_acc) < 1e-13, var
        assert _rel(ctx.av(v), ref) < 1e-13, var


def test_16site_gf_start_vector_and_lanczos_against_oracle():
    """A Green's-function run of the 16-site 4x4 Hubbard sector (8,8) at
    its stated size: the ground state (device eigensolver, NEV = 1,
    NCV = 20, tol 1e-10), its start vector c+_{0 up}|psi> in sector (9,8)
    (147,232,800 states) bit-exact against the oracle's
    SpinResolvedStorage::a_adag restatement, its weight <n> = 1/2 (half
    filling of a bipartite lattice at mu = U/2: particle-hole symmetry), and
    the first 8 Lanczos coefficients against the oracle's Lanczos::lanczos
    (OpenMP form) from the same normalised vector (1e-10)."""
    Ns = 16
    t = models.torus_hopping(4, 4)
    U, mu = np.full(Ns, 6.0), np.full(Ns, 3.0)
    om = eo.hubbard(Ns, t, U, mu)
    ctx = ea.Context(0)
    try:
        ctx.set_model(ea.HubbardModel(Ns, t, U, mu))
        ctx.set_sector(8, 8)
        e, vecs, _ = ctx.eigsh(1, 20, 1e-10, 1000, 20261202, vectors=True)
        psi = vecs[:, 0].copy()
        del vecs
        nxt, out = ctx.adag((8, 8), 0, False, psi)
        assert nxt == (9, 8) and out.size == 11440 * 12870
        nxt_o, out_o = eo.a_adag(om, (8, 8), 0, False, psi)
        assert nxt_o == nxt and np.array_equal(out, out_o)
        del out_o, psi
        norm2 = float(np.dot(out, out))
        assert abs(norm2 - 0.5) < 1e-8, norm2
        out /= np.sqrt(norm2)
        ctx.set_sector(*nxt)
        nl, a, b = ctx.lanczos(out, 8)
        srs = eo.SRS(om, *nxt)
        nl_o, a_o, b_o = srs.lanczos(out, 8, omp=True)
        assert nl == nl_o == 8
        assert np.abs(a - a_o).max() < 1e-10 * max(1.0, np.abs(a_o).max())
        assert np.abs(b[1:8] - b_o[1:8]).max() < 1e-10 * max(1.0, np.abs(b_o).max())
    finally:
        ctx.close()
