"""Batched engine and serving loop on the GPU.

* eigenpairs of every sector from one batch == the single-sector eigensolver
  (1e-10 relative) and the oracle's ARPACK run (scipy eigsh);
* batched Lanczos coefficients == the single-sector recurrence;
* a batched impurity solve (all sectors, then every Green's-function start
  vector) == the plain per-sector solve and the oracle's GreensFunction
  restatement (1e-8 absolute);
* the served result of a request is bit-identical alone and in a mix.
"""
import math

import numpy as np
import pytest

from models import oracle_model, product_model, siam_params

pytestmark = pytest.mark.gpu

SEED = 20261202


def _siam(Ns, ml=1, seed=SEED + 4):
    return siam_params(Ns, ml, np.random.default_rng(seed))


def test_batch_eigenpairs_match_single_sector_and_arpack():
    import edlib_amd as ea
    from edlib_amd.batch import Batch
    from oracle import edoracle as eo
    p = _siam(7)
    m = product_model("siam", p)
    om = oracle_model("siam", p)
    b = Batch(0)
    ctx = ea.Context(0)
    ctx.set_model(m)
    jobs = {}
    for i in range(8):
        for j in range(8):
            sid, n = b.add_sector(m, i, j)
            jobs[(i, j)] = b.submit_eig(sid, 4, 11, 1e-14, 1000, SEED + 8 * i + j)
    b.run()
    checked = 0
    for (i, j), jid in jobs.items():
        ev, vecs = b.eig_result(jid, vectors=True)
        n = ctx.set_sector(i, j)
        ref, _, _ = ctx.eigsh(4, 11, 1e-14, 1000, SEED + 8 * i + j, vectors=False)
        assert len(ev) == len(ref)
        np.testing.assert_allclose(ev, ref, rtol=1e-10, atol=1e-12)
        if n > 64:  # Krylov path: residuals and the ARPACK oracle
            srs = eo.SRS(om, i, j)
            for k in range(len(ev)):
                r = srs.av(vecs[:, k]) - ev[k] * vecs[:, k]
                assert np.linalg.norm(r) < 1e-8
            e_ar, _ = eo.storage_diag(srs, 4, 11)
            np.testing.assert_allclose(ev, e_ar, rtol=1e-10, atol=1e-12)
            checked += 1
    assert checked > 5
    ctx.close()
    b.close()


def test_batch_lanczos_matches_single_sector():
    import edlib_amd as ea
    from edlib_amd.batch import Batch
    from models import ring_hopping
    Ns = 8
    m = ea.HubbardModel(Ns, ring_hopping(Ns), np.full(Ns, 4.0), np.full(Ns, 2.0))
    ctx = ea.Context(0)
    ctx.set_model(m)
    b = Batch(0)
    runs = []
    for sec in [(4, 4), (3, 5), (2, 2)]:
        n = ctx.set_sector(*sec)
        v0 = np.random.default_rng(SEED + 1 + sec[0]).uniform(-1, 1, n)
        v0 /= np.linalg.norm(v0)
        sid, _ = b.add_sector(m, *sec)
        runs.append((sec, v0, b.submit_lanczos(sid, v0, 40)))
    b.run()
    for sec, v0, jid in runs:
        ctx.set_sector(*sec)
        nl_ref, a_ref, b_ref = ctx.lanczos(v0, 40)
        nl, a, bb = b.lanczos_result(jid, 40)
        assert nl == nl_ref
        k = min(nl, 20)
        np.testing.assert_allclose(a[:k], a_ref[:k], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(bb[1:k], b_ref[1:k], rtol=1e-10, atol=1e-12)
    ctx.close()
    b.close()


def test_grouped_hv_timing_reports_bytes_and_leaves_batch_usable():
    """edh_batch_time_hv: elements = sum of n, algorithmic bytes = sum of
    16 n + table bytes (the single-sector context's figure for each sector),
    a positive time; the batch's later sweeps (whose descriptors differ from
    the timed launch's captured graph) still give the single-sector Lanczos
    coefficients."""
    import edlib_amd as ea
    from edlib_amd.batch import Batch
    from models import ring_hopping
    Ns = 8
    m = ea.HubbardModel(Ns, ring_hopping(Ns), np.full(Ns, 4.0), np.full(Ns, 2.0))
    ctx = ea.Context(0)
    ctx.set_model(m)
    b = Batch(0)
    secs = [(4, 4), (3, 5), (2, 2), (0, 0)]
    sids, want_b, want_n = [], 0, 0
    for sec in secs:
        n = ctx.set_sector(*sec)
        _, _, tb = ctx.sector_info()
        sid, _ = b.add_sector(m, *sec)
        sids.append(sid)
        want_n += n
        want_b += 16 * n + (tb if n else 0)
    for graphs in (True, False):
        b.set_graphs(graphs)
        ms, nbytes, nel = b.time_hv(sids, 5, SEED)
        assert nel == want_n and nbytes == want_b
        assert ms > 0.0 and math.isfinite(ms)
    b.set_graphs(True)
    runs = []
    for sec, sid in zip(secs[:3], sids[:3]):
        n = ctx.set_sector(*sec)
        v0 = np.random.default_rng(SEED + 2 + sec[0]).uniform(-1, 1, n)
        v0 /= np.linalg.norm(v0)
        runs.append((sec, v0, b.submit_lanczos(sid, v0, 30)))
    b.run()
    for sec, v0, jid in runs:
        ctx.set_sector(*sec)
        nl_ref, a_ref, b_ref = ctx.lanczos(v0, 30)
        nl, a, bb = b.lanczos_result(jid, 30)
        assert nl == nl_ref
        k = min(nl, 20)
        np.testing.assert_allclose(a[:k], a_ref[:k], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(bb[1:k], b_ref[1:k], rtol=1e-10, atol=1e-12)
    ctx.close()
    b.close()


def test_large_operands_on_rows_kernel_match_single_sector():
    """Operands with long rows (D_dn > 4608, N >= 2^22) run on the
    single-sector rows kernel inside a batch sweep: eigenpairs and Lanczos
    coefficients equal the single-sector path's, and a run's coefficients
    are bit-identical alone and next to a grouped (small) operand."""
    import edlib_amd as ea
    from edlib_amd.batch import Batch
    from models import ring_hopping
    Ns = 16
    m = ea.HubbardModel(Ns, ring_hopping(Ns), np.full(Ns, 4.0), np.full(Ns, 2.0))
    big = (3, 6)  # 560 x 8008 = 4,484,480 states
    ctx = ea.Context(0)
    ctx.set_model(m)
    n = ctx.set_sector(*big)
    assert n == 560 * 8008
    ref_e, _, _ = ctx.eigsh(2, 8, 1e-12, 1000, SEED + 5, vectors=False)
    v0 = np.random.default_rng(SEED + 6).uniform(-1, 1, n)
    v0 /= np.linalg.norm(v0)
    nl_ref, a_ref, b_ref = ctx.lanczos(v0, 30)
    coef = []
    for mixed in (False, True):
        b = Batch(0)
        sid, _ = b.add_sector(m, *big)
        jid = b.submit_eig(sid, 2, 8, 1e-12, 1000, SEED + 5)
        lid = b.submit_lanczos(sid, v0, 30)
        if mixed:
            small, ns = b.add_sector(m, 2, 2)
            b.submit_eig(small, 2, 8, 1e-12, 1000, SEED + 7)
            b.submit_lanczos(small, np.full(ns, 1.0 / math.sqrt(ns)), 30)
        b.run()
        ev, _ = b.eig_result(jid)
        np.testing.assert_allclose(ev, ref_e, rtol=1e-10, atol=1e-12)
        nl, a, bb = b.lanczos_result(lid, 30)
        assert nl == nl_ref
        np.testing.assert_allclose(a[:nl], a_ref[:nl], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(bb[1:nl], b_ref[1:nl], rtol=1e-10, atol=1e-12)
        coef.append((a.copy(), bb.copy()))
        b.close()
    assert np.array_equal(coef[0][0], coef[1][0]) and np.array_equal(coef[0][1], coef[1][1])
    ctx.close()


def _params(**kw):
    from edlib_amd.solver import SolveParams
    base = dict(nev=4, beta=10.0, nomega=32, nlanc=100, orbitals=(0,), seed=SEED)
    base.update(kw)
    return SolveParams(**base)


def test_batched_solve_matches_plain_and_oracle():
    from edlib_amd.solver import solve_batched, solve_plain
    from oracle import edoracle as eo
    p = _siam(6)
    m = product_model("siam", p)
    sp = _params()
    rb = solve_batched(m, sp)
    rp = solve_plain(m, sp)
    assert abs(rb["E0"] - rp["E0"]) < 1e-10 * abs(rp["E0"])
    np.testing.assert_allclose(rb["energies"], rp["energies"], rtol=1e-10, atol=1e-10)
    assert rb["n_gf"] == rp["n_gf"] > 0
    assert np.abs(rb["G"] - rp["G"]).max() < 1e-8
    # oracle: Hamiltonian::diag + GreensFunction::compute restatement
    om = oracle_model("siam", p)
    prm = {"arpack.NEV": 4, "lanc.BETA": 10.0, "lanc.NOMEGA": 32, "lanc.NLANC": 100,
           "lanc.BOLTZMANN_CUTOFF": 1e-12}
    pairs = eo.hamiltonian_diag(om, prm)
    go = eo.greens_function(om, pairs, prm, gf_orbs=[0])
    assert abs(pairs[0].eigenvalue - rb["E0"]) < 1e-10 * abs(rb["E0"])
    assert np.abs(rb["G"][:, 0, :] - go["G"][:, 0, :]).max() < 1e-8
    # static observables (StaticObservables.h:106-158, 356-419): batched and
    # plain from device sums, against the oracle's per-state loop
    oo = eo.static_observables(om, pairs, prm)
    for key in ("N", "N_up", "N_dn", "M", "D_occ", "M_i M_j", "E"):
        assert np.abs(rb["observables"][key] - rp["observables"][key]).max() < 1e-10, key
        assert np.abs(rb["observables"][key] - oo[key]).max() < 1e-8, key
    assert abs(rb["observables"]["N_eff"][0] - oo["N_eff"][0]) < 1e-8 * oo["N_eff"][0]


def test_served_result_same_alone_and_in_mix():
    from edlib_amd.serve import Server, latency_stats
    pa = _siam(6, seed=SEED + 1000)
    pb = _siam(7, seed=SEED + 1001)
    pc = _siam(5, seed=SEED + 1002)
    sp = _params(nlanc=60)
    ma, mb, mc = product_model("siam", pa), product_model("siam", pb), product_model("siam", pc)
    s1 = Server(0)
    s1.submit("a", ma, sp)
    while s1.tick():
        pass
    alone = s1.done["a"]
    s1.close()
    s2 = Server(0, sweeps_per_tick=3)
    s2.submit("b", mb, sp)
    s2.tick()
    s2.tick()
    s2.submit("a", ma, sp)   # joins while b is running
    s2.tick()
    s2.submit("c", mc, sp)
    while s2.tick():
        pass
    mix = s2.done["a"]
    st = latency_stats(s2.done)
    s2.close()
    assert set(s2.done) == {"a", "b", "c"}
    assert np.array_equal(alone["energies"], mix["energies"])
    assert np.array_equal(alone["G"], mix["G"])
    assert st["n"] == 3 and st["p99_s"] >= st["p50_s"] > 0


def test_vanishing_start_vector_has_no_run():
    """A start vector c(+)|psi> of squared norm <= 1e-10 has no Lanczos run
    and no Green's-function term (GreensFunction.h:408-418): with the
    impurity decoupled and deep (mu = 10) its low-lying states are doubly
    occupied, so c^+_0 vanishes on them while c_0 has norm 1.  Batched and
    plain solves agree on which terms exist and on G."""
    from edlib_amd.batch import Batch
    from edlib_amd.solver import solve_batched, solve_plain
    p = _siam(5)
    p["Vk"][:] = 0.0
    p["mu"] = 10.0
    m = product_model("siam", p)
    b = Batch(0)
    sid, n = b.add_sector(m, 2, 2)
    assert n == 100
    jid = b.submit_eig(sid, 4, 11, 1e-14, 1000, SEED)
    b.run()
    up, _ = b.add_sector(m, 3, 2)
    dn, _ = b.add_sector(m, 1, 2)
    ids = b.submit_gf([(jid, 0, 0, 0, up), (jid, 0, 0, 1, dn)], 20)
    assert min(ids) >= 0
    b.run()
    n2 = b.lanczos_norm2(ids)
    assert n2[0] <= 1e-10, n2
    assert abs(n2[1] - 1.0) < 1e-10, n2
    assert b.lanczos_result(ids[0], 20)[0] == 0
    assert b.lanczos_result(ids[1], 20)[0] > 0
    b.close()
    sp = _params(nlanc=40)
    rb = solve_batched(m, sp)
    rp = solve_plain(m, sp)
    assert rb["n_gf"] == rp["n_gf"] > 0
    assert np.abs(rb["G"] - rp["G"]).max() < 1e-8


def test_graph_replayed_sweeps_match_eager():
    """Steady sweeps replay one captured hipGraph; a solve with replays is
    bit-identical to the same solve launched kernel by kernel, over many
    replays with the engine's own eager sweeps and reads in between.  The
    Green's-function runs are not held back by memory here (their starts,
    and so the sweep count, then follow the eigensolver runs only); with a
    budget that holds them back (started as others finish, which depends on
    timing) the result is still bit-identical."""
    from edlib_amd.batch import Batch
    from edlib_amd.solver import solve_batched
    p = _siam(7, seed=SEED + 77)
    m = product_model("siam", p)
    sp = _params(nlanc=40)
    out = {}
    for graphs in (True, False):
        b = Batch(0)
        b.set_graphs(graphs)
        out[graphs] = (solve_batched(m, sp, batch=b, gf_budget=float("inf")), b.sweep_counts())
        b.close()
    (rg, (sw_g, rep_g)), (re_, (sw_e, rep_e)) = out[True], out[False]
    assert rep_e == 0 and rep_g > 10, (sw_g, rep_g)
    assert sw_g == sw_e
    assert np.array_equal(rg["energies"], re_["energies"])
    assert np.array_equal(rg["G"], re_["G"])
    tight = solve_batched(m, sp, gf_budget=200_000)  # a few runs in flight at a time
    assert np.array_equal(tight["energies"], rg["energies"])
    assert np.array_equal(tight["G"], rg["G"])


def test_lanczos_sweeps_replay_graphs_bit_identical():
    """Lanczos runs resolve their buffer rotation and history entries from a
    device step counter, so their sweeps are steady from step to step and
    replay the captured graph; the coefficients are bit-identical to eager
    launches and equal the single-sector recurrence.  The run on the
    8-state sector (0, 1) stops early (|beta| < 1e-10 once the Krylov space
    is the whole sector) while the others go on."""
    import edlib_amd as ea
    from edlib_amd.batch import Batch
    from models import ring_hopping
    Ns = 8
    m = ea.HubbardModel(Ns, ring_hopping(Ns), np.full(Ns, 4.0), np.full(Ns, 2.0))
    ctx = ea.Context(0)
    ctx.set_model(m)
    secs = [(4, 4), (3, 5), (2, 3), (0, 1)]
    v0s = {}
    for sec in secs:
        n = ctx.set_sector(*sec)
        v = np.random.default_rng(SEED + 5 + sec[0]).uniform(-1, 1, n)
        v0s[sec] = v / np.linalg.norm(v)
    nl = 45
    res = {}
    for graphs in (True, False):
        b = Batch(0)
        b.set_graphs(graphs)
        jobs = []
        for sec in secs:
            sid, _ = b.add_sector(m, *sec)
            jobs.append(b.submit_lanczos(sid, v0s[sec], nl))
        b.run()
        res[graphs] = ([b.lanczos_result(j, nl) for j in jobs], b.sweep_counts())
        b.close()
    (rg, (sw_g, rep_g)), (re_, (sw_e, rep_e)) = res[True], res[False]
    assert sw_g == sw_e and rep_e == 0
    assert rep_g >= nl - 4, (sw_g, rep_g)  # every step but the first two and the last
    for (nl_g, a_g, b_g), (nl_e, a_e, b_e) in zip(rg, re_):
        assert nl_g == nl_e
        assert np.array_equal(a_g, a_e) and np.array_equal(b_g, b_e)
    assert rg[3][0] <= 8 and all(r[0] == nl for r in rg[:3])
    for sec, (nl_g, a_g, b_g) in zip(secs, rg):
        ctx.set_sector(*sec)
        nl_ref, a_ref, b_ref = ctx.lanczos(v0s[sec], nl)
        assert nl_g == nl_ref
        k = min(nl_g, 20)
        np.testing.assert_allclose(a_g[:k], a_ref[:k], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(b_g[1:k], b_ref[1:k], rtol=1e-10, atol=1e-12)
    ctx.close()


def test_aim12_all_sectors_batched_against_oracle():
    """The batched solve at its stated size: single-orbital Anderson impurity
    + 11 bath sites, all 169 (N_up, N_dn) sectors (dimensions 1..853,776),
    NEV = 4, G from 4 start vectors per pair (spin x particle/hole) on 256
    Matsubara points at beta = 10, NLANC = 60.  Eigenvalues against the
    oracle's ARPACK runs to 1e-10 relative, G against the oracle's
    GreensFunction restatement to 1e-8 (vectors written by
    tests/golden/make_aim12.py)."""
    import json
    import os
    from edlib_amd.solver import solve_batched
    fx = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "aim12_solve.json")
    with open(fx) as f:
        ref = json.load(f)
    m = product_model("siam", siam_params(12, 1, np.random.default_rng(SEED + 4)))
    sp = _params(nomega=256, nlanc=60)
    rb = solve_batched(m, sp)
    ev = np.array(ref["energies"])
    assert len(rb["energies"]) == len(ev) == 169 * 4 - 4 * 3  # the four 1-state sectors hold one pair each
    assert np.max(np.abs(rb["energies"] - ev) / np.maximum(1.0, np.abs(ev))) < 1e-10
    for s in range(2):
        G = np.array(ref["G_re"][s]) + 1j * np.array(ref["G_im"][s])
        assert np.max(np.abs(rb["G"][:, 0, s] - G)) < 1e-8


def test_kanamori_stream_to_14_sites_served():
    """The served form on the three-orbital Kanamori stream at its stated
    sizes: requests with bath sizes 6, 11 and 8 (9, 14 and 11 sites; the
    14-site half-filled sectors hold up to 11.8 M states), admitted by HBM
    footprint with a budget that fits the first two only, so the third
    waits for memory.  Each result is bit-identical alone and in the mix
    (the 9- and 11-site requests run alone again); every request's spectral
    weight <c c+> + <c+ c> is 1 for each orbital and spin, its Green's
    function is spin-symmetric (the model is) and i w G(i w) -> 1.  The
    14-site ground energy equals the oracle's (ARPACK on the oracle's
    H*v restatement in sector (8,9), kanamori14_e0.json) to 1e-10."""
    import json
    import os
    from edlib_amd.serve import Server, footprint_bytes
    from edlib_amd.solver import solve_batched
    baths = (6, 11, 8)
    models = [product_model("siam", siam_params(3 + b, 3, np.random.default_rng(SEED + 2000 + b))) for b in baths]
    sp = _params(nomega=256, nlanc=60, orbitals=(0, 1, 2))
    fp = [footprint_bytes(m.Ns, sp) for m in models]
    srv = Server(0, hbm_budget=fp[0] + fp[1] + (1 << 20), sweeps_per_tick=8)
    for r, m in enumerate(models):
        srv.submit(r, m, sp)
    while srv.tick():
        pass
    done = srv.done
    blocked = srv.blocked_by["memory"]
    srv.close()
    assert sorted(done) == [0, 1, 2] and not any("error" in d for d in done.values())
    assert blocked > 0 and done[2]["queued"] > 0
    wn = (2 * np.arange(256) + 1) * math.pi / 10.0
    for r in range(3):
        G, w = done[r]["G"], done[r]["weights"]
        assert np.max(np.abs(w - 1.0)) < 1e-10, (r, w)
        assert np.max(np.abs(G[:, :, 0] - G[:, :, 1])) < 1e-10
        assert np.max(np.abs(-wn[-1] * G[-1, :, :].imag - 1.0)) < 2e-2
    fx = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kanamori14_e0.json")
    with open(fx) as f:
        ref = json.load(f)
    e0 = float(np.min(done[1]["energies"]))
    assert abs(e0 - ref["E0"]) < 1e-10 * abs(ref["E0"]), (e0, ref["E0"])
    for r in (0, 2):
        alone = solve_batched(models[r], sp)
        assert np.array_equal(alone["energies"], done[r]["energies"])
        assert np.array_equal(alone["G"], done[r]["G"])


def test_one_wanted_eigenvalue_and_one_step_lanczos():
    """NEV = 1: the restart keeps half the basis (ARPACK dsaup2's guard for a
    single wanted eigenvalue), so after the first NCV steps every cycle adds
    NCV/2 H*v, in the single-sector and the batched eigensolver alike; both
    find the oracle's ARPACK ground energy.  A one-step Lanczos run (the
    first step is also the last: beta_1 is not stored, Lanczos.h:64) gives
    alpha_1 = <v|H|v> as the single-sector recurrence does."""
    import edlib_amd as ea
    from edlib_amd.batch import Batch
    from oracle import edoracle as eo
    from models import ring_hopping
    Ns, sec, ncv = 10, (5, 5), 10
    U = np.full(Ns, 4.0)
    m = ea.HubbardModel(Ns, ring_hopping(Ns), U, U / 2)
    ctx = ea.Context(0)
    ctx.set_model(m)
    n = ctx.set_sector(*sec)
    ev_p, _, nav_p = ctx.eigsh(1, ncv, 1e-12, 1000, SEED + 31, vectors=False)
    b = Batch(0)
    sid, _ = b.add_sector(m, *sec)
    jid = b.submit_eig(sid, 1, ncv, 1e-12, 1000, SEED + 31)
    v0 = np.random.default_rng(SEED + 32).uniform(-1, 1, n)
    v0 /= np.linalg.norm(v0)
    lid = b.submit_lanczos(sid, v0, 1)
    b.run()
    ev_b, _ = b.eig_result(jid)
    nav_b = b.job_info(jid)["n_av"]
    om = eo.hubbard(Ns, ring_hopping(Ns), U, U / 2)
    e_ar, _ = eo.storage_diag(eo.SRS(om, *sec), 1, ncv)
    for ev, nav in ((ev_p, nav_p), (ev_b, nav_b)):
        assert len(ev) == 1
        assert abs(ev[0] - e_ar[0]) <= 1e-10 * abs(e_ar[0])
        assert nav > ncv and (nav - ncv) % (ncv // 2) == 0, (nav, ncv)
    nl, a, bb = b.lanczos_result(lid, 1)
    nl_ref, a_ref, b_ref = ctx.lanczos(v0, 1)
    assert nl == nl_ref == 1
    np.testing.assert_allclose(a[0], a_ref[0], rtol=1e-12)
    assert bb[1] == 0.0
    ctx.close()
    b.close()


@pytest.mark.parametrize("case", ["kanamori14", "aim12"])
def test_lanczos_runs_sharing_a_rows_kernel_launch_are_bit_identical(case):
    """Lanczos runs on one sector routed to the rows kernel share
    multi-operand launches when several are active: each run's coefficients
    are bit-identical to the run alone (single-operand launch) whatever the
    group (2, 4, 4 + 1 operands), and equal the single-sector recurrence.
    Kanamori 3 + 11, sector (5,7): 2002 x 3432 = 6,870,864 states (with the
    impurity terms); Anderson 1 + 11, sector (6,6): 924 x 924 (a mid-size
    sector that Green's-function runs route to the rows kernel)."""
    import edlib_amd as ea
    from edlib_amd.batch import Batch
    from edlib_amd.workloads import siam_model
    if case == "kanamori14":
        m, sec, dims = siam_model(14, 3, SEED + 1001), (5, 7), (2002, 3432)
    else:
        m, sec, dims = siam_model(12, 1, SEED + 4), (6, 6), (924, 924)
    nl = 12
    ctx = ea.Context(0)
    ctx.set_model(m)
    n = ctx.set_sector(*sec)
    assert n == dims[0] * dims[1]
    rng = np.random.default_rng(SEED + 41)
    v0s = []
    for _ in range(5):
        v = rng.uniform(-1, 1, n)
        v0s.append(v / np.linalg.norm(v))

    def run(vs):
        b = Batch(0)
        sid, _ = b.add_sector(m, *sec)
        jobs = [b.submit_lanczos(sid, v, nl) for v in vs]
        b.run()
        out = [b.lanczos_result(j, nl) for j in jobs]
        b.close()
        return out

    alone = [run([v])[0] for v in v0s]
    for group in ([0, 1], [0, 1, 2, 3], [0, 1, 2, 3, 4]):
        res = run([v0s[i] for i in group])
        for i, (nl_g, a_g, b_g) in zip(group, res):
            nl_a, a_a, b_a = alone[i]
            assert nl_g == nl_a
            assert np.array_equal(a_g, a_a) and np.array_equal(b_g, b_a), (group, i)
    nl_ref, a_ref, b_ref = ctx.lanczos(v0s[0], nl)
    assert alone[0][0] == nl_ref
    np.testing.assert_allclose(alone[0][1][:nl], a_ref[:nl], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(alone[0][2][1:nl], b_ref[1:nl], rtol=1e-10, atol=1e-12)
    ctx.close()
