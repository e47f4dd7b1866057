"""GPU checks of the corners of the storage contract:

* constant_shift (SpinResolvedStorage.h:464-466): H v + shift v until the
  next fill, which discards it;
* the real-frequency mesh (MeshFactory.h:27-34, CommonUtils.h:18-20):
  G(w + i pi/beta) from the plain and the batched solve against the exact
  Lehmann sum at the same complex points;
* the 12-site plain solve at its stated size (3x4 torus t=1, t'=-0.3, U=8,
  half filling, sector (6,6) of 853,776 states): 6 eigenpairs, then G on
  512 Matsubara points at beta=20 from 150 Lanczos steps per start vector
  (from the two lowest pairs, see the test),
  against the oracle (scipy's ARPACK for the eigenpairs, the oracle's
  GreensFunction restatement for G), whose vectors are committed fixtures;
* the 14-site Kanamori sector (7,7) (11,778,624 states): H*v and a short
  Lanczos run against the oracle.
"""
import ctypes
import math

import numpy as np
import pytest

import edlib_amd as ea
from edlib_amd import _lib
from edlib_amd.solver import SolveParams, freq_points, solve_batched, solve_plain
from oracle import edoracle as eo
from tests import models
from tests.fixtures import hubbard_arrays, load_input

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b)))


def test_constant_shift():
    p = models.random_hubbard_params(6, np.random.default_rng(5))
    om, pm = models.oracle_model("hubbard", p), models.product_model("hubbard", p)
    ctx = ea.Context(0)
    try:
        ctx.set_model(pm)
        n = ctx.set_sector(3, 2)
        srs = eo.SRS(om, 3, 2)
        v = np.random.default_rng(6).uniform(-1, 1, n)
        shift = 0.8125
        _lib.check(_lib.lib().edh_constant_shift(ctx.handle, ctypes.c_double(shift)))
        assert _rel(ctx.av(v), srs.av(v) + shift * v) < 1e-13
        _lib.check(_lib.lib().edh_constant_shift(ctx.handle, ctypes.c_double(-0.25)))
        assert _rel(ctx.av(v), srs.av(v) + (shift - 0.25) * v) < 1e-13
        ctx.set_sector(3, 2)  # a refill discards the shift, as the reference's fill() does
        assert _rel(ctx.av(v), srs.av(v)) < 1e-13
    finally:
        ctx.close()


def _gfchi():
    d = load_input("GF_Chi_input.json")
    Ns, t, U, xmu, H, J = hubbard_arrays(d)
    return Ns, (t, U, xmu, H, J)


@pytest.mark.parametrize("driver", ["plain", "batched"])
def test_real_frequency_gf_matches_lehmann(driver):
    Ns, (t, U, xmu, H, J) = _gfchi()
    om = eo.hubbard(Ns, t, U, xmu, H, J)
    pm = ea.HubbardModel(Ns, t, U, xmu, H, J)
    sp = SolveParams(nev=100, beta=20.0, nomega=41, nlanc=100, orbitals=(0,), mesh="real", emin=-4.0, emax=4.0)
    res = (solve_plain if driver == "plain" else solve_batched)(pm, sp)
    z = freq_points(sp)
    assert np.allclose(z.real, np.linspace(-4.0, 4.0, 41)) and np.allclose(z.imag, math.pi / 20.0)
    for s in range(2):
        exact = eo.lehmann_gf_z(om, 20.0, z, 0, s)
        assert np.max(np.abs(res["G"][:, 0, s] - exact)) < 1e-8


def test_12site_plain_solve_against_oracle():
    import json
    import os
    p = dict(Ns=12, t=models.torus_hopping(3, 4, 1.0, -0.3), U=np.full(12, 8.0), xmu=np.full(12, 4.0))
    pm = models.product_model("hubbard", p)
    # cutoff 1e-4: G from the ground and first excited states; the next
    # pairs are single members of two-fold multiplets, which depend on the
    # eigensolver's start vector (ARPACK's as well)
    sp = SolveParams(nev=6, beta=20.0, nomega=512, nlanc=150, orbitals=(0,), sectors=[(6, 6)], seed=20261202 + 2,
                     cutoff=1e-4)
    res = solve_plain(pm, sp)
    # oracle vectors written by tests/golden/make_plain12.py
    fx = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plain12_solve.json")
    with open(fx) as f:
        ref = json.load(f)
    ev = np.array(ref["energies"])
    assert np.max(np.abs(res["energies"] - ev) / np.abs(ev)) < 1e-10
    for s in range(2):
        G = np.array(ref["G_re"][s]) + 1j * np.array(ref["G_im"][s])
        assert np.max(np.abs(res["G"][:, 0, s] - G)) < 1e-8


def test_kanamori_14site_sector_against_oracle():
    """The largest sector of the three-orbital Kanamori stream at its stated
    size: 3 orbitals + 11 bath sites (the served stream's 14-site request,
    same seed), sector (7,7) of 11,778,624 states.  H*v in accumulate and
    clear modes against the oracle's SpinResolvedStorage::av restatement
    (1e-13 relative to the largest entry) and the first 10 Lanczos
    coefficients against the oracle's Lanczos::lanczos (1e-10)."""
    p = models.siam_params(14, 3, np.random.default_rng(20261202 + 2011))
    om, pm = models.oracle_model("siam", p), models.product_model("siam", p)
    srs = eo.SRS(om, 7, 7)
    assert srs.n == 3432 ** 2
    rng = np.random.default_rng(20261202 + 14)
    v = rng.uniform(-1, 1, srs.n)
    w0 = rng.uniform(-1, 1, srs.n)
    ref_acc = srs.av(v, w0.copy(), clear=False, omp=True)
    ref = srs.av(v, omp=True)
    ctx = ea.Context(0)
    try:
        ctx.set_model(pm)
        assert ctx.set_sector(7, 7) == srs.n
        for var in (0, 1):
            ctx.hv_variant = var
            assert _rel(ctx.av(v, w0, clear=False), ref_acc) < 1e-13, var
            assert _rel(ctx.av(v), ref) < 1e-13, var
        v /= np.linalg.norm(v)
        nl_o, a_o, b_o = srs.lanczos(v, 10)
        nl_g, a_g, b_g = ctx.lanczos(v, 10)
        assert nl_o == nl_g == 10
        assert np.max(np.abs(a_o - a_g)) < 1e-10
        assert np.max(np.abs(b_o[1:10] - b_g[1:10])) < 1e-10
    finally:
        ctx.close()
