"""Host logic of the solves' non-local Green's function and self-energy,
without a device: the orbital lists and start-vector targets
(GreensFunction.h:65-102), non_local_gf's combination (GreensFunction.h:
336-359) on synthetic parts that obey <(c_i + c_j)(c_i + c_j)+> = G_ii +
G_jj + 2 G_ij, the Dyson step, the models' bare G0 against closed forms and
the argument checks."""
import numpy as np
import pytest

import edlib_amd as ea
from edlib_amd.solver import SolveParams, _check_pairs, _finish_gf, _gf_targets


class _Fake:
    def __init__(self, ml, g0):
        self.ml, self.g0 = ml, g0

    def interacting_orbitals(self):
        return self.ml

    def bare_gf(self, z):
        return self.g0


def _sym(rng, nw, ml):
    a = rng.normal(size=(nw, ml, ml, 2)) + 1j * rng.normal(size=(nw, ml, ml, 2))
    return a + np.swapaxes(a, 1, 2)


def test_orbital_list_and_targets():
    p = SolveParams(orbitals=(2, 0), pairs=((0, 1), (3, 2)))
    assert p.gf_orbitals == [2, 0, 1, 3]
    t = _gf_targets(4, (2, 1), p)
    ops = [x[0] for x in t]
    assert ops[:4] == [(2,)] * 4 and ops[-4:] == [(3, 2)] * 4
    assert len(t) == 6 * 4
    # a full up-spin level has no creation target, an empty one no annihilation
    t = _gf_targets(4, (4, 0), SolveParams(orbitals=(0,)))
    assert [(s, ann) for _, s, ann, _ in t] == [(0, True), (1, False)]


def test_nonlocal_combination_recovers_gij():
    rng = np.random.default_rng(1)
    nw, ml, Z = 7, 3, 2.5
    Gg_true, Gl_true = _sym(rng, nw, ml), _sym(rng, nw, ml)
    p = SolveParams(orbitals=(0, 1, 2), pairs=((0, 2), (1, 0)))
    orbs = p.gf_orbitals
    Gg = np.stack([Gg_true[:, o, o, :] for o in orbs], axis=1) * Z
    Gl = np.stack([Gl_true[:, o, o, :] for o in orbs], axis=1) * Z
    Ggij = np.zeros((nw, ml * ml, 2), complex)
    Glij = np.zeros_like(Ggij)
    for i, j in p.pairs:
        Ggij[:, i * ml + j] = Z * (Gg_true[:, i, i] + Gg_true[:, j, j] + 2 * Gg_true[:, i, j])
        Glij[:, i * ml + j] = Z * (Gl_true[:, i, i] + Gl_true[:, j, j] + 2 * Gl_true[:, i, j])
    out = _finish_gf(_Fake(ml, None), p, np.ones(nw), Gg, Gl, Ggij, Glij, Z)
    true = (Gg_true + Gl_true).reshape(nw, ml * ml, 2)
    for k in [0 * ml + 2, 1 * ml + 0, 0, 4, 8]:
        assert np.abs(out["G_ij"][:, k] - true[:, k]).max() < 1e-13, k
    for k in [0 * ml + 1, 2 * ml + 0]:  # not listed: stays zero
        assert np.abs(out["G_ij"][:, k]).max() == 0.0
    assert np.abs(out["G"] - np.stack([true[:, o * ml + o] for o in orbs], axis=1)).max() < 1e-13
    assert "Sigma" not in out


def test_dyson_step():
    rng = np.random.default_rng(2)
    nw, ml = 5, 2
    G = _sym(rng, nw, ml)
    S = _sym(rng, nw, ml)
    G0 = np.moveaxis(np.linalg.inv(np.linalg.inv(np.moveaxis(G, 3, 1)) + np.moveaxis(S, 3, 1)), 1, 3)
    p = SolveParams(orbitals=(0, 1), pairs=((0, 1), (1, 0)), selfenergy=True)
    Gg = np.stack([G[:, o, o, :] for o in (0, 1)], axis=1)
    Gl = np.zeros_like(Gg)
    Ggij = np.zeros((nw, 4, 2), complex)
    for i, j in p.pairs:
        Ggij[:, i * ml + j] = G[:, i, i] + G[:, j, j] + 2 * G[:, i, j]
    out = _finish_gf(_Fake(ml, G0.reshape(nw, 4, 2)), p, np.ones(nw), Gg, Gl, Ggij, np.zeros_like(Ggij), 1.0)
    assert np.abs(out["Sigma"] - S.reshape(nw, 4, 2)).max() < 1e-11


def test_bare_gf_closed_forms():
    rng = np.random.default_rng(3)
    Ns = 4
    t = rng.normal(size=(Ns, Ns))
    t = 0.5 * (t + t.T)
    xmu = rng.uniform(0, 1, Ns)
    z = np.array([0.3j, 1.0 + 0.1j])
    g = ea.HubbardModel(Ns, t, np.zeros(Ns), xmu).bare_gf(z)
    for q in range(2):
        ref = np.linalg.inv(z[q] * np.eye(Ns) + np.diag(xmu) + t).reshape(-1)
        for s in range(2):
            assert np.abs(g[q, :, s] - ref).max() < 1e-14
    ml, nb = 2, 3
    eps = rng.uniform(-1, 1, (nb, 2))
    vk = rng.uniform(0.1, 0.5, (ml, nb, 2))
    h0 = np.zeros((ml, ml, 2))
    h0[0, 0], h0[1, 1] = 0.1, -0.2
    U = np.zeros((2, 2, ml, ml, ml, ml))
    g = ea.SingleImpurityAndersonModel(ml + nb, ml, eps, vk, h0, U, 0.4).bare_gf(z)
    for q in range(2):
        for s in range(2):
            for i in range(ml):
                d = sum(vk[i, k, s] ** 2 / (z[q] - eps[k, s]) for k in range(nb))
                assert abs(g[q, i * ml + i, s] - 1 / (z[q] - h0[i, i, s] - d)) < 1e-14
            assert g[q, 1, s] == 0 and g[q, 2, s] == 0


def test_pair_checks():
    m = ea.HubbardModel(3, np.zeros((3, 3)), np.zeros(3), np.zeros(3))
    for bad in (((0, 0),), ((0, 3),), ((1,),)):
        with pytest.raises(ValueError):
            _check_pairs(m, SolveParams(pairs=bad))
    with pytest.raises(ValueError):
        _check_pairs(m, SolveParams(orbitals=(0, 1), selfenergy=True))
    _check_pairs(m, SolveParams(orbitals=(0, 1, 2), pairs=((2, 1),), selfenergy=True))
