"""C-ABI checks that need no GPU: exports, error channel, and the host-side
tables (transitions, per-spin CSR, decomposed diagonal) against the oracle."""
import ctypes

import numpy as np
import pytest

import edlib_amd as ea
from edlib_amd import _lib
from oracle import edoracle as eo
from tests import models


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    syms = _lib.header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s


def test_error_channel_without_device_or_bad_args():
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.edh_create(ctypes.byref(h), -1)
    assert rc < 0 and L.edh_last_error()
    with pytest.raises(ea.EdhError):
        ea.SingleImpurityAndersonModel(2, 3, np.zeros((0, 2)), np.zeros((3, 0, 2)), np.zeros((3, 3, 2)),
                                       np.zeros((2, 2, 3, 3, 3, 3)), 0.0)


def _both(kind, p):
    return models.oracle_model(kind, p), models.product_model(kind, p)


CASES = []
_rng = np.random.default_rng(7)
CASES.append(("hubbard", models.random_hubbard_params(5, _rng)))
CASES.append(("hubbard", dict(Ns=8, t=models.ring_hopping(8), U=np.full(8, 4.0), xmu=np.full(8, 2.0))))
CASES.append(("hubbard", dict(Ns=12, t=models.torus_hopping(3, 4, 1.0, -0.3), U=np.full(12, 8.0),
                              xmu=np.full(12, 4.0))))
CASES.append(("siam", models.siam_params(6, 3, np.random.default_rng(3), h0_offdiag=0.2)))
CASES.append(("siam", models.siam_params(8, 1, np.random.default_rng(4))))


@pytest.mark.parametrize("kind,p", CASES)
def test_transitions_match_oracle(kind, p):
    om, pm = _both(kind, p)
    for which in (0, 1):
        oi, ov = om.transitions(which)
        pi, pv = pm.transitions(which)
        assert np.array_equal(oi, pi)
        assert np.array_equal(ov, pv)


@pytest.mark.parametrize("kind,p", CASES)
def test_sector_tables_bit_exact(kind, p):
    om, pm = _both(kind, p)
    Ns = p["Ns"]
    secs = [(Ns // 2, Ns // 2), (Ns // 2 + 1, Ns // 2 - 1), (1, Ns - 1), (0, Ns), (Ns, 0)]
    for (nu, nd) in secs:
        srs = eo.SRS(om, nu, nd)
        for which in (0, 1):
            orp, oc, ov = srs.csr(which)
            prp, pc, pv = pm.sector_csr(nu, nd, which)
            assert np.array_equal(orp, prp)
            assert np.array_equal(oc, pc)
            assert np.array_equal(ov, pv)  # bit-exact values (same merge order)


@pytest.mark.parametrize("kind,p", CASES)
def test_decomposed_diagonal(kind, p):
    om, pm = _both(kind, p)
    Ns = p["Ns"]
    for (nu, nd) in [(Ns // 2, Ns // 2), (1, 2), (Ns, Ns - 1)]:
        srs = eo.SRS(om, nu, nd)
        ref = srs.diagonal()
        got = pm.sector_diagonal(nu, nd)
        assert np.max(np.abs(ref - got)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("kind,p", CASES[:1] + CASES[3:4])
def test_reference_order_diagonal_is_bit_exact(kind, p):
    om, pm = _both(kind, p)
    rng = np.random.default_rng(0)
    for st in rng.integers(0, 1 << (2 * p["Ns"]), 200):
        assert om.diagonal(int(st)) == pm.diagonal(int(st))
