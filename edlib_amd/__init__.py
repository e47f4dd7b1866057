"""edlib_amd: MI355X-native spin-resolved exact diagonalisation (Python plumbing).

The engine is libedhip.so (HIP kernels for gfx950 + host C++ behind the C ABI
in include/edhip.h); the reference-compatible C++ drop-in
(HIPSpinResolvedStorage<Model> for Hamiltonian<Storage>) lives in
include/edlib_amd/.  This module wraps the C ABI with numpy for tests and the
benchmark.  It never computes anything itself: a missing library raises.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import EDH_MEM_DEVICE, EDH_MEM_HOST, EDH_OP_N, EDH_OP_SZ, EdhError, check

__all__ = ["HubbardModel", "SingleImpurityAndersonModel", "Context", "EdhError", "binom"]


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


def binom(n, k):
    from math import comb
    return comb(n, k) if 0 <= k <= n else 0


class _Model:
    kind = None

    def __init__(self):
        self._h = ctypes.c_void_p()

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib._lib is not None:
            _lib._lib.edh_model_free(self._h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    def sites(self):
        ns, ml = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().edh_model_sites(self._h, ctypes.byref(ns), ctypes.byref(ml)))
        return ns.value, ml.value

    @property
    def Ns(self):
        return self.sites()[0]

    def interacting_orbitals(self):
        return self.sites()[1]

    def transitions(self, which):
        L = _lib.lib()
        n = ctypes.c_int()
        check(L.edh_model_transitions(self._h, which, ctypes.byref(n), None, None))
        ints = np.zeros((n.value, 7), dtype=np.int32)
        vals = np.zeros(n.value)
        if n.value:
            check(L.edh_model_transitions(self._h, which, ctypes.byref(n),
                                          ints.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(vals)))
        return ints, vals

    def diagonal(self, state):
        out = ctypes.c_double()
        check(_lib.lib().edh_model_diagonal(self._h, int(state), ctypes.byref(out)))
        return out.value

    def sector_csr(self, nup, ndn, which):
        L = _lib.lib()
        nr, nnz = ctypes.c_int64(), ctypes.c_int64()
        check(L.edh_sector_csr(self._h, nup, ndn, which, ctypes.byref(nr), ctypes.byref(nnz), None, None, None))
        rp = np.zeros(nr.value + 1, dtype=np.int32)
        col = np.zeros(max(nnz.value, 1), dtype=np.int32)
        val = np.zeros(max(nnz.value, 1))
        ip = ctypes.POINTER(ctypes.c_int)
        check(L.edh_sector_csr(self._h, nup, ndn, which, ctypes.byref(nr), ctypes.byref(nnz),
                               rp.ctypes.data_as(ip), col.ctypes.data_as(ip), _d(val)))
        return rp, col[:nnz.value], val[:nnz.value]

    def sector_diagonal(self, nup, ndn):
        Ns = self.Ns
        out = np.zeros(binom(Ns, nup) * binom(Ns, ndn))
        check(_lib.lib().edh_sector_diagonal(self._h, nup, ndn, _d(out)))
        return out


class HubbardModel(_Model):
    """HubbardModel<double> (HubbardModel.h:48-77) from input.h5-schema arrays."""
    kind = "hubbard"

    def __init__(self, Ns, t, U, xmu, Hmag=None, J=None):
        super().__init__()
        self._keep = [_f64(t, (Ns, Ns)), _f64(U, (Ns,)), _f64(xmu, (Ns,)),
                      _f64(np.zeros(Ns) if Hmag is None else Hmag, (Ns,)),
                      _f64(np.zeros((Ns, Ns)) if J is None else J, (Ns, Ns))]
        t_, U_, x_, H_, J_ = self._keep
        check(_lib.lib().edh_model_hubbard(ctypes.byref(self._h), Ns, _d(t_), _d(U_), _d(J_), _d(x_), _d(H_)))

    def bare_gf(self, z):
        """HubbardModel::bare_greens_function (HubbardModel.h:181-201): per
        frequency point and spin, G0 = (z + mu_I delta_IJ + t_IJ)^-1 (the
        model's on-site energies are zero); returns (len(z), Ns*Ns, 2) with
        entry I*Ns + J."""
        z = np.atleast_1d(np.asarray(z, dtype=complex))
        t_, x_ = self._keep[0], self._keep[2]
        Ns = t_.shape[0]
        m = t_[None, :, :] + np.einsum("w,ij->wij", z, np.eye(Ns)) + np.diag(x_)[None, :, :]
        g = np.linalg.inv(m).reshape(len(z), Ns * Ns)
        return np.repeat(g[:, :, None], 2, axis=2)


class SingleImpurityAndersonModel(_Model):
    """SingleImpurityAndersonModel<double> (SingleImpurityAndersonModel.h:195-275)."""
    kind = "siam"

    def __init__(self, Ns, ml, Epsk, Vk, H0, U, mu):
        super().__init__()
        nb = Ns - ml
        self._keep = [_f64(Epsk, (nb, 2)) if nb else np.zeros((0, 2)), _f64(Vk, (ml, nb, 2)), _f64(H0, (ml, ml, 2)),
                      _f64(U, (2, 2, ml, ml, ml, ml))]
        e, v, h, u = self._keep
        ep = _d(e) if nb else None
        check(_lib.lib().edh_model_siam(ctypes.byref(self._h), Ns, ml, ep, _d(v), _d(h), _d(u), float(mu)))

    def bare_gf(self, z):
        """SingleImpurityAndersonModel::bare_greens_function
        (SingleImpurityAndersonModel.h:378-391): per impurity orbital and
        spin, G0 = 1 / (z - H0_ii - sum_k V_ik^2 / (z - eps_k)) (the reference
        leaves mu out, so at U = 0 the self-energy is -mu); returned in the
        non-local layout (len(z), ml*ml, 2), diagonal entries i*ml + i set."""
        z = np.atleast_1d(np.asarray(z, dtype=complex))
        e, v, h = self._keep[0], self._keep[1], self._keep[2]
        ml = h.shape[0]
        g = np.zeros((len(z), ml * ml, 2), dtype=complex)
        for s in range(2):
            for i in range(ml):
                delta = np.zeros(len(z), dtype=complex)
                for k in range(e.shape[0]):
                    delta += v[i, k, s] * v[i, k, s] / (z - e[k, s])
                g[:, i * ml + i, s] = 1.0 / (z - h[i, i, s] - delta)
        return g


class DeviceBuffer:
    """Device memory owned by a Context (freed with it or explicitly)."""

    def __init__(self, ctx, n):
        self.ctx = ctx
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        check(_lib.lib().edh_alloc(ctx.handle, max(self.n, 1) * 8, ctypes.byref(self.ptr)))

    def upload(self, a):
        a = _f64(a).reshape(-1)
        assert a.size == self.n
        check(_lib.lib().edh_memcpy(self.ctx.handle, self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.size * 8, 0))

    def download(self):
        out = np.zeros(self.n)
        check(_lib.lib().edh_memcpy(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.n * 8, 1))
        return out

    def free(self):
        if self.ptr.value:
            check(_lib.lib().edh_free(self.ctx.handle, self.ptr))
            self.ptr = ctypes.c_void_p()


class Context:
    """One edh_ctx: a GPU, its stream, the model and the current sector."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        check(_lib.lib().edh_create(ctypes.byref(self._h), device))
        self.model = None
        self.sector = None
        self.n = 0

    def close(self):
        if self._h.value:
            check(_lib.lib().edh_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            if getattr(self, "_h", None) and self._h.value and _lib._lib is not None:
                _lib._lib.edh_destroy(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_model(self, model):
        check(_lib.lib().edh_set_model(self._h, model.handle))
        self.model = model
        self.sector = None

    def set_sector(self, nup, ndn):
        dim = ctypes.c_int64()
        check(_lib.lib().edh_set_sector(self._h, nup, ndn, ctypes.byref(dim)))
        self.sector = (nup, ndn)
        self.n = dim.value
        return self.n

    def sector_info(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(_lib.lib().edh_sector_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    @property
    def hv_variant(self):
        v = ctypes.c_int()
        check(_lib.lib().edh_get_hv_variant(self._h, ctypes.byref(v)))
        return v.value

    @hv_variant.setter
    def hv_variant(self, variant):
        check(_lib.lib().edh_set_hv_variant(self._h, int(variant)))

    def buffer(self, n):
        return DeviceBuffer(self, n)

    def synchronize(self):
        check(_lib.lib().edh_synchronize(self._h))

    # -- hot path ----------------------------------------------------------
    def av(self, v, w=None, clear=True):
        """w = H v + (clear ? 0 : w) on host numpy arrays (SpinResolvedStorage::av)."""
        v = _f64(v).reshape(-1)
        w = np.zeros(self.n) if w is None else np.array(w, dtype=np.float64, copy=True).reshape(-1)
        check(_lib.lib().edh_av(self._h, v.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p),
                                self.n, 1 if clear else 0, EDH_MEM_HOST))
        return w

    def static_observables(self, v, norb):
        """StaticObservables' per-eigenvector sums (StaticObservables.h:356-419)
        for a vector of the current sector: host numpy array or a device
        buffer.  Returns N, N_up, N_dn, M, D_occ, "M_i M_j", N_eff."""
        out = np.zeros(5 * norb + norb * norb + 1)
        if isinstance(v, DeviceBuffer):
            ptr, mem = v.ptr, EDH_MEM_DEVICE
        else:
            v = _f64(v).reshape(-1)
            ptr, mem = v.ctypes.data_as(ctypes.c_void_p), EDH_MEM_HOST
        check(_lib.lib().edh_static_observables(self._h, ptr, self.n, mem, norb,
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        k = norb
        return {"N": out[0:k], "N_up": out[k:2 * k], "N_dn": out[2 * k:3 * k], "M": out[3 * k:4 * k],
                "D_occ": out[4 * k:5 * k], "M_i M_j": out[5 * k:5 * k + k * k], "N_eff": out[5 * k + k * k:]}

    def av_device(self, vbuf, wbuf, clear=True):
        check(_lib.lib().edh_av(self._h, vbuf.ptr, wbuf.ptr, self.n, 1 if clear else 0, EDH_MEM_DEVICE))

    def av_time(self, vbuf, wbuf, clear=False, warmup=3, reps=20):
        ms = ctypes.c_double()
        check(_lib.lib().edh_av_time(self._h, vbuf.ptr, wbuf.ptr, 1 if clear else 0, warmup, reps, ctypes.byref(ms)))
        return ms.value

    def vv(self, a, b):
        a = _f64(a).reshape(-1)
        b = _f64(b).reshape(-1)
        out = ctypes.c_double()
        check(_lib.lib().edh_vv(self._h, a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
                                a.size, ctypes.byref(out), EDH_MEM_HOST))
        return out.value

    def lanczos(self, v0, nl):
        """Lanczos::lanczos (Lanczos.h:36-77): returns (nlanc, alpha[nl], beta[nl+1])."""
        v0 = _f64(v0).reshape(-1)
        a = np.zeros(nl)
        b = np.zeros(nl + 1)
        nlanc = ctypes.c_int()
        check(_lib.lib().edh_lanczos(self._h, v0.ctypes.data_as(ctypes.c_void_p), self.n, nl, _d(a), _d(b),
                                     ctypes.byref(nlanc), EDH_MEM_HOST))
        return nlanc.value, a, b

    def lanczos_time(self, v0buf, nl):
        a = np.zeros(nl)
        b = np.zeros(nl + 1)
        nlanc = ctypes.c_int()
        ms = ctypes.c_double()
        check(_lib.lib().edh_lanczos_time(self._h, v0buf.ptr, nl, _d(a), _d(b), ctypes.byref(nlanc),
                                          ctypes.byref(ms)))
        return nlanc.value, a, b, ms.value

    def adag(self, sector, orb, annihilate, vec):
        """SpinResolvedStorage::a_adag: returns (next_sector, out) or (next, None)."""
        nup, ndn = sector
        Ns = self.model.Ns
        spin = 1 if orb >= Ns else 0
        nxt = (nup + (0 if spin else (-1 if annihilate else 1)), ndn + ((-1 if annihilate else 1) if spin else 0))
        if not (0 <= nxt[0] <= Ns and 0 <= nxt[1] <= Ns):
            return nxt, None
        vec = _f64(vec).reshape(-1)
        out = np.zeros(binom(Ns, nxt[0]) * binom(Ns, nxt[1]))
        n = ctypes.c_int64()
        check(_lib.lib().edh_adag(self._h, nup, ndn, orb, 1 if annihilate else 0, vec.ctypes.data_as(ctypes.c_void_p),
                                  out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n), EDH_MEM_HOST))
        return nxt, out

    def local_op(self, sector, op, orbs, vec):
        vec = _f64(vec).reshape(-1)
        out = np.zeros_like(vec)
        o = np.ascontiguousarray(np.asarray(orbs, dtype=np.int32))
        check(_lib.lib().edh_local_op(self._h, sector[0], sector[1], op, len(o),
                                      o.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      vec.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                                      EDH_MEM_HOST))
        return out

    def cf_eval(self, z, ev, nlanc, isign, expb, alpha, beta):
        z = np.asarray(z, dtype=np.complex128).reshape(-1)
        zr, zi = _f64(z.real), _f64(z.imag)
        a, b = _f64(alpha), _f64(beta)
        outr, outi = np.zeros(len(z)), np.zeros(len(z))
        check(_lib.lib().edh_cf_eval(self._h, len(z), _d(zr), _d(zi), float(ev), int(nlanc), int(isign), float(expb),
                                     _d(a), _d(b), _d(outr), _d(outi)))
        return outr + 1j * outi

    def eigsh(self, nev, ncv=None, tol=1e-14, maxitr=1000, seed=20261202, vectors=True):
        """Storage::diag replacement: lowest nev eigenpairs of the current sector."""
        ncv = 2 * nev + 3 if ncv is None else ncv
        ncv_eff = min(ncv, self.n)
        nev_eff = min(nev, max(ncv_eff - 1, 1)) if self.n > 1 else 1
        ev = np.zeros(max(nev_eff, 1))
        vecs = np.zeros((nev_eff, self.n)) if vectors else None
        nconv, nav = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().edh_eigsh(self._h, nev, ncv, tol, maxitr, seed, _d(ev), _d(vecs) if vectors else None,
                                   ctypes.byref(nconv), ctypes.byref(nav)))
        k = nconv.value
        return ev[:k], (vecs[:k].T.copy() if vectors else None), nav.value
