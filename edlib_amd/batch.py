"""Batched engine (edh_batch_* in include/edhip.h) for numpy callers.

A Batch holds sectors of any number of models and the Krylov processes on
them (eigensolver runs and Green's-function Lanczos runs).  ``step()``
advances every active process by one Krylov step per sweep with grouped
kernel launches; processes can be submitted between sweeps.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import EDH_MEM_HOST, check

JOB_ACTIVE, JOB_DONE, JOB_FAILED = 0, 1, 2
JOB_EIG, JOB_LANCZOS = 0, 1


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Batch:
    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        check(_lib.lib().edh_batch_create(ctypes.byref(self._h), device))
        self.sector_key = {}  # sector id -> (model key, nup, ndn)

    def close(self):
        if self._h.value:
            check(_lib.lib().edh_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            if getattr(self, "_h", None) and self._h.value and _lib._lib is not None:
                _lib._lib.edh_batch_destroy(self._h)
        except Exception:
            pass

    def add_sector(self, model, nup, ndn):
        sid, dim = ctypes.c_int(), ctypes.c_int64()
        check(_lib.lib().edh_batch_add_sector(self._h, model.handle, nup, ndn, ctypes.byref(sid), ctypes.byref(dim)))
        return sid.value, dim.value

    def add_sectors(self, model, secs):
        """Many sectors of one model at once (tables built on several host
        threads); returns (ids, dims) in the order of secs."""
        k = len(secs)
        nu = np.ascontiguousarray([s[0] for s in secs], dtype=np.int32)
        nd = np.ascontiguousarray([s[1] for s in secs], dtype=np.int32)
        ids = np.zeros(max(k, 1), dtype=np.int32)
        dims = np.zeros(max(k, 1), dtype=np.int64)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        check(_lib.lib().edh_batch_add_sectors(self._h, model.handle, k, ip(nu), ip(nd), ip(ids),
                                               dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return ids[:k].tolist(), dims[:k].tolist()

    def remove_sector(self, sid):
        check(_lib.lib().edh_batch_remove_sector(self._h, sid))

    def submit_eig(self, sid, nev, ncv, tol=1e-14, maxitr=1000, seed=20261202):
        jid = ctypes.c_int()
        check(_lib.lib().edh_batch_submit_eig(self._h, sid, nev, ncv, tol, maxitr, seed, ctypes.byref(jid)))
        return jid.value

    def submit_lanczos(self, sid, v0, nl):
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        jid = ctypes.c_int()
        check(_lib.lib().edh_batch_submit_lanczos(self._h, sid, v0.ctypes.data_as(ctypes.c_void_p), nl, EDH_MEM_HOST,
                                                  ctypes.byref(jid)))
        return jid.value

    def submit_gf(self, requests, nl):
        """Green's-function start vectors c(+)_orb |psi_index>, one per request
        (eig_job, index, orb, annihilate, dst_sid[, orb2]), and their Lanczos
        runs; with orb2 >= 0 the start vector is (c(+)_orb + c(+)_orb2)|psi>
        (a non-local pair).  Returns the job ids (-1 where the target sector
        is empty); nothing waits for the device.  A run whose start vector
        vanishes (squared norm <= 1e-10) stops at once; the squared norms are
        known once the runs have finished (lanczos_norm2)."""
        k = len(requests)
        cols = [np.ascontiguousarray([int(r[c]) for r in requests], dtype=np.int32) for c in range(5)]
        ids = np.zeros(max(k, 1), dtype=np.int32)
        ip = [c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) for c in cols]
        out = ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        if any(len(r) > 5 and r[5] >= 0 for r in requests):
            o2 = np.ascontiguousarray([int(r[5]) if len(r) > 5 else -1 for r in requests], dtype=np.int32)
            check(_lib.lib().edh_batch_submit_gf2(self._h, k, ip[0], ip[1], ip[2],
                                                  o2.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ip[3], ip[4], nl,
                                                  out))
        else:
            check(_lib.lib().edh_batch_submit_gf(self._h, k, ip[0], ip[1], ip[2], ip[3], ip[4], nl, out))
        return ids[:k].tolist()

    def lanczos_norm2(self, jids):
        """Squared start-vector norms of finished Lanczos runs (NaN while
        running)."""
        ids = np.ascontiguousarray(jids, dtype=np.int32)
        out = np.zeros(max(len(ids), 1))
        check(_lib.lib().edh_batch_lanczos_norm2(self._h, len(ids), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                 _d(out)))
        return out[:len(ids)]

    def step(self, nsweeps=1):
        act = ctypes.c_int()
        check(_lib.lib().edh_batch_step(self._h, nsweeps, ctypes.byref(act)))
        return act.value

    def run(self, chunk=64):
        """Sweeps until no process is active; returns the number of sweep calls."""
        calls = 0
        while self.step(chunk) > 0:
            calls += 1
        return calls

    def job_info(self, jid):
        st, kind, cnt, nav = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        n = ctypes.c_int64()
        check(_lib.lib().edh_batch_job_info(self._h, jid, ctypes.byref(st), ctypes.byref(kind), ctypes.byref(n),
                                            ctypes.byref(cnt), ctypes.byref(nav)))
        return {"state": st.value, "kind": kind.value, "n": n.value, "count": cnt.value, "n_av": nav.value}

    def job_states(self, jids):
        """States of many jobs in one call (array of JOB_*)."""
        ids = np.ascontiguousarray(jids, dtype=np.int32)
        st = np.zeros(max(len(ids), 1), dtype=np.int32)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        check(_lib.lib().edh_batch_job_states(self._h, len(ids), ip(ids), ip(st)))
        return st[:len(ids)]

    def eig_bounds(self, jids):
        """Upper bounds of the lowest eigenvalue of eigensolver jobs (NaN
        before a running job's first restart cycle has ended)."""
        ids = np.ascontiguousarray(jids, dtype=np.int32)
        out = np.zeros(max(len(ids), 1))
        check(_lib.lib().edh_batch_eig_bounds(self._h, len(ids), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                              _d(out)))
        return out[:len(ids)]

    def eig_result(self, jid, vectors=False):
        info = self.job_info(jid)
        nev, n = info["count"], info["n"]
        ev = np.zeros(max(nev, 1))
        vecs = np.zeros((nev, n)) if (vectors and nev > 0) else None
        nconv = ctypes.c_int()
        check(_lib.lib().edh_batch_eig_result(self._h, jid, _d(ev), _d(vecs) if vecs is not None else None,
                                              ctypes.byref(nconv)))
        return ev[:nev], (vecs.T.copy() if vecs is not None else None)

    def static_observables(self, jid, index, norb):
        """Per-eigenvector static observables of eigenpair `index` of a
        finished eigensolver job, as edlib_amd.Context.static_observables."""
        out = np.zeros(5 * norb + norb * norb + 1)
        check(_lib.lib().edh_batch_static_observables(self._h, jid, index, norb, _d(out)))
        k = norb
        return {"N": out[0:k], "N_up": out[k:2 * k], "N_dn": out[2 * k:3 * k], "M": out[3 * k:4 * k],
                "D_occ": out[4 * k:5 * k], "M_i M_j": out[5 * k:5 * k + k * k], "N_eff": out[5 * k + k * k:]}

    def static_observables_many(self, pairs, norb):
        """static_observables for many (job id, index) pairs with one wait."""
        k = len(pairs)
        if k == 0:
            return []
        sz = 5 * norb + norb * norb + 1
        out = np.zeros(k * sz)
        ids = np.ascontiguousarray([p[0] for p in pairs], dtype=np.int32)
        idx = np.ascontiguousarray([p[1] for p in pairs], dtype=np.int32)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        check(_lib.lib().edh_batch_static_observables_many(self._h, k, ip(ids), ip(idx), norb, _d(out)))
        res = []
        for q in range(k):
            o = out[q * sz:(q + 1) * sz]
            res.append({"N": o[0:norb], "N_up": o[norb:2 * norb], "N_dn": o[2 * norb:3 * norb],
                        "M": o[3 * norb:4 * norb], "D_occ": o[4 * norb:5 * norb],
                        "M_i M_j": o[5 * norb:5 * norb + norb * norb], "N_eff": o[5 * norb + norb * norb:]})
        return res

    def lanczos_result(self, jid, nl):
        a, b = np.zeros(nl), np.zeros(nl + 1)
        nlanc = ctypes.c_int()
        check(_lib.lib().edh_batch_lanczos_result(self._h, jid, _d(a), _d(b), ctypes.byref(nlanc)))
        return nlanc.value, a, b

    def cancel(self, jid):
        check(_lib.lib().edh_batch_cancel(self._h, jid))

    def free_job(self, jid):
        check(_lib.lib().edh_batch_free_job(self._h, jid))

    def set_graphs(self, on):
        """Steady sweeps replay a captured hipGraph (default on)."""
        check(_lib.lib().edh_batch_set_graphs(self._h, 1 if on else 0))

    def sweep_counts(self):
        """(sweeps, sweeps run as graph replays)."""
        a, g = ctypes.c_int64(), ctypes.c_int64()
        check(_lib.lib().edh_batch_sweeps(self._h, ctypes.byref(a), ctypes.byref(g)))
        return a.value, g.value

    def memory(self):
        b = ctypes.c_size_t()
        check(_lib.lib().edh_batch_memory(self._h, ctypes.byref(b)))
        return b.value

    def autotune(self, sids):
        ids = np.ascontiguousarray(sids, dtype=np.int32)
        blk = ctypes.c_int()
        check(_lib.lib().edh_batch_autotune(self._h, len(ids), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            ctypes.byref(blk)))
        return blk.value

    def time_hv(self, sids, reps=20, seed=1):
        """Grouped H*v of an eigensolver sweep over one operand per sector
        (w = H v, seeded uniform v): (mean ms per launch, algorithmic bytes
        per launch = sum of 16 n + table bytes, elements)."""
        ids = np.ascontiguousarray(sids, dtype=np.int32)
        ms = ctypes.c_double()
        nb, ne = ctypes.c_int64(), ctypes.c_int64()
        check(_lib.lib().edh_batch_time_hv(self._h, len(ids), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           reps, seed, ctypes.byref(ms), ctypes.byref(nb), ctypes.byref(ne)))
        return ms.value, nb.value, ne.value

    def cf_many(self, job_ids, ev, isign, expb, excited, wn):
        """Continued fractions of finished Lanczos runs at z = p + excited*isign
        for each frequency point p: real `wn` are Matsubara frequencies
        (p = i wn), complex ones are the points themselves (the real axis's
        w + i pi/beta); returns a (len(job_ids), len(wn)) complex array."""
        k, nz = len(job_ids), len(wn)
        if k == 0:
            return np.zeros((0, nz), dtype=complex)
        ids = np.ascontiguousarray(job_ids, dtype=np.int32)
        isg = np.ascontiguousarray(isign, dtype=np.int32)
        evs, exb, exc = (np.ascontiguousarray(x, dtype=np.float64) for x in (ev, expb, excited))
        re, im = np.zeros((k, nz)), np.zeros((k, nz))
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        if np.iscomplexobj(wn):
            zr = np.ascontiguousarray(np.real(wn), dtype=np.float64)
            zi = np.ascontiguousarray(np.imag(wn), dtype=np.float64)
            check(_lib.lib().edh_batch_cf_many_z(self._h, k, ip(ids), _d(evs), ip(isg), _d(exb), _d(exc), nz, _d(zr),
                                                 _d(zi), _d(re), _d(im)))
        else:
            w = np.ascontiguousarray(wn, dtype=np.float64)
            check(_lib.lib().edh_batch_cf_many(self._h, k, ip(ids), _d(evs), ip(isg), _d(exb), _d(exc), nz, _d(w),
                                               _d(re), _d(im)))
        return re + 1j * im

    def cf_eval(self, z, ev, nlanc, isign, expb, alpha, beta):
        z = np.asarray(z, dtype=np.complex128).reshape(-1)
        zr, zi = np.ascontiguousarray(z.real), np.ascontiguousarray(z.imag)
        a, b = np.ascontiguousarray(alpha, dtype=np.float64), np.ascontiguousarray(beta, dtype=np.float64)
        outr, outi = np.zeros(len(z)), np.zeros(len(z))
        check(_lib.lib().edh_batch_cf_eval(self._h, len(z), _d(zr), _d(zi), float(ev), int(nlanc), int(isign),
                                           float(expb), _d(a), _d(b), _d(outr), _d(outi)))
        return outr + 1j * outi
