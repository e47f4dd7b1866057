"""Impurity solves: eigenpairs of every sector and the local Green's function.

Both drivers follow Hamiltonian::diag (Hamiltonian.h:54-101) and
GreensFunction::compute (GreensFunction.h:109-159, 241-282, 389-462) for the
local Green's function of the listed orbitals:

* every (N_up, N_dn) sector: lowest NEV eigenpairs (ncv = NCV or 2*NEV+3,
  tol, maxitr, which = "SA"; Storage.h:43-146);
* the eigenpairs ordered by energy, ties in sector order (EigenPair.h:35-41);
  E0 is the first, Z = sum_k exp(-beta (E_k - E0)) over every stored pair;
* for every pair whose Boltzmann factor passes the cutoff
  (GreensFunction.h:135), every listed orbital o and spin s:
  c+_{o s}|psi> contributes to G_g with isign +1 and c_{o s}|psi> to G_l
  with isign -1; each start vector is normalised, its squared norm is the
  weight, and a Lanczos run of NLANC steps feeds the continued fraction at
  z = i w_n + isign * E (Lanczos.h:82-100, 146-148, 201-224);
* G = (G_g + G_l) / Z on the fermionic Matsubara mesh w_n = (2n+1) pi / beta
  (or at w + i pi/beta on the real axis);
* for each listed orbital pair (i, j), i != j (GreensFunction.h:65-102,
  290-327, 336-359): start vectors (c+_i + c+_j)|psi> and (c_i + c_j)|psi>,
  whose fractions give G_ij = ((G_{i+j} - G_ii) - G_jj) / 2 per part, the
  diagonal entries copied from G, both parts summed; the members of every
  pair join the orbital list.  With `selfenergy`, the model's bare G0 and
  Sigma = G0^-1 - G_ij^-1 per frequency and spin (GreensFunction.h:202-223,
  FermionicModel.h:84-108) over the interacting orbitals.

``solve_plain`` runs one sector, then one start vector, at a time on a
Context -- what the drop-in storage does inside Hamiltonian and
GreensFunction.  ``solve_batched`` advances every sector's eigensolver run
together on a Batch, then every start vector's Lanczos run together.

Sector order: sectors are listed by (dimension, N_up, N_dn).  The reference's
SzSymmetry order (a std::sort with a non-strict comparator) only decides
which of several exactly degenerate pairs comes first, which changes no
result beyond rounding.
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import Context, binom
from .batch import JOB_DONE, Batch


@dataclass
class SolveParams:
    nev: int = 4
    ncv: int = 0            # 0: 2 * nev + 3
    tol: float = 1e-14
    maxitr: int = 1000
    beta: float = 10.0
    nomega: int = 32
    nlanc: int = 100
    cutoff: float = 1e-12
    orbitals: tuple = (0,)
    seed: int = 20261202
    sectors: list = field(default_factory=list)  # empty: every (N_up, N_dn)
    mesh: str = "matsubara"  # or "real": nomega points on [emin, emax], z = w + i pi/beta
    emin: float = -4.0
    emax: float = 4.0
    pairs: tuple = ()        # orbital pairs (i, j), i != j, of the non-local G_ij
    selfenergy: bool = False  # G0 and Sigma from G_ij (every entry of G_ij listed)

    @property
    def gf_orbitals(self):
        """The orbitals of the local G: the listed ones, then the members of
        the pairs that are not listed (GreensFunction.h:84-100)."""
        out = list(self.orbitals)
        for pr in self.pairs:
            for o in pr:
                if o not in out:
                    out.append(o)
        return out

    @property
    def ncv_eff(self):
        return self.ncv if self.ncv > 0 else 2 * self.nev + 3


def _interacting(model):
    """Interacting orbitals: the impurity orbitals of an Anderson model, every
    site of a Hubbard model."""
    ml = model.interacting_orbitals()
    return ml if ml > 0 else model.Ns


class _Observables:
    """StaticObservables::calculate_static_observables (StaticObservables.h:
    106-158): the per-eigenvector sums (from the device) averaged with the
    Boltzmann factors of the pairs that pass the cutoff, E included."""

    def __init__(self):
        self.sum = None
        self.z = 0.0

    def add(self, obs, e, bf):
        c = dict(obs)
        c["E"] = np.array([e])
        if self.sum is None:
            self.sum = {k: np.zeros_like(v) for k, v in c.items()}
        for k, v in c.items():
            self.sum[k] += v * bf
        self.z += bf

    def result(self):
        return {k: v / self.z for k, v in self.sum.items()} if self.sum else {}


def sector_list(Ns, p):
    secs = list(p.sectors) if p.sectors else [(i, j) for i in range(Ns + 1) for j in range(Ns + 1)]
    return sorted(secs, key=lambda s: (binom(Ns, s[0]) * binom(Ns, s[1]), s[0], s[1]))


def matsubara(beta, nomega):
    return (2 * np.arange(nomega) + 1) * math.pi / beta


def freq_points(p):
    """The Green's function's frequency points (MeshFactory.h:19-34,
    CommonUtils.h:14-20): i w_n = i (2n+1) pi / beta on the Matsubara mesh;
    w + i pi/beta on the real axis, w on the linear grid [emin, emax] of
    nomega points."""
    if p.mesh == "matsubara":
        return 1j * matsubara(p.beta, p.nomega)
    if p.mesh == "real":
        w = np.linspace(p.emin, p.emax, p.nomega) if p.nomega > 1 else np.array([p.emin])
        return w + 1j * math.pi / p.beta
    raise ValueError("mesh must be 'matsubara' or 'real'")


def _expb(beta, e, e0):
    return 0.0 if beta * (e - e0) > 25 else math.exp(-beta * (e - e0))


def _passes(bf, cutoff):
    return not (abs(cutoff - bf) > np.finfo(float).eps and bf < cutoff)


def _order_pairs(pairs):
    # stable sort by energy: ties keep the insertion (sector, index) order
    return sorted(pairs, key=lambda q: q[0])


def _gf_targets(Ns, sector, p):
    """(orbitals, spin, annihilate, next sector) of every start vector of one
    eigenpair: one orbital (o,) per entry of the local G, then the two of
    each non-local pair (i, j)."""
    out = []
    for ops in [(o,) for o in p.gf_orbitals] + [tuple(int(x) for x in pr) for pr in p.pairs]:
        for s in (0, 1):
            for ann in (False, True):
                n = sector[s]
                if (ann and n == 0) or ((not ann) and n >= Ns):
                    continue
                d = -1 if ann else 1
                nxt = (sector[0] + d, sector[1]) if s == 0 else (sector[0], sector[1] + d)
                out.append((ops, s, ann, nxt))
    return out


def _check_pairs(model, p):
    ml = _interacting(model)
    for pr in p.pairs:
        if len(pr) != 2 or pr[0] == pr[1] or not all(0 <= o < ml for o in pr):
            raise ValueError("orbital pairs are two different interacting orbitals: %r" % (pr,))
    if p.selfenergy and sorted(set(p.gf_orbitals)) != list(range(ml)):
        raise ValueError("the self-energy needs the local G of every interacting orbital (0..%d)" % (ml - 1))


def _finish_gf(model, p, zpts, Gg, Gl, Ggij, Glij, Z):
    """Normalisation, non_local_gf (GreensFunction.h:336-359) and, on
    request, the self-energy; returns the result entries."""
    Gg, Gl = Gg / Z, Gl / Z
    out = {"G": Gg + Gl}
    if not p.pairs and not p.selfenergy:
        return out
    ml = _interacting(model)
    orbs = p.gf_orbitals
    Ggij, Glij = Ggij / Z, Glij / Z
    for i, j in p.pairs:
        k = ml * i + j
        for o in (i, j):
            Ggij[:, k, :] -= Gg[:, orbs.index(o), :]
            Glij[:, k, :] -= Gl[:, orbs.index(o), :]
        Ggij[:, k, :] *= 0.5
        Glij[:, k, :] *= 0.5
    for q, o in enumerate(orbs):
        Ggij[:, ml * o + o, :] = Gg[:, q, :]
        Glij[:, ml * o + o, :] = Gl[:, q, :]
    Gij = Ggij + Glij
    out["G_ij"] = Gij
    if p.selfenergy:
        G0 = model.bare_gf(zpts)
        a = np.moveaxis(G0.reshape(len(zpts), ml, ml, 2), 3, 1)
        b = np.moveaxis(Gij.reshape(len(zpts), ml, ml, 2), 3, 1)
        sig = np.linalg.inv(a) - np.linalg.inv(b)
        out["G0"] = G0
        out["Sigma"] = np.moveaxis(sig, 1, 3).reshape(len(zpts), ml * ml, 2)
    return out


def solve_plain(model, p, device=0, ctx=None):
    """One sector and one start vector at a time (the drop-in path)."""
    Ns = model.Ns
    _check_pairs(model, p)
    own = ctx is None
    ctx = Context(device) if own else ctx
    t0 = time.perf_counter()
    ctx.set_model(model)
    secs = sector_list(Ns, p)
    pairs = []
    n_av = 0
    for i, sec in enumerate(secs):
        if binom(Ns, sec[0]) * binom(Ns, sec[1]) == 0:
            continue
        ctx.set_sector(*sec)
        ev, vecs, nav = ctx.eigsh(p.nev, p.ncv_eff, p.tol, p.maxitr, p.seed + i, vectors=True)
        n_av += nav
        for k in range(len(ev)):
            pairs.append((float(ev[k]), sec, vecs[:, k].copy()))
    t_eig = time.perf_counter() - t0
    pairs = _order_pairs(pairs)
    e0 = pairs[0][0]
    Z = sum(math.exp(-(e - e0) * p.beta) for e, _, _ in pairs)
    zpts = freq_points(p)
    orbs = p.gf_orbitals
    ml = _interacting(model)
    Gg = np.zeros((p.nomega, len(orbs), 2), dtype=complex)
    Gl = np.zeros_like(Gg)
    Ggij = np.zeros((p.nomega, ml * ml, 2), dtype=complex)
    Glij = np.zeros_like(Ggij)
    wsum = np.zeros((len(orbs), 2))
    n_gf = 0
    obs = _Observables()
    for e, sec, vec in pairs:
        bf = math.exp(-(e - e0) * p.beta)
        if not _passes(bf, p.cutoff):
            continue
        ctx.set_sector(*sec)
        obs.add(ctx.static_observables(vec, ml), e, bf)
        for ops, s, ann, nxt in _gf_targets(Ns, sec, p):
            out = None
            for o in ops:  # the pair's vectors summed in orbital order
                _, x = ctx.adag(sec, o + s * Ns, ann, vec)
                out = x if out is None else x + out
            norm = ctx.vv(out, out)
            if len(ops) == 1:
                wsum[orbs.index(ops[0]), s] += bf * norm
            if not abs(norm) > 1e-10:
                continue
            out /= math.sqrt(norm)
            ctx.set_sector(*nxt)
            nlanc, a, b = ctx.lanczos(out, p.nlanc)
            isign = -1 if ann else 1
            g = ctx.cf_eval(zpts + e * isign, norm, nlanc, isign, _expb(p.beta, e, e0), a, b)
            if len(ops) == 1:
                (Gl if ann else Gg)[:, orbs.index(ops[0]), s] += g
            else:
                (Glij if ann else Ggij)[:, ml * ops[0] + ops[1], s] += g
            n_gf += 1
    res = _finish_gf(model, p, zpts, Gg, Gl, Ggij, Glij, Z)
    t_all = time.perf_counter() - t0
    if own:
        ctx.close()
    res.update({"E0": e0, "Z": Z, "energies": np.array([q[0] for q in pairs]), "n_gf": n_gf, "n_av": n_av,
                "observables": obs.result(), "t_eig": t_eig, "t_total": t_all, "weights": wsum / Z})
    return res


class BatchedSolve:
    """One impurity solve on a shared Batch, driven phase by phase so that
    several solves can share the batch's sweeps (see serve.py).

    Green's-function runs start while slower sectors are still converging:
    once a sector's eigenpairs are known, every pair within the Boltzmann
    window of the current bound on the ground energy gets its start vectors.
    The bound is the lowest of the finished sectors' eigenvalues and of the
    running sectors' lowest Ritz values (each an upper bound of its sector's
    lowest eigenvalue), so it can only fall towards the ground energy: every
    pair that passes the final cutoff has been started, and few pairs are
    started that later fall out of the window (their runs are cancelled).
    Contributions are summed in the pair order of the final eigenpair list,
    so the result does not depend on when sectors finished.

    Green's-function runs hold three vectors of their target sector each
    while they run; they are started in window order as long as the runs in
    flight stay within `gf_budget` bytes (by default the allowance the
    server's footprint estimate makes for them), the rest wait for runs to
    finish.  A run's arithmetic does not depend on when it starts.
    """

    def __init__(self, batch, model, p, gf_budget=None):
        self.b = batch
        self.model = model
        self.p = p
        self.Ns = model.Ns
        self.secs = [s for s in sector_list(self.Ns, p) if binom(self.Ns, s[0]) * binom(self.Ns, s[1]) > 0]
        self.sid = {}
        self.eig_jobs = []
        self.known = []      # (E, sector index, sector, eig job, k) of finished sectors
        self.seen = set()    # finished eig jobs
        self.gf = {}         # (eig job, k, orbitals, spin, annihilate) -> gf job or -1
        self.gf_queue = []   # (key, request, bytes) waiting for memory
        self.gf_queued = set()
        self.gf_live = {}    # running gf job -> bytes
        big = max(binom(self.Ns, a) * binom(self.Ns, b) for a, b in self.secs) if self.secs else 0
        nops = len(p.gf_orbitals) + len(p.pairs)
        self.gf_budget = gf_budget if gf_budget is not None else 3 * 8 * big * (4 * nops * p.nev * 4)
        self.phase = "new"
        self.result = None
        c = p.cutoff
        self.window = math.inf if c <= 0 else (-math.log(c) / p.beta) * (1 + 1e-9) + 1e-9

    def _request(self, key, nxt):
        j, k, ops, s, ann = key
        req = (j, k, ops[0] + s * self.Ns, 1 if ann else 0, self.sid[nxt])
        if len(ops) > 1:
            req += (ops[1] + s * self.Ns,)
        nbytes = 3 * 8 * binom(self.Ns, nxt[0]) * binom(self.Ns, nxt[1]) + 4096
        return (key, req, nbytes)

    def start(self):
        _check_pairs(self.model, self.p)
        ids, _ = self.b.add_sectors(self.model, self.secs)
        self.sid = dict(zip(self.secs, ids))
        for i, sec in enumerate(sector_list(self.Ns, self.p)):
            if sec in self.sid:
                self.eig_jobs.append((sec, self.b.submit_eig(self.sid[sec], self.p.nev, self.p.ncv_eff, self.p.tol,
                                                             self.p.maxitr, self.p.seed + i)))
        self.phase = "eig"

    def _collect(self):
        new = False
        pend = [(idx, sec, j) for idx, (sec, j) in enumerate(self.eig_jobs) if j not in self.seen]
        states = self.b.job_states([j for _, _, j in pend]) if pend else []
        for (idx, sec, j), st in zip(pend, states):
            if st == 0:
                continue
            self.seen.add(j)
            ev, _ = self.b.eig_result(j)
            for k in range(len(ev)):
                self.known.append((float(ev[k]), idx, sec, j, k))
            new = True
        return new

    def _submit_window(self):
        if not self.known:
            return
        e0 = min(q[0] for q in self.known)
        running = [j for _, j in self.eig_jobs if j not in self.seen]
        if running:
            bounds = self.b.eig_bounds(running)
            bounds = bounds[~np.isnan(bounds)]
            if bounds.size:
                e0 = min(e0, float(bounds.min()))
        for e, _, sec, j, k in sorted(self.known, key=lambda q: q[0]):
            if e - e0 > self.window:
                continue
            for ops, s, ann, nxt in _gf_targets(self.Ns, sec, self.p):
                key = (j, k, ops, s, ann)
                if key not in self.gf and key not in self.gf_queued:
                    self.gf_queue.append(self._request(key, nxt))
                    self.gf_queued.add(key)
        self._pump()

    def _pump(self, needed=None):
        """Starts queued Green's-function runs while the runs in flight fit
        the budget (at least one run at a time); with `needed`, queued runs
        outside that set are dropped."""
        if self.gf_live:
            ids = list(self.gf_live)
            for jid, st in zip(ids, self.b.job_states(ids)):
                if st != 0:
                    del self.gf_live[jid]
        if needed is not None:
            self.gf_queue = [q for q in self.gf_queue if q[0] in needed]
        used = sum(self.gf_live.values())
        reqs, keys, sizes = [], [], []
        while self.gf_queue:
            key, req, nbytes = self.gf_queue[0]
            if (self.gf_live or reqs) and used + nbytes > self.gf_budget:
                break
            self.gf_queue.pop(0)
            self.gf_queued.discard(key)
            reqs.append(req)
            keys.append(key)
            sizes.append(nbytes)
            used += nbytes
        if reqs:
            for key, jid, nbytes in zip(keys, self.b.submit_gf(reqs, self.p.nlanc), sizes):
                self.gf[key] = jid
                if jid >= 0:
                    self.gf_live[jid] = nbytes

    def abort(self):
        """Stops and frees every process and sector of this solve (after a
        failure), leaving the batch's other solves untouched."""
        jobs = [j for _, j in self.eig_jobs] + [j for j in self.gf.values() if j >= 0]
        for j in jobs:
            for fn in (self.b.cancel, self.b.free_job):
                try:
                    fn(j)
                except Exception:  # noqa: BLE001 -- already finished or freed
                    pass
        for sid in self.sid.values():
            try:
                self.b.remove_sector(sid)
            except Exception:  # noqa: BLE001 -- already removed
                pass
        self.eig_jobs, self.gf, self.sid = [], {}, {}
        self.phase = "failed"

    def advance(self):
        """Moves the solve on; returns True when it is complete."""
        p = self.p
        if self.phase == "eig":
            if self._collect():
                self._submit_window()
            elif self.gf_queue:
                self._pump()
            if len(self.seen) == len(self.eig_jobs):
                # final eigenpair order: energy, ties in sector order
                pairs = _order_pairs([(e, idx, sec, j, k) for e, idx, sec, j, k in
                                      sorted(self.known, key=lambda q: (q[1], q[4]))])
                self.pairs = pairs
                self.e0 = pairs[0][0]
                self.Z = sum(math.exp(-(q[0] - self.e0) * p.beta) for q in pairs)
                self.used = []
                for e, _, sec, j, k in pairs:
                    if not _passes(math.exp(-(e - self.e0) * p.beta), p.cutoff):
                        continue
                    for ops, s, ann, nxt in _gf_targets(self.Ns, sec, p):
                        self.used.append(((j, k, ops, s, ann), e))
                needed = {key for key, _ in self.used}
                for key, jid in self.gf.items():
                    if key not in needed and jid >= 0:
                        self.b.cancel(jid)
                        self.b.free_job(jid)
                        self.gf_live.pop(jid, None)
                # every needed run not started yet (possible only when the
                # memory budget held it back) joins the queue
                for e, _, sec, j, k in pairs:
                    if not _passes(math.exp(-(e - self.e0) * p.beta), p.cutoff):
                        continue
                    for ops, s_, ann, nxt in _gf_targets(self.Ns, sec, p):
                        key = (j, k, ops, s_, ann)
                        if key not in self.gf and key not in self.gf_queued:
                            self.gf_queue.append(self._request(key, nxt))
                            self.gf_queued.add(key)
                self.gf_needed = needed
                self.phase = "gf"
        if self.phase == "gf":
            self._pump(self.gf_needed)
            pending = bool(self.gf_queue) or any(key not in self.gf for key, _ in self.used)
            live = [self.gf[key] for key, _ in self.used if key in self.gf and self.gf[key] >= 0]
            if not pending and (not live or not np.any(self.b.job_states(live) == 0)):
                wn = freq_points(p) if p.mesh != "matsubara" else matsubara(p.beta, p.nomega)
                orbs = p.gf_orbitals
                ml = _interacting(self.model)
                Gg = np.zeros((p.nomega, len(orbs), 2), dtype=complex)
                Gl = np.zeros_like(Gg)
                Ggij = np.zeros((p.nomega, ml * ml, 2), dtype=complex)
                Glij = np.zeros_like(Ggij)
                n2 = dict(zip(live, self.b.lanczos_norm2(live))) if live else {}
                # a vanishing start vector (<out|out> <= 1e-10) has no run
                # (GreensFunction.h:408-418)
                rows = [(self.gf[key], key, e) for key, e in self.used
                        if self.gf[key] >= 0 and abs(n2[self.gf[key]]) > 1e-10]
                g = self.b.cf_many([r[0] for r in rows], [n2[r[0]] for r in rows],
                                   [-1 if r[1][4] else 1 for r in rows], [_expb(p.beta, r[2], self.e0) for r in rows],
                                   [r[2] for r in rows], wn)
                for q, (jid, (_, _, ops, s, ann), _) in enumerate(rows):
                    if len(ops) == 1:
                        (Gl if ann else Gg)[:, orbs.index(ops[0]), s] += g[q]
                    else:
                        (Glij if ann else Ggij)[:, ml * ops[0] + ops[1], s] += g[q]
                # spectral weight <c c+> + <c+ c> per orbital and spin
                # (1 in every eigenstate: the anticommutator)
                wsum = np.zeros((len(orbs), 2))
                for key, e in self.used:
                    jid = self.gf[key]
                    if jid >= 0 and len(key[2]) == 1:
                        wsum[orbs.index(key[2][0]), key[3]] += math.exp(-(e - self.e0) * p.beta) * n2[jid]
                # H*v applications per sector (eigensolver runs, then
                # Green's-function steps in their target sectors)
                work = {}
                jsec = {}
                for sec, j in self.eig_jobs:
                    jsec[j] = sec
                    work[sec] = work.get(sec, 0) + self.b.job_info(j)["n_av"]
                for key, jid in self.gf.items():
                    if jid < 0 or jid not in live:
                        continue
                    j, _, _, s_, ann = key
                    sec, d = jsec[j], (-1 if ann else 1)
                    tgt = (sec[0] + d, sec[1]) if s_ == 0 else (sec[0], sec[1] + d)
                    work[tgt] = work.get(tgt, 0) + self.b.job_info(jid)["count"]
                for jid in live:
                    self.b.free_job(jid)
                obs = _Observables()
                inside = [(e, j, k, math.exp(-(e - self.e0) * p.beta)) for e, _, _, j, k in self.pairs]
                inside = [q for q in inside if _passes(q[3], p.cutoff)]
                sums = self.b.static_observables_many([(j, k) for _, j, k, _ in inside], _interacting(self.model))
                for (e, _, _, bf), o in zip(inside, sums):
                    obs.add(o, e, bf)
                for _, j in self.eig_jobs:
                    self.b.free_job(j)
                for sid in self.sid.values():
                    self.b.remove_sector(sid)
                zpts = freq_points(p)
                self.result = _finish_gf(self.model, p, zpts, Gg, Gl, Ggij, Glij, self.Z)
                self.result.update({"E0": self.e0, "Z": self.Z, "energies": np.array([q[0] for q in self.pairs]),
                                    "n_gf": len(rows), "observables": obs.result(), "hv_by_sector": work,
                                    "weights": wsum / self.Z})
                self.phase = "done"
        return self.phase == "done"


def solve_batched(model, p, device=0, batch=None, autotune=True, gf_budget=None):
    """Every sector's eigensolver run advanced together, then every start
    vector (gf_budget: bytes of Green's-function runs in flight, see
    BatchedSolve)."""
    own = batch is None
    b = Batch(device) if own else batch
    t0 = time.perf_counter()
    s = BatchedSolve(b, model, p, gf_budget=gf_budget)
    s.start()
    if autotune:
        b.autotune(list(s.sid.values()))
    sweeps = 0
    while not s.advance():
        b.step(16)
        sweeps += 1
    out = dict(s.result)
    out["t_total"] = time.perf_counter() - t0
    if own:
        b.close()
    return out
