"""CPU oracle for the spin-resolved ED hot path -- TEST INFRASTRUCTURE ONLY.

Restates the reference EDLib algorithms on the CPU: the integer/byte-heavy
parts (basis, model terms, per-spin tables, SpinResolvedStorage::fill/av,
a_adag, SOCRS baseline, Lanczos recurrence, continued fraction) live in
``oracle.cpp`` (built into ``oracle/_build/liboracle.so``); the drivers
(Storage::diag via scipy's ARPACK, Hamiltonian's sector loop, GreensFunction,
ChiLoc, StaticObservables) are restated here with numpy.  A dense brute force
(Lehmann sums over the full Fock space) pins the drivers independently.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module.  The product (``edlib_amd``) never does.

Pinning status (checked in tests/test_oracle_*.py):
  * basis / rank / sector order: closed-form rank, round trips
    (test/SzSymmetry_Test.cpp:27-39, test/NSymmetry_Test.cpp:9-21), libstdc++ sort;
  * model terms, av: dense brute-force Hamiltonian (numpy);
  * eigenvalues: dense eigh and 4-ring E0 (test/HubbardModel_Test.cpp:54);
  * G / chi: gom1.dat, xiats.dat, xiatd.dat (test/Lanczos_Test.cpp:115-119);
  * static observables: test/StaticObservables_Test.cpp:63-72.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile oracle.cpp with the committed Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        L.orc_cnk.restype = ctypes.c_longlong
        L.orc_basis.restype = ctypes.c_int
        L.orc_nsym_index.restype = ctypes.c_int
        L.orc_nsym_index.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_longlong]
        L.orc_sector_order.restype = ctypes.c_int
        L.orc_model_hubbard.restype = vp
        L.orc_model_hubbard.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_model_siam.restype = vp
        L.orc_model_siam.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double]
        L.orc_model_free.argtypes = [vp]
        L.orc_model_ntrans.argtypes = [vp, ctypes.c_int]
        L.orc_model_trans.argtypes = [vp, ctypes.c_int, _ip, _dp]
        L.orc_model_diagonal.restype = ctypes.c_double
        L.orc_model_diagonal.argtypes = [vp, ctypes.c_longlong]
        L.orc_model_apply.restype = ctypes.c_int
        L.orc_model_apply.argtypes = [vp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong), _dp, ctypes.c_int]
        L.orc_srs_fill.restype = vp
        L.orc_srs_fill.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_srs_free.argtypes = [vp]
        L.orc_srs_dim.restype = ctypes.c_longlong
        L.orc_srs_dim.argtypes = [vp]
        L.orc_srs_av.argtypes = [vp, _dp, _dp, ctypes.c_int]
        L.orc_srs_av_omp.argtypes = [vp, _dp, _dp, ctypes.c_int]
        L.orc_srs_diag.argtypes = [vp, _dp]
        L.orc_srs_csr.restype = ctypes.c_longlong
        L.orc_srs_csr.argtypes = [vp, ctypes.c_int, _ip, _ip, _dp]
        L.orc_srs_int_start.restype = ctypes.c_longlong
        L.orc_srs_int_start.argtypes = [vp]
        L.orc_socrs_fill.restype = vp
        L.orc_socrs_fill.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_socrs_free.argtypes = [vp]
        L.orc_socrs_av.argtypes = [vp, _dp, _dp, ctypes.c_int]
        L.orc_adag.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, _dp, _dp]
        L.orc_occupations.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(ctypes.c_int8)]
        L.orc_lanczos.restype = ctypes.c_int
        L.orc_lanczos.argtypes = [vp, _dp, ctypes.c_int, _dp, _dp]
        L.orc_lanczos_omp.restype = ctypes.c_int
        L.orc_lanczos_omp.argtypes = [vp, _dp, ctypes.c_int, _dp, _dp]
        L.orc_frac_points.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_eigenpair_order.argtypes = [ctypes.c_int, _dp, _ip]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


# ---------------------------------------------------------------------------
# Basis (Combination.h, SzSymmetry.h, NSymmetry.h)
# ---------------------------------------------------------------------------
def cnk(n, k):
    """Binomial as the reference's c_n_k for 0 <= k <= n (Combination.h:23-27)."""
    return int(lib().orc_cnk(n, k))


def basis(Ns, k):
    """Per-spin basis bit patterns in the reference order (SzSymmetry.h:191-202)."""
    out = np.zeros(cnk(Ns, k), dtype=np.int32)
    lib().orc_basis(Ns, k, _i(out))
    return out


def nsym_index(Ns, nel, bits):
    """NSymmetry::index (NSymmetry.h:94-96) on a per-spin bit pattern."""
    return int(lib().orc_nsym_index(Ns, nel, int(bits)))


def sector_order(Ns, sector_list=None):
    """SzSymmetry sector list after std::sort (SzSymmetry.h:72-98)."""
    n = (Ns + 1) ** 2 if sector_list is None else len(sector_list)
    nu = np.zeros(n, dtype=np.int32)
    nd = np.zeros(n, dtype=np.int32)
    if sector_list is None:
        lib().orc_sector_order(Ns, 0, None, _i(nu), _i(nd))
    else:
        flat = np.ascontiguousarray(np.asarray(sector_list, dtype=np.int32).reshape(-1))
        lib().orc_sector_order(Ns, n, _i(flat), _i(nu), _i(nd))
    return [(int(a), int(b)) for a, b in zip(nu, nd)]


def eigenpair_order(evals):
    """Iteration order of std::set<EigenPair> (EigenPair.h:39-41)."""
    e = np.ascontiguousarray(evals, dtype=np.float64)
    ids = np.zeros(len(e), dtype=np.int32)
    lib().orc_eigenpair_order(len(e), _d(e), _i(ids))
    return [int(x) for x in ids]


# ---------------------------------------------------------------------------
# Models (HubbardModel.h, SingleImpurityAndersonModel.h)
# ---------------------------------------------------------------------------
class Model:
    """Holds an oracle model handle plus the parameters it was built from."""

    def __init__(self, kind, Ns, handle, params):
        self.kind = kind
        self.Ns = Ns
        self._h = handle
        self.params = params
        self.ml = params.get("ml", Ns)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_model_free(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def interacting_orbitals(self):
        return self.Ns if self.kind == "hubbard" else self.ml

    def transitions(self, which):
        n = lib().orc_model_ntrans(self._h, which)
        ints = np.zeros((n, 7), dtype=np.int32)
        vals = np.zeros(n, dtype=np.float64)
        if n:
            lib().orc_model_trans(self._h, which, _i(ints), _d(vals))
        return ints, vals

    def diagonal(self, state):
        return lib().orc_model_diagonal(self._h, int(state))

    def apply(self, state, cap=4096):
        t = (ctypes.c_longlong * cap)()
        c = np.zeros(cap)
        n = lib().orc_model_apply(self._h, int(state), t, _d(c), cap)
        return [int(t[q]) for q in range(n)], c[:n].copy()


def hubbard(Ns, t, U, xmu, Hmag=None, J=None):
    t = np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(Ns, Ns))
    U = np.ascontiguousarray(np.asarray(U, dtype=np.float64).reshape(Ns))
    xmu = np.ascontiguousarray(np.asarray(xmu, dtype=np.float64).reshape(Ns))
    Hmag = np.zeros(Ns) if Hmag is None else np.ascontiguousarray(np.asarray(Hmag, dtype=np.float64))
    J = np.zeros((Ns, Ns)) if J is None else np.ascontiguousarray(np.asarray(J, dtype=np.float64))
    h = lib().orc_model_hubbard(Ns, _d(t), _d(U), _d(J), _d(xmu), _d(Hmag))
    return Model("hubbard", Ns, h, dict(t=t, U=U, xmu=xmu, Hmag=Hmag, J=J))


def siam(Ns, ml, Epsk, Vk, H0, Utensor, mu):
    nb = Ns - ml
    Epsk = np.ascontiguousarray(np.asarray(Epsk, dtype=np.float64).reshape(nb, 2))
    Vk = np.ascontiguousarray(np.asarray(Vk, dtype=np.float64).reshape(ml, nb, 2))
    H0 = np.ascontiguousarray(np.asarray(H0, dtype=np.float64).reshape(ml, ml, 2))
    Ut = np.ascontiguousarray(np.asarray(Utensor, dtype=np.float64).reshape(2, 2, ml, ml, ml, ml))
    h = lib().orc_model_siam(Ns, ml, _d(Epsk), _d(Vk), _d(H0), _d(Ut), float(mu))
    return Model("siam", Ns, h, dict(ml=ml, Epsk=Epsk, Vk=Vk, H0=H0, U=Ut, mu=mu))


# ---------------------------------------------------------------------------
# SpinResolvedStorage restatement
# ---------------------------------------------------------------------------
class SRS:
    """fill() + av() of SpinResolvedStorage for one (nup, ndn) sector."""

    def __init__(self, model, nup, ndn):
        self.model = model
        self.nup, self.ndn = nup, ndn
        self._h = lib().orc_srs_fill(model.handle, nup, ndn)
        self.Dup = cnk(model.Ns, nup)
        self.Ddn = cnk(model.Ns, ndn)
        self.n = int(lib().orc_srs_dim(self._h))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_srs_free(self._h)
            self._h = None

    def av(self, v, w=None, clear=True, omp=False):
        v = np.ascontiguousarray(v, dtype=np.float64)
        w = np.zeros(self.n) if w is None else np.array(w, dtype=np.float64, copy=True)
        (lib().orc_srs_av_omp if omp else lib().orc_srs_av)(self._h, _d(v), _d(w), 1 if clear else 0)
        return w

    def diagonal(self):
        out = np.zeros(self.n)
        lib().orc_srs_diag(self._h, _d(out))
        return out

    def csr(self, which):
        """which: 0 = H_up, 1 = H_down, 2 = H_loc; returns (row_ptr, col, val)."""
        nrows = self.Dup if which == 0 else (self.Ddn if which == 1 else self.n)
        nnz = lib().orc_srs_csr(self._h, which, None, None, None)
        rp = np.zeros(nrows + 1, dtype=np.int32)
        col = np.zeros(max(nnz, 1), dtype=np.int32)
        val = np.zeros(max(nnz, 1), dtype=np.float64)
        lib().orc_srs_csr(self._h, which, _i(rp), _i(col), _d(val))
        return rp, col[:nnz], val[:nnz]

    def int_start(self):
        return int(lib().orc_srs_int_start(self._h))

    def lanczos(self, v, nlanc_max, omp=False):
        """Lanczos::lanczos (Lanczos.h:36-77): returns (nlanc, alfa, beta)
        (omp: the OpenMP av and vector loops, for large sectors)."""
        v = np.array(v, dtype=np.float64, copy=True)
        a = np.zeros(nlanc_max)
        b = np.zeros(nlanc_max + 1)
        fn = lib().orc_lanczos_omp if omp else lib().orc_lanczos
        nl = fn(self._h, _d(v), nlanc_max, _d(a), _d(b))
        return nl, a, b

    def dense(self):
        """Dense matrix whose action equals av (columns = av of unit vectors)."""
        H = np.zeros((self.n, self.n))
        e = np.zeros(self.n)
        for j in range(self.n):
            e[:] = 0.0
            e[j] = 1.0
            H[:, j] = self.av(e)
        return H


class SOCRS:
    """SOCRSStorage fill/av (OpenMP) -- the timed CPU baseline."""

    def __init__(self, model, nup, ndn):
        self._h = lib().orc_socrs_fill(model.handle, nup, ndn)
        self.n = cnk(model.Ns, nup) * cnk(model.Ns, ndn)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_socrs_free(self._h)
            self._h = None

    def av(self, v, w=None, clear=True):
        v = np.ascontiguousarray(v, dtype=np.float64)
        w = np.zeros(self.n) if w is None else np.array(w, dtype=np.float64, copy=True)
        lib().orc_socrs_av(self._h, _d(v), _d(w), 1 if clear else 0)
        return w


def a_adag(model, sector, orb, annihilate, invec):
    """SpinResolvedStorage::a_adag (SpinResolvedStorage.h:293-381).

    ``orb`` is the orbital index site + spin*Ns.  Returns (next_sector, outvec).
    """
    nup, ndn = sector
    Ns = model.Ns
    spin = 1 if orb >= Ns else 0
    if annihilate:
        nxt = (nup - (1 - spin), ndn - spin)
    else:
        nxt = (nup + (1 - spin), ndn + spin)
    if not (0 <= nxt[0] <= Ns and 0 <= nxt[1] <= Ns):
        return nxt, None
    out = np.zeros(cnk(Ns, nxt[0]) * cnk(Ns, nxt[1]))
    inv = np.ascontiguousarray(invec, dtype=np.float64)
    lib().orc_adag(model.handle, nup, ndn, nxt[0], nxt[1], orb, 1 if annihilate else 0, _d(inv), _d(out))
    return nxt, out


def occupations(model, sector, norb):
    nup, ndn = sector
    n = cnk(model.Ns, nup) * cnk(model.Ns, ndn)
    nu = np.zeros((n, norb), dtype=np.int8)
    nd = np.zeros((n, norb), dtype=np.int8)
    lib().orc_occupations(model.handle, nup, ndn, norb, nu.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                          nd.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
    return nu, nd


def frac_points(z, ev, nlanc, isign, expb, alfa, beta):
    """Lanczos::get_frac_point (Lanczos.h:201-224) at complex energies z."""
    z = np.asarray(z, dtype=np.complex128).reshape(-1)
    zr = np.ascontiguousarray(z.real)
    zi = np.ascontiguousarray(z.imag)
    a = np.ascontiguousarray(alfa, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    outr = np.zeros(len(z))
    outi = np.zeros(len(z))
    lib().orc_frac_points(len(z), _d(zr), _d(zi), float(ev), int(nlanc), int(isign), float(expb), _d(a), _d(b),
                          _d(outr), _d(outi))
    return outr + 1j * outi


# ---------------------------------------------------------------------------
# Meshes (MeshFactory.h:19-34, CommonUtils.h:14-20, Lanczos.h:146-148)
# ---------------------------------------------------------------------------
def matsubara_points(beta, nomega, fermionic=True):
    n = np.arange(nomega)
    return ((2 * n + 1) if fermionic else (2 * n)) * math.pi / beta


# ---------------------------------------------------------------------------
# Storage::diag restatement (Storage.h:43-146) with scipy's ARPACK.
# ---------------------------------------------------------------------------
# The av of the ARPACK loop on host threads (the restatement of
# SpinResolvedStorage::av over up-rows; same sums, row by row).
AV_OMP = False


def storage_diag(srs, nev_param, ncv_param=None, dense_limit=None):
    """Lowest eigenpairs of one sector with the reference's ARPACK parameters.

    ncv = min(NCV or 2*NEV+3, ntot), nev = min(NEV, ncv-1) (Storage.h:53-54),
    which='SA', tol=1e-14, maxitr=1000.  ntot == 1 -> (diag[0], [1]).
    When ncv == ntot the Krylov space is the whole sector and ARPACK returns
    exact eigenpairs, so a dense eigh gives the same numbers.
    """
    n = srs.n
    if n == 0:
        return np.zeros(0), np.zeros((0, 0))
    if n == 1:
        return np.array([srs.diagonal()[0]]), np.ones((1, 1))
    ncv0 = ncv_param if ncv_param is not None else 2 * nev_param + 3
    ncv = min(ncv0, n)
    nev = min(nev_param, ncv - 1)
    if ncv >= n or (dense_limit is not None and n <= dense_limit):
        H = srs.dense()
        e, V = np.linalg.eigh(0.5 * (H + H.T))
        return e[:nev].copy(), V[:, :nev].copy()
    from scipy.sparse.linalg import LinearOperator, eigsh
    op = LinearOperator((n, n), matvec=lambda x: srs.av(np.asarray(x).reshape(-1), omp=AV_OMP), dtype=np.float64)
    e, V = eigsh(op, k=nev, ncv=ncv, which="SA", tol=1e-14, maxiter=1000)
    order = np.argsort(e, kind="stable")
    return e[order], V[:, order]


class EigenPair:
    def __init__(self, e, vec, pid, sector):
        self.eigenvalue = float(e)
        self.eigenvector = vec
        self.id = pid
        self.sector = sector


def hamiltonian_diag(model, params, sectors=None, dense_limit=None):
    """Hamiltonian::diag (Hamiltonian.h:54-101): sector loop + ordered eigenpairs."""
    Ns = model.Ns
    order = sector_order(Ns, sectors)
    pairs = []
    k = 0
    for sec in order:
        srs = SRS(model, *sec)
        e, V = storage_diag(srs, params["arpack.NEV"], params.get("arpack.NCV"), dense_limit)
        for i in range(len(e)):
            pairs.append(EigenPair(e[i], V[:, i].copy(), k, sec))
            k += 1
    ids = eigenpair_order([p.eigenvalue for p in pairs])
    byid = {p.id: p for p in pairs}
    return [byid[i] for i in ids]


def _passes_cutoff(boltzmann_f, cutoff):
    # GreensFunction.h:135 / ChiLoc.h:208 / StaticObservables.h:129
    return not (abs(cutoff - boltzmann_f) > np.finfo(np.float64).eps and boltzmann_f < cutoff)


def _expb(beta, e, e0):
    # Lanczos.h:87-90
    return 0.0 if beta * (e - e0) > 25 else math.exp(-beta * (e - e0))


def greens_function(model, pairs, params, gf_orbs=None, gf_pairs=()):
    """GreensFunction::compute (GreensFunction.h:109-159) on the Matsubara mesh.

    Returns dict with G (nomega, norb, 2), G_g, G_l, G_ij (nomega, norb*norb, 2), Z.
    """
    beta = float(params["lanc.BETA"])
    nomega = int(params["lanc.NOMEGA"])
    nlanc_max = int(params["lanc.NLANC"])
    cutoff = float(params.get("lanc.BOLTZMANN_CUTOFF", 1e-12))
    Ns = model.Ns
    norb = model.interacting_orbitals()
    if gf_orbs is None:
        gf_orbs = list(range(norb))
    gf_orbs = sorted(set(gf_orbs))
    for p in gf_pairs:
        for o in p:
            if o not in gf_orbs:
                gf_orbs.append(o)
    wn = matsubara_points(beta, nomega, True)
    z = 1j * wn
    Gg = np.zeros((nomega, norb, 2), complex)
    Gl = np.zeros((nomega, norb, 2), complex)
    Ggij = np.zeros((nomega, norb * norb, 2), complex)
    Glij = np.zeros((nomega, norb * norb, 2), complex)
    if not pairs:
        return dict(G=Gg + Gl, G_g=Gg, G_l=Gl, G_ij=Ggij + Glij, Z=0.0)
    e0 = pairs[0].eigenvalue
    Z = sum(math.exp(-(p.eigenvalue - e0) * beta) for p in pairs)

    def start_vector(pair, orbs, spin, annihilate):
        out = None
        nxt = None
        for orb in orbs:
            nxt, o = a_adag(model, pair.sector, orb + spin * Ns, annihilate, pair.eigenvector)
            if o is None:
                return None, None, 0.0
            out = o if out is None else out + o
        norm = float(np.dot(out, out))
        if not abs(norm) > 1e-10:
            return None, None, norm
        return nxt, out / math.sqrt(norm), norm

    def contribute(target, idx, spin, pair, orbs, annihilate):
        nxt, vec, ev = start_vector(pair, orbs, spin, annihilate)
        if vec is None:
            return
        srs = SRS(model, *nxt)
        nl, a, b = srs.lanczos(vec, nlanc_max)
        isign = -1 if annihilate else 1
        ener = z + pair.eigenvalue * isign
        target[:, idx, spin] += frac_points(ener, ev, nl, isign, _expb(beta, pair.eigenvalue, e0), a, b)

    for pair in pairs:
        bf = math.exp(-(pair.eigenvalue - e0) * beta)
        if not _passes_cutoff(bf, cutoff):
            continue
        for orb in gf_orbs:
            for spin in range(2):
                contribute(Gg, orb, spin, pair, [orb], False)
                contribute(Gl, orb, spin, pair, [orb], True)
        for (i, j) in gf_pairs:
            for spin in range(2):
                contribute(Ggij, norb * i + j, spin, pair, [i, j], False)
                contribute(Glij, norb * i + j, spin, pair, [i, j], True)
    Gg /= Z
    Gl /= Z
    Ggij /= Z
    Glij /= Z
    # non_local_gf, GreensFunction.h:336-359
    for (i, j) in gf_pairs:
        for o in (i, j):
            Ggij[:, norb * i + j, :] -= Gg[:, o, :]
            Glij[:, norb * i + j, :] -= Gl[:, o, :]
        Ggij[:, norb * i + j, :] *= 0.5
        Glij[:, norb * i + j, :] *= 0.5
    for o in gf_orbs:
        Ggij[:, norb * o + o, :] = Gg[:, o, :]
        Glij[:, norb * o + o, :] = Gl[:, o, :]
    return dict(G=Gg + Gl, G_g=Gg, G_l=Gl, G_ij=Ggij + Glij, Z=Z)


def _op_action(kind, nu, nd):
    # SzOperator::action (ChiLoc.h:57-61) / NOperator::action (:85-89)
    if kind == "Sz":
        return 0.5 * (nu.astype(np.float64) - nd)
    return nu.astype(np.float64) + nd


def chi_loc(model, pairs, params, kind="Sz", avg=None, orbs=None, orb_pairs=()):
    """ChiLoc::compute<Op> (ChiLoc.h:180-233) on the bosonic Matsubara mesh."""
    beta = float(params["lanc.BETA"])
    nomega = int(params["lanc.NOMEGA"])
    nlanc_max = int(params["lanc.NLANC"])
    cutoff = float(params.get("lanc.BOLTZMANN_CUTOFF", 1e-12))
    norb = model.interacting_orbitals()
    if avg is None:
        avg = 0.0 if kind == "Sz" else 1.0
    if orbs is None:
        orbs = list(range(norb))
    orbs = sorted(set(orbs))
    wn = matsubara_points(beta, nomega, False)
    G = np.zeros((nomega, norb), complex)
    Gij = np.zeros((nomega, norb * norb), complex)
    if not pairs:
        return dict(G=G, G_ij=Gij, Z=0.0)
    e0 = pairs[0].eigenvalue
    Z = sum(math.exp(-(p.eigenvalue - e0) * beta) for p in pairs)

    def contribute(target, idx, pair, olist, require_norm):
        nu, nd = occupations(model, pair.sector, norb)
        act = sum(_op_action(kind, nu[:, o], nd[:, o]) for o in olist)
        out = act * pair.eigenvector
        norm = float(np.dot(out, out))
        out = out / math.sqrt(norm) if norm != 0 else out
        if require_norm and not norm > 1e-9:
            return
        srs = SRS(model, *pair.sector)
        nl, a, b = srs.lanczos(out, nlanc_max)
        expb = _expb(beta, pair.eigenvalue, e0)
        # compute_sym_continued_fraction + update_static (Lanczos.h:105-124,165-169)
        target[0, idx] -= norm * beta * expb
        z1 = 1j * wn[1:] + pair.eigenvalue
        z2 = -1j * wn[1:] + pair.eigenvalue
        target[1:, idx] += frac_points(z1, norm, nl, 1, expb, a, b) + frac_points(z2, norm, nl, 1, expb, a, b)

    for pair in pairs:
        bf = math.exp(-(pair.eigenvalue - e0) * beta)
        if not _passes_cutoff(bf, cutoff):
            continue
        for o in orbs:
            contribute(G, o, pair, [o], True)
        for (i, j) in orb_pairs:
            contribute(Gij, i * norb + j, pair, [i, j], False)
    G /= Z
    Gij /= Z

    def zero_freq(Gv, i):
        # zero_freq_contribution + get_tail, ChiLoc.h:358-411
        def tail(f):
            om1, om2 = wn[f], wn[f - 1]
            o1, o2 = om1 * om1, om2 * om2
            g1, g2 = Gv[f, i].real, Gv[f - 1, i].real
            c2 = -(g2 * o2 * o2 - g1 * o1 * o1) / (o1 - o2)
            c4 = -(g1 * o1 * o1 * o2 - g2 * o2 * o2 * o1) / (o1 - o2)
            return c2, c4, c2 * beta * beta / 24.0 + c4 * beta ** 4 / 1440.0
        c2, c4, t = tail(nomega - 1)
        om = wn[1:]
        s = np.sum(Gv[1:, i].real - c2 / (om * om) - c4 / (om ** 4))
        Gv[0, i] -= 2 * s + 2 * t - avg * avg * beta

    for o in orbs:
        zero_freq(G, o)
    for (i, j) in orb_pairs:
        zero_freq(Gij, i * norb + j)
        for o in (i, j):
            Gij[:, i * norb + j] -= G[:, o]
        Gij[:, i * norb + j] *= 0.5
    for o in orbs:
        Gij[:, o * norb + o] = G[:, o]
    return dict(G=G, G_ij=Gij, Z=Z)


def static_observables(model, pairs, params):
    """StaticObservables::calculate_static_observables (StaticObservables.h:106-158,356-421)."""
    beta = float(params["lanc.BETA"])
    cutoff = float(params.get("lanc.BOLTZMANN_CUTOFF", 1e-12))
    norb = model.interacting_orbitals()
    keys = dict(N=norb, N_up=norb, N_dn=norb, M=norb, D_occ=norb, **{"M_i M_j": norb * norb}, N_eff=1, E=1)
    avg = {k: np.zeros(v) for k, v in keys.items()}
    e0 = pairs[0].eigenvalue
    total = 0.0
    for pair in pairs:
        bf = math.exp(-(pair.eigenvalue - e0) * beta)
        if not _passes_cutoff(bf, cutoff):
            continue
        nu_all, nd_all = occupations(model, pair.sector, model.Ns)
        nu = nu_all[:, :norb].astype(np.float64)
        ndd = nd_all[:, :norb].astype(np.float64)
        wgt = pair.eigenvector ** 2
        m = nu - ndd
        c = dict(N=wgt @ (nu + ndd), N_up=wgt @ nu, N_dn=wgt @ ndd, M=wgt @ m, D_occ=wgt @ (nu * ndd))
        # StaticObservables.h:378-381: the second factor's down occupation is
        # read at orbital orb2 + interacting_orbitals() (orbital o < Ns is the
        # up spin of site o, o >= Ns the down spin of site o - Ns), i.e. the
        # down spin of orb2 only when every orbital interacts
        occ = np.concatenate([nu_all, nd_all], axis=1).astype(np.float64)
        m2 = nu - occ[:, norb:2 * norb]
        c["M_i M_j"] = np.einsum("s,si,sj->ij", wgt, m, m2).reshape(-1)
        c["N_eff"] = np.array([1.0 / np.sum(wgt * wgt)])
        c["E"] = np.array([pair.eigenvalue])
        for k in avg:
            avg[k] += c[k] * bf
        total += bf
    for k in avg:
        avg[k] /= total
    return avg


# ---------------------------------------------------------------------------
# Dense brute force over the full Fock space (independent of the storages).
# ---------------------------------------------------------------------------
def fock_hamiltonian(model):
    """Dense H over all 4^Ns states, indexed by the full 2Ns-bit pattern.

    Matrix element convention of the reference storages: row = source state,
    column = target state, value = sign * t (SpinResolvedStorage.h:129-146,
    CRSMatrix.h:46-78), plus the model diagonal on the diagonal.
    """
    Ns = model.Ns
    dim = 1 << (2 * Ns)
    H = np.zeros((dim, dim))
    for s in range(dim):
        H[s, s] += model.diagonal(s)
        targets, coef = model.apply(s)
        for t, c in zip(targets, coef):
            H[s, t] += c
    return H


def lehmann_gf(model, beta, wn, orb, spin, nev_cut=None):
    """Exact G(iw) = sum_nm (e^-bEn + e^-bEm)|<m|c^+|n>|^2/(iw + En - Em) / Z."""
    return lehmann_gf_z(model, beta, 1j * np.asarray(wn, dtype=float), orb, spin)


def lehmann_gf_z(model, beta, z, orb, spin):
    """The same Lehmann sum at arbitrary complex points z (the real axis's
    w + i pi/beta of CommonUtils.h:18-20 included)."""
    Ns = model.Ns
    H = fock_hamiltonian(model)
    e, V = np.linalg.eigh(0.5 * (H + H.T))
    e0 = e[0]
    w = np.exp(-beta * (e - e0))
    Z = w.sum()
    o = orb + spin * Ns
    pos = 2 * Ns - 1 - o
    dim = 1 << (2 * Ns)
    C = np.zeros((dim, dim))  # c^+_o as a matrix acting on column vectors
    for s in range(dim):
        if not (s >> pos) & 1:
            cnt = bin(s >> (pos + 1)).count("1")
            C[s | (1 << pos), s] = -1.0 if cnt % 2 else 1.0
    M = V.T @ C @ V  # M[m, n] = <m|c^+|n>
    G = np.zeros(len(z), complex)
    for i, x in enumerate(z):
        den = x + e[None, :] - e[:, None]
        G[i] = np.sum((w[None, :] + w[:, None]) * M ** 2 / den) / Z
    return G
