"""Model factories shared by the tests: the same parameters build the oracle
model and the product model."""
import numpy as np

from oracle import edoracle as eo


def torus_hopping(L1, L2, t=1.0, tp=0.0):
    """Hopping matrix of an L1 x L2 periodic cluster (values as stored in
    hopping/values; H_hop = -sum t_ij c^+_j c_i)."""
    Ns = L1 * L2
    T = np.zeros((Ns, Ns))
    for r in range(L1):
        for c in range(L2):
            i = r * L2 + c
            for (dr, dc, val) in ((1, 0, t), (0, 1, t), (1, 1, tp), (1, -1, tp)):
                if val == 0.0:
                    continue
                j = ((r + dr) % L1) * L2 + (c + dc) % L2
                if j != i:
                    T[i, j] = val
                    T[j, i] = val
    return T


def ring_hopping(Ns, t=1.0):
    T = np.zeros((Ns, Ns))
    for i in range(Ns):
        T[i, (i + 1) % Ns] = t
        T[(i + 1) % Ns, i] = t
    return T


def kanamori_tensor(ml, U, J):
    """scripts/generate_anderson_model_input.py:7-17, all four spin blocks."""
    Ut = np.zeros((ml, ml, ml, ml))
    for i in range(ml):
        Ut[i, i, i, i] = U
        for j in range(ml):
            if i != j:
                Ut[i, j, i, j] = U - 2 * J
                Ut[i, j, j, i] = J
                Ut[i, i, j, j] = J
    return np.broadcast_to(Ut, (2, 2) + Ut.shape).copy()


def siam_params(Ns, ml, rng, U=4.0, J=0.7, mu=None, h0_offdiag=0.0):
    nb = Ns - ml
    eps = np.repeat(rng.uniform(-2, 2, nb)[:, None], 2, axis=1)
    Vk = np.zeros((ml, nb, 2))
    for im in range(ml):
        Vk[im] = rng.uniform(0.1, 0.8, nb)[:, None]
    H0 = np.zeros((ml, ml, 2))
    for im in range(ml):
        H0[im, im, :] = rng.uniform(-0.3, 0.3)
    if ml > 1 and h0_offdiag:
        H0[1, 0, :] = h0_offdiag
        H0[0, 1, :] = h0_offdiag
    Ut = kanamori_tensor(ml, U, J) if ml > 1 else np.full((2, 2, 1, 1, 1, 1), U)
    if mu is None:
        mu = (5 * U - 10 * J) / 2 if ml == 3 else U / 2
    return dict(Ns=Ns, ml=ml, Epsk=eps, Vk=Vk, H0=H0, U=Ut, mu=mu)


def random_hubbard_params(Ns, rng, with_J=True, uniform_U=False):
    t = rng.normal(size=(Ns, Ns))
    t[rng.random((Ns, Ns)) < 0.4] = 0.0
    t = 0.5 * (t + t.T)
    np.fill_diagonal(t, 0.0)
    U = np.full(Ns, 3.0) if uniform_U else rng.uniform(1, 5, Ns)
    return dict(Ns=Ns, t=t, U=U, xmu=rng.uniform(0, 2, Ns), Hmag=rng.uniform(-0.3, 0.3, Ns),
                J=(0.2 * rng.normal(size=(Ns, Ns)) if with_J else np.zeros((Ns, Ns))))


def oracle_model(kind, p):
    if kind == "hubbard":
        return eo.hubbard(p["Ns"], p["t"], p["U"], p["xmu"], p.get("Hmag"), p.get("J"))
    return eo.siam(p["Ns"], p["ml"], p["Epsk"], p["Vk"], p["H0"], p["U"], p["mu"])


def product_model(kind, p):
    import edlib_amd as ea
    if kind == "hubbard":
        return ea.HubbardModel(p["Ns"], p["t"], p["U"], p["xmu"], p.get("Hmag"), p.get("J"))
    return ea.SingleImpurityAndersonModel(p["Ns"], p["ml"], p["Epsk"], p["Vk"], p["H0"], p["U"], p["mu"])
