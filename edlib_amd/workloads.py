"""Synthetic model builders shared by the benchmark and the profiling
workloads: periodic Hubbard clusters and Anderson impurity models with seeded
random baths (numpy in, product models out)."""
import numpy as np


def torus(L1, L2, t=1.0):
    """Nearest-neighbour hopping matrix of an L1 x L2 periodic cluster."""
    Ns = L1 * L2
    T = np.zeros((Ns, Ns))
    for r in range(L1):
        for c in range(L2):
            i = r * L2 + c
            for j in (((r + 1) % L1) * L2 + c, r * L2 + (c + 1) % L2):
                T[i, j] = T[j, i] = t
    return T


def siam_model(Ns, ml, seed, U=4.0, J=0.7):
    """Anderson impurity with ml orbitals and Ns - ml bath sites: bath
    energies U[-2, 2], hybridisations U[0.1, 0.8] (spin-symmetric), impurity
    levels U[-0.3, 0.3]; Kanamori interaction (U, J, U' = U - 2J) for ml > 1,
    mu = U/2 (ml = 1) or (5U - 10J)/2 (ml = 3, half filling)."""
    import edlib_amd as ea
    return ea.SingleImpurityAndersonModel(*siam_arrays(Ns, ml, seed, U, J))


def siam_arrays(Ns, ml, seed, U=4.0, J=0.7):
    """(Ns, ml, Epsk, Vk, H0, Utensor, mu) of siam_model."""
    rng = np.random.default_rng(seed)
    nb = Ns - ml
    eps = np.repeat(rng.uniform(-2, 2, nb)[:, None], 2, axis=1)
    Vk = np.zeros((ml, nb, 2))
    for im in range(ml):
        Vk[im] = rng.uniform(0.1, 0.8, nb)[:, None]
    H0 = np.zeros((ml, ml, 2))
    for im in range(ml):
        H0[im, im, :] = rng.uniform(-0.3, 0.3)
    if ml > 1:
        Ut = np.zeros((ml, ml, ml, ml))
        for i in range(ml):
            Ut[i, i, i, i] = U
            for j in range(ml):
                if i != j:
                    Ut[i, j, i, j] = U - 2 * J
                    Ut[i, j, j, i] = J
                    Ut[i, i, j, j] = J
        Ut = np.broadcast_to(Ut, (2, 2) + Ut.shape).copy()
        mu = (5 * U - 10 * J) / 2
    else:
        Ut = np.full((2, 2, 1, 1, 1, 1), U)
        mu = U / 2
    return Ns, ml, eps, Vk, H0, Ut, mu
