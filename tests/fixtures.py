"""Loaders for the committed golden fixtures (tests/golden)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_input(name):
    with open(os.path.join(GOLDEN, name)) as f:
        d = json.load(f)
    out = {}
    for k, v in d.items():
        a = np.array(v["data"], dtype=np.float64)
        out[k] = a.reshape(v["shape"]) if v["shape"] else a.reshape(())
    return out


def hubbard_arrays(d):
    """(Ns, t, U, xmu, Hmag, J) from an input.h5-schema dict (HubbardModel.h:55-66)."""
    t = d["hopping/values"]
    Ns = t.shape[0]
    U = d["interaction/values"]
    xmu = d["chemical_potential/values"]
    H = d.get("magnetic_field/values", np.zeros(Ns))
    J = d.get("exchange/values", np.zeros((Ns, Ns)))
    return Ns, t, U, xmu, H, J


def load_dat(name):
    return np.loadtxt(os.path.join(GOLDEN, name))
