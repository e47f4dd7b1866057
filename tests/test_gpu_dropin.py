"""The C++ drop-in (HIPSpinResolvedStorage in Hamiltonian, GreensFunction,
ChiLoc, StaticObservables) run like the reference's own tests:
test/Lanczos_Test.cpp (G to 1e-10, chi to 1e-9), test/HubbardModel_Test.cpp,
test/StaticObservables_Test.cpp; the self-energy from the Dyson equation
against what the mathematics fixes (Sigma = 0 at U = 0 with every orbital
pair; Sigma(i w) -> U <n> at high frequency; Sigma = -mu for the Anderson
model at U = 0, the reference's bare function having no mu)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "edlib_amd", "bin", "edtest")
GOLDEN = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["gfchi", "4ring", "4ring_nomag", "dyson_hubbard0", "dyson_hubbard", "dyson_siam0",
                                  "storage"])
def test_dropin_reference_tests(case):
    assert os.path.exists(BIN), "edtest driver not built (run __graft_entry__.build())"
    r = subprocess.run([BIN, case, GOLDEN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0 and res["ok"], res


EDSOLVE = os.path.join(ROOT, "edlib_amd", "bin", "edsolve")


def _edsolve(args, tmp_path):
    if not os.path.exists(EDSOLVE):
        pytest.skip("HDF5 driver not built (no HDF5 C library)")
    out = str(tmp_path / "out.h5")
    r = subprocess.run([EDSOLVE] + args + ["--OUTPUT_FILE=" + out], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    d = subprocess.run([EDSOLVE, "--dump", out], capture_output=True, text=True, check=True, timeout=60).stdout
    return json.loads(d)


def test_hdf5_driver_4ring(tmp_path):
    """examples/Hubbard.cpp's steps on the reference's test/input/4ring/input.h5,
    results in the reference's HDF5 paths: E0 -11.8443 in sector (2, 2)
    (HubbardModel_Test.cpp:54), G, chi_Sz, chi_N, static observables."""
    res = _edsolve(["hubbard", "--NSITES=4", "--NSPINS=2", "--INPUT_FILE=" + os.path.join(GOLDEN, "4ring_input.h5"),
                    "--arpack.SECTOR=0", "--storage.MAX_DIM=36", "--arpack.NEV=1", "--lanc.NOMEGA=64"], tmp_path)
    ev = res["results/eigenvalues/data"]["data"]
    assert res["results/eigenvalues/N"]["data"][0] == len(ev) > 0
    assert abs(ev[0] - (-11.8443)) < 1e-4
    assert res["results/eigenvalues/sectors/0/nup"]["data"] == [2]
    assert res["results/eigenvalues/sectors/0/ndown"]["data"] == [2]
    assert res["results/eigenvalues/sectors/0/size"]["data"] == [36]
    assert res["results/@Statsum"]["data"][0] >= 1.0
    assert res["results/G_omega/data"]["shape"] == [64, 4, 2, 2]
    assert res["results/G_omega/mesh/1/N"]["data"] == [64]
    assert res["results/G_omega/mesh/1/kind"]["string"] == "MATSUBARA"
    assert res["results/G_omega/mesh/2/kind"]["string"] == "INDEX"
    assert res["results/G_omega/data/@__complex__"]["data"] == [1]
    assert res["results/ChiSz_omega/data"]["shape"] == [64, 4, 2]
    assert res["results/ChiN_omega/data"]["shape"] == [64, 4, 2]
    n = res["results/static_observables/N"]["data"]
    assert len(n) == 4 and max(abs(x - 1.0) for x in n) < 1e-8


def test_hdf5_driver_gf_chi(tmp_path):
    """The reference's test/input/GF_Chi/input.h5 through the HDF5 driver: G on
    the first 200 Matsubara points equals gom1.dat to 1e-10
    (Lanczos_Test.cpp:115-119)."""
    res = _edsolve(["hubbard", "--NSITES=4", "--NSPINS=2", "--INPUT_FILE=" + os.path.join(GOLDEN, "GF_Chi_input.h5"),
                    "--arpack.SECTOR=0", "--storage.MAX_DIM=36", "--arpack.NEV=100", "--lanc.BETA=20",
                    "--lanc.NOMEGA=200"], tmp_path)
    g = res["results/G_omega/data"]
    nw, norb, ns, _ = g["shape"]
    data = g["data"]
    ref = []
    with open(os.path.join(GOLDEN, "gom1.dat")) as f:
        for line in f:
            ref.append([float(x) for x in line.split()])
    err = 0.0
    for w in range(200):
        for s in range(ns):
            q = ((w * norb + 0) * ns + s) * 2
            err = max(err, abs(complex(data[q], data[q + 1]) - complex(ref[w][1], ref[w][2])))
    assert err < 1e-10, err


def test_dropin_one_state_pairs_match_oracle_diagonal():
    """The storage corner case's one-state pairs (sectors (0,0), (0,4),
    (4,0), (4,4) of the 4-site ring with field and chemical potential) equal
    the oracle's HubbardModel diagonal of the single state (HubbardModel.h:
    116-132, pinned by the 4-ring E0 and the brute-force tests)."""
    from oracle import edoracle as eo
    from tests.fixtures import hubbard_arrays, load_input
    r = subprocess.run([BIN, "storage", GOLDEN], capture_output=True, text=True, timeout=120)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0 and res["ok"], res
    Ns, t, U, xmu, H, J = hubbard_arrays(load_input("4ring_input.json"))
    om = eo.hubbard(Ns, t, U, xmu, H, J)
    got = {(a, b): e for a, b, e in res["one_state_pairs"]}
    assert set(got) == {(0, 0), (0, 4), (4, 0), (4, 4)}
    for (a, b), e in got.items():
        st = ((0xF if a else 0) << Ns) | (0xF if b else 0)
        assert abs(e - om.diagonal(st)) < 1e-12, ((a, b), e, om.diagonal(st))
