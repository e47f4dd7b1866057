"""The C++ drop-in headers compile (host C++, g++) and the driver links
against the in-tree library; without a GPU it must fail loudly, never fall
back to a CPU path."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_driver_builds_and_refuses_without_device():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "edlib_amd", "drivers")])
    binp = os.path.join(ROOT, "edlib_amd", "bin", "edtest")
    assert os.path.exists(binp)
    if os.path.exists("/dev/kfd"):
        return  # a GPU is present: the gpu-marked test runs it
    r = subprocess.run([binp, "4ring", os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True)
    assert r.returncode != 0
    assert "device" in r.stderr


EDSOLVE = os.path.join(ROOT, "edlib_amd", "bin", "edsolve")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _hdf5_driver():
    import pytest
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "edlib_amd", "drivers")])
    if not os.path.exists(EDSOLVE):
        pytest.skip("HDF5 C library not found: the HDF5 driver is not built")
    return EDSOLVE


def test_hdf5_input_reader_matches_fixtures():
    """The drop-in reads the reference's own input.h5 files (test/input/*)
    directly: every dataset equals the committed JSON fixture made from the
    same file by the h5dump tool (shape and values)."""
    import json
    exe = _hdf5_driver()
    for case in ("4ring", "4ring_nomagfield", "GF_Chi"):
        out = subprocess.run([exe, "--dump", os.path.join(GOLDEN, case + "_input.h5")], capture_output=True,
                             text=True, check=True).stdout
        got = json.loads(out)
        with open(os.path.join(GOLDEN, case + "_input.json")) as f:
            ref = json.load(f)
        assert set(got) == set(ref), case
        for k, v in ref.items():
            assert got[k]["shape"] == v["shape"], (case, k)
            assert got[k]["data"] == v["data"], (case, k)


def test_hdf5_driver_refuses_without_device():
    exe = _hdf5_driver()
    if os.path.exists("/dev/kfd"):
        return  # a GPU is present: the gpu-marked test runs it
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([exe, "hubbard", "--NSITES=4", "--INPUT_FILE=" + os.path.join(GOLDEN, "4ring_input.h5"),
                            "--OUTPUT_FILE=" + os.path.join(d, "out.h5")], capture_output=True, text=True)
    assert r.returncode != 0
    assert "device" in r.stderr


def test_gf_result_layout_is_h5gf():
    """Green's functions are written in the layout of ALPSCore's gf library
    (the H5GF format its save() writes; GreensFunction.h:167-194 calls it):
    version group, data with the __complex__ attribute and a trailing
    re/im dimension, mesh/N and one descriptor group per index (MATSUBARA
    with statistics, beta, positive_only and points; REAL_FREQUENCY; INDEX
    for orbitals and spins).  No device is needed to write the sample."""
    import json
    import tempfile
    exe = _hdf5_driver()
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "gf.h5")
        subprocess.check_call([exe, "--write-gf-sample", out])
        res = json.loads(subprocess.run([exe, "--dump", out], capture_output=True, text=True, check=True).stdout)
    g = "results/G_omega/"
    assert res[g + "data"]["shape"] == [4, 2, 2, 2]
    assert res[g + "data/@__complex__"]["data"] == [1]
    assert res[g + "mesh/N"]["data"] == [3]
    assert res[g + "mesh/1/kind"]["string"] == "MATSUBARA"
    assert res[g + "mesh/1/statistics"]["data"] == [1] and res[g + "mesh/1/positive_only"]["data"] == [1]
    assert res[g + "mesh/1/beta"]["data"] == [10.0] and res[g + "mesh/1/N"]["data"] == [4]
    assert abs(res[g + "mesh/1/points"]["data"][1] - 3 * 3.141592653589793 / 10) < 1e-15
    for i, n in ((2, 2), (3, 2)):
        assert res[g + "mesh/%d/kind" % i]["string"] == "INDEX" and res[g + "mesh/%d/N" % i]["data"] == [n]
    assert res[g + "version/major"]["data"] == [0] and "string" in res[g + "version/originator"]
    # data[w][orbital][spin][re/im] of the sample: (w + o/2) - i (s/4 + w)
    data = res[g + "data"]["data"]
    assert data[((1 * 2 + 1) * 2 + 1) * 2: ((1 * 2 + 1) * 2 + 1) * 2 + 2] == [1.5, -1.25]
    c = "results/ChiSz_omega/"
    assert res[c + "data"]["shape"] == [4, 2, 2] and res[c + "mesh/N"]["data"] == [2]
    assert res[c + "mesh/1/statistics"]["data"] == [0] and res[c + "mesh/1/points"]["data"][0] == 0.0
    r = "results/G_real/"
    assert res[r + "mesh/1/kind"]["string"] == "REAL_FREQUENCY"
    assert res[r + "mesh/1/points"]["data"][0] == -2.0 and res[r + "mesh/1/points"]["data"][-1] == 2.0
