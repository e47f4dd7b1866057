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
