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
