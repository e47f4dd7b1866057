"""The C++ drop-in (HIPSpinResolvedStorage in Hamiltonian, GreensFunction,
ChiLoc, StaticObservables) run like the reference's own tests:
test/Lanczos_Test.cpp (G to 1e-10, chi to 1e-9), test/HubbardModel_Test.cpp,
test/StaticObservables_Test.cpp."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "edlib_amd", "bin", "edtest")
GOLDEN = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["gfchi", "4ring", "4ring_nomag"])
def test_dropin_reference_tests(case):
    assert os.path.exists(BIN), "edtest driver not built (run __graft_entry__.build())"
    r = subprocess.run([BIN, case, GOLDEN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0 and res["ok"], res
