"""The reference's PatchMatch suite (tests/test_patchmatch.cpp:72-539) against
the C++ host API on an MI355X: runPatchMatchMVS on the 36-view turntable
with exact SfM cameras and points, including the end-to-end gates (sphere
retention >= 50 %, total retention < 80 %, mean relative error < 0.15).
The program is tests/cpp/test_patchmatch.cpp; its exit status is the number
of failed tests."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "test_patchmatch")


def test_reference_suite_on_host_api(pkg):
    assert os.path.exists(BIN), "tests/cpp/bin/test_patchmatch is not built (run build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "0 test(s) failed" in r.stdout
