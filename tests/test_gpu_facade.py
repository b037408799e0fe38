"""GPU: the C++ façade (include/cgp/cgp.hpp) runs the reference's known-answer
validators through the C-ABI (cgp_amd/bin/facade_check, built by make)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_facade_known_answers(gpu):
    exe = os.path.join(ROOT, "cgp_amd", "bin", "facade_check")
    assert os.path.exists(exe), "facade_check not built (make -C cgp_amd/csrc)"
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True,
                       timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASS" in r.stdout and "[FAIL]" not in r.stdout
