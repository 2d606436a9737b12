"""The CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer: `make
-C oracle sanitize` builds oracle/sanitize_check.c (a driver of every oracle
entry point on inputs with empty columns, ties, NA padding and short
buffers) with both sanitizers and runs it; any finding aborts the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_clean_under_asan_and_ubsan():
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ran clean" in r.stdout
