"""Host sanitizer builds (ASan + UBSan, TSan) of the serving loop's host logic
-- the arena free list and the admission of queued requests into sets, under
many producer and worker threads -- and of the round-prediction table, the
same headers the library is built from (tests/sanitize/)."""
import shutil
import subprocess
from pathlib import Path

import pytest

SAN = Path(__file__).resolve().parent / "sanitize"


@pytest.fixture(scope="module")
def built():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++/make not available")
    r = subprocess.run(["make", "-C", str(SAN), "-j2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return SAN / "_build"


@pytest.mark.parametrize("variant", ["asan", "tsan"])
def test_server_core_under_sanitizer(built, variant):
    r = subprocess.run([str(built / f"stress_{variant}")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "admission threads: ok" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
