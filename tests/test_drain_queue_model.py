"""CPU model of the switched-off persistent drain's tile queue
(tests/queue_model/drain_queue_model.cpp): host threads stand in for waves and
each simulated XCD serves relaxed loads from its own stale cache.  The removed
pop's compare-and-swap retry loop spins without progress in the 3-tile case
that hung the GPU (and drains when loads are coherent); the bounded ticket
queue drains that case and 40 random workloads with no activation lost."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "queue_model" / "drain_queue_model.cpp"


def test_drain_queue_model(tmp_path):
    if shutil.which("g++") is None or not SRC.exists():
        pytest.skip("g++ or the model source not available")
    exe = tmp_path / "dqm"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", str(SRC), "-o", str(exe)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "stale relaxed loads: LIVELOCK" in r.stdout
    assert "coherent relaxed loads: drained" in r.stdout
    assert "40 random workloads: drained" in r.stdout
