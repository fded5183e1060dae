"""The header-only C++ drop-in (include/nscale_dropin.hpp): the check program
tests/cpp/dropin_check (built by __graft_entry__.build()) runs every nscale::
function on a fixed synthetic image; its outputs must equal the oracle's."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
PROG = REPO / "tests" / "cpp" / "dropin_check"
R, C = 61, 83


def inputs():
    y, x = np.mgrid[0:R, 0:C]
    mask = ((x * 37 + y * 91 + ((x * y) >> 3)) & 255).astype(np.uint8)
    marker = np.where(mask > 30, mask.astype(np.int32) - 30, 0).astype(np.uint8)
    binimg = np.where(mask > 128, 255, 0).astype(np.uint8)
    return marker, mask, binimg


def seeds_of():
    y, x = np.mgrid[0:R, 0:C]
    return ((x * 13 + y * 7) % 97 == 0).astype(np.uint8)


def test_dropin_header_compiles_and_fails_cleanly_without_gpu(tmp_path):
    """Compiled against the in-tree library; without a GPU the first call
    throws the library's error (exit 3), with one it runs (exit 0)."""
    import shutil
    import torch
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    lib = REPO / "emory-cci-fast-ia-gsoc_amd" / "nscale_hip"
    exe = tmp_path / "dropin_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{REPO / 'include'}",
                    f"-I{REPO / 'tests' / 'cpp'}", str(REPO / "tests" / "cpp" / "dropin_check.cpp"), f"-L{lib}",
                    "-lnscale_hip", f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    if torch.cuda.is_available():
        return   # the gpu test below runs the program
    res = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    assert res.returncode == 3 and "no-gpu" in res.stdout


@pytest.mark.gpu
def test_dropin_matches_oracle(orc, tmp_path):
    assert PROG.exists(), "tests/cpp/dropin_check is built by __graft_entry__.build()"
    res = subprocess.run([str(PROG), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    marker, mask, binimg = inputs()

    def load(name, dt):
        return np.fromfile(tmp_path / name, dtype=dt).reshape(R, C)

    assert np.array_equal(load("recon8.bin", np.uint8), orc.imreconstruct(marker, mask, 8))
    assert np.array_equal(load("recon4.bin", np.uint8), orc.imreconstruct(marker, mask, 4))
    assert np.array_equal(load("reconbin8.bin", np.uint8), orc.imreconstruct_binary(marker, binimg, 8))
    assert np.array_equal(load("fill4.bin", np.uint8), orc.fill_holes(mask, False, 4))
    assert np.array_equal(load("fillbin8.bin", np.uint8), orc.fill_holes(binimg, True, 8))
    seeds = seeds_of()
    assert np.array_equal(load("imfill4.bin", np.uint8), orc.imfill(mask, seeds, False, 4))
    assert np.array_equal(load("imfillbin8.bin", np.uint8), orc.imfill(binimg, seeds, True, 8))
    assert np.array_equal(load("bwselect8.bin", np.uint8), orc.bwselect(binimg, seeds, 8))
    assert np.array_equal(load("hmin8.bin", np.uint8), orc.imhmin(mask, 20, 8))
    assert np.array_equal(load("lmax8.bin", np.uint8), orc.local_extrema(mask, 8, False))
    assert np.array_equal(load("lmin4.bin", np.uint8), orc.local_extrema(mask, 4, True))
    want_nn, want_dist = orc.distance_transform(binimg)
    assert np.array_equal(load("dist.bin", np.float32), want_dist)
    assert np.array_equal(load("distnn.bin", np.int32), want_nn)
    assert np.array_equal(load("label8.bin", np.int32), orc.bwlabel2(binimg, 8, True)[0])
    want_area, want_n = orc.bwareaopen2(binimg, False, False, 3, 40, 8)
    assert np.array_equal(load("area8.bin", np.int32), want_area)
    assert f"count {want_n}" in res.stdout
