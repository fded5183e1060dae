"""CPU checks of the C ABI library: it loads, exports every declared symbol,
and its synthetic generators are deterministic."""
import hashlib
import re
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (REPO / "include" / "nscale_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nsx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(nh):
    lib = nh._lib.load()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(nh._lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_version(nh):
    assert b"gfx950" in nh._lib.load().nsx_version()


def test_splitmix64_known_values(nh):
    import ctypes as C
    lib = nh._lib.load()
    s = C.c_uint64(0)
    got = [lib.nsx_splitmix64_next(C.byref(s)) for _ in range(3)]
    # SplitMix64 reference outputs for seed 0
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_disk_generator_frozen(nh):
    a = nh.gen_disks_u8(512, 512, 200, 0x5EED0000)
    b = nh.gen_disks_u8(512, 512, 200, 0x5EED0000)
    assert np.array_equal(a, b)
    assert a.dtype == np.uint8 and a.shape == (512, 512)
    digest = hashlib.sha256(a.tobytes()).hexdigest()
    golden = (REPO / "tests" / "golden" / "c0_mask.sha256").read_text().split()[0]
    assert digest == golden
    m = nh.gen_marker_u8(a, 30)
    assert np.array_equal(m, np.clip(a.astype(int) - 30, 0, 255))


def test_maze_generator(nh):
    mask, marker = nh.gen_maze_u8(96, 120, 0x5EED0003)
    assert set(np.unique(mask)) == {20, 200}
    assert (marker[:5, :5] == 200).all() and marker.sum() == 25 * 200
    # a perfect maze: corridors form one 4-connected component
    from scipy import ndimage
    _, n = ndimage.label(mask == 200)
    assert n == 1


def test_ctx_without_gpu_fails_cleanly(nh):
    import torch
    if torch.cuda.is_available():
        return
    import ctypes as C
    h = C.c_void_p()
    st = nh._lib.load().nsx_ctx_create(C.byref(h), 0, 0)
    assert st != 0


def test_image_check_host_only(nh):
    """nsx_image_check applies the entry points' descriptor checks without a
    context or a device (runs here, without a GPU)."""
    import ctypes as C
    L = nh._lib
    lib = L.load()
    buf = np.zeros(64, np.uint8)
    why = C.c_char_p()

    def check(rows, cols, pitch, dtype=L.NSX_U8, mem=L.NSX_MEM_DEVICE, data=buf.ctypes.data):
        return lib.nsx_image_check(C.byref(L.NsxImg(data, pitch, rows, cols, dtype, mem)), C.byref(why))

    assert check(8, 8, 8) == L.NSX_OK and why.value == b""
    assert check(0, 5, 5, data=None) == L.NSX_OK            # empty images need no data
    assert check(50000, 50000, 50000) == L.NSX_ERR_INVALID and b"2^31" in why.value
    assert check(46341, 46341, 46341, mem=L.NSX_MEM_HOST) == L.NSX_ERR_INVALID       # 2^31 + 1 pixels
    assert check(46340, 46340, 46340) == L.NSX_OK                                      # 2147395600 pixels
    assert check(4, 8, 7) == L.NSX_ERR_INVALID and b"pitch" in why.value
    assert check(4, 8, 31, dtype=L.NSX_I32) == L.NSX_ERR_INVALID
    assert check(4, 8, 8, dtype=9) == L.NSX_ERR_INVALID and b"dtype" in why.value
    assert check(4, 8, 8, mem=5) == L.NSX_ERR_INVALID and b"memory" in why.value
    assert check(4, 8, 8, data=None) == L.NSX_ERR_INVALID and b"NULL" in why.value
    assert check(-1, 8, 8) == L.NSX_ERR_INVALID
    assert lib.nsx_image_check(None, None) == L.NSX_ERR_INVALID
