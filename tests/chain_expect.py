"""Expected output of the segmentation chain for any dtype, composed from the
oracle's single-operation functions (the order of or_chain_u8:
imreconstruct -> imfillHoles -> threshold -> bwareaopen2 -> relabel)."""
import numpy as np


def expected_chain(orc, marker, mask, thresh, recon_conn=8, fill_conn=4, ccl_conn=8, min_size=11, max_size=1000,
                   i32_saturate=False):
    r = orc.imreconstruct(marker, mask, recon_conn)
    f = orc.fill_holes(r, False, fill_conn, i32_saturate=i32_saturate)
    b = np.where(f > thresh, 255, 0).astype(np.uint8)
    lab, _ = orc.bwareaopen2(b, False, False, min_size, max_size, ccl_conn)
    canon, n = orc.relabel(lab, -1)
    return r, f, canon, n


def request_images(nh, side, kind, seed):
    """A serving request (nscale_hip.workloads.request_images)."""
    from nscale_hip.workloads import request_images as gen
    return gen(side, kind, seed)
