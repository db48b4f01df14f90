"""The highlight scene of the reference's tests (tests/test_pipeline.c:
4349-4415 of cyanreg/libcraw), restated: two Gaussian blobs of constant
chroma on a dim gradient, in balanced camera RGB; the first peaks at
G = 1.4 with R and B under their clips (only G clips), the second at G = 4.0
(its core blown in every channel). Inverted to raw codes with the
reference's black/white levels and white balance, saturating at white, with
its deterministic +-8 count hash noise."""
import ctypes

import numpy as np

from libcraw_amd import capi

W, H = 256, 160
BLACK = np.float32(256.0 / 65535.0)
WHITE = np.float32(63743.0 / 65535.0)
WB = np.array([2.02539, 1.0, 1.66797], np.float32)
PART = np.array([1.30, 1.0, 0.90], np.float32)
FULL = np.array([1.10, 1.0, 1.05], np.float32)
PART_X, FULL_X, BLOB_Y = 64.0, 184.0, 80.0


def identity_profile():
    """tests/test_pipeline.c:77-107 with an identity colour stage."""
    p = capi.CRAWProfile()
    capi.lib().craw_profile_init(ctypes.byref(p))
    p.cfa = capi.CRAW_CFA_RGGB
    for i in range(4):
        p.black_level[i] = BLACK
        p.white_level[i] = WHITE
    for c in range(3):
        p.as_shot_wb[c] = WB[c]
        for k in range(3):
            p.as_shot_matrix[c][k] = 1.0 if c == k else 0.0
    p.as_shot_cct = 5500.0
    p.gain = 1.0
    return p


def truth(w=W, h=H):
    """Balanced camera RGB, (h, w, 3) float32."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    fx, fy = x / np.float32(w - 1), y / np.float32(h - 1)
    dp = (x - PART_X) ** 2 + (y - BLOB_Y) ** 2
    df = (x - FULL_X) ** 2 + (y - BLOB_Y) ** 2
    gp = np.float32(1.4) * np.exp(-dp / np.float32(2 * 22.0 * 22.0))
    gf = np.float32(4.0) * np.exp(-df / np.float32(2 * 18.0 * 18.0))
    rgb = np.empty((h, w, 3), np.float32)
    rgb[..., 0] = 0.02 + 0.02 * fx
    rgb[..., 1] = 0.02 + 0.02 * fy
    rgb[..., 2] = 0.03 + 0.01 * (fx + fy)
    part = gp >= 0.05
    full = (gf >= 0.05) & (gf >= gp)
    rgb[part] = PART * gp[part][:, None]
    rgb[full] = FULL * gf[full][:, None]
    return rgb


def _hash(v):
    v = v.astype(np.uint64)
    m = np.uint64(0xffffffff)
    v ^= v >> np.uint64(16)
    v = (v * np.uint64(0x7feb352d)) & m
    v ^= v >> np.uint64(15)
    v = (v * np.uint64(0x846ca68b)) & m
    v ^= v >> np.uint64(16)
    return v


def mosaic(scale=1.0, w=W, h=H):
    rgb = truth(w, h)
    y, x = np.mgrid[0:h, 0:w]
    chan = np.where((y & 1) == 0, np.where((x & 1) == 0, 0, 1), np.where((x & 1) == 0, 1, 2))
    v = np.take_along_axis(rgb, chan[..., None], axis=2)[..., 0]
    u = np.minimum(v * np.float32(scale) / WB[chan], np.float32(1.0))
    iv = np.rint((BLACK + u * (WHITE - BLACK)) * np.float32(65535.0)).astype(np.int64)
    idx = ((y * w + x).astype(np.uint64) * np.uint64(2654435761)) & np.uint64(0xffffffff)
    iv += (_hash(idx) % np.uint64(17)).astype(np.int64) - 8
    return np.clip(iv, 0, 65535).astype(np.uint16)


def u_of(m):
    """u of the definition test (tests/test_pipeline.c:4418-4423)."""
    raw = m.astype(np.float32) * np.float32(1.0 / 65535.0)
    return np.maximum((raw - BLACK) * (np.float32(1.0) / (WHITE - BLACK)), np.float32(0.0))


def channels(h=H, w=W):
    y, x = np.mgrid[0:h, 0:w]
    return np.where((y & 1) == 0, np.where((x & 1) == 0, 0, 1), np.where((x & 1) == 0, 1, 2))
