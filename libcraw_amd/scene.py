"""Seeded synthetic raw frames and the synthetic camera profile.

The workloads are the ones the benchmark and the tests name: RGGB uint16
mosaics holding 12-bit data (black 512, white 4095), developed from a
scene-linear [0, 1] scene of smooth colour gradients, hard-edged rectangles
and sinusoidal texture (periods 2 to 40 px), with per-channel white balance
gains inverted into the raw and Gaussian noise of variance a * u + b per
sample (the Gaussian approximation of Poisson-Gaussian noise the reference's
tests use, tests/test_pipeline.c:846-861 of cyanreg/libcraw). Noise comes
from numpy's Philox, a counter-based generator, so a seed names one frame on
any machine.
"""
import numpy as np

from . import capi

BLACK_12 = 512.0 / 65535.0
WHITE_12 = 4095.0 / 65535.0

# Camera RGB (balanced) to XYZ D65: rows sum to the XYZ of D65 white.
CAMERA_MATRIX = np.array([[0.6300, 0.2400, 0.0805],
                          [0.2600, 0.8300, -0.0900],
                          [0.0300, -0.1000, 1.1590]], dtype=np.float32)


def make_profile(wb=(2.0, 1.0, 1.6), black=BLACK_12, white=WHITE_12, matrix=CAMERA_MATRIX,
                 name=b"synthetic 12-bit RGGB"):
    """A CRAWProfile: levels, as-shot gains and matrix; no tables."""
    p = capi.CRAWProfile()
    p.cfa = capi.CRAW_CFA_RGGB
    for s in range(4):
        p.black_level[s] = black
        p.white_level[s] = white
    for c in range(3):
        p.as_shot_wb[c] = wb[c]
        for k in range(3):
            p.as_shot_matrix[c][k] = float(matrix[c][k])
    p.gain = 1.0
    p.name = name
    return p


def scene_rgb(width, height, seed, x_off=0, y_off=0):
    """Scene-linear RGB, float64 (H, W, 3), deterministic in `seed`.

    The texture is six plane waves sin(kx x + ky y + phase); each is split as
    sin(kx x) cos(ky y + phase) + cos(kx x) sin(ky y + phase), so that the
    sum of all six for one channel is one (H x 12) by (12 x W) product."""
    rng = np.random.Generator(np.random.Philox(seed))
    x = np.arange(width, dtype=np.float64) + x_off
    y = np.arange(height, dtype=np.float64) + y_off
    fx = x / max(width - 1, 1)
    fy = y / max(height - 1, 1)
    rgb = np.empty((3, height, width))
    rgb[0] = 0.08 + 0.55 * fx[None, :] + 0.10 * fy[:, None]
    rgb[1] = 0.10 + 0.45 * fy[:, None] + 0.08 * fx[None, :]
    rgb[2] = 0.12 + 0.30 * fx[None, :] + 0.30 * (1.0 - fy[:, None])

    # sinusoidal texture, periods 2..40 px, mixed orientations
    cols, rows, tints = [], [], []
    for _ in range(6):
        period = rng.uniform(2.0, 40.0)
        theta = rng.uniform(0.0, np.pi)
        amp = rng.uniform(0.02, 0.08)
        phase = rng.uniform(0.0, 2 * np.pi)
        ax = 2 * np.pi * x * np.cos(theta) / period
        by = 2 * np.pi * y * np.sin(theta) / period + phase
        cols += [np.sin(ax), np.cos(ax)]
        rows += [amp * np.cos(by), amp * np.sin(by)]
        tint = rng.uniform(0.5, 1.0, size=3)
        tints += [tint, tint]
    cols, rows, tints = np.array(cols), np.array(rows), np.array(tints)
    for c in range(3):
        rgb[c] += (rows.T * tints[:, c]) @ cols

    # hard-edged rectangles, one of them above the clip
    n_rect = 12
    for i in range(n_rect):
        w = int(rng.integers(max(width // 20, 2), max(width // 5, 3)))
        h = int(rng.integers(max(height // 20, 2), max(height // 5, 3)))
        x0 = int(rng.integers(0, max(width - w, 1)))
        y0 = int(rng.integers(0, max(height - h, 1)))
        col = rng.uniform(0.02, 0.9, size=3)
        if i == 0:
            col = np.array([1.3, 1.2, 1.25])
        rgb[:, y0:y0 + h, x0:x0 + w] = col[:, None, None]
    np.clip(rgb, 0.0, 1.5, out=rgb)
    return np.ascontiguousarray(np.moveaxis(rgb, 0, -1))


def mosaic_from_scene(rgb, wb=(2.0, 1.0, 1.6), a=2e-4, b=4e-6, seed=0, black=BLACK_12,
                      white=WHITE_12, max_code=4095):
    """RGGB uint16 mosaic: the scene with the gains divided out, noise
    N(0, a*u + b) added, re-encoded to the 12-bit range."""
    height, width, _ = rgb.shape
    u = np.empty((height, width))
    for s, c in enumerate((0, 1, 1, 2)):
        u[s >> 1::2, s & 1::2] = rgb[s >> 1::2, s & 1::2, c] / float(wb[c])
    rng = np.random.Generator(np.random.Philox(seed ^ 0x5EED))
    sigma = np.sqrt(a * np.maximum(u, 0.0) + b)
    u = u + sigma * rng.standard_normal(u.shape)
    raw = np.rint((black + u * (white - black)) * 65535.0)
    return np.clip(raw, 0, max_code).astype(np.uint16)


def make_frame(width, height, seed=1, wb=(2.0, 1.0, 1.6), a=2e-4, b=4e-6):
    return mosaic_from_scene(scene_rgb(width, height, seed), wb=wb, a=a, b=b, seed=seed)


def clip_frame(width, height, index, seed=3, wb=(2.0, 1.0, 1.6), a=4e-4, b=1.6e-5,
               square=400, speed=8):
    """Frame `index` of a clip: a static seeded background and a textured
    square of `square` px translating `speed` px/frame, fresh noise each
    frame."""
    bg = scene_rgb(width, height, seed)
    tex = scene_rgb(square, square, seed + 1000)
    sq = min(square, width // 2, height // 2)
    x0 = (index * speed) % max(width - sq, 1)
    y0 = (height - sq) // 2
    bg[y0:y0 + sq, x0:x0 + sq] = tex[:sq, :sq]
    return mosaic_from_scene(bg, wb=wb, a=a, b=b, seed=seed * 7919 + index)
