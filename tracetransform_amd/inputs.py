"""Seeded synthetic images (bit-identical on every machine for the uniform
generator; the blob generator is deterministic up to libm's exp).

All images are returned as float32 arrays in the reference's column-major
layout: a flat array where element (row, col) sits at row + col*rows, or
equivalently a [cols, rows] C-ordered array.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """SplitMix64 outputs number offset..offset+count-1 of stream ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """float32 in [0, 1): top 24 bits of SplitMix64, exact in float."""
    u = splitmix64(seed, count, offset) >> np.uint64(40)
    return (u.astype(np.float64) * (1.0 / (1 << 24))).astype(np.float32)


def uniform_image(side: int, seed: int, rows: int | None = None) -> np.ndarray:
    rows = side if rows is None else rows
    return uniform(seed, rows * side).reshape(side, rows)


def blob_image(side: int, seed: int, nblobs: int = 32, rows: int | None = None,
               sigma=(8.0, 64.0)) -> np.ndarray:
    """Sum of Gaussian blobs: centre uniform over the image, sigma uniform in
    ``sigma``, amplitude uniform in [0, 1); evaluated in float64, stored as
    float32.  Column-major [cols, rows]."""
    rows = side if rows is None else rows
    u = uniform(seed, 4 * nblobs).astype(np.float64).reshape(nblobs, 4)
    cx = u[:, 0] * side
    cy = u[:, 1] * rows
    sg = sigma[0] + u[:, 2] * (sigma[1] - sigma[0])
    amp = u[:, 3]
    col = np.arange(side, dtype=np.float64)[:, None]
    row = np.arange(rows, dtype=np.float64)[None, :]
    img = np.zeros((side, rows), np.float64)
    for b in range(nblobs):
        img += amp[b] * np.exp(-((col - cx[b]) ** 2 + (row - cy[b]) ** 2)
                               / (2.0 * sg[b] ** 2))
    return img.astype(np.float32)


def batch_sides(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """``count`` sides drawn uniformly from {lo..hi}."""
    u = splitmix64(seed, count, offset=1 << 20)
    return (lo + (u % np.uint64(hi - lo + 1))).astype(np.int64)
