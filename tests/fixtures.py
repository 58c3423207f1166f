"""Deterministic inputs used by the reference's own tests, regenerated here.

- jittered_cloud: the LCG cloud of LODBucketingTests.swift:12-27
  (state = state * 6364136223846793005 + 1442695040888963407,
  next = Float(state >> 40) / 2^24).
- lattice: GPUPackerParityTests.swift:55-73 (distinct Morton keys).
- solid_cube: MultiCloudDepthOrderTests.swift:32-49.
"""
import numpy as np

MASK64 = (1 << 64) - 1


def jittered_cloud(count: int = 3000, seed: int = 0x5EED):
    state = seed
    f32 = np.float32

    def nxt():
        nonlocal state
        state = (state * 6364136223846793005 + 1442695040888963407) & MASK64
        return f32(state >> 40) / f32(1 << 24)

    pos = np.zeros((count, 3), np.float32)
    col = np.zeros((count, 4), np.float32)
    for i in range(count):
        x = nxt() * f32(10) - f32(5)
        y = nxt() * f32(10) - f32(5)
        z = nxt() * f32(10) - f32(5)
        pos[i] = (x, y, z)
        col[i] = (nxt(), nxt(), nxt(), f32(1))
    return pos, col


def lattice(n: int, extent: float = 1.0):
    idx = np.arange(n, dtype=np.float32)
    f = idx / np.float32(n - 1)
    z, y, x = np.meshgrid(f, f, f, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.float32)
    col = np.concatenate([pos.copy(), np.ones((pos.shape[0], 1), np.float32)], axis=1)
    pos = (pos - np.float32(0.5)) * np.float32(extent)
    return pos.astype(np.float32), col.astype(np.float32)


def solid_cube(color, n: int = 16):
    pos, _ = lattice(n)
    col = np.tile(np.asarray(color, np.float32), (pos.shape[0], 1))
    return pos, col
