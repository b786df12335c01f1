"""Named synthetic workloads (sizes, tilt schemes, phantom seeds).

Each workload fixes a 64-bit master seed; the ellipsoid stream uses the master
seed and the noise stream ``master ^ NOISE_SALT``.  Phantom: K ellipsoids with
density U[0.1,1], centre uniform in the inner-80% ellipsoid of the volume box,
semi-axes U[0.03,0.25]*min(W,H,D), uniform random rotation.  Projections are
analytic FP64 line integrals plus N(0, noise_rel * max) noise.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

NOISE_SALT = 0x5DEECE66D


@dataclass(frozen=True)
class Workload:
    name: str
    W: int
    H: int
    D: int
    tilts: List[float] = field(hash=False)
    ellipsoids: int
    noise_rel: float
    seed: int
    iterations: int
    relaxation: float
    update: int          # projections per subset (0 = all, SIRT)
    wbp: bool = False

    @property
    def n(self) -> int:
        return len(self.tilts)

    @property
    def nvox(self) -> int:
        return self.W * self.H * self.D

    @property
    def npx(self) -> int:
        return self.W * self.H

    @property
    def noise_seed(self) -> int:
        return self.seed ^ NOISE_SALT

    def voxel_ray_updates_per_iteration(self) -> int:
        """One forward sample plus one backprojection update per (projection, voxel)."""
        return 2 * self.n * self.nvox

    def subsets(self) -> int:
        u = self.n if self.update == 0 else self.update
        return -(-self.n // u)


def tilt_range(lo: float, hi: float, step: float) -> List[float]:
    n = int(round((hi - lo) / step)) + 1
    return [lo + k * step for k in range(n)]


# Plain SIRT on the size TomoJ's CPU path handles: 256x256x128, 61 tilts -60..60 step 2.
SIRT_256 = Workload("sirt_256x256x128_61", 256, 256, 128, tilt_range(-60, 60, 2), 12, 0.01,
                    0x70A1_0001, 10, 1.0, 0)
# Production size: 512x512x256, 121 tilts -60..60 step 1, OS-SART 10 subsets (13 per subset),
# 20 iterations, plus one weighted backprojection.
OSSART_512 = Workload("ossart_512x512x256_121", 512, 512, 256, tilt_range(-60, 60, 1), 40, 0.01,
                      0x70A1_0002, 20, 1.0, 13, wbp=True)
# The literal 512^3 volume with the same tilt series.
SIRT_512_CUBE = Workload("sirt_512x512x512_121", 512, 512, 512, tilt_range(-60, 60, 1), 40, 0.01,
                         0x70A1_0003, 10, 1.0, 0)
# Precision form: 1024x1024x512, 121 tilts, 80 ellipsoids, 2% noise.
SIRT_1024 = Workload("sirt_1024x1024x512_121", 1024, 1024, 512, tilt_range(-60, 60, 1), 80, 0.02,
                     0x70A1_0004, 30, 1.0, 0)

ALL = {w.name: w for w in (SIRT_256, OSSART_512, SIRT_512_CUBE, SIRT_1024)}


def small(W=32, H=24, D=16, step=10.0, lo=-60.0, hi=60.0, ellipsoids=5, seed=0x5EED, update=0,
          iterations=2) -> Workload:
    """A small workload for parity tests."""
    return Workload(f"small_{W}x{H}x{D}", W, H, D, tilt_range(lo, hi, step), ellipsoids, 0.01, seed,
                    iterations, 1.0, update)


def algorithmic_bytes_per_iteration(w: Workload, s_v: int = 4, s_p: int = 4) -> int:
    """Per-projection slab-streaming model: every projection streams the volume
    once in the forward projection, every subset update reads and writes the
    volume once, and each projection pixel is read (measured), written
    (residual) and read again (backprojection)."""
    return w.n * w.nvox * s_v + w.subsets() * 2 * w.nvox * s_v + w.n * w.npx * (s_p + 8)


def compulsory_bytes_per_iteration(w: Workload, s_v: int = 4, s_p: int = 4) -> int:
    """Lower bound independent of the design: one read of the volume for the
    forward pass and one read-modify-write per subset update, plus the images."""
    return w.nvox * s_v * (1 + 2 * w.subsets()) + w.n * w.npx * (s_p + 8)


def eulers_for(w: Workload) -> np.ndarray:
    from .reconstruction import tilt_eulers
    return tilt_eulers(w.tilts)
