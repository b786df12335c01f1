"""Y-slab reconstruction on one device.

ReconstructionThreadManagerGPU.doGPURec (src/TomoJ_GPU/.../tomography/
ReconstructionThreadManagerGPU.java:181-297) splits the volume into slabs of
rows that fit the device budget (computeMaximumSizeForVolume, :184-191) and
reconstructs each on its own: with the tilt axis vertical, volume row y meets
only detector row y - vcy + pcy (and, with a zero weight, the row after it), so
a slab, its detector rows and one halo row below form a complete problem.  ``reconstruct_in_slabs`` does the same with one projector
context per slab (centres shifted by the slab's first row), after supplying
the full images once (getPixels: alignment, normalisation, threshold; and the
WBP weighting, a 2-D filter of whole images).  It is also the unit a
multi-device run would hand out, one slab per device.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Tuple

import numpy as np

from ._lib import call
from .parameters import (BAYESIAN, BGART, BP, OSSART, TVM, WBP, DEFINED_PROJECTIONS, ReconstructionParameters,
                         TiltSeriesData, supply)
from .projector import VoxelProjector3DHIP


def max_slab_rows(params: ReconstructionParameters, device: int = 0) -> int:
    """Rows per slab within computeMaximumSizeForVolume's budget (:184-191)."""
    b = C.c_size_t()
    call("tj_max_volume_bytes", int(device), C.byref(b))
    return max(1, int(b.value // (4 * params.width * params.depth)))


def reconstruct_in_slabs(ts: TiltSeriesData, params: ReconstructionParameters, slab_rows: Optional[int] = None,
                         device: int = 0, first_index: int = 0) -> Tuple[np.ndarray, List[np.ndarray]]:
    """The volume (D, H, W) and each OS-SART slab's err rows.  Supported:
    BP, WBP and the OS-SART family with the vertical tilt axis, scale 1,
    detector height = volume height and no long-object compensation (its norm
    is a whole-image maximum); TVM, Bayesian and BgART couple neighbouring
    rows (BgART also takes its background from the whole volume's histogram)
    and are refused."""
    W, H, D = params.width, params.height, params.depth
    t = params.reconstructionType
    if t in (TVM, BAYESIAN, BGART):
        raise ValueError("TVM, Bayesian and BgART regularisation couple neighbouring rows: no slabs")
    if t not in (BP, WBP, OSSART, DEFINED_PROJECTIONS):
        raise ValueError(f"reconstruction type {t} has no slab form")
    if not ts.axis_vertical or ts.height != H or ts.width != W or params.longObjectCompensation:
        raise ValueError("slabs need the vertical axis, a W x H detector and no long-object compensation")
    rows = slab_rows or max_slab_rows(params, device)
    # the supplied (and, for WBP, weighted) full images
    with VoxelProjector3DHIP((W, H, 1), (W, H), device=device) as S:
        supply(S, ts)
        imgs = np.stack([S.get_loaded_projection(k) for k in range(len(ts.tilts))])
        if t == WBP:
            S.setWeightingFilter(params.weightingRadius * ts.width, ts.tilts, ts.tilt_axes())
            imgs = np.stack([S.weighting(im) for im in imgs])
    E = ts.eulers()
    vc = params.volume_centers()
    pc = [(W - 1) / 2.0, (H - 1) / 2.0, (W - 1) / 2.0, (H - 1) / 2.0]
    avail = params.getAvailableIndexes(ts)
    full = len(avail) == len(ts.tilts) and avail == list(range(len(avail)))
    vol = np.zeros((D, H, W), np.float32)
    errs: List[np.ndarray] = []
    for y0 in range(0, H, rows):
        n = min(rows, H - y0)
        # one halo row below the slab: TomoJ's sampler takes the last volume
        # row (and the backprojection the last image row) by its nearest value,
        # so the slab's own last row must not be the context's last
        nc = min(n + 1, H - y0)
        with VoxelProjector3DHIP((W, nc, D), (W, nc), device=device,
                                 positivity=(t == OSSART and params.positivityConstraint),
                                 proj_center=(pc[0], pc[1] - y0, pc[2], pc[3] - y0),
                                 vol_center=(vc[0], vc[1] - y0, vc[2])) as P:
            P.load_tilt_series(np.ascontiguousarray(imgs[:, y0:y0 + nc, :]), E)
            P.setAvailableIndexes(None if full else avail)
            if params.fista:
                P.setFista(True)
            if t in (BP, WBP):
                P.run_wbp()
            else:
                errs.append(P.run_ossart(params.nbIterations, params.relaxationCoefficient, params.updateNb,
                                         first_index=first_index))
            vol[:, y0:y0 + n, :] = P.get_volume()[:, :n, :]
    return vol, errs
