"""Measurements of the precision, batched and serving forms on one GPU.

Each function returns a plain dict (what bench.py prints next to its main
line).  Work units: a voxel-ray update is one forward sample or one
backprojected voxel, 2 * N * Nvox per SIRT iteration.
"""
from __future__ import annotations

import time
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import batch as B
from . import serving as S
from .projector import VoxelProjector3DHIP, phantom_ellipsoids
from .reconstruction import tilt_eulers

GIB = 1 << 30

# (name, storage, FP32 master volume, FP32 projections)
PRECISION_MODES = (
    ("f32", "f32", False, False),
    ("bf16+master+p32", "bf16", True, True),
    ("bf16+master", "bf16", True, False),
    ("bf16", "bf16", False, False),
    ("f16+master+p32", "f16", True, True),
    ("f16", "f16", False, False),
)


def _rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    num = 0.0
    den = 0.0
    for k in range(a.shape[0]):                       # slice by slice: bounded host memory
        d = a[k].astype(np.float64) - b[k]
        num += float(np.dot(d.ravel(), d.ravel()))
        den += float(np.dot(b[k].ravel(), b[k].ravel()))
    return float(np.sqrt(num / den)) if den > 0 else float("nan")


def plain_sirt_rate(W: int, H: int, D: int, tilts: Sequence[float], seed: int, k: int, noise: float,
                    iters: int = 3) -> float:
    """Iterations/s of TomoJ's SIRT through the bit-exact projector (FP32)."""
    P = VoxelProjector3DHIP((W, H, D))
    try:
        P.synth_tilt_series(phantom_ellipsoids(seed, k, W, H, D), tilt_eulers(tilts), noise, seed ^ 0x77)
        P.run_ossart(1, 1.0, 0)
        t = time.perf_counter()
        P.run_ossart(iters, 1.0, 0)
        return iters / (time.perf_counter() - t)
    finally:
        P.close()


def precision_study(W: int, H: int, D: int, n: int, iterations: int = 30, refine: int = 2,
                    ellipsoids: int = 80, noise: float = 0.02, seed: int = 0xC3C3_1024,
                    modes=PRECISION_MODES, reference: bool = True, tune: bool = True) -> Dict:
    """SIRT of one W x H x D job from n tilts (+-60 deg): `iterations` in each
    storage mode, then `refine` FP32 iterations against the FP32 projections;
    rel-L2 against the FP64 twin after the same number of iterations, and
    iterations/s of each mode next to the bit-exact FP32 path."""
    tilts = list(np.linspace(-60.0, 60.0, n))
    ell = phantom_ellipsoids(seed, ellipsoids, W, H, D)
    nseed = seed ^ 0x5A5A
    upd = 2 * n * W * H * D
    out: Dict = {"volume": [W, H, D], "projections": n, "iterations": iterations, "refine": refine,
                 "ellipsoids": ellipsoids, "noise_rel": noise, "modes": {}}
    plain = plain_sirt_rate(W, H, D, tilts, seed, ellipsoids, noise)
    out["plain_f32_it_s"] = plain
    out["plain_f32_gups"] = plain * upd / 1e9
    ref = None
    if reference:
        with B.JobEngine(256 * GIB) as eng:
            j = eng.add(B.JobSpec(W, H, D, tilts, storage="f64", max_iterations=iterations + refine))
            eng.synth(j, ell, noise, nseed)
            t = time.perf_counter()
            eng.run([j], iterations + refine)
            out["f64_it_s"] = (iterations + refine) / (time.perf_counter() - t)
            ref = eng.get_volume(j)
    tuning = None
    for name, st, master, p32 in modes:
        with B.JobEngine(256 * GIB) as eng:
            j = eng.add(B.JobSpec(W, H, D, tilts, storage=st, fp32_master=master, fp32_projections=p32,
                                  max_iterations=iterations + refine))
            eng.synth(j, ell, noise, nseed)
            if tune:
                tuning = eng.autotune([j])
                eng.reset_volume(j)
            ms = eng.run([j], iterations, graph=True)
            rate = iterations / (ms * 1e-3)
            if st != "f32":
                eng.convert(j, "f32")
                eng.synth(j, ell, noise, nseed)            # the FP32 projections again
            eng.run([j], refine)
            rec = {"it_s": rate, "gups": rate * upd / 1e9, "vs_plain": rate / plain, "tuning": tuning}
            if ref is not None:
                rec["rel_l2_vs_f64"] = _rel_l2(eng.get_volume(j), ref)
            out["modes"][name] = rec
    return out


def batched_run(n_jobs: int = 64, iterations: int = 15, storage: str = "f32", master: bool = False,
                seed: int = 0x5B70_0064, budget: int = 200 * GIB, tune_cache: Optional[str] = None) -> Dict:
    """The batched form: n_jobs uneven subtomograms, `iterations` SIRT
    iterations of all of them in one graph launch (timed after a warm-up
    launch of the same graph on reset volumes)."""
    jobs = B.subtomogram_jobs(n_jobs, seed=seed, storage=storage)
    for jb in jobs:
        jb.spec.fp32_master = master
        jb.spec.fp32_projections = master
        jb.spec.max_iterations = 2 * iterations
    with B.JobEngine(budget) as eng:
        t = time.perf_counter()
        ids = [B.load_job(eng, jb) for jb in jobs]
        setup = time.perf_counter() - t
        tun = eng.autotune(ids, cache=tune_cache)
        eng.run(ids, iterations, graph=True)                  # capture + warm-up
        for j in ids:
            eng.reset_volume(j)
        ms = eng.run(ids, iterations, graph=True)             # replay of the same graph
        upd = sum(jb.spec.voxel_ray_updates_per_iteration() for jb in jobs) * iterations
        return {"jobs": n_jobs, "iterations": iterations, "storage": storage, "fp32_master": master,
                "sizes": sorted({jb.spec.W for jb in jobs}), "setup_s": setup, "ms": ms,
                "gups": upd / (ms * 1e-3) / 1e9, "jobs_per_s": n_jobs / (ms * 1e-3),
                "device_bytes": eng.used, "tuning": tun}


def serving_run(n_requests: int = 200, rate: float = 20.0, budget: int = 96 * GIB, seed: int = 0x5E21_0020,
                widths: Sequence[int] = (128, 256, 512, 1024), graph: bool = True) -> Dict:
    """The serving form: a seeded Poisson stream served in real time (graph
    ticks by default)."""
    reqs = S.poisson_requests(n_requests, rate=rate, seed=seed, widths=widths)
    with B.JobEngine(budget) as eng:
        res = S.ServeLoop(eng, graph=graph).run(reqs, realtime=True)
    res.pop("done")
    res.update({"rate_per_s": rate, "budget_bytes": budget, "widths": list(widths),
                "storages": ["f32", "bf16"], "graph": graph})
    return res
