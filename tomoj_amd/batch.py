"""Host side of the job engine (``tj_batch_*`` in include/tomoj_hip.h).

A ``JobEngine`` holds independent SIRT reconstructions resident on one GPU
within an HBM budget and advances any subset of them together: one persistent
forward launch (residual fused), one FP64 error reduction and one persistent
backprojection per storage type and iteration, optionally captured in one HIP
graph.  This is what the precision form (BF16/FP16 storage closed by FP32
refinement), the batched form (many uneven subtomograms in one graph launch)
and the serving loop (tomoj_amd.serving) are built on.

TomoJ has no counterpart: it reconstructs one tilt series at a time
(ReconstructionThreadManager.java:79-167) with SIRT = osartIteration with
update = N (TomoReconstruction2.java:636-733).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import call, tj_batch_tuning, tj_job_desc
from .reconstruction import tilt_eulers

STORAGE = _lib.STORAGE_NAMES
STORAGE_BY_CODE = {v: k for k, v in STORAGE.items()}


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class JobSpec:
    """One reconstruction job: a tilt series about the vertical axis with
    TomoJ's default centres, reconstructed by SIRT."""
    W: int
    H: int
    D: int
    tilts: Sequence[float]
    storage: str = "f32"
    relaxation: float = 1.0
    positivity: bool = False
    max_iterations: int = 256
    pw: Optional[int] = None
    ph: Optional[int] = None
    fp32_master: bool = False
    fp32_projections: bool = False

    def desc(self) -> tj_job_desc:
        d = tj_job_desc()
        d.vol_w, d.vol_h, d.vol_d = int(self.W), int(self.H), int(self.D)
        d.proj_w = int(self.pw if self.pw is not None else self.W)
        d.proj_h = int(self.ph if self.ph is not None else self.H)
        d.n_proj = len(self.tilts)
        d.storage = STORAGE[self.storage]
        d.positivity = int(bool(self.positivity))
        d.relaxation = float(self.relaxation)
        d.max_iterations = int(self.max_iterations)
        d.fp32_master = int(bool(self.fp32_master))
        d.fp32_projections = int(bool(self.fp32_projections))
        return d

    def bytes(self) -> int:
        """Device bytes the job takes (what admission charges against the budget)."""
        n = _lib.load().tj_job_bytes(C.byref(self.desc()))
        if n == 0:
            raise _lib.TjError(_lib.TJ_EINVAL, f"invalid job description: {_lib.last_error()}")
        return int(n)

    def voxel_ray_updates_per_iteration(self) -> int:
        return 2 * len(self.tilts) * self.W * self.H * self.D


class JobEngine:
    """Resident SIRT jobs on one device (tj_batch)."""

    def __init__(self, hbm_budget: int, device: int = 0):
        self._b = C.c_void_p()
        call("tj_batch_open", int(device), C.c_size_t(int(hbm_budget)), C.byref(self._b))
        self.specs: Dict[int, JobSpec] = {}

    def close(self) -> None:
        if self._b:
            call("tj_batch_close", self._b)
            self._b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- memory
    @property
    def used(self) -> int:
        return int(_lib.load().tj_batch_used(self._b))

    @property
    def budget(self) -> int:
        return int(_lib.load().tj_batch_budget(self._b))

    def fits(self, spec: JobSpec) -> bool:
        return self.used + spec.bytes() <= self.budget

    # -- jobs
    def add(self, spec: JobSpec, eulers: Optional[np.ndarray] = None) -> int:
        E = tilt_eulers(spec.tilts) if eulers is None else np.asarray(eulers, np.float64)
        E = np.ascontiguousarray(E.reshape(-1, 9), np.float64)
        jid = C.c_int()
        call("tj_batch_add_job", self._b, C.byref(spec.desc()), _ptr(E), C.byref(jid))
        self.specs[jid.value] = spec
        return jid.value

    def remove(self, job: int) -> None:
        call("tj_batch_remove_job", self._b, int(job))
        self.specs.pop(job, None)

    def set_projections(self, job: int, stack: np.ndarray) -> None:
        s = self.specs[job]
        d = s.desc()
        stack = np.ascontiguousarray(stack, np.float32)
        if stack.shape != (d.n_proj, d.proj_h, d.proj_w):
            raise ValueError(f"stack shape {stack.shape} != {(d.n_proj, d.proj_h, d.proj_w)}")
        call("tj_batch_set_projections", self._b, int(job), _ptr(stack))

    def synth(self, job: int, ellipsoids: np.ndarray, noise_rel: float, seed: int) -> None:
        ell = np.ascontiguousarray(ellipsoids, np.float64).reshape(-1, 16)
        call("tj_batch_synth_job", self._b, int(job), _ptr(ell), ell.shape[0], float(noise_rel),
             C.c_uint64(int(seed)))

    def get_projection(self, job: int, index: int) -> np.ndarray:
        d = self.specs[job].desc()
        out = np.empty((d.proj_h, d.proj_w), np.float32)
        call("tj_batch_get_projection", self._b, int(job), int(index), _ptr(out))
        return out

    def get_difference(self, job: int, index: int) -> np.ndarray:
        d = self.specs[job].desc()
        out = np.empty((d.proj_h, d.proj_w), np.float64)
        call("tj_batch_get_difference", self._b, int(job), int(index), _ptr(out))
        return out

    def set_volume(self, job: int, vol_zyx: np.ndarray) -> None:
        s = self.specs[job]
        v = np.ascontiguousarray(vol_zyx, np.float32)
        if v.shape != (s.D, s.H, s.W):
            raise ValueError(f"volume shape {v.shape} != {(s.D, s.H, s.W)}")
        call("tj_batch_set_volume", self._b, int(job), _ptr(v))

    def get_volume(self, job: int) -> np.ndarray:
        s = self.specs[job]
        if s.storage == "f64":
            out64 = np.empty((s.D, s.H, s.W), np.float64)
            call("tj_batch_get_volume_f64", self._b, int(job), _ptr(out64))
            return out64
        out = np.empty((s.D, s.H, s.W), np.float32)
        call("tj_batch_get_volume", self._b, int(job), _ptr(out))
        return out

    def reset_volume(self, job: int) -> None:
        call("tj_batch_reset_volume", self._b, int(job))

    def convert(self, job: int, storage: str) -> None:
        call("tj_batch_convert", self._b, int(job), STORAGE[storage])
        self.specs[job].storage = storage

    # -- running
    def run(self, jobs: Iterable[int], iterations: int, graph: bool = False) -> float:
        """`iterations` SIRT iterations of every job in `jobs`; returns the
        device milliseconds of the run (HIP events around it)."""
        ids = np.ascontiguousarray(list(jobs), np.int32)
        call("tj_batch_run", self._b, _ptr(ids), len(ids), int(iterations),
             _lib.TJ_RUN_GRAPH if graph else 0)
        return float(_lib.load().tj_batch_last_run_ms(self._b))

    def errors(self, job: int) -> np.ndarray:
        """Sum of squared differences per iteration done (FP64)."""
        cap = self.specs[job].max_iterations
        out = np.zeros(cap, np.float64)
        n = call("tj_batch_errors", self._b, int(job), _ptr(out), cap)
        return out[: min(n, cap)]

    def iterations_done(self, job: int) -> int:
        return call("tj_batch_errors", self._b, int(job), None, 0)

    # -- tuning
    def get_tuning(self) -> dict:
        t = tj_batch_tuning()
        call("tj_batch_get_tuning", self._b, C.byref(t))
        return {k: getattr(t, k) for k, _ in tj_batch_tuning._fields_}

    def set_tuning(self, **kw) -> None:
        cur = self.get_tuning()
        cur.update(kw)
        call("tj_batch_set_tuning", self._b, C.byref(tj_batch_tuning(**cur)))

    def autotune(self, jobs: Iterable[int], cache: Optional[str] = None) -> dict:
        """Picks tile/slab shapes for this job mix by timing candidates on the
        jobs themselves (volumes restored).  With `cache`, a JSON file keyed by
        the mix's shape signature, a known signature is reused without timing."""
        ids = list(jobs)
        key = mix_signature(self.specs[j] for j in ids)
        if cache and os.path.exists(cache):
            with open(cache) as fh:
                known = json.load(fh)
            if key in known:
                self.set_tuning(**known[key]["tuning"])
                return known[key]["tuning"]
        arr = np.ascontiguousarray(ids, np.int32)
        best = tj_batch_tuning()
        ms = C.c_double()
        call("tj_batch_autotune", self._b, _ptr(arr), len(ids), C.byref(best), C.byref(ms))
        tun = {k: getattr(best, k) for k, _ in tj_batch_tuning._fields_}
        if cache:
            known = {}
            if os.path.exists(cache):
                with open(cache) as fh:
                    known = json.load(fh)
            known[key] = {"tuning": tun, "ms": ms.value}
            with open(cache, "w") as fh:
                json.dump(known, fh, indent=1, sort_keys=True)
        return tun


def mix_signature(specs: Iterable[JobSpec]) -> str:
    """Shape signature of a job mix: storage types with the voxel-ray work
    bucketed by powers of two and the largest width (what the tile shapes
    depend on)."""
    by: Dict[str, List[JobSpec]] = {}
    for s in specs:
        by.setdefault(s.storage, []).append(s)
    parts = []
    for st in sorted(by):
        ss = by[st]
        work = sum(x.voxel_ray_updates_per_iteration() for x in ss)
        parts.append(f"{st}:n{len(ss)}:w{max(x.W for x in ss)}:u2^{int(np.log2(max(work, 1)))}")
    return "|".join(parts)


# ----------------------------------------------------------------------------
# workloads of the batched and serving forms
# ----------------------------------------------------------------------------
class SplitMix64:
    """SplitMix64 stream (Steele, Lea, Flood 2014), as the library's phantom generator."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def u01(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.u01() * (hi - lo + 1))

    def choice(self, seq):
        return seq[int(self.u01() * len(seq))]


@dataclass
class SubtomoJob:
    spec: JobSpec
    ellipsoids: int
    seed: int
    noise_rel: float = 0.01


def subtomogram_jobs(n: int = 64, seed: int = 0x5B70_0064, storage: str = "f32",
                     sizes=(64, 96, 128, 192, 256)) -> List[SubtomoJob]:
    """The batched workload: cubic sizes uniform in `sizes`, tilt counts
    uniform in [41, 81], tilt ranges +-U[45, 70] deg, 4-20 ellipsoids and 1%
    noise per job, all from one SplitMix64 stream."""
    r = SplitMix64(seed)
    out = []
    for _ in range(n):
        L = r.choice(sizes)
        nt = r.randint(41, 81)
        rng = 45.0 + 25.0 * r.u01()
        tilts = list(np.linspace(-rng, rng, nt))
        k = r.randint(4, 20)
        out.append(SubtomoJob(JobSpec(L, L, L, tilts, storage=storage, max_iterations=64), k, r.next()))
    return out


def load_job(engine: JobEngine, job: SubtomoJob) -> int:
    """Admits a job and generates its phantom projections on the device."""
    from .projector import phantom_ellipsoids
    s = job.spec
    jid = engine.add(s)
    ell = phantom_ellipsoids(job.seed, job.ellipsoids, s.W, s.H, s.D)
    engine.synth(jid, ell, job.noise_rel, job.seed ^ 0xA5A5A5A5)
    return jid
