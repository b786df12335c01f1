"""Serving loop: reconstruction requests admitted by HBM budget and advanced
together on one GPU (continuous batching).

Requests arrive over time (a seeded Poisson stream in the benchmark).  The
loop keeps every admitted job resident in the job engine (tomoj_amd.batch)
and advances them in ticks.  A tick gives every active job about the same
amount of work: a job whose one SIRT iteration is small gets several
iterations (about one "quantum" of voxel-ray updates, by default the largest
active job's one iteration, and never more than it has left), so a small job
is not held behind the iterations of a large one.  Jobs with the same
iteration count share persistent launches (optionally one HIP graph per
group, cached by the engine); the groups run smallest remaining work first,
and a job that finishes is retired (memory freed, volume read back) as soon
as its group's launch ends.  Between ticks waiting requests are admitted in
arrival order while they fit the budget.  A request's latency is its
completion time minus its arrival time, including queueing, setup (phantom
projections generated on the device) and the read-back of its volume.

No TomoJ counterpart: TomoJ reconstructs one tilt series per run.  Each job's
arithmetic is independent of the other jobs in its launches, so a result is
the same whether the job ran alone or in the mix.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .batch import JobEngine, JobSpec, SplitMix64


@dataclass
class Request:
    rid: int
    arrival: float          # seconds from the start of the stream
    spec: JobSpec
    iterations: int
    ellipsoids: int
    seed: int
    noise_rel: float = 0.01


@dataclass
class Done:
    rid: int
    arrival: float          # when the request became visible to the loop
    admitted: float
    finished: float
    iterations: int
    updates: int            # voxel-ray updates of the job
    volume: Optional[np.ndarray] = None
    width: int = 0
    storage: str = ""

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def poisson_requests(n: int, rate: float = 20.0, seed: int = 0x5E21_0020,
                     widths: Sequence[int] = (128, 256, 512, 1024), tilt_counts: Sequence[int] = (41, 61, 121),
                     iterations: Sequence[int] = (10, 30), storages: Sequence[str] = ("f32", "bf16"),
                     tilt_range: float = 60.0) -> List[Request]:
    """The serving workload: exponential inter-arrival times (mean 1/rate),
    detector W x W with W from `widths` and volume depth W/2, tilt counts from
    `tilt_counts` over +-tilt_range, SIRT iterations uniform in `iterations`,
    storage from `storages`, 4-20 ellipsoids and 1% noise per job."""
    r = SplitMix64(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += -math.log(1.0 - r.u01()) / rate
        W = r.choice(widths)
        nt = r.choice(tilt_counts)
        it = r.randint(iterations[0], iterations[1])
        st = r.choice(storages)
        tilts = list(np.linspace(-tilt_range, tilt_range, nt))
        spec = JobSpec(W, W, max(W // 2, 1), tilts, storage=st, max_iterations=max(it, 1))
        out.append(Request(k, t, spec, it, r.randint(4, 20), r.next()))
    return out


def percentile(values: Sequence[float], q: float) -> float:
    """Nearest-rank percentile (q in [0, 100])."""
    if not values:
        return float("nan")
    v = sorted(values)
    k = max(1, int(math.ceil(q / 100.0 * len(v))))
    return v[min(k, len(v)) - 1]


class ServeLoop:
    """Continuous batching over a JobEngine with a fixed HBM budget."""

    def __init__(self, engine: JobEngine, keep_volumes: bool = False, graph: bool = False,
                 clock: Callable[[], float] = time.perf_counter, quantum: Optional[int] = None):
        self.engine = engine
        self.keep_volumes = keep_volumes
        self.graph = graph
        self.clock = clock
        self.quantum = quantum      # voxel-ray updates per job per tick (None: largest active iteration)

    def _setup(self, req: Request) -> int:
        from .projector import phantom_ellipsoids
        s = req.spec
        jid = self.engine.add(s)
        ell = phantom_ellipsoids(req.seed, req.ellipsoids, s.W, s.H, s.D)
        self.engine.synth(jid, ell, req.noise_rel, req.seed ^ 0xA5A5A5A5)
        return jid

    def plan_tick(self, active: Dict[int, Request], left: Dict[int, int]) -> List[Tuple[int, List[int]]]:
        """The launches of one tick: (iterations, jobs) groups.  Each job gets
        max(1, quantum // its per-iteration updates) iterations, capped by what
        it has left; jobs with equal counts share a group; groups are ordered
        by the largest remaining work of their jobs, smallest first."""
        upd = {j: active[j].spec.voxel_ray_updates_per_iteration() for j in active}
        quantum = self.quantum if self.quantum else max(upd.values())
        groups: Dict[int, List[int]] = {}
        for j in active:
            k = max(1, min(left[j], quantum // max(upd[j], 1)))
            groups.setdefault(k, []).append(j)
        return sorted(((k, sorted(js)) for k, js in groups.items()),
                      key=lambda kv: (max(upd[j] * left[j] for j in kv[1]), kv[0]))

    def run(self, requests: Sequence[Request], realtime: bool = True) -> Dict:
        """Serves `requests` (sorted by arrival).  With realtime, a request is
        visible only once the wall clock has passed its arrival time; without,
        every request is visible (and its latency counted) from the start.
        Returns per-request records and aggregate statistics."""
        pending = deque(sorted(requests, key=lambda q: q.arrival))
        for q in pending:
            if q.iterations < 1:
                raise ValueError(f"request {q.rid} asks for {q.iterations} iterations; at least 1 is needed")
            if q.spec.bytes() > self.engine.budget:
                raise ValueError(f"request {q.rid} needs {q.spec.bytes()} bytes, more than the whole budget")
        active: Dict[int, Request] = {}
        left: Dict[int, int] = {}
        admitted_at: Dict[int, float] = {}
        done: List[Done] = []
        t0 = self.clock()
        ticks = 0
        launches = 0
        max_active = 0
        busy = 0.0
        while pending or active:
            now = self.clock() - t0
            # admission: arrived requests, in order, while they fit
            while pending and (pending[0].arrival <= now or not realtime) and self.engine.fits(pending[0].spec):
                q = pending.popleft()
                jid = self._setup(q)
                active[jid] = q
                left[jid] = q.iterations
                admitted_at[jid] = self.clock() - t0
            if not active:
                if pending and realtime:
                    time.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
                continue
            max_active = max(max_active, len(active))
            for k, ids in self.plan_tick(active, left):
                tb = self.clock()
                self.engine.run(ids, k, graph=self.graph)
                busy += self.clock() - tb
                launches += 1
                for jid in ids:
                    left[jid] -= k
                    if left[jid] <= 0:
                        q = active.pop(jid)
                        vol = self.engine.get_volume(jid) if self.keep_volumes else None
                        self.engine.remove(jid)
                        done.append(Done(q.rid, q.arrival if realtime else 0.0, admitted_at.pop(jid),
                                         self.clock() - t0, q.iterations,
                                         q.spec.voxel_ray_updates_per_iteration() * q.iterations, vol,
                                         q.spec.W, q.spec.storage))
                        del left[jid]
            ticks += 1
        wall = self.clock() - t0
        lat = [d.latency for d in done]
        updates = sum(d.updates for d in done)
        by_width = {}
        for w in sorted({d.width for d in done}):
            lw = [d.latency for d in done if d.width == w]
            by_width[str(w)] = {"n": len(lw), "p50_s": percentile(lw, 50), "p99_s": percentile(lw, 99)}
        return {
            "done": sorted(done, key=lambda d: d.rid),
            "requests": len(done),
            "wall_s": wall,
            "ticks": ticks,
            "launches": launches,
            "max_active": max_active,
            "busy_s": busy,
            "p50_s": percentile(lat, 50),
            "p99_s": percentile(lat, 99),
            "mean_s": float(np.mean(lat)) if lat else float("nan"),
            "by_width": by_width,
            "jobs_per_s": len(done) / wall if wall > 0 else float("nan"),
            # completed jobs per second of device-busy time: the highest
            # arrival rate of this request mix the loop can sustain
            "capacity_jobs_per_s": len(done) / busy if busy > 0 else float("nan"),
            "gups": updates / wall / 1e9 if wall > 0 else float("nan"),
        }


def run_alone(engine: JobEngine, req: Request) -> np.ndarray:
    """The same request reconstructed with no other job on the device."""
    loop = ServeLoop(engine, keep_volumes=True)
    out = loop.run([Request(req.rid, 0.0, req.spec, req.iterations, req.ellipsoids, req.seed, req.noise_rel)],
                   realtime=False)
    return out["done"][0].volume
