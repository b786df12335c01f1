"""Host logic of the batched and serving forms (CPU only): workload
generators, statistics and the admission/continuous-batching state machine
(driven here by an engine stand-in that records what the loop asks for)."""
import numpy as np
import pytest

from tomoj_amd import batch as B
from tomoj_amd import serving as S


def test_splitmix64_known_answers():
    # SplitMix64 (Steele, Lea, Flood 2014) from state 0: published first outputs
    r = B.SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_splitmix64_matches_library_phantom_stream():
    """The Python stream is the library's: the first ellipsoid density is
    0.1 + 0.9 * u01(first output) in tj_phantom_ellipsoids."""
    from tomoj_amd.projector import phantom_ellipsoids
    for seed in (1, 0x5EED, 2 ** 63 + 5):
        ell = phantom_ellipsoids(seed, 1, 64, 64, 32)
        assert ell[0, 0] == 0.1 + 0.9 * B.SplitMix64(seed).u01()


def test_poisson_requests_shape_and_rate():
    reqs = S.poisson_requests(4000, rate=20.0, seed=7)
    t = np.array([q.arrival for q in reqs])
    assert (np.diff(t) > 0).all()
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.05 / 20.0 * 4      # 4000 exponential gaps
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1           # exponential: cv = 1
    for q in reqs[:200]:
        s = q.spec
        assert s.W in (128, 256, 512, 1024) and s.H == s.W and s.D == s.W // 2
        assert len(s.tilts) in (41, 61, 121) and s.tilts[0] == -60.0 and s.tilts[-1] == 60.0
        assert 10 <= q.iterations <= 30 and s.storage in ("f32", "bf16") and 4 <= q.ellipsoids <= 20
    again = S.poisson_requests(50, rate=20.0, seed=7)
    assert [(q.arrival, q.spec.W, q.iterations) for q in again] == [(q.arrival, q.spec.W, q.iterations)
                                                                     for q in reqs[:50]]


def test_subtomogram_jobs_ranges():
    jobs = B.subtomogram_jobs(64)
    assert len(jobs) == 64
    sizes = {j.spec.W for j in jobs}
    assert sizes <= {64, 96, 128, 192, 256} and len(sizes) >= 4
    for j in jobs:
        s = j.spec
        assert s.W == s.H == s.D
        assert 41 <= len(s.tilts) <= 81
        assert 45.0 <= s.tilts[-1] <= 70.0 and s.tilts[0] == -s.tilts[-1]
        assert 4 <= j.ellipsoids <= 20


def test_percentile_nearest_rank():
    v = list(range(1, 101))
    assert S.percentile(v, 50) == 50 and S.percentile(v, 99) == 99 and S.percentile(v, 100) == 100
    assert S.percentile([3.0], 99) == 3.0
    assert np.isnan(S.percentile([], 50))


class FakeEngine:
    """Stands in for JobEngine: memory accounting and one tick of fixed cost."""

    def __init__(self, budget, clock, tick_s=1.0):
        self.budget, self.used, self.clock, self.tick_s = budget, 0, clock, tick_s
        self.jobs, self.next_id, self.runs = {}, 1, []

    def fits(self, spec):
        return self.used + spec.bytes() <= self.budget

    def add(self, spec):
        jid = self.next_id
        self.next_id += 1
        self.jobs[jid] = spec
        self.used += spec.bytes()
        return jid

    def synth(self, *a):
        pass

    def run(self, ids, iterations, graph=False):
        self.runs.append(list(ids))
        self.clock.t += self.tick_s

    def remove(self, jid):
        self.used -= self.jobs.pop(jid).bytes()

    def get_volume(self, jid):
        return np.zeros(1)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def test_serve_loop_admission_and_continuous_batching():
    spec = B.JobSpec(32, 32, 16, list(np.linspace(-60, 60, 11)))
    need = spec.bytes()
    clock = Clock()
    eng = FakeEngine(2 * need, clock)
    reqs = [S.Request(0, 0.0, spec, 3, 4, 1), S.Request(1, 0.0, spec, 1, 4, 2),
            S.Request(2, 0.0, spec, 2, 4, 3), S.Request(3, 0.5, spec, 1, 4, 4)]
    out = S.ServeLoop(eng, clock=clock).run(reqs, realtime=False)
    # two fit; the third joins as soon as the second finishes (tick 1), the
    # fourth when the first or third leaves
    assert eng.runs[0] == [1, 2]
    assert eng.runs[1] == [1, 3]
    assert 4 in eng.runs[2] + eng.runs[3]
    assert out["max_active"] == 2 and out["requests"] == 4 and eng.used == 0
    lat = {d.rid: d.latency for d in out["done"]}
    assert lat[1] == 1.0 and lat[0] == 3.0 and lat[2] == 3.0
    assert out["p99_s"] == max(lat.values())


def test_serve_loop_rejects_oversized_request():
    spec = B.JobSpec(64, 64, 32, list(np.linspace(-60, 60, 11)))
    eng = FakeEngine(spec.bytes() - 1, Clock())
    with pytest.raises(ValueError):
        S.ServeLoop(eng).run([S.Request(0, 0.0, spec, 1, 4, 1)], realtime=False)


def test_mix_signature_is_order_free():
    a = B.JobSpec(64, 64, 64, [0.0] * 41)
    b = B.JobSpec(128, 128, 128, [0.0] * 61, storage="bf16")
    assert B.mix_signature([a, b]) == B.mix_signature([b, a])
    assert B.mix_signature([a]) != B.mix_signature([a, b])


def test_reconstruction_parameters_factories():
    from tomoj_amd import parameters as RP
    p = RP.ReconstructionParameters.createOSSARTParameters(64, 32, 16, 10, 1.0, 13)
    assert (p.reconstructionType, p.nbIterations, p.relaxationCoefficient, p.updateNb) == (RP.OSSART, 10, 1.0, 13)
    assert np.isnan(p.weightingRadius) and p.rescaleData and not p.fista and not p.positivityConstraint
    assert p.volume_centers() == [31.5, 15.5, 7.5]
    q = RP.ReconstructionParameters.createOSSARTParameters(64, 32, 16, 1, 1.0, 0, [1.0, -2.0, 0.5])
    assert q.volume_centers() == [32.5, 13.5, 8.0]
    w = RP.ReconstructionParameters.createWBPParameters(64, 32, 16, 0.5)
    assert w.reconstructionType == RP.WBP and w.nbIterations == 0 and np.isnan(w.relaxationCoefficient)
    for r in (float("nan"), 0.0, -1.0):
        assert RP.ReconstructionParameters.createWBPParameters(64, 32, 16, r).reconstructionType == RP.BP
    s = RP.ReconstructionParameters.createOSSARTParameters(8, 8, 8, 3, -1.0, 0)
    assert [s.getRelaxationCoefficient() for _ in range(3)] == [1.0, 0.5, 1.0 / 3]
    s.resetTmpNbCall()
    assert s.getRelaxationCoefficient() == 1.0


def test_bgart_parameters_and_slab_refusal():
    """createBgARTParameters (AdvancedReconstructionParameters.java:145-155)
    keeps the OS-SART settings and the BgART flags; the slab driver refuses
    the regularised types before touching a device."""
    from tomoj_amd import parameters as RP
    from tomoj_amd import slabs
    p = RP.AdvancedReconstructionParameters.createBgARTParameters(64, 32, 16, 7, 0.6, 5, 1.25, False, True)
    assert p.reconstructionType == RP.BGART == 6
    assert (p.nbIterations, p.relaxationCoefficient, p.updateNb) == (7, 0.6, 5)
    assert (p.bgartK, p.bgartMedian, p.bgartDarkBG) == (1.25, False, True)
    assert np.isnan(p.weightingRadius)
    q = RP.AdvancedReconstructionParameters(8, 8, 8)
    assert (q.bgartK, q.bgartMedian, q.bgartDarkBG) == (1.0, True, True)   # :29-31
    ts = RP.TiltSeriesData(np.zeros((3, 32, 64), np.float32), [-10.0, 0.0, 10.0])
    for t in (RP.TVM, RP.BAYESIAN, RP.BGART):
        p.reconstructionType = t
        with pytest.raises(ValueError):
            slabs.reconstruct_in_slabs(ts, p, slab_rows=8)


class _FakeEngine:
    """Stands in for JobEngine on the CPU: records runs, counts iterations."""

    def __init__(self, budget):
        self.budget = budget
        self.used = 0
        self.specs = {}
        self.iters = {}
        self.calls = []
        self._next = 1

    def fits(self, spec):
        return self.used + spec.bytes() <= self.budget

    def add(self, spec):
        j = self._next
        self._next += 1
        self.specs[j] = spec
        self.iters[j] = 0
        self.used += spec.bytes()
        return j

    def synth(self, *a):
        pass

    def run(self, ids, k, graph=False):
        self.calls.append((tuple(ids), k))
        for j in ids:
            self.iters[j] += k
        return 0.0

    def get_volume(self, j):
        return np.full(1, self.iters[j])

    def remove(self, j):
        self.used -= self.specs.pop(j).bytes()


def _spec(W, n=41):
    from tomoj_amd.batch import JobSpec

    class S(JobSpec):
        def bytes(self):
            return self.W * self.H * self.D * 4
    return S(W, W, W // 2, list(np.linspace(-60, 60, n)))


def test_serving_tick_gives_small_jobs_several_iterations():
    """A tick gives each job about one quantum of voxel-ray updates (the
    largest job's iteration), never more iterations than it has left; the
    group with the least remaining work runs first."""
    big, mid, small = _spec(256), _spec(64), _spec(32)
    reqs = [S.Request(0, 0.0, big, 5, 4, 1), S.Request(1, 0.0, mid, 30, 4, 2), S.Request(2, 0.0, small, 3, 4, 3)]
    loop = S.ServeLoop(_FakeEngine(1 << 40))
    active = {1: reqs[0], 2: reqs[1], 3: reqs[2]}
    left = {1: 5, 2: 30, 3: 3}
    plan = loop.plan_tick(active, left)
    ks = {j: k for k, js in plan for j in js}
    assert ks[1] == 1                       # the largest job: one iteration
    assert ks[2] == 30                      # 64 = 256/4: 64 iterations fit the quantum, 30 left
    assert ks[3] == 3                       # capped by what it has left
    assert plan[0][1] == [3]                # least remaining work first


def test_serving_loop_runs_each_request_exactly_its_iterations():
    reqs = [S.Request(i, 0.0, _spec(w), it, 4, i) for i, (w, it) in
            enumerate([(256, 4), (32, 17), (64, 9), (128, 2), (32, 1)])]
    eng = _FakeEngine(1 << 40)
    out = S.ServeLoop(eng, keep_volumes=True).run(reqs, realtime=False)
    assert out["requests"] == len(reqs) and eng.used == 0
    for d in out["done"]:
        assert int(d.volume[0]) == reqs[d.rid].iterations
    assert set(out["by_width"]) == {"32", "64", "128", "256"}
    # the small jobs finish before the large one
    fin = {d.rid: d.finished for d in out["done"]}
    assert fin[1] <= fin[0] and fin[4] <= fin[0]


def test_serving_rejects_requests_without_iterations():
    reqs = [S.Request(0, 0.0, _spec(32), 0, 4, 1)]
    with pytest.raises(ValueError):
        S.ServeLoop(_FakeEngine(1 << 40)).run(reqs, realtime=False)
