"""The benchmark configurations at their own sizes on the GPU:

  512x512x256 / 121 tilts   the weighted backprojection pass of the production
                            workload (Radermacher weighting against the oracle,
                            ramp x Hamming against its closed form)
  1024x1024x512 / 121 tilts the precision form: BF16 and FP16 storage with an
                            FP32 master volume and FP32 projections, 30
                            iterations then 2 FP32 refinement iterations,
                            against the FP64 twin on the device
  64 subtomograms           the batched form: one graph launch over uneven jobs,
                            each job bit-identical alone and in the mix, the
                            smallest FP32 jobs against the FP64 oracle
  Poisson stream            the serving form at detector widths 128..1024 in
                            FP32 and BF16, every result identical alone and in
                            the mix
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def rel_l2(a, b):
    num = den = 0.0
    for k in range(a.shape[0]):                 # slice by slice: bounded host memory
        d = np.asarray(a[k], np.float64) - np.asarray(b[k], np.float64)
        num += float(np.vdot(d, d))
        den += float(np.vdot(np.asarray(b[k], np.float64), np.asarray(b[k], np.float64)))
    return float(np.sqrt(num / max(den, 1e-300)))


@pytest.fixture(scope="module")
def production(tj):
    """The 512x512x256, 121-tilt tilt series of the production workload
    (seeded 40-ellipsoid phantom, analytic projections + 1% noise, generated on
    the device), read back once."""
    from tomoj_amd import workloads as WL
    w = WL.OSSART_512
    E = WL.eulers_for(w)
    with tj.VoxelProjector3DHIP((w.W, w.H, w.D)) as P:
        P.synth_tilt_series(tj.phantom_ellipsoids(w.seed, w.ellipsoids, w.W, w.H, w.D), E, w.noise_rel,
                            w.noise_seed)
        exps = np.stack([P.get_loaded_projection(k) for k in range(w.n)])
    return w, E, exps


def test_wbp_radermacher_production_size(tj, oracle, production):
    """TomoReconstruction2.WBP with FFTWeighting (diameter 0.5 W) over the 121
    projections of the 512x512x256 workload: the device volume within 1e-5
    (rel-L2) of the FP32 oracle (FP32 transforms on both sides; parallelcolt's
    own rounding is not available here)."""
    w, E, exps = production
    dia = 0.5 * w.W
    print(f"WBP Radermacher {w.W}x{w.H}x{w.D}, {w.n} tilts", flush=True)
    with tj.VoxelProjector3DHIP((w.W, w.H, w.D)) as P:
        P.load_tilt_series(exps, E)
        P.setWeightingFilter(dia, w.tilts)
        mask = P.getWeightingMask()
        P.run_wbp()
        got = P.get_volume()
    ref_mask = oracle.weighting_mask(w.W, w.H, w.tilts, dia)
    assert rel_l2(mask[None], ref_mask[None]) < 1e-7
    g = oracle.make_geom(w.W, w.H, w.D, nthreads=16)
    ref = np.zeros((w.D, w.H, w.W), np.float32)
    oracle.wbp(g, ref, exps, E, ref_mask)
    e = rel_l2(got, ref)
    print(f"WBP Radermacher rel-L2 vs the FP32 oracle {e:.2e}")
    assert e <= 1e-5, e


def test_wbp_ramp_hamming_production_size(tj, oracle, production):
    """Ramp x Hamming weighting over the whole pass: each image row filtered by
    H(k) = (pi/N) f (0.54 + 0.46 cos(pi f / fc)), f = |k|/W <= fc = 0.5
    (numpy FP64 FFT), then the oracle's FP64 backprojection of the filtered
    images; the device volume (FP32 transforms) within 1e-5."""
    w, E, exps = production
    print(f"WBP ramp x Hamming {w.W}x{w.H}x{w.D}, {w.n} tilts", flush=True)
    with tj.VoxelProjector3DHIP((w.W, w.H, w.D)) as P:
        P.load_tilt_series(exps, E)
        P.setRampFilter(hamming=True, cutoff=1.0)
        P.run_wbp()
        got = P.get_volume()
    k = np.arange(w.W)
    ci = np.where(k > w.W / 2.0, k - w.W, k)
    f = np.abs(ci) / w.W
    fc = 0.5
    Hk = np.where(f <= fc, (np.pi / w.n) * f * (0.54 + 0.46 * np.cos(np.pi * f / fc)), 0.0)
    filt = np.real(np.fft.ifft(np.fft.fft(exps.astype(np.float64), axis=2) * Hk, axis=2))
    g = oracle.make_geom(w.W, w.H, w.D, nthreads=16)
    ref = np.zeros((w.D, w.H, w.W), np.float64)
    oracle.wbp(g, ref, filt, E, None, f64=True)
    e = rel_l2(got, ref)
    print(f"WBP ramp x Hamming rel-L2 vs the closed form {e:.2e}")
    assert e <= 1e-5, e


def test_precision_form_1024x1024x512(tj):
    """The precision configuration in full: 1024x1024x512, 121 tilts, 80
    ellipsoids, 2% noise; 30 SIRT iterations in 16-bit storage, then 2 FP32
    iterations; rel-L2 against the FP64 twin after the same 32 iterations.
    BF16 and FP16 with an FP32 master volume and FP32 projections must meet
    1e-3; pure BF16 storage is reported (its 2^-9 rounding of the iterate sets a
    floor above 1e-3)."""
    from tomoj_amd import forms
    modes = [m for m in forms.PRECISION_MODES if m[0] in ("bf16+master+p32", "f16+master+p32", "bf16")]
    print("precision form 1024x1024x512, 121 tilts, 30 + 2 iterations", flush=True)
    res = forms.precision_study(1024, 1024, 512, 121, iterations=30, refine=2, ellipsoids=80, noise=0.02,
                                modes=modes, reference=True, tune=False)
    for name, rec in res["modes"].items():
        print(f"{name}: rel-L2 vs FP64 {rec['rel_l2_vs_f64']:.2e}, {rec['it_s']:.1f} it/s "
              f"({rec['vs_plain']:.2f}x the plain FP32 path)")
    assert res["modes"]["bf16+master+p32"]["rel_l2_vs_f64"] <= 1e-3
    assert res["modes"]["f16+master+p32"]["rel_l2_vs_f64"] <= 1e-3
    assert np.isfinite(res["modes"]["bf16"]["rel_l2_vs_f64"])


def _alone(B, job, iterations):
    with B.JobEngine(8 * GIB) as eng:
        jid = B.load_job(eng, job)
        eng.run([jid], iterations, graph=True)
        return eng.get_volume(jid)


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_batched_64_subtomograms(tj, oracle, storage):
    """The batched configuration: 64 subtomograms of {64..256}^3 with 41-81
    tilts over +-45..70 deg, 15 SIRT iterations of all of them in one graph
    launch (the autotuner's tile shapes).  The 4 smallest and the 2 largest
    jobs are bit-identical alone; in FP32, three 64^3 jobs are within 1e-5 of
    the FP64 oracle."""
    from tomoj_amd import batch as B
    iters = 15
    jobs = B.subtomogram_jobs(64, storage=storage)
    for jb in jobs:
        jb.spec.max_iterations = 2 * iters
    print(f"batched 64 jobs, {storage}", flush=True)
    with B.JobEngine(200 * GIB) as eng:
        ids = [B.load_job(eng, jb) for jb in jobs]
        eng.autotune(ids)
        for j in ids:
            eng.reset_volume(j)
        eng.run(ids, iters, graph=True)
        work = [jb.spec.voxel_ray_updates_per_iteration() for jb in jobs]
        order = sorted(range(len(jobs)), key=lambda k: work[k])
        picks = order[:4] + order[-2:]
        mixed = {k: eng.get_volume(ids[k]) for k in picks}
        small64 = [k for k in order if jobs[k].spec.W == 64][:3] if storage == "f32" else []
        projs = {k: np.stack([eng.get_projection(ids[k], p) for p in range(len(jobs[k].spec.tilts))])
                 for k in small64}
        for k in small64:
            mixed.setdefault(k, eng.get_volume(ids[k]))
    for k in picks:
        assert np.array_equal(_alone(B, jobs[k], iters), mixed[k]), (k, jobs[k].spec.W)
    for k in small64:
        s = jobs[k].spec
        E = oracle.tilt_eulers(s.tilts)
        g = oracle.make_geom(s.W, s.H, s.D, nthreads=16)
        ref = np.zeros((s.D, s.H, s.W), np.float64)
        oracle.ossart(g, ref, projs[k], E, iters, 1.0, len(s.tilts), first_index=0, f64=True)
        e = rel_l2(mixed[k], ref)
        print(f"job {k} ({s.W}^3, {len(s.tilts)} tilts): rel-L2 vs FP64 oracle {e:.2e}")
        assert e <= 1e-5, e


def test_serving_stream_all_widths(tj):
    """The serving configuration's stream (seeded Poisson arrivals, detectors
    128..1024 with depth W/2, 41/61/121 tilts, 10-30 iterations, FP32 and BF16)
    served with graph ticks under a 96 GiB budget: every request's volume is
    bit-identical to the same request reconstructed alone."""
    from tomoj_amd import batch as B, serving as S
    reqs = S.poisson_requests(40, rate=20.0)
    widths = {1024: 0, 512: 0, 256: 0, 128: 0}
    chosen = []
    for q in reqs:                          # the first two requests of each width, and of each storage
        if widths[q.spec.W] < 2:
            widths[q.spec.W] += 1
            chosen.append(q)
    assert {q.spec.storage for q in chosen} == {"f32", "bf16"}
    assert sorted({q.spec.W for q in chosen}) == [128, 256, 512, 1024]
    print("serving " + ", ".join(f"{q.spec.W}/{len(q.spec.tilts)}/{q.iterations}/{q.spec.storage}"
                                 for q in chosen), flush=True)
    with B.JobEngine(96 * GIB) as eng:
        out = S.ServeLoop(eng, keep_volumes=True, graph=True).run(chosen, realtime=False)
        assert eng.used == 0
    assert out["requests"] == len(chosen)
    for d in out["done"]:
        q = reqs[d.rid]
        with B.JobEngine(96 * GIB) as eng:
            alone = S.run_alone(eng, q)
        assert np.array_equal(alone, d.volume), (q.rid, q.spec.W, q.spec.storage)
