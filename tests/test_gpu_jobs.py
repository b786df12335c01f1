"""Job engine (precision, batched and serving forms) against the oracle (GPU tests).

The job kernels evaluate TomoJ's SIRT with closed-form positions in the
arithmetic type, so they are held to tolerances against the oracle's FP64
twin rather than bit for bit:
  FP64 storage           rel-L2 <= 1e-9  (same algorithm, FP64 throughout)
  FP32 storage           rel-L2 <= 1e-5  (the parity bound for FP results)
  BF16 / FP16 storage    rel-L2 <= 1e-3 after 2 FP32 refinement iterations
Everything that does not change the arithmetic (other jobs in the launch, tile
shapes, graph capture) must leave a job's result bit-identical.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def batch(tj):
    from tomoj_amd import batch as B
    return B


def phantom_case(oracle, W=48, H=20, D=24, tilts=None, seed=3, k=6):
    tilts = list(np.arange(-60, 61, 6.0)) if tilts is None else list(tilts)
    E = oracle.tilt_eulers(tilts)
    ell = oracle.ellipsoids(seed, k, W, H, D)
    exps = oracle.phantom_project(ell, E, W, H).astype(np.float32)
    return tilts, E, exps


def oracle_sirt(oracle, W, H, D, exps, E, iters, f64=True, positivity=False, vol0=None):
    g = oracle.make_geom(W, H, D, positivity=positivity, nthreads=8)
    dt = np.float64 if f64 else np.float32
    vol = np.zeros((D, H, W), dt) if vol0 is None else np.array(vol0, dt)
    err = oracle.ossart(g, vol, exps, E, iters, 1.0, len(E), first_index=0, f64=f64)
    return vol, err


def run_job(batch, W, H, D, tilts, exps, iters, storage="f32", positivity=False, engine=None, graph=False):
    own = engine is None
    eng = batch.JobEngine(4 * GIB) if own else engine
    try:
        jid = eng.add(batch.JobSpec(W, H, D, tilts, storage=storage, positivity=positivity))
        eng.set_projections(jid, exps)
        eng.run([jid], iters, graph=graph)
        return eng.get_volume(jid), eng.errors(jid), jid
    finally:
        if own:
            eng.close()


def test_f64_storage_matches_fp64_oracle(batch, oracle):
    tilts, E, exps = phantom_case(oracle)
    got, err, _ = run_job(batch, 48, 20, 24, tilts, exps, 4, storage="f64")
    ref, ref_err = oracle_sirt(oracle, 48, 20, 24, exps, E, 4)
    assert rel_l2(got, ref) <= 1e-9, rel_l2(got, ref)
    som = ref_err[:, 0] ** 2 * len(E) * 48 * 20
    assert np.allclose(err, som, rtol=1e-9)


@pytest.mark.parametrize("shape", [(48, 20, 24), (300, 6, 17), (33, 9, 40)])
def test_f32_storage_within_parity_bound(batch, oracle, shape):
    W, H, D = shape
    tilts, E, exps = phantom_case(oracle, W, H, D, tilts=np.arange(-57, 58, 9.5))
    got, err, _ = run_job(batch, W, H, D, tilts, exps, 3, storage="f32")
    ref, _ = oracle_sirt(oracle, W, H, D, exps, E, 3)
    assert rel_l2(got, ref) <= 1e-5, rel_l2(got, ref)


@pytest.mark.parametrize("shape", [(3071, 4, 3), (3073, 3, 2)], ids=["3071-shared", "3073-one-group"])
def test_f32_storage_at_the_shared_forward_width_limit(batch, oracle, shape):
    """3071 columns: the widest shared FP32 forward slab (one slice per slab);
    3073: the one-group kernel takes the launch.  Both forms give
    bit-identical volumes.  The FP32 positions round to an ulp of 2^-12 px at
    x ~ 3000, ten times that of 300 columns, so the bound against the FP64
    oracle is 1e-4 here (3.5e-5 and 4.5e-5 measured)."""
    W, H, D = shape
    tilts, E, exps = phantom_case(oracle, W, H, D, tilts=np.arange(-57, 58, 9.5))
    vols = []
    for sh in (1, 0):
        with batch.JobEngine(GIB) as eng:
            eng.set_tuning(fp_shared=sh)
            jid = eng.add(batch.JobSpec(W, H, D, tilts, storage="f32"))
            eng.set_projections(jid, exps)
            eng.run([jid], 3)
            vols.append(eng.get_volume(jid))
    assert np.array_equal(vols[0], vols[1])
    ref, _ = oracle_sirt(oracle, W, H, D, exps, E, 3)
    assert rel_l2(vols[0], ref) <= 1e-4, rel_l2(vols[0], ref)


def test_positivity_and_taller_volume(batch, oracle):
    """Positivity after each image; a volume taller than the detector (rows
    outside the detector stay untouched, rays through them are NaN)."""
    W, H, D = 40, 14, 16
    tilts = list(np.arange(-45, 46, 15.0))
    E = oracle.tilt_eulers(tilts)
    ell = oracle.ellipsoids(11, 5, W, 10, D)
    exps = oracle.phantom_project(ell, E, W, 10).astype(np.float32) - np.float32(2.0)
    with batch.JobEngine(GIB) as eng:
        jid = eng.add(batch.JobSpec(W, H, D, tilts, storage="f64", positivity=True, pw=W, ph=10))
        eng.set_projections(jid, exps)
        eng.run([jid], 3)
        got = eng.get_volume(jid)
    g = oracle.make_geom(W, H, D, W, 10, positivity=True, nthreads=8)
    ref = np.zeros((D, H, W), np.float64)
    oracle.ossart(g, ref, exps, E, 3, 1.0, len(E), f64=True)
    assert rel_l2(got, ref) <= 1e-9
    assert (got >= 0).all()


def test_differences_match_oracle(batch, oracle):
    """The fused forward projection + residual gives the oracle's difference images."""
    W, H, D = 64, 8, 30
    tilts, E, exps = phantom_case(oracle, W, H, D, tilts=[-50.0, -12.5, 0.0, 33.0])
    rng = np.random.default_rng(2)
    vol = rng.random((D, H, W)).astype(np.float32)
    with batch.JobEngine(GIB) as eng:
        jid = eng.add(batch.JobSpec(W, H, D, tilts, storage="f64"))
        eng.set_projections(jid, exps)
        eng.set_volume(jid, vol)
        eng.run([jid], 1)
        diffs = np.stack([eng.get_difference(jid, p) for p in range(len(tilts))])
    g = oracle.make_geom(W, H, D, nthreads=8)
    factor = 1.0 / len(tilts)   # the FP64 twin keeps lambda/N in double
    ref, _, _, _ = oracle.project_and_difference(g, vol.astype(np.float64), exps, E, factor, f64=True)
    assert rel_l2(diffs, ref) <= 1e-12, rel_l2(diffs, ref)


# (storage, FP32 master volume, FP32 projections, bound on rel-L2 after refinement)
HALF_MODES = [
    ("bf16", True, True, 1e-3),    # 16-bit volume read by the forward projection
    ("f16", True, True, 1e-3),
    ("f16", True, False, 1e-3),    # volume and projections 16-bit
    ("f16", False, False, 1e-3),
    ("bf16", True, False, 3e-3),   # BF16 projections: 2^-9 data rounding
    ("bf16", False, False, 1e-2),  # BF16 iterate: 2^-9 volume rounding every iteration
]


@pytest.mark.parametrize("mode", HALF_MODES, ids=[f"{m[0]}{'-master' if m[1] else ''}{'-p32' if m[2] else ''}"
                                                  for m in HALF_MODES])
def test_half_storage_with_refinement(batch, oracle, mode):
    """Half-precision storage for 10 iterations, then 2 FP32 iterations against
    the FP32 projections, compared with 12 FP64 iterations of the oracle."""
    storage, master, p32, bound = mode
    W, H, D = 64, 16, 32
    tilts, E, exps = phantom_case(oracle, W, H, D, tilts=np.arange(-60, 61, 4.0), k=8)
    with batch.JobEngine(GIB) as eng:
        jid = eng.add(batch.JobSpec(W, H, D, tilts, storage=storage, fp32_master=master, fp32_projections=p32))
        eng.set_projections(jid, exps)
        eng.run([jid], 10)
        half = eng.get_volume(jid)
        eng.convert(jid, "f32")
        eng.set_projections(jid, exps)
        eng.run([jid], 2)
        got = eng.get_volume(jid)
        assert eng.iterations_done(jid) == 12
    ref, _ = oracle_sirt(oracle, W, H, D, exps, E, 12)
    ref10, _ = oracle_sirt(oracle, W, H, D, exps, E, 10)
    e_half, e_ref = rel_l2(half, ref10), rel_l2(got, ref)
    print(f"{mode}: rel-L2 after 10 half iterations {e_half:.2e}, after refinement {e_ref:.2e}")
    assert e_ref <= bound, e_ref


# 16-bit forms on shapes that take the other code paths of their kernels:
# (W, H, D, pw, ph, storage, positivity)
AWKWARD = [
    (45, 13, 22, 45, 13, "bf16", False),  # odd width (element staging), rows past the volume in the last tile
    (64, 14, 30, 64, 10, "f16", True),    # detector shorter than the volume, positivity after each image
    (50, 16, 17, 40, 16, "bf16", True),   # detector narrower than the volume, positivity
    (96, 24, 40, 96, 24, "f16", False),   # word staging, several row tiles
    (2048, 6, 5, 2048, 6, "bf16", False),  # widest double-buffered slab: one slice per slab
    (2050, 5, 4, 2050, 5, "f16", False),   # wider: the single-buffer kernel takes both forms
]


@pytest.mark.parametrize("case", AWKWARD, ids=[f"{c[0]}x{c[1]}x{c[2]}-{c[3]}x{c[4]}-{c[5]}{'-pos' if c[6] else ''}"
                                              for c in AWKWARD])
def test_half_storage_awkward_shapes(batch, oracle, case):
    """BF16/FP16 storage with FP32 master volume and projections, 10 iterations
    then 2 FP32 iterations, against 12 FP64 oracle iterations: rel-L2 <= 1e-3."""
    W, H, D, pw, ph, storage, pos = case
    tilts = list(np.arange(-60, 61, 5.0))
    E = oracle.tilt_eulers(tilts)
    ell = oracle.ellipsoids(17, 6, pw, ph, D)
    exps = oracle.phantom_project(ell, E, pw, ph).astype(np.float32)
    if pos:
        exps -= np.float32(0.5 * exps.mean())
    vols = []
    for packed in (1, 2):   # both packed forward kernels (form 2 only takes even widths)
        with batch.JobEngine(GIB) as eng:
            eng.set_tuning(fp_packed=packed)
            jid = eng.add(batch.JobSpec(W, H, D, tilts, storage=storage, positivity=pos, fp32_master=True,
                                        fp32_projections=True, pw=pw, ph=ph))
            eng.set_projections(jid, exps)
            eng.run([jid], 10)
            eng.convert(jid, "f32")
            eng.set_projections(jid, exps)
            eng.run([jid], 2)
            vols.append(eng.get_volume(jid))
    assert np.array_equal(vols[0], vols[1])
    got = vols[1]
    g = oracle.make_geom(W, H, D, pw, ph, positivity=pos, nthreads=8)
    ref = np.zeros((D, H, W), np.float64)
    oracle.ossart(g, ref, exps, E, 12, 1.0, len(E), first_index=0, f64=True)
    e = rel_l2(got, ref)
    print(f"{case}: rel-L2 {e:.2e}")
    assert e <= 1e-3, e
    if pos:
        assert (got >= 0).all()


def test_half_tile_shapes_do_not_change_results(batch, oracle):
    """The 16-bit kernels' tile and slab shapes partition the work without
    changing any voxel's arithmetic: bit-identical volumes."""
    W, H, D = 300, 21, 27   # two 256-column tiles (one shared block with fp_cols=2)
    tilts, E, exps = phantom_case(oracle, W, H, D, tilts=np.arange(-60, 61, 10.0))
    shapes = [dict(fp_rays=2, fp_lds_kb=16, bp_rows=2, bp_zchunk=8, blocks_per_cu=1),
              dict(fp_rays=4, fp_lds_kb=64, bp_rows=4, bp_zchunk=16, blocks_per_cu=8),
              dict(fp_rays=4, fp_lds_kb=128, bp_rows=8, bp_zchunk=8, blocks_per_cu=2),
              dict(fp_rays=2, fp_lds_kb=32, bp_rows=8, bp_zchunk=16, blocks_per_cu=4),
              dict(fp_rays=4, fp_lds_kb=48, bp_rows=4, bp_zchunk=4, blocks_per_cu=3),
              dict(fp_rays=4, fp_lds_kb=64, fp_cols=2, bp_rows=4, bp_zchunk=16, blocks_per_cu=4),
              dict(fp_rays=2, fp_lds_kb=32, fp_cols=2, bp_rows=8, bp_zchunk=8, blocks_per_cu=2),
              # the double-buffered packed kernel (four projection groups per block)
              dict(fp_packed=2, fp_rays=2, fp_lds_kb=152, bp_rows=4, bp_zchunk=8, blocks_per_cu=1),
              dict(fp_packed=2, fp_rays=2, fp_lds_kb=64, bp_rows=4, bp_zchunk=8, blocks_per_cu=2)]
    results = []
    for sh in shapes:
        with batch.JobEngine(GIB) as eng:
            eng.set_tuning(**{"fp_packed": 1, **sh})
            jid = eng.add(batch.JobSpec(W, H, D, tilts, storage="bf16", fp32_master=True, fp32_projections=True))
            eng.set_projections(jid, exps)
            eng.run([jid], 3)
            results.append(eng.get_volume(jid))
    for r in results[1:]:
        assert np.array_equal(results[0], r)


def test_result_same_alone_and_in_mix(batch, oracle):
    """A job's volume is bit-identical alone and in a mixed launch (other sizes
    and storage types), with and without graph capture."""
    W, H, D = 50, 12, 20
    tilts, E, exps = phantom_case(oracle, W, H, D, tilts=np.arange(-48, 49, 8.0))
    alone, _, _ = run_job(batch, W, H, D, tilts, exps, 4)
    with batch.JobEngine(4 * GIB) as eng:
        others = []
        for (w, h, d, st, seed) in [(96, 30, 40, "bf16", 1), (33, 7, 65, "f32", 2), (64, 64, 32, "f16", 3)]:
            t2 = list(np.linspace(-55, 55, 13))
            jid = eng.add(batch.JobSpec(w, h, d, t2, storage=st))
            ell = oracle.ellipsoids(seed, 5, w, h, d)
            eng.set_projections(jid, oracle.phantom_project(ell, oracle.tilt_eulers(t2), w, h).astype(np.float32))
            others.append(jid)
        me = eng.add(batch.JobSpec(W, H, D, tilts))
        eng.set_projections(me, exps)
        eng.run(others + [me], 2, graph=True)
        eng.run(others + [me], 2, graph=True)   # graph replay
        mixed = eng.get_volume(me)
    assert np.array_equal(alone, mixed)


@pytest.mark.parametrize("size", [(70, 10, 27), (300, 9, 37)], ids=["one-slab", "partial-slabs"])
def test_tile_shapes_do_not_change_results(batch, oracle, size):
    """FP32 storage: the one-group kernel (k_job_fp) in several shapes and the
    shared double-buffered one (k_job_fq; at 300 columns its slabs hold 10
    slices, so 37 slices end in a partial slab) give bit-identical volumes."""
    W, H, D = size
    tilts, E, exps = phantom_case(oracle, W, H, D, tilts=np.arange(-60, 61, 10.0))
    results = []
    shapes = [dict(fp_shared=0, fp_rays=2, fp_lds_kb=16, bp_rows=2, bp_zchunk=8, blocks_per_cu=1),
              dict(fp_shared=0, fp_rays=4, fp_lds_kb=64, bp_rows=4, bp_zchunk=8, blocks_per_cu=4),
              dict(fp_shared=0, fp_rays=4, fp_lds_kb=32, bp_rows=4, bp_zchunk=4, blocks_per_cu=8),
              dict(fp_shared=0, fp_rays=2, fp_lds_kb=128, bp_rows=2, bp_zchunk=16, blocks_per_cu=2),
              dict(fp_shared=1, bp_rows=4, bp_zchunk=8, blocks_per_cu=4)]
    for sh in shapes:
        with batch.JobEngine(GIB) as eng:
            eng.set_tuning(**sh)
            jid = eng.add(batch.JobSpec(W, H, D, tilts))
            eng.set_projections(jid, exps)
            eng.run([jid], 3)
            results.append(eng.get_volume(jid))
    for r in results[1:]:
        assert np.array_equal(results[0], r)


def test_f32_forward_forms_on_shorter_narrower_detector(batch, oracle):
    """FP32 storage with positivity on a detector shorter and narrower than the
    volume (rows outside the detector stay untouched, rays that miss the
    volume read zero cells): the shared double-buffered forward (k_job_fq,
    fp_shared 1) and the one-group one (k_job_fp, fp_shared 0) give
    bit-identical volumes, within the FP32 bound of the FP64 oracle."""
    W, H, D, pw, ph = 52, 14, 20, 44, 10
    tilts = list(np.arange(-50, 51, 10.0))
    E = oracle.tilt_eulers(tilts)
    ell = oracle.ellipsoids(13, 5, pw, ph, D)
    exps = oracle.phantom_project(ell, E, pw, ph).astype(np.float32) - np.float32(1.0)
    got = []
    for sh in (1, 0):
        with batch.JobEngine(GIB) as eng:
            eng.set_tuning(fp_shared=sh)
            jid = eng.add(batch.JobSpec(W, H, D, tilts, storage="f32", positivity=True, pw=pw, ph=ph))
            eng.set_projections(jid, exps)
            eng.run([jid], 3)
            got.append(eng.get_volume(jid))
    assert np.array_equal(got[0], got[1])
    g = oracle.make_geom(W, H, D, pw, ph, positivity=True, nthreads=8)
    ref = np.zeros((D, H, W), np.float64)
    oracle.ossart(g, ref, exps, E, 3, 1.0, len(E), f64=True)
    assert rel_l2(got[0], ref) <= 1e-5, rel_l2(got[0], ref)
    assert (got[0] >= 0).all()


def test_autotune_keeps_volumes_and_caches(batch, oracle, tmp_path):
    W, H, D = 64, 16, 24
    tilts, E, exps = phantom_case(oracle, W, H, D)
    with batch.JobEngine(GIB) as eng:
        jid = eng.add(batch.JobSpec(W, H, D, tilts))
        eng.set_projections(jid, exps)
        eng.run([jid], 2)
        before = eng.get_volume(jid)
        cache = str(tmp_path / "tune.json")
        tun = eng.autotune([jid], cache=cache)
        assert np.array_equal(before, eng.get_volume(jid))
        assert eng.get_tuning() == tun
        assert eng.autotune([jid], cache=cache) == tun   # from the cache
        eng.run([jid], 1)
        assert eng.iterations_done(jid) == 3


def test_budget_admission(batch):
    spec = batch.JobSpec(64, 64, 32, list(np.linspace(-60, 60, 41)))
    need = spec.bytes()
    with batch.JobEngine(2 * need + need // 2) as eng:
        a = eng.add(spec)
        b = eng.add(spec)
        assert eng.used == 2 * need
        with pytest.raises(batch._lib.TjError) as ei:
            eng.add(spec)
        assert ei.value.code == batch._lib.TJ_ENOMEM
        eng.remove(a)
        c = eng.add(spec)
        assert eng.used == 2 * need
        eng.run([b, c], 1)


def test_synthetic_job_matches_oracle_phantom(batch, oracle):
    W, H, D = 40, 12, 20
    tilts = list(np.linspace(-50, 50, 9))
    E = oracle.tilt_eulers(tilts)
    ell = oracle.ellipsoids(21, 7, W, H, D)
    clean = oracle.phantom_project(ell, E, W, H)
    with batch.JobEngine(GIB) as eng:
        jid = eng.add(batch.JobSpec(W, H, D, tilts))
        eng.synth(jid, ell, 0.0, 5)
        got = np.stack([eng.get_projection(jid, p) for p in range(len(tilts))])
    assert np.allclose(got, clean.astype(np.float32), rtol=1e-6, atol=1e-5)


def test_tunings_out_of_range_are_refused(batch, tj):
    """tj_batch_set_tuning and tj_set_tuning refuse values outside their ranges
    and keep the tuning they had; a new engine reports the documented defaults
    (packed double-buffered 16-bit forward, shared FP32 forward)."""
    from tomoj_amd import TjError
    with batch.JobEngine(GIB) as eng:
        before = eng.get_tuning()
        assert before["fp_packed"] == 2 and before["fp_shared"] == 1
        for bad in (dict(fp_shared=2), dict(fp_shared=-1), dict(fp_packed=3), dict(fp_lds_kb=200),
                    dict(blocks_per_cu=17)):
            with pytest.raises(TjError):
                eng.set_tuning(**bad)
            assert eng.get_tuning() == before
    with tj.VoxelProjector3DHIP((16, 8, 8)) as P:
        before = P.get_tuning()
        for bad in (dict(fp_prefetch=2), dict(fp_groups=3)):
            with pytest.raises(TjError):
                P.set_tuning(**bad)
            assert P.get_tuning() == before


def test_bad_jobs_are_rejected(batch, oracle):
    with batch.JobEngine(GIB) as eng:
        with pytest.raises(batch._lib.TjError):
            eng.add(batch.JobSpec(32, 16, 8, [0.0, 10.0], pw=32, ph=15))   # odd row shift
        tilted = oracle.tilt_eulers([0.0, 20.0], tilt_axis_deg=10.0, axis_vertical=False)
        with pytest.raises(batch._lib.TjError):
            eng.add(batch.JobSpec(32, 16, 8, [0.0, 20.0]), eulers=tilted)
        with pytest.raises(batch._lib.TjError):
            eng.run([12345], 1)
