"""The GPU manager's lifecycle (ProjectorGPU: initForBP, initForIterative,
createErrorVolume, updateFromGPU with the usable rows, copyInGPUBuffer,
releaseCL_Memory) and ReconstructionThreadManagerGPU.doGPURec's slab loop over
the HIP library: with slabs much smaller than the volume, every result equals
the whole-volume reconstruction of the oracle bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(oracle, W=40, H=17, D=12, seed=9):
    from tomoj_amd import parameters as RP
    tilts = list(np.arange(-60, 61, 10.0))
    E = oracle.tilt_eulers(tilts)
    exps = oracle.phantom_project(oracle.ellipsoids(seed, 5, W, H, D), E, W, H).astype(np.float32)
    return RP.TiltSeriesData(exps, tilts), E, exps


def _manager(ts, W, H, D, weighting=None):
    from tomoj_amd import gpu_manager as GM
    rec = np.zeros((D, H, W), np.float32)
    return rec, GM.VoxelProjector3DHIPGPU(ts, rec, [(W - 1) / 2.0, (H - 1) / 2.0, (D - 1) / 2.0],
                                          weighting_diameter=weighting)


@pytest.mark.parametrize("rows,positivity", [(5, False), (8, True)])
def test_dogpurec_ossart_slabs(tj, oracle, rows, positivity):
    """OS-SART through doGPURec in slabs of `rows` rows (17 = 5+5+5+2 or
    8+8+1): volume bit-exact with the whole-volume oracle; each slab's errors
    cover its own detector rows, so their squares add up to the whole error."""
    from tomoj_amd import gpu_manager as GM, parameters as RP
    W, H, D = 40, 17, 12
    ts, E, exps = _case(oracle, W, H, D)
    p = RP.ReconstructionParameters.createOSSARTParameters(W, H, D, 2, 1.0, 4)
    p.positivityConstraint = positivity
    rec, P = _manager(ts, W, H, D)
    slab_err = []
    print(f"doGPURec OS-SART, slabs of {rows} rows", flush=True)
    errs = GM.doGPURec(rec, ts, [P], p, first_index=3, max_bytes=rows * W * D * 4, slab_errors=slab_err)
    g = oracle.make_geom(W, H, D, positivity=positivity, nthreads=8)
    ref = np.zeros((D, H, W), np.float32)
    ref_err = oracle.ossart(g, ref, exps, E, 2, 1.0, 4, first_index=3)
    assert np.array_equal(rec, ref)
    assert len(slab_err) == -(-H // rows)
    whole2 = sum(e[:, 0] ** 2 for e in slab_err)
    assert np.allclose(whole2, ref_err[:, 0] ** 2, rtol=1e-12)
    assert np.allclose(errs, sum(e[:, 0] for e in slab_err), rtol=1e-15)
    assert P.P is None                                  # releaseCL_Memory after every slab


@pytest.mark.parametrize("weighted", [False, True])
def test_dogpurec_wbp_slabs(tj, oracle, weighted):
    """BP and WBP (Radermacher, diameter 0.5 W, applied to the whole images
    before the slabs are cut) through initForBP: identical to the
    whole-volume pass."""
    from tomoj_amd import gpu_manager as GM, parameters as RP
    W, H, D = 40, 17, 12
    ts, E, exps = _case(oracle, W, H, D, seed=2)
    p = RP.ReconstructionParameters.createWBPParameters(W, H, D, 0.5 if weighted else float("nan"))
    rec, P = _manager(ts, W, H, D, weighting=0.5 * W if weighted else None)
    GM.doGPURec(rec, ts, [P], p, max_bytes=6 * W * D * 4)
    with tj.VoxelProjector3DHIP((W, H, D)) as Q:
        Q.load_tilt_series(exps, E)
        if weighted:
            Q.setWeightingFilter(0.5 * W, ts.tilts)
        Q.run_wbp()
        whole = Q.get_volume()
    if weighted:
        assert np.allclose(rec, whole, rtol=1e-6, atol=1e-6 * np.abs(whole).max())
    else:
        assert np.array_equal(rec, whole)
        ref = np.zeros((D, H, W), np.float32)
        oracle.wbp(oracle.make_geom(W, H, D, nthreads=8), ref, exps, E)
        assert np.array_equal(rec, ref)


def test_dogpurec_error_volume_and_fista(tj, oracle):
    """doGPURec's saveError branch (n - 1 iterations, createErrorVolume, the
    last one) in slabs: the assembled error volume of the last iteration and
    the volume equal the oracle's whole-volume ones; FISTA run inside the
    library over slabs equals the whole-volume FISTA loop."""
    from tomoj_amd import gpu_manager as GM, parameters as RP
    W, H, D = 40, 17, 12
    ts, E, exps = _case(oracle, W, H, D, seed=4)
    g = oracle.make_geom(W, H, D, nthreads=8)
    p = RP.ReconstructionParameters.createOSSARTParameters(W, H, D, 3, 1.0, 6)
    rec, P = _manager(ts, W, H, D)
    GM.doGPURec(rec, ts, [P], p, saveError=True, max_bytes=7 * W * D * 4)
    ref = np.zeros((D, H, W), np.float32)
    oracle.ossart(g, ref, exps, E, 2, 1.0, 6)
    ev = np.zeros((D, H, W), np.float32)
    oracle.ossart(g, ref, exps, E, 1, 1.0, 6, errvol=ev)
    assert np.array_equal(rec, ref)
    assert sorted(P.errorVolumes) == [3]               # each slab's last iteration is its third
    assert np.array_equal(P.errorVolume, ev)

    f = RP.ReconstructionParameters.createOSSARTParameters(W, H, D, 3, 1.0, 0)
    rec2, P2 = _manager(ts, W, H, D)
    P2.setFista(True)
    GM.doGPURec(rec2, ts, [P2], f, max_bytes=5 * W * D * 4)
    ref2 = np.zeros((D, H, W), np.float32)
    t = 0.0
    for _ in range(3):
        prev = ref2.copy()
        oracle.ossart(g, ref2, exps, E, 1, 1.0, 0)
        t = oracle.fista_step(ref2, prev, t)
    assert np.array_equal(rec2, ref2)
