"""Serving loop on the GPU: continuous batching under an HBM budget, and each
request's volume identical to the same request reconstructed alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_requests(S, n=8):
    reqs = S.poisson_requests(n, rate=50.0, seed=11, widths=(32, 48, 64), tilt_counts=(11, 21),
                              iterations=(2, 4), storages=("f32", "bf16", "f16"))
    return reqs


def test_results_same_alone_and_in_mix(tj):
    from tomoj_amd import batch as B, serving as S
    reqs = small_requests(S)
    budget = 3 * max(q.spec.bytes() for q in reqs)     # forces queueing
    with B.JobEngine(budget) as eng:
        out = S.ServeLoop(eng, keep_volumes=True).run(reqs, realtime=False)
        assert eng.used == 0
    assert out["requests"] == len(reqs)
    assert 1 < out["max_active"] < len(reqs)           # shared launches, and a queue
    assert np.isfinite(out["p50_s"]) and out["p99_s"] >= out["p50_s"] > 0
    for d in out["done"]:
        q = reqs[d.rid]
        with B.JobEngine(budget) as eng:
            alone = S.run_alone(eng, q)
        assert np.array_equal(alone, d.volume), q.rid


def test_graph_ticks_match_plain_ticks(tj):
    from tomoj_amd import batch as B, serving as S
    reqs = small_requests(S, 5)
    budget = 8 * max(q.spec.bytes() for q in reqs)
    vols = []
    for graph in (False, True):
        with B.JobEngine(budget) as eng:
            out = S.ServeLoop(eng, keep_volumes=True, graph=graph).run(reqs, realtime=False)
        vols.append([d.volume for d in out["done"]])
    for a, b in zip(*vols):
        assert np.array_equal(a, b)


def test_realtime_arrivals(tj):
    from tomoj_amd import batch as B, serving as S
    reqs = small_requests(S, 6)
    with B.JobEngine(1 << 30) as eng:
        out = S.ServeLoop(eng).run(reqs, realtime=True)
    for d in out["done"]:
        assert d.admitted >= d.arrival - 1e-9 and d.finished > d.admitted
