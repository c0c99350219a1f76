"""Continuous batching: requests admitted while others run, repacked on the
device, give the drag history they give alone (the flow state is carried over
bit for bit; the force sums differ only in summation order)."""
import numpy as np
import pytest

from mortontrie_amr_amd import amr, cases as C, serve

pytestmark = pytest.mark.gpu


def small_request(i, arrival, rng, max_steps=300):
    dims = (48, 24, 24)
    r = float(rng.uniform(3.0, 5.0))
    case = C.sphere_amr(dims, radius=r, center=(16.0, 12.0, 12.0), shells=(3.0,), U=0.04,
                        Re=float(rng.uniform(20, 60)), options=2)
    return serve.Request(i, case, arrival, max_steps=max_steps, tol=0.0)


def check_against_alone(reqs, done, precision):
    assert sorted(r.id for r in done) == [r.id for r in reqs]
    for r in done:
        alone = serve.run_alone(r, precision=precision, chunk=100)
        assert r.steps == alone.steps
        assert np.allclose(r.cd, alone.cd, rtol=1e-9, atol=1e-12), r.id


def test_staggered_admission_matches_alone():
    """Arrivals at chunk 0, 1.5 and 2.5 force two repacks while members run."""
    rng = np.random.default_rng(5)
    reqs = [small_request(0, 0.0, rng), small_request(1, 1.5, rng), small_request(2, 2.5, rng)]
    srv = serve.Server(precision=amr.FP32_SHIFTED, chunk=100, chunk_seconds=1.0)
    done = srv.run(reqs)
    assert srv.repacks >= 3
    for r in done:
        assert r.admitted >= r.arrival and r.done > r.admitted
    check_against_alone(reqs, done, amr.FP32_SHIFTED)


def test_budget_limits_concurrency():
    """A budget that fits one request serves them one after another; the results
    do not change."""
    rng = np.random.default_rng(6)
    reqs = [small_request(i, 0.0, rng, max_steps=200) for i in range(2)]
    one = serve.footprint_bytes(reqs[0].case, amr.FP64)
    srv = serve.Server(precision=amr.FP64, chunk=100, hbm_budget=one + 1, chunk_seconds=1.0)
    done = srv.run(reqs)
    a = [r for r in done if r.id == 0][0]
    b = [r for r in done if r.id == 1][0]
    assert b.admitted >= a.done
    check_against_alone(reqs, done, amr.FP64)
