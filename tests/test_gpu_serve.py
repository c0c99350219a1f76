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


def test_staggered_admission_reference_explosion_matches_alone():
    """With the reference's explosion (sub-step A only) the explosions also run
    inside the updates; a repack copies each member's state, C2F copies and
    ghost targets included, and a member still computes what it computes alone
    (Re 8-12, tau0 about 0.6: stable under these semantics)."""
    rng = np.random.default_rng(12)
    reqs = []
    for i, t in enumerate((0.0, 1.5, 2.5)):
        r = float(rng.uniform(3.0, 5.0))
        case = C.sphere_amr((48, 24, 24), radius=r, center=(16.0, 12.0, 12.0), shells=(3.0,), U=0.04,
                            Re=float(rng.uniform(8, 12)), options=0)
        reqs.append(serve.Request(i, case, t, max_steps=300, tol=0.0))
    srv = serve.Server(precision=amr.FP32_SHIFTED, chunk=100, chunk_seconds=1.0)
    done = srv.run(reqs)
    assert srv.repacks >= 3
    g = amr.AmrCase(reqs[0].case)
    assert g.schedule() == "merged"
    g.close()
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


def test_batched_admission_and_lazy_retirement_match_alone():
    """Admission every 3 chunks: arrivals wait for the next repack, finished
    requests keep stepping with the batch until it, and every result is still
    the one computed alone."""
    rng = np.random.default_rng(7)
    reqs = [small_request(0, 0.0, rng, max_steps=500), small_request(1, 0.5, rng),
            small_request(2, 1.2, rng, max_steps=400), small_request(3, 4.5, rng, max_steps=200)]
    srv = serve.Server(precision=amr.FP32_SHIFTED, chunk=100, chunk_seconds=1.0, admit_every=3)
    done = srv.run(reqs)
    by_id = {r.id: r for r in done}
    # requests 1 and 2 (arrived at 0.5 and 1.2) wait for the repack after three
    # chunks; request 3 (4.5) for the one three chunks later, which also drops
    # request 0 (finished at 5.0, stepping on meanwhile)
    assert by_id[1].admitted == 3.0 and by_id[2].admitted == 3.0
    assert by_id[3].admitted == 6.0
    assert by_id[0].done == 5.0
    # fewer repacks than membership changes
    assert srv.repacks < 2 * len(reqs)
    for r in done:
        assert r.admitted >= r.arrival and r.done > r.admitted
    check_against_alone(reqs, done, amr.FP32_SHIFTED)


def test_cases_rebuilt_on_cached_memory_compute_the_same():
    """Device memory and streams come from process-wide caches: a case built on
    blocks another case has just released, and two cases alive at once, give
    the same bits as a case built on fresh memory."""
    case = C.sphere_amr((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5), options=2)
    ref = amr.AmrCase(case)
    ref.step(3)
    want = [ref.f(L) for L in range(case["Lmax"] + 1)]
    ref.close()
    for _ in range(2):
        a = amr.AmrCase(case)
        b = amr.AmrCase(C.cavity(12))
        a.step(2)
        b.step(5)
        a.step(1)
        for L in range(case["Lmax"] + 1):
            assert np.array_equal(a.f(L), want[L])
        b.close()
        a.close()


def test_native_server_matches_alone():
    """The library's own serving loop (mt_server_*): requests submitted while
    others run finish with the steps and drag they give alone."""
    rng = np.random.default_rng(8)
    reqs = [small_request(i, 0.0, rng, max_steps=300 + 100 * (i % 2)) for i in range(4)]
    srv = serve.NativeServer(precision=amr.FP32_SHIFTED, chunk=100, admit_every=2)
    srv.submit(0, reqs[0].case, reqs[0].max_steps, 0.0)
    srv.submit(1, reqs[1].case, reqs[1].max_steps, 0.0)
    got = {}
    srv.step()
    srv.submit(2, reqs[2].case, reqs[2].max_steps, 0.0)
    srv.step()
    srv.submit(3, reqs[3].case, reqs[3].max_steps, 0.0)
    for _ in range(50):
        nr, nw = srv.step()
        for rid, steps, cd, diverged in srv.poll():
            assert not diverged
            got[rid] = (steps, cd)
        if nr == 0 and nw == 0:
            break
    for rid, steps, cd, diverged in srv.poll():
        got[rid] = (steps, cd)
    srv.close()
    assert sorted(got) == [0, 1, 2, 3]
    for r in reqs:
        alone = serve.run_alone(r, precision=amr.FP32_SHIFTED, chunk=100)
        assert got[r.id][0] == alone.steps
        assert np.isclose(got[r.id][1], alone.cd[-1], rtol=1e-9, atol=1e-12), r.id


def poisoned(req):
    """The request with a non-finite initial velocity: every population of it is NaN."""
    req.case = dict(req.case, u0=(float("nan"), 0.0, 0.0))
    return req


def test_diverged_request_is_flagged_and_evicted():
    """A request whose state goes non-finite is found by the per-member device scan
    after its first chunk and finishes as failed; with about half the batch's
    footprint and 5 chunks left to run, evicting it pays, so it leaves the
    batch at the next chunk's repack; the requests beside it compute what they
    compute alone."""
    rng = np.random.default_rng(10)
    reqs = [small_request(0, 0.0, rng, max_steps=600), poisoned(small_request(1, 0.0, rng)),
            small_request(2, 1.5, rng)]
    srv = serve.Server(precision=amr.FP32_SHIFTED, chunk=100, chunk_seconds=1.0, admit_every=4)
    done = srv.run(reqs)
    by_id = {r.id: r for r in done}
    assert sorted(by_id) == [0, 1, 2]
    bad = by_id[1]
    assert bad.failed and bad.steps == 100 and bad.done == 1.0
    # the eviction repack at 1.0 is due although the admission interval is 4 chunks
    assert srv.repack_log[0] == (0.0, [0, 1])
    assert srv.repack_log[1] == (1.0, [0])
    good = [by_id[0], by_id[2]]
    assert not any(r.failed for r in good)
    check_against_alone([reqs[0], reqs[2]], good, amr.FP32_SHIFTED)


def test_native_server_flags_and_evicts_a_diverged_request():
    rng = np.random.default_rng(11)
    reqs = [small_request(0, 0.0, rng), poisoned(small_request(1, 0.0, rng)), small_request(2, 0.0, rng)]
    srv = serve.NativeServer(precision=amr.FP32_SHIFTED, chunk=100, admit_every=4)
    for r in reqs:
        srv.submit(r.id, r.case, r.max_steps, 0.0)
    got = {}
    for _ in range(50):
        nr, nw = srv.step()
        for rid, steps, cd, diverged in srv.poll():
            got[rid] = (steps, cd, diverged)
        if nr == 0 and nw == 0:
            break
    srv.close()
    assert sorted(got) == [0, 1, 2]
    assert got[1][2] and got[1][0] == 100
    for r in (reqs[0], reqs[2]):
        alone = serve.run_alone(r, precision=amr.FP32_SHIFTED, chunk=100)
        assert not got[r.id][2] and not alone.failed
        assert got[r.id][0] == alone.steps
        assert np.isclose(got[r.id][1], alone.cd[-1], rtol=1e-9, atol=1e-12), r.id


def test_native_server_rejects_bad_requests():
    srv = serve.NativeServer(precision=amr.FP32_SHIFTED)
    big = C.sphere_amr((128, 32, 32), radius=4.0, center=(20.0, 16.0, 16.0), shells=(3.0,), options=2)
    with pytest.raises(amr.MtError, match="slot"):
        srv.submit(0, big, 100, 0.0)
    rng = np.random.default_rng(9)
    srv.submit(1, small_request(1, 0.0, rng).case, 100, 0.0)
    other = C.sphere_amr((48, 24, 24), radius=3.0, center=(16.0, 12.0, 12.0), shells=(), options=2)
    with pytest.raises(amr.MtError, match="share"):
        srv.submit(2, other, 100, 0.0)
    srv.close()
