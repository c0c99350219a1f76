"""The benchmark workloads at their full size, through the C-ABI.

* The 64^3 lid-driven cavity for its stated 1000 steps against the dense oracle
  (the oracle's map mode, same bits, on arrays with threads) at 1e-12.
* The 3-level sphere for its stated 4000 coarse steps against samples of the
  dense oracle's run kept in tests/golden/sphere_4000.npz (made by
  tests/golden/parity_curves.py).
* The 256-cavity packed batch (seed 0): members packed equal members alone, bit
  for bit, over the workload's 2000 steps in FP32-shifted storage, and small
  members of the FP64 batch equal the oracle.
* The Poisson serving stream (32 requests, 20/s, seed 1) in virtual time with a
  shortened run: every served result equals the request run alone.
"""
import os

import numpy as np
import pytest

from mortontrie_amr_amd import amr, cases as C, serve
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b, floor=1e-300):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), floor))


def test_cavity_64_stated_steps_against_dense_oracle():
    """cfg1 at full size for the stated 1000 steps (LBM_Simulator.cpp:8-18)."""
    case = C.cavity(64)
    g = amr.AmrCase(case)
    g.step(1000)
    o = O.DenseOracleCase(case)
    o.step(1000)
    assert rel(g.f(0), o.f(0)) < 1e-12
    assert rel(g.rho(0), o.rho(0)) < 1e-12
    u = o.u(0)
    assert rel(g.u(0), u) < 1e-12
    assert np.abs(u).max() > 0.01  # the lid has set the cavity in motion


def test_sphere_stated_steps_against_oracle_samples():
    """cfg2 at full size for the stated 4000 coarse steps: density and velocity of
    8192 sampled fluid cells per level against the dense oracle's run."""
    path = os.path.join(GOLDEN, "sphere_4000.npz")
    if not os.path.exists(path):
        pytest.skip("tests/golden/sphere_4000.npz has not been generated")
    gold = np.load(path)
    case = C.sphere_amr(options=2)
    g = amr.AmrCase(case)
    done = 0
    for s in gold["checkpoints"]:
        g.step(int(s) - done)
        done = int(s)
        for L in range(case["Lmax"] + 1):
            k = gold[f"keys_L{L}"]
            tol = gold[f"tol_s{s}"]
            assert rel(g.rho(L, k), gold[f"rho_L{L}_s{s}"]) < tol[0], (s, L, "rho")
            assert rel(g.u(L, k), gold[f"u_L{L}_s{s}"]) < tol[1], (s, L, "u")


def member_state(g, packed, i, m):
    out = []
    for L in range(m["Lmax"] + 1):
        keys = C.member_keys(packed, i, L, m["levels"][L]["keys"])
        out += [g.f(L, keys), g.rho(L, keys), g.u(L, keys)]
    return out


def alone_state(g, m):
    out = []
    for L in range(m["Lmax"] + 1):
        keys = m["levels"][L]["keys"]
        out += [g.f(L, keys), g.rho(L, keys), g.u(L, keys)]
    return out


@pytest.fixture(scope="module")
def batch256():
    ms = C.cavity_batch(256, seed=0)
    return ms, C.pack(ms)


def test_cavity_batch_256_members_equal_alone(batch256):
    """cfg4: the 256 seed-0 cavities packed in one case, FP32-shifted, 2000 steps
    under graphs; sampled members (both edge extremes, one and two levels) are bit
    for bit what they compute alone, and the per-member scan flags exactly the
    members that diverge alone."""
    ms, packed = batch256
    gb = amr.AmrCase(packed, precision=amr.FP32_SHIFTED)
    gb.step(2000)
    flags = gb.scan_members()
    edges = np.array([m["dims0"][0] for m in ms])
    picks = [int(np.nonzero(edges == 16)[0][0]), int(np.nonzero(edges == 64)[0][0]),
             int(np.nonzero((edges == 24) & (np.arange(256) % 2 == 1))[0][0]),
             int(np.nonzero((edges == 48) & (np.arange(256) % 2 == 1))[0][0])]
    bad = [int(i) for i in np.nonzero(flags)[0]]
    if bad:
        picks.append(bad[0])
    for i in picks:
        ga = amr.AmrCase(ms[i], precision=amr.FP32_SHIFTED)
        ga.step(2000)
        assert (ga.scan_members()[0] > 0) == (flags[i] > 0), i
        for a, b in zip(member_state(gb, packed, i, ms[i]), alone_state(ga, ms[i])):
            assert np.array_equal(a, b, equal_nan=True), i
        ga.close()
    gb.close()


def test_cavity_batch_small_members_against_oracle(batch256):
    """The same 256-member batch in FP64: the small members equal the oracle run
    of each alone at 1e-12 (200 steps)."""
    ms, packed = batch256
    gb = amr.AmrCase(packed, precision=amr.FP64)
    gb.step(200)
    edges = np.array([m["dims0"][0] for m in ms])
    picks = [int(np.nonzero(edges == 16)[0][0]), int(np.nonzero((edges == 16) & (np.arange(256) % 2 == 1))[0][0])]
    for i in picks:
        m = ms[i]
        o = O.DenseOracleCase(m)
        o.step(200)
        for L in range(m["Lmax"] + 1):
            keys = C.member_keys(packed, i, L, m["levels"][L]["keys"])
            assert rel(gb.f(L, keys), o.f(L)) < 1e-12, (i, L)
            fl = (m["levels"][L]["kinds"] & (C.KIND_FLUID | C.KIND_F2C)) != 0
            assert rel(gb.rho(L, keys[fl]), o.rho(L)[fl]) < 1e-12, (i, L)
    gb.close()


def test_poisson_stream_each_result_equals_alone():
    """cfg5: the 32-request Poisson stream (20/s, seed 1) served FP32-shifted with
    admission every 4 chunks in virtual time, runs shortened to 300 steps: every
    request's drag history is the one it has alone and a request that diverges
    alone diverges served."""
    reqs = serve.poisson_requests(32, 20.0, seed=1, max_steps=300, tol=0.0)
    srv = serve.Server(precision=amr.FP32_SHIFTED, chunk=100, chunk_seconds=0.2, admit_every=4)
    done = srv.run(reqs)
    assert sorted(r.id for r in done) == list(range(32))
    assert srv.repacks >= 3
    for r in done:
        alone = serve.run_alone(r, precision=amr.FP32_SHIFTED, chunk=100)
        assert r.failed == alone.failed, r.id
        assert r.steps == alone.steps, r.id
        if not r.failed:
            assert np.allclose(r.cd, alone.cd, rtol=1e-9, atol=1e-12), r.id
