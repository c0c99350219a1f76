"""Lagrangian-marker forcing on the device against the oracle.

Bar: FP64 fields within 1e-12 relative L2 after the stated number of steps; the
body force sums within 1e-12 relative.  The device's step runs one collision
ahead (the next step's first collision, with its marker forcing, is done by the
last fused update), so after n device steps the force log holds one more
finest-level call than the oracle's after n steps, until a call that returns
the case to the reference's buffers (reading fcol with explosions inside the
updates, writing a field, moving markers) drops that collision with its force."""
import numpy as np
import pytest

from mortontrie_amr_amd import amr, cases as C
from oracle import pyoracle as O
from tests.test_gpu_parity import compare, rel

pytestmark = pytest.mark.gpu


def cavity_with_markers(n=16):
    """A fixed sphere of markers inside a lid-driven cavity (one level)."""
    case = C.cavity(n, tau=0.7, lid=0.05)
    R = 3.0
    X = np.array([n / 2.0, n / 2.0 - 1.3, n / 2.0 + 0.4]) + R * C.fibonacci_sphere(120)
    keys, w = C.marker_stencils(X, 0)
    case["markers"] = {"keys": keys, "w": w, "dv": np.full(120, 4 * np.pi * R * R / 120), "ub": np.zeros(3)}
    return case


def small_marker_sphere(**kw):
    return C.sphere_markers((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.5, 2.0), n_markers=1800, **kw)


def check_force(g, o, case, steps):
    calls_per_step = 2 ** case["Lmax"]  # finest-level collisions per coarse step
    s, n = g.ib_force()
    assert n in (steps * calls_per_step, steps * calls_per_step + 1)
    o.step(1)
    want = o.force_log()[:n].sum(0)
    assert rel(s, want) < 1e-12


def test_cavity_markers_200_steps():
    case = cavity_with_markers()
    g, o = amr.AmrCase(case), O.OracleCase(case)
    g.step(200)
    o.step(200)
    compare(g, o, case)
    check_force(g, o, case, 200)
    g.close()


def test_sphere_markers_three_levels():
    """Three levels, D3Q27 cumulant with Smagorinsky LES, markers on the finest."""
    case = small_marker_sphere()
    g, o = amr.AmrCase(case), O.OracleCase(case)
    g.step(3)
    o.step(3)
    compare(g, o, case)
    check_force(g, o, case, 3)
    g.close()


def test_sphere_markers_graph_replay_and_split_steps():
    """Graph replays and eager launches, in different step splits, give the same bits."""
    case = small_marker_sphere()
    a, b = amr.AmrCase(case, graphs=True), amr.AmrCase(case, graphs=False)
    a.step(4)
    for n in (1, 2, 1):
        b.step(n)
    for L in range(case["Lmax"] + 1):
        assert np.array_equal(a.f(L), b.f(L))
        assert np.array_equal(a.fcol(L), b.fcol(L))
    assert np.array_equal(a.ib_force()[0], b.ib_force()[0])
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [amr.FP32_SHIFTED, amr.FP16_STORAGE])
def test_sphere_markers_schedules_same_bits_reduced_precision(precision):
    """In FP32 and FP16 storage too, the merged schedule, per-level streams and
    the serial order give the same bits (fields and marker force sums)."""
    case = small_marker_sphere()
    runs = []
    for sched in ("merged", "streams", "serial"):
        g = amr.AmrCase(case, precision=precision)
        g.set_schedule(sched)
        g.step(2)
        g.step(3)
        runs.append(g)
    a = runs[0]
    for b in runs[1:]:
        for L in range(case["Lmax"] + 1):
            assert np.array_equal(a.f(L), b.f(L))
            assert np.array_equal(a.fcol(L), b.fcol(L))
        assert np.array_equal(a.ib_force()[0], b.ib_force()[0])
    for g in runs:
        g.close()


def test_markers_on_interface_rejected():
    """A marker whose kernel reaches a cell that is not a streamed fluid cell of
    the finest level (here the C2F layer) is rejected at finalize."""
    case = C.cavity(16, refine_box=((5, 5, 5), (11, 11, 11)))
    # finest-level cells cover 10..21; put the marker on the layer's edge
    keys, w = C.marker_stencils(np.array([[10.6, 16.0, 16.0]]), 1)
    case["markers"] = {"keys": keys, "w": w, "dv": np.ones(1), "ub": np.zeros(3)}
    with pytest.raises(amr.MtError, match="not a streamed fluid cell"):
        amr.AmrCase(case)


def test_markers_need_one_member():
    m = cavity_with_markers(12)
    p = C.pack([C.cavity(8), C.cavity(8)])
    p["markers"] = m["markers"]
    with pytest.raises(amr.MtError, match="single-member"):
        amr.AmrCase(p)


def test_seam_collide_applies_markers():
    """mt_collide on the finest level applies the markers, so driving the seams
    in LBMkernel's order matches the oracle's step (one level)."""
    case = cavity_with_markers(12)
    g, o = amr.AmrCase(case, graphs=False), O.OracleCase(case)
    for _ in range(3):
        g.macros(0)
        g.collide(0)
        g.stream(0)
        g.bc_after_stream()
        g.macros(0)
        o.step(1)
        compare(g, o, case)
    g.close()
