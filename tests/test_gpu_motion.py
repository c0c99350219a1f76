"""Moving markers on the device (mt_markers_move / mt_markers_rigid).

A move rebuilds each marker's kernel cells, weights and the band on the device.
The bar is bit-equality with a case whose markers were built on the host at the
new positions (cases.marker_stencils), from the same state: before the first
step, in the middle of a run (the state carried over through mt_get_field /
mt_set_field), for moves that keep every marker's cells (new weights in place)
and for moves that change them (the band re-sorted).  A rigid motion's
positions and velocities are checked against Rodrigues' formula in numpy."""
import copy

import numpy as np
import pytest

from mortontrie_amr_amd import amr, cases as C
from tests.test_gpu_markers import cavity_with_markers, small_marker_sphere

pytestmark = pytest.mark.gpu


def at(case, xyz, ub=None):
    """The case with its markers rebuilt on the host at xyz."""
    c = copy.deepcopy(case)
    keys, w = C.marker_stencils(xyz, c["Lmax"])
    c["markers"] = dict(c["markers"], xyz=np.asarray(xyz, np.float64), keys=keys, w=w)
    if ub is not None:
        c["markers"]["ub"] = np.asarray(ub, np.float64)
    return c


def state(g, case):
    return [{name: g.get(L, fld) for name, fld in (("f", amr.FIELD_F), ("fcol", amr.FIELD_FCOL),
                                                   ("rho", amr.FIELD_RHO), ("u", amr.FIELD_U))}
            for L in range(case["Lmax"] + 1)]


def put(g, case, st):
    for L, s in enumerate(st):
        g.set(L, amr.FIELD_F, s["f"])
        g.set(L, amr.FIELD_FCOL, s["fcol"])
        g.set(L, amr.FIELD_RHO, s["rho"])
        g.set(L, amr.FIELD_U, s["u"])


def assert_same(a, b, case):
    for L in range(case["Lmax"] + 1):
        for fld in (amr.FIELD_F, amr.FIELD_RHO, amr.FIELD_U):
            x, y = a.get(L, fld), b.get(L, fld)
            assert np.array_equal(x, y, equal_nan=True), (L, fld, np.nanmax(np.abs(x - y)))
    fa, na = a.ib_force()
    fb, nb = b.ib_force()
    assert na == nb and np.array_equal(fa, fb)


def nearest(xyz):
    return np.floor(np.asarray(xyz) - 0.5 + 0.5)


def shifts(case):
    """A small move that keeps every marker's nearest node, and one that does not."""
    x = case["markers"]["xyz"]
    c = x - 0.5
    room = 0.5 - np.abs(c - np.round(c))           # distance to the next half-integer of c
    keep = x + 0.4 * np.minimum(room, 0.05)         # stays on the same side
    assert np.array_equal(nearest(keep), nearest(x))
    far = x + np.array([0.37, -0.21, 0.13])
    assert not np.array_equal(nearest(far), nearest(x))
    return keep, far


@pytest.mark.parametrize("which", ["reweight", "rebuild"])
def test_move_before_first_step_equals_building_there(which):
    case = small_marker_sphere()
    x2 = shifts(case)[which == "rebuild"]
    g = amr.AmrCase(case)
    g.move_markers(x2)
    h = amr.AmrCase(at(case, x2))
    g.step(3)
    h.step(3)
    assert_same(g, h, case)
    xg, _ = g.markers()
    assert np.array_equal(xg, x2)
    g.close()
    h.close()


@pytest.mark.parametrize("which", ["reweight", "rebuild"])
def test_move_mid_run_equals_rebuild_with_state(which):
    case = small_marker_sphere()
    x2 = shifts(case)[which == "rebuild"]
    g = amr.AmrCase(case)
    g.move_markers(case["markers"]["xyz"])   # positions known; cells unchanged
    g.step(2)
    st = state(g, case)
    g.move_markers(x2, np.full((len(x2), 3), 0.003))
    g.ib_force(reset=True)
    g.step(2)
    h = amr.AmrCase(at(case, x2, ub=np.full(3, 0.003)))
    put(h, case, st)
    h.step(2)
    assert_same(g, h, case)
    g.close()
    h.close()


def rodrigues(x, c, v, w, dt):
    r = x - c
    wn = np.linalg.norm(w)
    if wn == 0:
        p = r
    else:
        k = w / wn
        th = wn * dt
        p = r * np.cos(th) + np.cross(k, r) * np.sin(th) + np.outer(r @ k, k) * (1 - np.cos(th))
    return c + v * dt + p, -(v + np.cross(w, p))


def test_rigid_motion_positions_and_rebuild():
    case = small_marker_sphere()
    s = 2.0 ** case["Lmax"]
    ctr = np.array([12.0, 12.0, 12.0]) * s
    v, w = np.array([0.01, -0.004, 0.002]), np.array([0.0, 0.0015, 0.001])
    g = amr.AmrCase(case)
    g.move_markers(case["markers"]["xyz"])
    x0 = case["markers"]["xyz"]
    for k in range(3):
        g.move_markers_rigid(20.0, ctr + k * 20.0 * v, v, w)
        x0, u0 = rodrigues(x0, ctr + k * 20.0 * v, v, w, 20.0)
    xg, ug = g.markers()
    assert np.max(np.abs(xg - x0)) < 1e-11
    assert np.max(np.abs(ug - u0)) < 1e-14
    g.step(2)
    h = amr.AmrCase(at(case, xg, ub=None))
    h.move_markers(xg, ug)
    h.step(2)
    assert_same(g, h, case)
    g.close()
    h.close()


def test_invalid_move_changes_nothing():
    case = small_marker_sphere()
    g, h = amr.AmrCase(case), amr.AmrCase(case)
    g.move_markers(case["markers"]["xyz"])
    h.move_markers(case["markers"]["xyz"])
    bad = case["markers"]["xyz"].copy()
    bad[7] = [2.0, 2.0, 2.0]                     # outside the finest level
    with pytest.raises(amr.MtError, match="marker 7"):
        g.move_markers(bad)
    with pytest.raises(amr.MtError):
        g.move_markers_rigid(1.0, [0, 0, 0], [1e6, 0, 0], [0, 0, 0])
    with pytest.raises(amr.MtError):
        g.move_markers(np.full_like(bad, np.nan))
    g.step(2)
    h.step(2)
    assert_same(g, h, case)
    assert np.array_equal(g.markers()[0], case["markers"]["xyz"])
    g.close()
    h.close()


def test_body_moving_through_still_fluid_is_resisted():
    """Markers of a sphere moving along +x through a closed box at rest: the fluid
    force on the body points along -x."""
    R, ctr = 3.0, np.array([9.0, 12.0, 12.0])
    X = ctr + R * C.fibonacci_sphere(120)
    case = C.cavity(24, tau=0.7, lid=0.0)
    keys, w = C.marker_stencils(X, 0)
    case["markers"] = {"xyz": X, "keys": keys, "w": w, "dv": np.full(120, 4 * np.pi * R * R / 120),
                       "ub": np.zeros(3)}
    g = amr.AmrCase(case)
    g.move_markers(X)
    v = np.array([0.02, 0.0, 0.0])
    for k in range(60):
        g.move_markers_rigid(1.0, ctr, v, np.zeros(3))
        ctr = ctr + v
        g.step(1)
    f, n = g.ib_force()
    # one marker call per step, plus the collision computed ahead by the last step
    # (a move discards the one computed ahead, with its force)
    assert n == 61
    assert f[0] < 0 and abs(f[1]) < 0.2 * abs(f[0]) and abs(f[2]) < 0.2 * abs(f[0])
    assert np.isfinite(g.get(0, amr.FIELD_F)).all()
    g.close()


def test_marker_body_carried_with_remeshing():
    """A marker sphere carried by the stream at the stream's speed, the refinement
    rebuilt around it as it goes (motion.MarkerBody).  Uniform flow past a body
    moving with it is an exact solution, so the marker forces stay at rounding
    level and the velocity stays uniform through every move and remesh; each
    remesh copies, explodes and coalesces cells (a marker body has no solid
    cells, so none is filled).  Held still, the same sphere feels the stream's
    drag (positive along the flow) at once."""
    from mortontrie_amr_amd.motion import MarkerBody, RigidBody

    def build(center):
        return C.sphere_markers((32, 24, 24), radius=3.0, center=tuple(center), shells=(5.5, 2.0), n_markers=1800,
                                builder="gpu")

    still = MarkerBody(build, RigidBody((12.0, 12.0, 12.0)))
    still.step(4)
    assert still.forces[0][0] > 1.0 and not still.remeshes
    still.close()

    U = -still.case["cd"]["U"]  # the stream's physical speed
    mb = MarkerBody(build, RigidBody((12.0, 12.0, 12.0), velocity=(U, 0.0, 0.0)), remesh_dist=0.5)
    mb.step(40)
    f = np.array(mb.forces)
    assert np.max(np.abs(f)) < 1e-8, np.max(np.abs(f))
    assert len(mb.remeshes) >= 3
    assert all(r["copied"] > 0 and r["exploded"] > 0 and r["coalesced"] > 0 and r["filled"] == 0
               for r in mb.remeshes), mb.remeshes
    assert abs(mb.body.center[0] - (12.0 + 40 * U)) < 1e-12
    for L, lv in enumerate(mb.case["levels"]):
        keys = lv["keys"][(lv["kinds"] & (C.KIND_FLUID | C.KIND_F2C)) != 0]
        u = mb.g.u(L, keys)
        assert np.max(np.abs(u - np.array([[-U], [0.0], [0.0]]))) < 1e-10, L
    x, _ = mb.g.markers()
    c = mb.body.center * 2 ** mb.Lmax
    assert np.allclose(np.linalg.norm(x - c, axis=1), 3.0 * 2 ** mb.Lmax, rtol=1e-12)
    mb.close()
