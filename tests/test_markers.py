"""Lagrangian markers (explicit direct forcing of the finest level): pins of the
kernel, the stencils and the oracle's forcing against closed forms and
conservation laws.  The forcing has no counterpart in the reference; its
definition is the one in include/mt_amr.h (mt_case_upload_markers)."""
import numpy as np
import pytest

from mortontrie_amr_amd import cases as C
from oracle import pyoracle as O

RNG = np.random.default_rng(77)


def test_roma3_moment_conditions():
    """Roma's 3-point kernel on unit-spaced nodes: sum 1, first moment 0 and
    sum of squares 1/2 for every offset (the three conditions that define it)."""
    for r in RNG.random(50):
        j = np.arange(-3, 4)
        phi = C.roma3(j - r)
        assert abs(phi.sum() - 1) < 1e-15
        assert abs((phi * (j - r)).sum()) < 1e-15
        assert abs((phi * phi).sum() - 0.5) < 1e-15
    assert C.roma3(np.array([1.5, 2.0, -1.5]))[:].max() == 0.0


def test_marker_stencils_interpolate_linear_fields():
    """Each marker's cells (decoded from the keys) and weights reproduce constants
    and linear functions of the position: sum w = 1, sum w (x - X) = 0 on every
    axis (a swapped axis or an off-by-one cell centre fails)."""
    Lmax = 2
    X = np.array([[10.3, 17.61, 33.92], [8.5, 8.5, 8.5], [40.0, 12.25, 9.77]])
    keys, w = C.marker_stencils(X, Lmax)
    assert keys.shape == (3, 27) and w.shape == (3, 27)
    for m in range(3):
        x, y, z = C.decode_key(keys[m], Lmax, Lmax)
        pos = np.stack([x, y, z], axis=1).astype(float) + 0.5
        assert abs(w[m].sum() - 1) < 1e-14
        assert np.abs((w[m][:, None] * (pos - X[m])).sum(0)).max() < 1e-13
        assert (w[m] >= 0).all()
        # a marker on a cell centre touches 3 cells per axis (the outer two at 1/6)
        if m == 1:
            assert np.count_nonzero(w[m]) == 27


def sphere_marker_sizes():
    c = C.sphere_markers((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.5, 2.0), n_markers=1800)
    fluid = c["levels"][c["Lmax"]]
    ok = set(fluid["keys"][(fluid["kinds"] & C.KIND_FLUID) != 0].tolist())
    return c, ok


def test_sphere_marker_support_is_finest_fluid():
    """Every kernel cell of the sphere's markers is a fluid cell of the finest
    level, and the surface elements sum to the sphere's area."""
    c, ok = sphere_marker_sizes()
    mk = c["markers"]
    used = mk["keys"][mk["w"] != 0]
    assert all(int(k) in ok for k in np.unique(used))
    with pytest.raises(ValueError, match="widen"):
        C.sphere_markers((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5), n_markers=1800)
    R = 3.0 * 4
    assert abs(mk["dv"].sum() - 4 * np.pi * R * R) < 1e-9
    assert np.allclose(mk["w"].sum(1), 1.0, atol=1e-14)


def single_marker_case(X, ub, dv=0.7):
    case = C.cavity(12)
    keys, w = C.marker_stencils(np.array([X]), 0)
    case["markers"] = {"keys": keys, "w": w, "dv": np.array([dv]), "ub": np.asarray(ub, dtype=float)}
    return case


def set_linear_fcol(o, case, rho, a, B):
    keys = case["levels"][0]["keys"]
    x, y, z = C.decode_key(keys, 0, 0)
    pos = np.stack([x, y, z], axis=1).astype(float) + 0.5
    u = a[None, :] + pos @ B.T
    f = np.stack([O.feq(19, rho, ui) for ui in u], axis=1)
    o.set(0, 1, f.ravel())
    return pos, u


def gain(o, before):
    """Mass and momentum the finest level's fcol gained since `before`."""
    E = np.stack(C.velocities(19), axis=1).astype(float)
    d = (o.fcol(0) - before).sum(1)
    return d.sum(), E.T @ d


def test_single_marker_force_in_linear_flow():
    """One marker in a linear velocity field: the force is dV rho (ub - u(X))
    exactly (the kernel interpolates linear fields) and is the logged force (the
    body's reaction with the reported velocities' sign flipped back), the fluid
    gains its momentum and keeps its mass, and only the marker's 27 cells
    change."""
    X = np.array([5.3, 6.2, 5.8])
    ub = np.array([0.01, -0.02, 0.015])
    dv, rho = 0.7, 1.03
    a = np.array([0.02, -0.01, 0.005])
    B = 1e-3 * np.array([[1.0, 2.0, -1.0], [0.5, -1.5, 2.5], [-2.0, 1.0, 0.7]])
    case = single_marker_case(X, ub, dv)
    o = O.OracleCase(case)
    set_linear_fcol(o, case, rho, a, B)
    before = o.fcol(0).copy()
    o.marker_correct()
    F_expect = dv * rho * (ub - (a + B @ X))
    log = o.force_log()
    assert log.shape == (1, 3)
    assert np.abs(log[0] - F_expect).max() < 1e-16
    dm, dp = gain(o, before)
    # rounding of 27 x 19 differences of O(0.1) values; a lost term is O(1e-3)
    assert abs(dm) < 1e-14
    assert np.abs(dp - F_expect).max() < 1e-14
    changed = np.nonzero(np.abs(o.fcol(0) - before).max(0) > 0)[0]
    assert changed.size == 27
    keys = case["levels"][0]["keys"]
    assert set(keys[changed].tolist()) == set(case["markers"]["keys"][0].tolist())


def test_marker_moving_with_the_flow_does_nothing():
    """A marker whose velocity equals the fluid's exerts no force."""
    u0 = np.array([0.03, 0.01, -0.02])
    case = single_marker_case(np.array([6.1, 5.4, 5.5]), u0)
    o = O.OracleCase(case)
    set_linear_fcol(o, case, 0.98, u0, np.zeros((3, 3)))
    before = o.fcol(0).copy()
    o.marker_correct()
    # only the moments' rounding remains (a force of the relative velocity 0.03
    # would be ~1e-2)
    assert np.abs(o.force_log()[0]).max() < 1e-15
    assert np.abs(o.fcol(0) - before).max() < 1e-15


def test_overlapping_markers_superpose():
    """Two markers sharing cells: each cell's momentum gain is the weighted sum
    of both forces, so the total gain is F1 + F2 (no double counting or loss)."""
    X = np.array([[5.3, 6.2, 5.8], [6.1, 6.4, 5.2]])
    keys, w = C.marker_stencils(X, 0)
    case = C.cavity(12)
    case["markers"] = {"keys": keys, "w": w, "dv": np.array([0.5, 0.9]), "ub": np.zeros(3)}
    o = O.OracleCase(case)
    a = np.array([0.02, -0.01, 0.005])
    B = 1e-3 * np.eye(3)
    set_linear_fcol(o, case, 1.0, a, B)
    before = o.fcol(0).copy()
    o.marker_correct()
    F = sum(case["markers"]["dv"][m] * 1.0 * (0 - (a + B @ X[m])) for m in range(2))
    _, dp = gain(o, before)
    assert np.abs(dp - F).max() < 1e-14
    assert np.abs(o.force_log()[0] - F).max() < 1e-16


@pytest.mark.parametrize("steps", [2])
def test_sphere_markers_oracle_runs(steps):
    """The small marker sphere steps in the oracle: finite fields, one force
    entry per finest-level collision, drag along the flow."""
    c, _ = sphere_marker_sizes()
    o = O.OracleCase(c)
    o.step(steps)
    log = o.force_log()
    assert log.shape == (4 * steps, 3)
    assert np.isfinite(log).all()
    # a drag along the physical flow (+x) is positive, as for the link-wise force
    assert log[-1, 0] > 0
