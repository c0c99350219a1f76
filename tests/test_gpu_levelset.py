"""The level sets built on the GPU (mt_levelset_build) against the host builder.

For the analytic sphere the device applies the host builder's rules with the same
arithmetic, so every level's keys and kinds and every link fraction must be equal.
For a body given as a triangle mesh the host builder runs the same rules in numpy
(cases.MeshBody), on a case small enough for it.  A mesh-defined case then steps on
the device against the oracle, and is served.
"""
import numpy as np
import pytest

from mortontrie_amr_amd import amr, cases as C, serve
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def same_sets(a, b):
    assert len(a["levels"]) == len(b["levels"])
    for L, (x, y) in enumerate(zip(a["levels"], b["levels"])):
        assert np.array_equal(x["keys"], y["keys"]), L
        assert np.array_equal(x["kinds"], y["kinds"]), L
    assert np.array_equal(a["dis_keys"], b["dis_keys"])
    assert np.array_equal(a["dis_vals"], b["dis_vals"])


@pytest.mark.parametrize("kw", [
    dict(),  # the 3-level benchmark sphere, 64x32x32 coarse
    dict(coarse=(32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5), Q=19),
    dict(coarse=(48, 24, 24), radius=5.0, center=(16, 12, 12), shells=(3.0,), periodic_yz=False),
    dict(coarse=(32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.5, 2.0), with_sphere=False),
    dict(coarse=(40, 32, 32), radius=2.0, center=(16.3, 15.7, 16.2), shells=(7.0, 4.0, 2.0)),
])
def test_sphere_sets_equal_host_builder(kw):
    same_sets(C.sphere_amr(builder="gpu", **kw), C.sphere_amr(**kw))


def test_reservoir_layout_equals_host_builder():
    """The compiled reference's layout: a frozen lat_bc shell with face and edge bits."""
    ref = C.reference_sphere((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5))
    lv, dk, dv = amr.build_levels(19, 2, (32, 24, 24), ref["faces"], (5.0, 1.5), center=(12, 12, 12), radius=3.0,
                                  reservoir=True)
    same_sets(ref, dict(levels=lv, dis_keys=dk, dis_vals=dv))


def test_mesh_sets_equal_host_builder():
    """An icosphere (320 triangles) off the grid's symmetry: winding-number inside,
    distance shells and Moller-Trumbore link fractions on the device equal numpy's."""
    T = C.icosphere((12.3, 11.8, 12.1), 3.2, 2)
    a = C.mesh_amr(T, coarse=(32, 24, 24), shells=(5.0, 1.5), builder="gpu")
    b = C.mesh_amr(T, coarse=(32, 24, 24), shells=(5.0, 1.5), builder="host")
    for L in range(3):
        assert np.array_equal(a["levels"][L]["keys"], b["levels"][L]["keys"]), L
        assert np.array_equal(a["levels"][L]["kinds"], b["levels"][L]["kinds"]), L
    assert np.array_equal(a["dis_keys"], b["dis_keys"])
    np.testing.assert_allclose(a["dis_vals"], b["dis_vals"], rtol=0, atol=1e-14)
    assert a["dis_keys"].size > 0


@pytest.mark.parametrize("kw,frag", [
    (dict(coarse=(32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(11.0, 1.5)), "domain edge"),
    (dict(coarse=(40, 32, 32), radius=3.5, center=(14.3, 15.1, 16.7), shells=(6.0, 3.5, 1.5)), "C2F band"),
])
def test_levelset_errors(kw, frag):
    """Layouts the host builder refuses, the device refuses with the same reason."""
    with pytest.raises(ValueError, match=frag):
        C.sphere_amr(**kw)
    with pytest.raises(amr.MtError, match=frag):
        C.sphere_amr(builder="gpu", **kw)


def test_mesh_case_steps_like_the_oracle_and_is_served():
    """A body known only as a triangle mesh: its level sets built on the device, the
    update matches the oracle at 1e-12, and the serving loop admits it."""
    T = C.icosphere((16.2, 12.0, 12.0), 4.0, 3)
    case = C.mesh_amr(T, coarse=(48, 24, 24), shells=(3.0,), U=0.04, Re=60.0)
    g, o = amr.AmrCase(case), O.OracleCase(case)
    g.step(3)
    o.step(3)
    for L in range(case["Lmax"] + 1):
        assert np.linalg.norm(g.f(L) - o.f(L)) / np.linalg.norm(o.f(L)) < 1e-12
    srv = serve.NativeServer(precision=amr.FP32_SHIFTED, chunk=100)
    srv.submit(0, case, 200, 0.0)
    got = []
    for _ in range(10):
        nr, nw = srv.step()
        got += srv.poll()
        if nr == 0 and nw == 0:
            break
    srv.close()
    assert len(got) == 1 and got[0][1] == 200 and np.isfinite(got[0][2]) and not got[0][3]
