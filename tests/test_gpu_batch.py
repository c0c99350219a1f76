"""Packed batching: independent members sharing every launch give, member by
member, the results they give alone (bit for bit for the fields; the force sums
to rounding, their per-brick order differs)."""
import numpy as np
import pytest

from mortontrie_amr_amd import amr, cases as C

pytestmark = pytest.mark.gpu


def member_fields(g, packed, i, member, L):
    lv = member["levels"][L]
    keys = C.member_keys(packed, i, L, lv["keys"])
    return g.f(L, keys), g.rho(L, keys), g.u(L, keys)


def alone_fields(g, member, L):
    keys = member["levels"][L]["keys"]
    return g.f(L, keys), g.rho(L, keys), g.u(L, keys)


def run_and_compare(members, steps, precision=amr.FP64):
    packed = C.pack(members)
    gb = amr.AmrCase(packed, precision=precision)
    gb.step(steps)
    for i, m in enumerate(members):
        ga = amr.AmrCase(m, precision=precision)
        ga.step(steps)
        for L in range(m["Lmax"] + 1):
            for a, b in zip(member_fields(gb, packed, i, m, L), alone_fields(ga, m, L)):
                assert np.array_equal(a, b), f"member {i} level {L}"
        ga.close()
    return gb, packed


def test_packed_cavities_equal_alone():
    """One- and two-level cavities of different size, tau and lid speed."""
    ms = [C.cavity(12), C.cavity(16, refine_box=((5, 5, 5), (11, 11, 11))),
          C.cavity(8, tau=0.8, lid=0.08), C.cavity(20, tau=0.55, lid=0.03)]
    gb, _ = run_and_compare(ms, 40)
    gb.close()


def test_packed_cavities_fp32_equal_alone():
    ms = [C.cavity(12), C.cavity(16, refine_box=((5, 5, 5), (11, 11, 11)))]
    gb, _ = run_and_compare(ms, 30, precision=amr.FP32_SHIFTED)
    gb.close()


def test_packed_spheres_equal_alone():
    """Three-level spheres with link IB, periodic y/z inside each member's box,
    different inflow speeds and Reynolds numbers."""
    mk = lambda **kw: C.sphere_amr((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5),
                                   options=2, **kw)
    ms = [mk(), mk(U=0.05, Re=60), mk(U=0.03, Re=150)]
    gb, packed = run_and_compare(ms, 6)
    fm, n = gb.ib_force_members()
    for i, m in enumerate(ms):
        ga = amr.AmrCase(m)
        ga.step(6)
        fa, na = ga.ib_force()
        assert n == na
        # components that vanish by symmetry carry rounding noise of the sum order
        assert np.allclose(fm[i], fa, rtol=1e-12, atol=1e-12 * np.abs(fa).max())
        ga.close()
    gb.close()


def test_poisoned_member_flagged_others_unchanged():
    """A member whose populations are non-finite (NaN initial velocity) is the only
    one the per-member device scan flags (the reference's NaN checks,
    LBM_Simulator.cpp:33-41), and the members beside it stay bit for bit what they
    compute alone."""
    ms = [C.cavity(12), C.cavity(16, refine_box=((5, 5, 5), (11, 11, 11))), C.cavity(12, lid=0.06)]
    bad = dict(ms[1], u0=(float("nan"), 0.0, 0.0))
    packed = C.pack([ms[0], bad, ms[2]])
    gb = amr.AmrCase(packed, precision=amr.FP32_SHIFTED)
    assert np.array_equal(gb.scan_members() > 0, [False, True, False])
    gb.step(20)
    flags = gb.scan_members()
    assert flags[0] == 0 and flags[2] == 0 and flags[1] > 0
    for i in (0, 2):
        ga = amr.AmrCase(ms[i], precision=amr.FP32_SHIFTED)
        ga.step(20)
        assert ga.scan_members()[0] == 0
        for L in range(ms[i]["Lmax"] + 1):
            for a, b in zip(member_fields(gb, packed, i, ms[i], L), alone_fields(ga, ms[i], L)):
                assert np.array_equal(a, b), f"member {i} level {L}"
        ga.close()
    gb.close()


def test_member_box_validation():
    """Overlapping member boxes and misaligned origins are rejected."""
    import ctypes
    p = C.pack([C.cavity(8), C.cavity(8)])
    p["members"][1]["origin0"] = (4, 0, 0)
    with pytest.raises(amr.MtError, match="share a brick"):
        amr.AmrCase(p)
    p["members"][1]["origin0"] = (14, 0, 0)
    with pytest.raises(amr.MtError, match="multiple of 4"):
        amr.AmrCase(p)


def test_autotune_keeps_results():
    """The autotuner's candidates compute the same values: a tuned run equals an
    untuned run of the same number of steps, bit for bit."""
    case = C.sphere_amr((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5), options=2)
    a, b = amr.AmrCase(case), amr.AmrCase(case)
    ms = a.autotune(2)
    assert np.all(ms > 0)
    a.step(3)
    b.step(8 * 3 + 3)
    for L in range(case["Lmax"] + 1):
        assert np.array_equal(a.f(L), b.f(L))
        assert np.array_equal(a.fcol(L), b.fcol(L))
