"""Remeshing a moving solid's refinement (mt_case_remap) on the device.

A case built for the moved geometry takes the old case's state cell by cell:
same key copied, else the nearest coarser ancestor exploded, else the finer
children coalesced, else filled with feq(rho0, u_fill).  Checked here: the
identity remap reproduces the run bit for bit, including the steps after it;
for a sphere moved by one coarse cell every cell's category and value follow
the rule (the transfers computed in numpy from the source's populations); the
moved case then runs stably."""
import numpy as np
import pytest

from mortontrie_amr_amd import amr, cases as C

pytestmark = pytest.mark.gpu

GEOM = dict(coarse=(32, 24, 24), radius=3.0, shells=(5.0, 1.5), Q=19, collision=C.COLLIDE_LBGK)
LIVE = C.KIND_SOLID | C.KIND_BCRES


def sphere(center):
    return C.sphere_amr(center=center, **GEOM)


def weights(Q):
    if Q == 19:
        return np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    return np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)


def transfer(f, scale, Q, rho0=1.0):
    """feq_HL + scale (f - feq_HL) per column of f [Q][n] (LBM_Manager.cpp:171-203)."""
    ex, ey, ez = C.velocities(Q)
    rho = f.sum(0)
    rs = np.where(np.abs(rho) < 1e-6, np.where(rho < 0, -1e-6, 1e-6), rho)
    u = np.stack([(f * e[:, None]).sum(0) for e in (ex, ey, ez)]) / rs
    eu = ex[:, None] * u[0] + ey[:, None] * u[1] + ez[:, None] * u[2]
    uu = (u * u).sum(0)
    fe = weights(Q)[:, None] * (rho + rho0 * (3 * eu + 4.5 * eu * eu - 1.5 * uu))
    return fe + scale * (f - fe)


def test_identity_remap_continues_the_run():
    case = sphere((12.0, 12.0, 12.0))
    g, h = amr.AmrCase(case), amr.AmrCase(case)
    g.step(5)
    cnt = h.remap_from(g)
    live = sum(int(((lv["kinds"] & LIVE) == 0).sum()) for lv in case["levels"])
    assert cnt == {"copied": live, "exploded": 0, "coalesced": 0, "filled": 0}
    for L, lv in enumerate(case["levels"]):
        keys = lv["keys"][(lv["kinds"] & LIVE) == 0]
        assert np.array_equal(g.f(L, keys), h.f(L, keys))
        assert np.array_equal(g.fcol(L, keys), h.fcol(L, keys))
        mk = lv["keys"][(lv["kinds"] & (C.KIND_FLUID | C.KIND_F2C)) != 0]
        assert np.array_equal(g.rho(L, mk), h.rho(L, mk))
        assert np.array_equal(g.u(L, mk), h.u(L, mk))
    g.step(5)
    h.step(5)
    for L, lv in enumerate(case["levels"]):
        assert np.array_equal(g.f(L), h.f(L), equal_nan=True)
    g.close()
    h.close()


def test_moved_sphere_remap_follows_the_rule():
    a_case, b_case = sphere((12.0, 12.0, 12.0)), sphere((13.0, 12.0, 12.0))
    Q, Lmax = a_case["Q"], a_case["Lmax"]
    a, b = amr.AmrCase(a_case), amr.AmrCase(b_case)
    a.step(5)
    fill = np.array([-0.01, 0.0, 0.0])
    cnt = b.remap_from(a, u_fill=fill)
    tau = C.taus(b_case)
    xs = [0.0] + [2 * tau[L] / tau[L - 1] for L in range(1, Lmax + 1)]
    cs = [tau[L] / (2 * tau[L + 1]) for L in range(Lmax)] + [0.0]
    src = []
    for L, lv in enumerate(a_case["levels"]):
        ok = (lv["kinds"] & C.KIND_SOLID) == 0
        keys = lv["keys"][ok]
        src.append((dict(zip(keys.tolist(), range(keys.size))), a.f(L, keys), a.fcol(L, keys)))
    want = {"copied": 0, "exploded": 0, "coalesced": 0, "filled": 0}
    ex, ey, ez = C.velocities(Q)
    for L, lv in enumerate(b_case["levels"]):
        keys = lv["keys"][(lv["kinds"] & LIVE) == 0]
        got_f, got_g = b.f(L, keys), b.fcol(L, keys)
        exp_f, exp_g = np.empty_like(got_f), np.empty_like(got_g)
        for i, k in enumerate(keys.tolist()):
            idx, sf, sg = src[L]
            if k in idx:
                want["copied"] += 1
                exp_f[:, i], exp_g[:, i] = sf[:, idx[k]], sg[:, idx[k]]
                continue
            done = False
            scale = 1.0
            for A in range(L - 1, -1, -1):
                scale *= xs[A + 1]
                pk = k & ~((1 << (3 * (Lmax - A))) - 1)
                ia, fa, ga = src[A]
                if pk in ia:
                    want["exploded"] += 1
                    exp_f[:, i] = transfer(fa[:, [ia[pk]]], scale, Q)[:, 0]
                    exp_g[:, i] = transfer(ga[:, [ia[pk]]], scale, Q)[:, 0]
                    done = True
                    break
            if done:
                continue
            if L < Lmax:
                idn, fd, gd = src[L + 1]
                ch = [idn[k | (c << (3 * (Lmax - L - 1)))] for c in range(8)
                      if (k | (c << (3 * (Lmax - L - 1)))) in idn]
                if ch:
                    want["coalesced"] += 1
                    exp_f[:, i] = transfer(fd[:, ch].sum(1, keepdims=True) * (1.0 / len(ch)), cs[L], Q)[:, 0]
                    exp_g[:, i] = transfer(gd[:, ch].sum(1, keepdims=True) * (1.0 / len(ch)), cs[L], Q)[:, 0]
                    continue
            want["filled"] += 1
            eu = ex * fill[0] + ey * fill[1] + ez * fill[2]
            exp_f[:, i] = exp_g[:, i] = weights(Q) * (1.0 + 3 * eu + 4.5 * eu * eu - 1.5 * (fill @ fill))
        assert np.allclose(got_f, exp_f, rtol=1e-13, atol=1e-15), L
        assert np.allclose(got_g, exp_g, rtol=1e-13, atol=1e-15), L
    assert cnt == want
    assert cnt["exploded"] > 0 and cnt["coalesced"] > 0 and cnt["filled"] > 0
    b.step(40)
    for L in range(Lmax + 1):
        assert np.isfinite(b.f(L)).all()
    a.close()
    b.close()


def test_remap_rejects_mismatched_cases():
    a = amr.AmrCase(sphere((12.0, 12.0, 12.0)))
    q27 = amr.AmrCase(C.sphere_amr(center=(12.0, 12.0, 12.0), **dict(GEOM, Q=27)))
    one = amr.AmrCase(C.cavity(16))
    for dst, src in ((q27, a), (one, a), (a, a)):
        with pytest.raises(amr.MtError):
            dst.remap_from(src)
    for c in (a, q27, one):
        c.close()
