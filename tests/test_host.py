"""Host-side checks that need no GPU: the case builders' set contract, the
C-ABI library's exports and argument validation, the C++ driver header."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from mortontrie_amr_amd import amr, cases as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- case builders
def parents(keys, L, Lmax):
    x, y, z = C.decode_key(keys, L, Lmax)
    return C.encode_key(x // 2, y // 2, z // 2, L - 1, Lmax)


@pytest.mark.parametrize("mk", [
    lambda: C.sphere_amr((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5)),
    lambda: C.cavity(16, refine_box=((5, 5, 5), (11, 11, 11))),
    lambda: C.reference_sphere((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5)),
])
def test_level_set_contract(mk):
    """The Lat_Manager set relations (header/Lat_Manager.hpp:245-316,
    Lat_Voxelize.cpp:525-733): unique keys; every C2F cell's parent is F2C and every
    F2C cell has all 8 children as C2F; ghost_C2F cells are Q-neighbours of C2F
    whose parents are ghost_F2C fluid cells; F2C is not in lat_f."""
    case = mk()
    Lmax = case["Lmax"]
    for L, lv in enumerate(case["levels"]):
        assert np.unique(lv["keys"]).size == lv["keys"].size
        k = lv["kinds"]
        assert not ((k & C.KIND_FLUID) & ((k & C.KIND_F2C) >> 1)).any()
        if L == 0:
            continue
        prev = case["levels"][L - 1]
        pk = dict(zip(prev["keys"].tolist(), prev["kinds"].tolist()))
        c2f = lv["keys"][(k & C.KIND_C2F) != 0]
        assert all(pk[int(p)] & C.KIND_F2C for p in parents(c2f, L, Lmax))
        n_f2c = int(((prev["kinds"] & C.KIND_F2C) != 0).sum())
        assert c2f.size == 8 * n_f2c
        ghost = lv["keys"][(k & C.KIND_GHOST_C2F) != 0]
        assert all(pk[int(p)] & C.KIND_GHOST_F2C for p in parents(ghost, L, Lmax))


def test_sphere_workload_counts():
    """The refined-sphere workload at full size: fluid cells per level
    (coarse 64x32x32, r=8, shells 6 and 3)."""
    case = C.sphere_amr()
    cnt = C.level_counts(case)
    assert cnt[0]["fluid"] == 54000
    assert [c["fluid"] + c["c2f"] for c in cnt] == [54000, 47568, 220384]
    assert C.updates_per_coarse_step(case) == 54000 + 2 * 47568 + 4 * 220384
    assert case["Lmax"] == 2 and case["Q"] == 27
    assert np.isclose(case["tau0"], 0.5192)


def test_dis_fractions_analytic():
    """dis: first crossing of the link with the sphere, in [0,1] (Lat_Voxelize.cpp:1252-1285)."""
    case = C.sphere_amr((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5))
    v = case["dis_vals"]
    assert ((v == -1) | ((v >= 0) & (v <= 1))).all()
    Lmax = case["Lmax"]
    x, y, z = C.decode_key(case["dis_keys"], Lmax, Lmax)
    h = 1.0 / (1 << Lmax)
    ex, ey, ez = C.velocities(27)
    for i in range(0, len(x), 97):
        c = np.array([x[i] + 0.5, y[i] + 0.5, z[i] + 0.5]) * h
        for q in range(1, 27):
            t = v[i, q - 1]
            if t < 0:
                continue
            p = c + t * h * np.array([ex[q], ey[q], ez[q]])
            assert abs(np.linalg.norm(p - 12.0) - 3.0) < 1e-9


def test_reported_velocity_convention():
    """Inflow cases store -u_physical (the reference's pull reports minus the
    physical velocity, lbm_kernel/Stream_Manager.cpp:53-62)."""
    case = C.sphere_amr((32, 24, 24), radius=3.0, center=(12, 12, 12), shells=(5.0, 1.5), U=0.04)
    assert case["faces"][0][1] == -0.04 and case["u0"][0] == -0.04


# ---------------------------------------------------------------- C-ABI library
def header_symbols():
    txt = open(os.path.join(ROOT, "include", "mt_amr.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(mt_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(amr.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing
    assert sorted(syms) == sorted(amr.EXPORTS)


def _desc(**kw):
    d = amr.CaseDesc()
    d.Q, d.Lmax, d.collision, d.ib, d.precision = 19, 0, 0, 0, 0
    d.dims0[:] = [8, 8, 8]
    d.tau0, d.rho0 = 0.6, 1.0
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("kw,frag", [
    (dict(Q=20), "Q must be 19 or 27"),
    (dict(collision=1), "cumulant collision needs Q = 27"),
    (dict(tau0=0.5), "tau0"),
    (dict(Lmax=9), "Lmax"),
    (dict(precision=7), "precision"),
])
def test_create_rejects_bad_descriptors(kw, frag):
    """Argument validation happens before any device call (the reference's
    log_error paths, e.g. Collision_Cumulant.cpp:24-26)."""
    lib = amr.load()
    h = ctypes.c_void_p()
    rc = lib.mt_case_create(ctypes.byref(_desc(**kw)), ctypes.byref(h))
    assert rc == -1
    assert frag in lib.mt_last_error().decode()


def test_periodic_faces_must_pair():
    lib = amr.load()
    d = _desc()
    d.face_kind[0] = 4
    h = ctypes.c_void_p()
    assert lib.mt_case_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert "pairs" in lib.mt_last_error().decode()


def test_null_handles_are_rejected():
    lib = amr.load()
    assert lib.mt_step(None, 1) == -1
    assert lib.mt_macros(None, 0) == -1
    assert lib.mt_case_destroy(None) == 0


def test_cpp_driver_header_compiles():
    """The LBM_Manager-shaped C++ driver (include/mt_lbm_manager.hpp) and the
    driver check program compile against the C-ABI."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = os.path.join(ROOT, "tests", "cpp", "driver_check.cpp")
    subprocess.check_call([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"), src])


# ---------------------------------------------------------------- packed cases
def test_pack_layout_and_key_translation():
    """pack(): members keep their cells, boxes are 4-aligned and disjoint, and
    member keys translate into the packed coordinates level by level."""
    ms = [C.cavity(12), C.cavity(16, refine_box=((5, 5, 5), (11, 11, 11))), C.cavity(8, tau=0.8, lid=0.08)]
    p = C.pack(ms)
    assert p["Lmax"] == 1 and len(p["members"]) == 3
    for L in range(p["Lmax"] + 1):
        n = sum(m["levels"][L]["keys"].size for m in ms if L <= m["Lmax"])
        assert p["levels"][L]["keys"].size == n
        assert np.unique(p["levels"][L]["keys"]).size == n
    xs = []
    for i, (m, md) in enumerate(zip(ms, p["members"])):
        assert md["origin0"][0] % 4 == 0
        xs.append((md["origin0"][0], md["origin0"][0] + md["dims0"][0]))
        for L in range(m["Lmax"] + 1):
            k = C.member_keys(p, i, L, m["levels"][L]["keys"])
            assert np.isin(k, p["levels"][L]["keys"]).all()
            x, y, z = C.decode_key(k, L, p["Lmax"])
            assert ((x >> L) >= md["origin0"][0]).all() and ((x >> L) < md["origin0"][0] + md["dims0"][0]).all()
    xs.sort()
    assert all((a1 + 3) // 4 * 4 <= b0 for (a0, a1), (b0, b1) in zip(xs, xs[1:]))


def test_pack_rejects_mixed_models():
    with pytest.raises(ValueError):
        C.pack([C.cavity(8), C.cavity(8, Q=27)])


def test_mesh_body_geometry(tmp_path):
    """The host mesh rules the device builder mirrors: winding-number inside,
    distance to the surface and Moller-Trumbore link crossings on an icosphere
    against the sphere it approximates; the ASCII STL reader on the same mesh."""
    Cc = np.array([5.2, 4.9, 5.1])
    T = C.icosphere(Cc, 2.0, 3)
    body = C.MeshBody(T)
    rng = np.random.default_rng(2)
    p = Cc + rng.uniform(-4, 4, (2000, 3))
    r = np.linalg.norm(p - Cc, axis=1)
    keep = np.abs(r - 2.0) > 0.05  # the mesh lies within 2 * (1 - cos 9.7 deg) < 0.03 of the sphere
    ins = body.inside(p[:, 0], p[:, 1], p[:, 2])
    assert np.array_equal(ins[keep], (r < 2.0)[keep])
    d = np.sqrt(body.dist2(p[:, 0], p[:, 1], p[:, 2]))
    assert np.abs(d - np.abs(r - 2.0)).max() < 0.03
    # a link from outside straight at the centre crosses at r - 2 (within the facet sag)
    q = Cc + np.array([3.0, 0.0, 0.0])
    t = body.link(q[0], q[1], q[2], -1.0, 0.0, 0.0, 2.0)
    assert abs(float(t) * 2.0 - 1.0) < 0.03
    assert float(body.link(q[0], q[1], q[2], 1.0, 0.0, 0.0, 2.0)) > 1.0  # pointing away: no crossing
    f = tmp_path / "body.stl"
    with open(f, "w") as fh:
        fh.write("solid s\n")
        for tri in T:
            fh.write(" facet normal 0 0 0\n  outer loop\n")
            for v in tri:
                fh.write(f"   vertex {float(v[0])!r} {float(v[1])!r} {float(v[2])!r}\n")
            fh.write("  endloop\n endfacet\n")
        fh.write("endsolid s\n")
    assert np.array_equal(C.read_stl(str(f)), T)


def test_rigid_body_free_motion_closed_form():
    """Constant impulse per step and gravity: V_n = V_0 + n (J/m + g), x_n = x_0 + sum V_k."""
    from mortontrie_amr_amd.motion import RigidBody
    b = RigidBody((10.0, 8.0, 8.0), velocity=(0.01, 0.0, 0.0), mass=4.0, gravity=(0.0, 0.0, -1e-4))
    J = np.array([0.002, -0.004, 0.0])
    for _ in range(5):
        b.advance(J)
    dv = J / 4.0 + np.array([0.0, 0.0, -1e-4])
    assert np.allclose(b.velocity, np.array([0.01, 0.0, 0.0]) + 5 * dv, rtol=0, atol=1e-15)
    want = np.array([10.0, 8.0, 8.0]) + sum(np.array([0.01, 0.0, 0.0]) + k * dv for k in range(1, 6))
    assert np.allclose(b.center, want, rtol=0, atol=1e-14)
    p = RigidBody((1.0, 2.0, 3.0), velocity=(0.5, 0.0, 0.0))
    p.advance(J, dt=2.0)
    assert np.array_equal(p.velocity, [0.5, 0.0, 0.0]) and np.allclose(p.center, [2.0, 2.0, 3.0])


def test_serving_eviction_repack_pays_for_itself():
    """An eviction-only repack is due when the finished members' share of the
    batch's footprint times the chunks the batch may still run reaches the
    repack cost (in chunks); a batch with nothing running or nothing finished
    never asks for one."""
    from mortontrie_amr_amd import serve

    def req(cells, max_steps, steps, done):
        case = {"Q": 27, "levels": [{"keys": np.zeros(cells, dtype=np.uint64)}]}
        r = serve.Request(0, case, 0.0, max_steps=max_steps)
        r.steps, r.done = steps, (1.0 if done else None)
        return r

    srv = serve.Server(chunk=100, repack_chunks=1.5)
    # a quarter of the footprint finished, 5 chunks left: 1.25 < 1.5
    srv.members = [req(100, 300, 300, True), req(300, 800, 300, False)]
    assert not srv._evict_due()
    # 6 chunks left (a partial chunk counts whole): 1.5 >= 1.5
    srv.members = [req(100, 300, 300, True), req(300, 850, 300, False)]
    assert srv._evict_due()
    # half finished, 2 chunks left: 1.0 < 1.5; a third member with 4 left: 2.0
    srv.members = [req(200, 100, 100, True), req(200, 500, 300, False)]
    assert not srv._evict_due()
    srv.members.append(req(0, 700, 300, False))
    assert srv._evict_due()
    srv.members = [req(200, 100, 100, True)]
    assert not srv._evict_due()
    srv.members = [req(200, 900, 100, False)]
    assert not srv._evict_due()


def test_graph_length_is_even():
    """mt_step replays graphs of mt_steps_per_graph() coarse steps; level 0's
    buffers flip once per coarse step, so an even length is what lets a timed
    call of one graph's length replay the graph an earlier call captured."""
    G = amr.steps_per_graph()
    assert G >= 2 and G % 2 == 0
