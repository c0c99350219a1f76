"""Python front end of libmt_amr.so (the C-ABI in include/mt_amr.h).

`AmrCase` mirrors the state and the methods LBM_Manager drives (LBM_Manager.hpp:58-88):
calculateMacros / collide / stream / AMR_transDDF_explosion / coalescence /
applyBoundaryCondition / treat, plus `step(n)`, the fluidSimulate loop body.
The library must be built (``python -m mortontrie_amr_amd.build``); there is no
fallback path: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import cases as _cases

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmt_amr.so")

FP64, FP32_SHIFTED, FP16_STORAGE = 0, 1, 2
FIELD_F, FIELD_FCOL, FIELD_RHO, FIELD_U = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_u16p = C.POINTER(C.c_uint16)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class CaseDesc(C.Structure):
    _fields_ = [
        ("Q", C.c_int32), ("Lmax", C.c_int32), ("collision", C.c_int32), ("ib", C.c_int32),
        ("precision", C.c_int32), ("dims0", C.c_int32 * 3), ("face_kind", C.c_int32 * 6),
        ("face_u", (C.c_double * 3) * 6), ("tau0", C.c_double), ("les_C", C.c_double),
        ("rho0", C.c_double), ("u0", C.c_double * 3), ("options", C.c_int32),
    ]


class MemberDesc(C.Structure):
    _fields_ = [
        ("origin0", C.c_int32 * 3), ("dims0", C.c_int32 * 3), ("face_kind", C.c_int32 * 6),
        ("face_u", (C.c_double * 3) * 6), ("tau0", C.c_double), ("les_C", C.c_double),
        ("rho0", C.c_double), ("u0", C.c_double * 3),
    ]


EXPORTS = [
    "mt_last_error", "mt_case_create", "mt_case_destroy", "mt_case_upload_level", "mt_case_upload_dis",
    "mt_case_finalize", "mt_case_level_info", "mt_find", "mt_find_noalloc", "mt_get_field", "mt_set_field", "mt_macros",
    "mt_collide", "mt_stream", "mt_explode", "mt_coalesce", "mt_apply_bc", "mt_ib_treat", "mt_step",
    "mt_set_graphs", "mt_set_tau", "mt_set_schedule", "mt_get_schedule", "mt_steps_per_graph", "mt_set_timing", "mt_get_timing", "mt_sync", "mt_reduce", "mt_ib_force",
    "mt_case_set_members", "mt_ib_force_members", "mt_autotune", "mt_case_copy_member",
    "mt_case_upload_markers", "mt_markers_move", "mt_markers_rigid", "mt_markers_get", "mt_case_remap", "mt_scan_members", "mt_set_diagnostics", "mt_diagnostics", "mt_levelset_build",
    "mt_levelset_counts", "mt_levelset_copy", "mt_levelset_copy_dis", "mt_levelset_destroy", "mt_snapshot",
    "mt_write_tecplot", "mt_write_nodes", "mt_server_create", "mt_server_submit", "mt_server_step",
    "mt_server_poll",
    "mt_server_destroy",
]

_lib = None


# MT_FIND_NOALLOC_MAX_KEYS of include/mt_amr.h: the largest call of mt_find_noalloc
FIND_NOALLOC_MAX_KEYS = 1 << 18


class Diag(C.Structure):
    """mt_diag of include/mt_amr.h."""
    _fields_ = [("velocity_error", C.c_double), ("max_v", C.c_double * 8), ("n_big", C.c_int64 * 8),
                ("max_f", C.c_double), ("max_fcol", C.c_double), ("max_f_key", C.c_uint64),
                ("max_fcol_key", C.c_uint64), ("max_f_q", C.c_int32), ("max_fcol_q", C.c_int32),
                ("min_rho_f", C.c_double), ("min_rho_fcol", C.c_double), ("n_neg_f", C.c_int64),
                ("n_neg_fcol", C.c_int64), ("n_low_f", C.c_int64), ("worst_key", C.c_uint64),
                ("worst_is_dis", C.c_int32), ("valid_v_old", C.c_int32)]


class LevelSetDesc(C.Structure):
    """mt_levelset_desc of include/mt_amr.h."""
    _fields_ = [("Q", C.c_int32), ("Lmax", C.c_int32), ("dims0", C.c_int32 * 3), ("face_kind", C.c_int32 * 6),
                ("reservoir", C.c_int32), ("shape", C.c_int32), ("solid", C.c_int32), ("shells", C.c_double * 8),
                ("center", C.c_double * 3), ("radius", C.c_double), ("tri", C.POINTER(C.c_double)),
                ("n_tri", C.c_int64)]


SHAPE_SPHERE, SHAPE_MESH = 0, 1


class MtError(RuntimeError):
    pass


def steps_per_graph():
    """Coarse steps one captured HIP graph of mt_step holds."""
    return int(load().mt_steps_per_graph())


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MtError(f"{LIB_PATH} is missing: build it with `python -m mortontrie_amr_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.mt_last_error.restype = C.c_char_p
    L.mt_case_create.argtypes = [C.POINTER(CaseDesc), C.POINTER(C.c_void_p)]
    L.mt_case_destroy.argtypes = [C.c_void_p]
    L.mt_case_upload_level.argtypes = [C.c_void_p, C.c_int, _u64p, _u16p, C.c_int64]
    L.mt_case_upload_dis.argtypes = [C.c_void_p, _u64p, _dp, C.c_int64]
    L.mt_case_finalize.argtypes = [C.c_void_p]
    L.mt_case_level_info.argtypes = [C.c_void_p, C.c_int, _i64p, _i64p]
    L.mt_find.argtypes = [C.c_void_p, C.c_int, _u64p, C.c_int64, _i32p]
    L.mt_find_noalloc.argtypes = [C.c_void_p, C.c_int, _u64p, C.c_int64, _i32p]
    L.mt_get_field.argtypes = [C.c_void_p, C.c_int, C.c_int, _u64p, C.c_int64, _dp]
    L.mt_set_field.argtypes = [C.c_void_p, C.c_int, C.c_int, _u64p, C.c_int64, _dp]
    for n in ("mt_macros", "mt_collide", "mt_stream", "mt_explode", "mt_coalesce"):
        getattr(L, n).argtypes = [C.c_void_p, C.c_int]
    L.mt_apply_bc.argtypes = [C.c_void_p]
    L.mt_ib_treat.argtypes = [C.c_void_p, _dp]
    L.mt_step.argtypes = [C.c_void_p, C.c_int]
    L.mt_set_graphs.argtypes = [C.c_void_p, C.c_int]
    L.mt_set_schedule.argtypes = [C.c_void_p, C.c_int]
    L.mt_get_schedule.argtypes = [C.c_void_p, _i32p]
    L.mt_steps_per_graph.argtypes = []
    L.mt_steps_per_graph.restype = C.c_int
    L.mt_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.mt_get_timing.argtypes = [C.c_void_p, _dp, _i64p, C.c_int]
    L.mt_sync.argtypes = [C.c_void_p]
    L.mt_reduce.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _i64p]
    L.mt_ib_force.argtypes = [C.c_void_p, _dp, _i64p, C.c_int]
    L.mt_case_set_members.argtypes = [C.c_void_p, C.POINTER(MemberDesc), C.c_int32]
    L.mt_ib_force_members.argtypes = [C.c_void_p, _dp, C.c_int32, _i64p, C.c_int]
    L.mt_autotune.argtypes = [C.c_void_p, C.c_int, _dp]
    L.mt_case_copy_member.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
    L.mt_case_upload_markers.argtypes = [C.c_void_p, C.c_int64, _u64p, _dp, _dp, _dp]
    L.mt_server_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]
    L.mt_server_submit.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_double,
                                   C.c_double]
    L.mt_server_step.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.mt_server_poll.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), _dp,
                                 C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]
    L.mt_scan_members.argtypes = [C.c_void_p, C.c_int32, _i64p]
    L.mt_set_diagnostics.argtypes = [C.c_void_p, C.c_int]
    L.mt_levelset_build.argtypes = [C.POINTER(LevelSetDesc), C.POINTER(C.c_void_p)]
    L.mt_levelset_counts.argtypes = [C.c_void_p, _i64p, _i64p]
    L.mt_levelset_copy.argtypes = [C.c_void_p, C.c_int, _u64p, _u16p]
    L.mt_levelset_copy_dis.argtypes = [C.c_void_p, _u64p, _dp]
    L.mt_levelset_destroy.argtypes = [C.c_void_p]
    L.mt_snapshot.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _i64p, _u64p, _dp, _dp]
    L.mt_write_tecplot.argtypes = [C.c_void_p, C.c_char_p, C.c_double, C.c_double]
    L.mt_write_nodes.argtypes = [C.c_void_p, C.c_char_p]
    L.mt_diagnostics.argtypes = [C.c_void_p, C.POINTER(Diag)]
    L.mt_set_tau.argtypes = [C.c_void_p, C.c_int32, _dp, C.c_int32]
    L.mt_markers_move.argtypes = [C.c_void_p, C.c_int64, _dp, _dp]
    L.mt_case_remap.argtypes = [C.c_void_p, C.c_void_p, _dp, C.POINTER(C.c_int64)]
    L.mt_markers_rigid.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
    L.mt_markers_get.argtypes = [C.c_void_p, C.c_int64, _dp, _dp]
    L.mt_server_destroy.argtypes = [C.c_void_p]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise MtError(f"mt error {rc}: {load().mt_last_error().decode()}")


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(_u64p)


class AmrCase:
    """A device-resident AMR case built from a `cases` dictionary."""

    def __init__(self, case, precision=FP64, graphs=True, finalize=True):
        """finalize=False stops after the uploads (a request for mt_server_submit, see detach())."""
        L = load()
        self.case = case
        self.Q = case["Q"]
        self.Lmax = case["Lmax"]
        d = CaseDesc()
        d.Q, d.Lmax, d.collision, d.ib, d.precision = self.Q, self.Lmax, case["collision"], case["ib"], precision
        for a in range(3):
            d.dims0[a] = case["dims0"][a]
            d.u0[a] = case["u0"][a]
        for f in range(6):
            d.face_kind[f] = case["faces"][f][0]
            for a in range(3):
                d.face_u[f][a] = case["faces"][f][1 + a]
        d.tau0, d.les_C, d.rho0 = case["tau0"], case["les_C"], case["rho0"]
        d.options = case.get("options", 0)
        h = C.c_void_p()
        _check(L.mt_case_create(C.byref(d), C.byref(h)))
        self.h = h
        members = case.get("members")
        if members:
            md = (MemberDesc * len(members))()
            for i, m in enumerate(members):
                for a in range(3):
                    md[i].origin0[a] = m["origin0"][a]
                    md[i].dims0[a] = m["dims0"][a]
                    md[i].u0[a] = m["u0"][a]
                for f in range(6):
                    md[i].face_kind[f] = m["faces"][f][0]
                    for a in range(3):
                        md[i].face_u[f][a] = m["faces"][f][1 + a]
                md[i].tau0, md[i].les_C, md[i].rho0 = m["tau0"], m["les_C"], m["rho0"]
            _check(L.mt_case_set_members(self.h, md, len(members)))
        for lvl, lv in enumerate(case["levels"]):
            keys, pk = _u64(lv["keys"])
            kinds = np.ascontiguousarray(lv["kinds"], dtype=np.uint16)
            _check(L.mt_case_upload_level(self.h, lvl, pk, kinds.ctypes.data_as(_u16p), keys.size))
        if len(case["dis_keys"]):
            dk, pdk = _u64(case["dis_keys"])
            dv = np.ascontiguousarray(case["dis_vals"], dtype=np.float64)
            _check(L.mt_case_upload_dis(self.h, pdk, dv.ctypes.data_as(_dp), dk.size))
        mk = case.get("markers")
        if mk is not None:
            mkeys, pmk = _u64(mk["keys"])
            w = np.ascontiguousarray(mk["w"], dtype=np.float64)
            dvm = np.ascontiguousarray(mk["dv"], dtype=np.float64)
            ub = np.ascontiguousarray(mk["ub"], dtype=np.float64)
            if mkeys.shape != (dvm.size, 27) or w.shape != mkeys.shape or ub.size != 3:
                raise MtError("markers: keys and w are [n][27], dv [n], ub [3]")
            _check(L.mt_case_upload_markers(self.h, dvm.size, pmk, w.ctypes.data_as(_dp),
                                            dvm.ctypes.data_as(_dp), ub.ctypes.data_as(_dp)))
        if finalize:
            _check(L.mt_case_finalize(self.h))
            _check(L.mt_set_graphs(self.h, 1 if graphs else 0))

    def detach(self):
        """Give up the handle (its new owner frees it)."""
        h, self.h = self.h, None
        return h

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.mt_case_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- info / lookup
    def level_info(self, level):
        n = C.c_int64()
        b = C.c_int64()
        _check(load().mt_case_level_info(self.h, level, C.byref(n), C.byref(b)))
        return n.value, b.value

    def find(self, level, keys):
        keys, pk = _u64(keys)
        out = np.empty(keys.size, dtype=np.int32)
        _check(load().mt_find(self.h, level, pk, keys.size, out.ctypes.data_as(_i32p)))
        return out

    def find_noalloc(self, level, keys):
        """mt_find through the allocation-free entry point (fixed per-GPU scratch)."""
        keys, pk = _u64(keys)
        out = np.empty(keys.size, dtype=np.int32)
        _check(load().mt_find_noalloc(self.h, level, pk, keys.size, out.ctypes.data_as(_i32p)))
        return out

    # ---- fields
    def get(self, level, field, keys=None):
        if keys is None:
            keys = self.case["levels"][level]["keys"]
        keys, pk = _u64(keys)
        rows = {FIELD_F: self.Q, FIELD_FCOL: self.Q, FIELD_RHO: 1, FIELD_U: 3}[field]
        out = np.empty(rows * keys.size)
        _check(load().mt_get_field(self.h, level, field, pk, keys.size, out.ctypes.data_as(_dp)))
        return out.reshape(rows, keys.size) if rows > 1 else out

    def set(self, level, field, values, keys=None):
        if keys is None:
            keys = self.case["levels"][level]["keys"]
        keys, pk = _u64(keys)
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(load().mt_set_field(self.h, level, field, pk, keys.size, v.ctypes.data_as(_dp)))

    def f(self, level, keys=None):
        return self.get(level, FIELD_F, keys)

    def fcol(self, level, keys=None):
        return self.get(level, FIELD_FCOL, keys)

    def rho(self, level, keys=None):
        return self.get(level, FIELD_RHO, keys)

    def u(self, level, keys=None):
        return self.get(level, FIELD_U, keys)

    # ---- seams (LBM_Manager / strategy method names)
    def macros(self, level):
        _check(load().mt_macros(self.h, level))

    def collide(self, level):
        _check(load().mt_collide(self.h, level))

    def stream(self, level):
        _check(load().mt_stream(self.h, level))

    def explode(self, level):
        _check(load().mt_explode(self.h, level))

    def coalesce(self, level):
        _check(load().mt_coalesce(self.h, level))

    def bc_after_stream(self):
        _check(load().mt_apply_bc(self.h))

    def ib_treat(self):
        F = np.zeros(3)
        _check(load().mt_ib_treat(self.h, F.ctypes.data_as(_dp)))
        return F

    # ---- schedule
    def step(self, n=1):
        _check(load().mt_step(self.h, n))

    def sync(self):
        _check(load().mt_sync(self.h))

    def set_graphs(self, on):
        _check(load().mt_set_graphs(self.h, 1 if on else 0))

    SCHEDULES = {"merged": 0, "streams": 1, "serial": 2}

    def set_tau(self, tau, member=0):
        """Per-level relaxation times of one member as given (mt_set_tau)."""
        t = np.ascontiguousarray(tau, np.float64)
        _check(load().mt_set_tau(self.h, member, t.ctypes.data_as(_dp), t.size))

    def set_schedule(self, name):
        """How mt_step issues the levels' updates: "merged" (one launch per finest
        sub-step, the default), "streams" (a stream per level) or "serial"."""
        _check(load().mt_set_schedule(self.h, self.SCHEDULES[name]))

    def schedule(self):
        """The schedule mt_step runs: the one set, except "merged" where the
        explosions do not run inside the updates, which runs as "serial" ("streams"
        when no schedule was set)."""
        v = np.zeros(1, dtype=np.int32)
        _check(load().mt_get_schedule(self.h, v.ctypes.data_as(_i32p)))
        return {i: n for n, i in self.SCHEDULES.items()}[int(v[0])]

    def copy_member_from(self, member, src, src_member):
        """This case's member `member` takes the state of `src`'s member `src_member`."""
        _check(load().mt_case_copy_member(self.h, member, src.h, src_member))

    def autotune(self, steps=5):
        """Pick the fastest launch configuration (advances the run by 8 (steps + 1) coarse steps);
        returns each candidate's ms per coarse step."""
        ms = np.zeros(4)
        _check(load().mt_autotune(self.h, steps, ms.ctypes.data_as(_dp)))
        return ms

    def set_timing(self, on):
        _check(load().mt_set_timing(self.h, 1 if on else 0))

    def timing(self):
        ms = np.zeros(8)
        calls = np.zeros(8, dtype=np.int64)
        _check(load().mt_get_timing(self.h, ms.ctypes.data_as(_dp), calls.ctypes.data_as(_i64p), 8))
        return ms[: self.Lmax + 1], calls[: self.Lmax + 1]

    def reduce(self, level):
        m = C.c_double()
        p = np.zeros(3)
        n = C.c_int64()
        _check(load().mt_reduce(self.h, level, C.byref(m), p.ctypes.data_as(_dp), C.byref(n)))
        return m.value, p, n.value

    def ib_force(self, reset=False):
        s = np.zeros(3)
        n = C.c_int64()
        _check(load().mt_ib_force(self.h, s.ctypes.data_as(_dp), C.byref(n), 1 if reset else 0))
        return s, n.value

    def ib_force_members(self, reset=False):
        """Per-member force sums of a packed case (mt_ib_force_members): (n, 3) and the call count."""
        nm = len(self.case.get("members") or [0])
        s = np.zeros((nm, 3))
        n = C.c_int64()
        _check(load().mt_ib_force_members(self.h, s.ctypes.data_as(_dp), nm, C.byref(n), 1 if reset else 0))
        return s, n.value

    def set_diagnostics(self, on=True):
        """Every update of step() leaves its macros, so diagnostics() has v_old."""
        _check(load().mt_set_diagnostics(self.h, 1 if on else 0))

    def diagnostics(self):
        """infoPrint's velocity error and per-level max |v| / n(|v| > 1), and the
        finest level's diagL2MaxF / diagL2RhoMin, as device reductions."""
        d = Diag()
        _check(load().mt_diagnostics(self.h, C.byref(d)))
        n = self.case["Lmax"] + 1
        return {"velocity_error": d.velocity_error, "max_v": np.array(d.max_v[:n]),
                "n_big": np.array(d.n_big[:n], dtype=np.int64), "max_f": d.max_f, "max_fcol": d.max_fcol,
                "min_rho_f": d.min_rho_f, "min_rho_fcol": d.min_rho_fcol, "n_neg_f": d.n_neg_f,
                "n_neg_fcol": d.n_neg_fcol, "n_low_f": d.n_low_f, "worst_is_dis": d.worst_is_dis,
                "max_f_key": d.max_f_key, "max_fcol_key": d.max_fcol_key, "worst_key": d.worst_key,
                "max_f_q": d.max_f_q, "max_fcol_q": d.max_fcol_q, "valid_v_old": d.valid_v_old}

    def snapshot(self, level, kind_mask=0):
        """(keys, rho, u[3][n]) of the level's cells of the given kinds (default
        lat_f and F2C), compacted on the device in ascending key order."""
        n = C.c_int64()
        _check(load().mt_snapshot(self.h, level, kind_mask, C.byref(n), None, None, None))
        keys = np.zeros(n.value, np.uint64)
        rho = np.zeros(n.value)
        u = np.zeros((3, n.value))
        _check(load().mt_snapshot(self.h, level, kind_mask, C.byref(n), keys.ctypes.data_as(_u64p),
                                  rho.ctypes.data_as(_dp), u.ctypes.data_as(_dp)))
        return keys, rho, u

    def write_tecplot(self, path, dx0=1.0, convert_u=1.0):
        """The reference's Tecplot FEBRICK flow file (IO_TECPLOTH::write)."""
        _check(load().mt_write_tecplot(self.h, str(path).encode(), float(dx0), float(convert_u)))

    def write_nodes(self, path):
        """The finest level's keys as 64-bit binary strings (the AMR node dump)."""
        _check(load().mt_write_nodes(self.h, str(path).encode()))

    def remap_from(self, src, u_fill=(0.0, 0.0, 0.0)):
        """This case's cells take the state of `src` (mt_case_remap); returns the
        cells copied, exploded, coalesced and filled."""
        u = np.ascontiguousarray(u_fill, np.float64).reshape(3)
        cnt = (C.c_int64 * 4)()
        _check(load().mt_case_remap(self.h, src.h, u.ctypes.data_as(_dp), cnt))
        return {"copied": cnt[0], "exploded": cnt[1], "coalesced": cnt[2], "filled": cnt[3]}

    def move_markers(self, xyz, ub=None):
        """New marker positions [n][3] (finest cells) and velocities (reported frame;
        None keeps them); kernel cells and band rebuilt on the device."""
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        v = None if ub is None else np.ascontiguousarray(ub, np.float64).reshape(-1, 3)
        _check(load().mt_markers_move(self.h, x.shape[0], x.ctypes.data_as(_dp),
                                      None if v is None else v.ctypes.data_as(_dp)))

    def move_markers_rigid(self, dt, center, vel=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0)):
        """Rigid motion of every marker over dt finest time steps (mt_markers_rigid);
        vel and omega physical, lattice units."""
        a = [np.ascontiguousarray(v, np.float64).reshape(3) for v in (center, vel, omega)]
        _check(load().mt_markers_rigid(self.h, float(dt), *[v.ctypes.data_as(_dp) for v in a]))

    def markers(self):
        """(positions [n][3], velocities [n][3] in the reported frame)."""
        n = len(self.case["markers"]["dv"])
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        _check(load().mt_markers_get(self.h, n, x.ctypes.data_as(_dp), v.ctypes.data_as(_dp)))
        return x, v

    def scan_members(self):
        """Per member: fluid/F2C cells (all levels) holding a non-finite population (mt_scan_members)."""
        nm = len(self.case.get("members") or [0])
        out = np.zeros(nm, dtype=np.int64)
        _check(load().mt_scan_members(self.h, nm, out.ctypes.data_as(_i64p)))
        return out

    def drag_coefficients(self, reset=True):
        """Cd of every member of a packed case (nan for members without a solid boundary)."""
        s, n = self.ib_force_members(reset=reset)
        out = []
        for i, m in enumerate(self.case.get("members") or [{"cd": self.case["cd"]}]):
            cd = m.get("cd") or {}
            if n == 0 or "U" not in cd:
                out.append(float("nan"))
            else:
                out.append(2.0 * (s[i, 0] / n) / (cd["U"] ** 2 * np.pi * cd["d"] ** 2 / 4.0))
        return np.array(out)

    def drag_coefficient(self, reset=True):
        """Cd = 2 Fx / (rho U^2 pi d^2 / 4), Fx the mean IB force per treat() call
        (LBM_Manager.cpp:964-979); U in level-0 units, d in finest-level cells."""
        s, n = self.ib_force(reset=reset)
        if n == 0:
            return float("nan")
        U, d = self.case["cd"]["U"], self.case["cd"]["d"]
        return 2.0 * (s[0] / n) / (U * U * np.pi * d * d / 4.0)


def taus(case):
    return _cases.taus(case)


def build_levels(Q, Lmax, dims0, faces, shells, center=None, radius=None, triangles=None, solid=True,
                 reservoir=False):
    """The per-level cell sets and link fractions built on the GPU (mt_levelset_*)
    for an analytic sphere (center, radius) or a closed triangle mesh
    (triangles: [n][3][3], level-0 units).  Returns (levels, dis_keys, dis_vals)
    in the form of cases.build_nested."""
    L = load()
    d = LevelSetDesc()
    d.Q, d.Lmax = Q, Lmax
    d.dims0[:] = [int(v) for v in dims0]
    d.face_kind[:] = [int(f[0]) for f in faces]
    d.reservoir, d.solid = int(bool(reservoir)), int(bool(solid))
    for i, v in enumerate(shells):
        d.shells[i] = float(v)
    tri = None
    if triangles is None:
        d.shape = SHAPE_SPHERE
        d.center[:] = [float(v) for v in center]
        d.radius = float(radius)
    else:
        d.shape = SHAPE_MESH
        tri = np.ascontiguousarray(triangles, dtype=np.float64).reshape(-1, 9)
        d.tri = tri.ctypes.data_as(_dp)
        d.n_tri = tri.shape[0]
    h = C.c_void_p()
    _check(L.mt_levelset_build(C.byref(d), C.byref(h)))
    try:
        n = np.zeros(Lmax + 1, dtype=np.int64)
        nd = C.c_int64()
        _check(L.mt_levelset_counts(h, n.ctypes.data_as(_i64p), C.byref(nd)))
        levels = []
        for lv in range(Lmax + 1):
            keys = np.zeros(n[lv], dtype=np.uint64)
            kinds = np.zeros(n[lv], dtype=np.uint16)
            _check(L.mt_levelset_copy(h, lv, keys.ctypes.data_as(_u64p), kinds.ctypes.data_as(_u16p)))
            levels.append({"keys": keys, "kinds": kinds})
        dk = np.zeros(nd.value, dtype=np.uint64)
        dv = np.zeros((nd.value, Q - 1))
        _check(L.mt_levelset_copy_dis(h, dk.ctypes.data_as(_u64p), dv.ctypes.data_as(_dp)))
    finally:
        L.mt_levelset_destroy(h)
    return levels, dk, dv
