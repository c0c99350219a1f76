"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this,
as the checker.  It loads oracle/_build/libmt_oracle.so (built by oracle/Makefile).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MT_ORACLE_LIB points the checker at another build of the oracle (mutation runs)
LIB_PATH = os.environ.get("MT_ORACLE_LIB") or os.path.join(HERE, "_build", "libmt_oracle.so")

_lib = None

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_u16p = C.POINTER(C.c_uint16)
_ip = C.POINTER(C.c_int)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.mto_create.restype = C.c_void_p
        L.mto_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, _dp, _ip, _ip,
                                 _dp, C.c_int, C.c_int]
        L.mto_destroy.argtypes = [C.c_void_p]
        L.mto_set_options.argtypes = [C.c_void_p, C.c_int]
        L.mto_set_tau.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.mto_set_mdf_sweeps.argtypes = [C.c_void_p, C.c_int]
        L.mto_diagnostics.argtypes = [C.c_void_p, _dp, _u64p, _ip]
        L.mto_dense_create.restype = C.c_void_p
        L.mto_dense_create.argtypes = [C.c_void_p, C.c_int]
        L.mto_dense_destroy.argtypes = [C.c_void_p]
        L.mto_dense_step.argtypes = [C.c_void_p, C.c_int]
        L.mto_dense_get.argtypes = [C.c_void_p, C.c_int, C.c_int, _u64p, C.c_int64, _dp]
        L.mto_dense_set.argtypes = [C.c_void_p, C.c_int, C.c_int, _u64p, C.c_int64, _dp]
        L.mto_dense_force_log.restype = C.c_int64
        L.mto_dense_force_log.argtypes = [C.c_void_p, _dp, C.c_int64]
        L.mto_add_level.argtypes = [C.c_void_p, C.c_int, _u64p, _u16p, C.c_int64]
        L.mto_set_dis.argtypes = [C.c_void_p, _u64p, _dp, C.c_int64]
        L.mto_set_markers.argtypes = [C.c_void_p, C.c_int64, _u64p, _dp, _dp, _dp]
        L.mto_init.argtypes = [C.c_void_p]
        L.mto_step.argtypes = [C.c_void_p, C.c_int]
        L.mto_get.argtypes = [C.c_void_p, C.c_int, C.c_int, _u64p, C.c_int64, _dp]
        L.mto_set.argtypes = [C.c_void_p, C.c_int, C.c_int, _u64p, C.c_int64, _dp]
        for name in ("mto_macros", "mto_collide", "mto_stream", "mto_explode", "mto_coalesce"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int]
        L.mto_bc_after_stream.argtypes = [C.c_void_p]
        L.mto_ib_treat.argtypes = [C.c_void_p, _dp]
        L.mto_marker_correct.argtypes = [C.c_void_p]
        L.mto_force_log.restype = C.c_int64
        L.mto_force_log.argtypes = [C.c_void_p, _dp, C.c_int64]
        L.mto_feq.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.mto_feq_heluo.argtypes = [C.c_int, C.c_double, _dp, C.c_double, _dp]
        L.mto_macros_cell.argtypes = [C.c_int, _dp, _dp, _dp]
        L.mto_cumulant.argtypes = [_dp, C.c_double, _dp, C.c_double, _dp]
        for name in ("mto_central_fwd", "mto_central_bwd"):
            getattr(L, name).argtypes = [_dp, _dp, _dp]
        for name in ("mto_to_cumulants", "mto_from_cumulants"):
            getattr(L, name).argtypes = [_dp, C.c_double, _dp]
        L.mto_stress_norm.restype = C.c_double
        L.mto_stress_norm.argtypes = [C.c_int, _dp, _dp]
        L.mto_smag_tau.restype = C.c_double
        L.mto_smag_tau.argtypes = [C.c_double, C.c_double, C.c_double]
        L.mto_find_neighbor.restype = C.c_uint64
        L.mto_find_neighbor.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _k(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(_u64p)


class OracleCase:
    """A case held by the oracle; mirrors LBM_Manager's state and loop."""

    def __init__(self, case, nee_sequential=True, _init=True):
        L = lib()
        self.case = case
        self.Q = case["Q"]
        u0, pu0 = _d(case["u0"])
        dims = np.ascontiguousarray(case["dims0"], dtype=np.int32)
        fk = np.ascontiguousarray([f[0] for f in case["faces"]], dtype=np.int32)
        fu, pfu = _d([v for f in case["faces"] for v in f[1:4]])
        self.h = L.mto_create(self.Q, case["Lmax"], case["tau0"], case["collision"], case["les_C"],
                              case["rho0"], pu0, dims.ctypes.data_as(_ip), fk.ctypes.data_as(_ip), pfu,
                              case["ib"], 1 if nee_sequential else 0)
        L.mto_set_options(self.h, case.get("options", 0))
        for lvl, lv in enumerate(case["levels"]):
            keys, pk = _k(lv["keys"])
            kinds = np.ascontiguousarray(lv["kinds"], dtype=np.uint16)
            L.mto_add_level(self.h, lvl, pk, kinds.ctypes.data_as(_u16p), keys.size)
        if len(case["dis_keys"]):
            dk, pdk = _k(case["dis_keys"])
            dv, pdv = _d(case["dis_vals"])
            L.mto_set_dis(self.h, pdk, pdv, dk.size)
        mk = case.get("markers")
        if mk is not None:
            keys = np.ascontiguousarray(mk["keys"], dtype=np.uint64)
            w = np.ascontiguousarray(mk["w"], dtype=np.float64)
            dv = np.ascontiguousarray(mk["dv"], dtype=np.float64)
            ub = np.ascontiguousarray(mk["ub"], dtype=np.float64)
            L.mto_set_markers(self.h, dv.size, keys.ctypes.data_as(_u64p), w.ctypes.data_as(_dp),
                              dv.ctypes.data_as(_dp), ub.ctypes.data_as(_dp))
        if _init:
            L.mto_init(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mto_destroy(self.h)
            self.h = None

    def step(self, n=1):
        lib().mto_step(self.h, n)

    def diagnostics(self):
        """infoPrint's velocity error and per-level max |v|, diagL2MaxF and
        diagL2RhoMin of the current state (the map mode keeps v_old as the
        reference does)."""
        vals = np.zeros(25)
        keys = np.zeros(3, np.uint64)
        qs = np.zeros(2, np.int32)
        lib().mto_diagnostics(self.h, vals.ctypes.data_as(_dp), keys.ctypes.data_as(_u64p),
                              qs.ctypes.data_as(_ip))
        n = self.case["Lmax"] + 1
        return {"velocity_error": vals[0], "max_v": vals[1:1 + n], "n_big": vals[9:9 + n].astype(np.int64),
                "max_f": vals[17], "max_fcol": vals[18], "min_rho_f": vals[19], "min_rho_fcol": vals[20],
                "n_neg_f": int(vals[21]), "n_neg_fcol": int(vals[22]), "n_low_f": int(vals[23]),
                "worst_is_dis": int(vals[24]), "max_f_key": int(keys[0]), "max_fcol_key": int(keys[1]),
                "worst_key": int(keys[2]), "max_f_q": int(qs[0]), "max_fcol_q": int(qs[1])}

    def set_mdf_sweeps(self, n):
        """Number of multi-direct forcing sweeps treat() runs (the reference's 8)."""
        lib().mto_set_mdf_sweeps(self.h, int(n))

    def set_tau(self, taus):
        """Replaces tau_L (after init) with the given per-level values."""
        for L, t in enumerate(taus):
            if lib().mto_set_tau(self.h, L, float(t)):
                raise ValueError(f"no level {L}")

    def get(self, level, what, keys=None):
        if keys is None:
            keys = self.case["levels"][level]["keys"]
        keys, pk = _k(keys)
        n = keys.size
        rows = {0: self.Q, 1: self.Q, 2: 1, 3: 3, 4: 3}[what]
        out = np.empty(rows * n)
        lib().mto_get(self.h, level, what, pk, n, out.ctypes.data_as(_dp))
        return out.reshape(rows, n) if rows > 1 else out

    def set(self, level, what, values, keys=None):
        if keys is None:
            keys = self.case["levels"][level]["keys"]
        keys, pk = _k(keys)
        v, pv = _d(values)
        lib().mto_set(self.h, level, what, pk, keys.size, pv)

    def f(self, level, keys=None):
        return self.get(level, 0, keys)

    def fcol(self, level, keys=None):
        return self.get(level, 1, keys)

    def rho(self, level, keys=None):
        return self.get(level, 2, keys)

    def u(self, level, keys=None):
        return self.get(level, 3, keys)

    # seams, named as in LBM_Manager / the strategy classes
    def macros(self, L):
        lib().mto_macros(self.h, L)

    def collide(self, L):
        lib().mto_collide(self.h, L)

    def stream(self, L):
        lib().mto_stream(self.h, L)

    def explode(self, L):
        lib().mto_explode(self.h, L)

    def coalesce(self, L):
        lib().mto_coalesce(self.h, L)

    def bc_after_stream(self):
        lib().mto_bc_after_stream(self.h)

    def ib_treat(self):
        F = np.zeros(3)
        if lib().mto_ib_treat(self.h, F.ctypes.data_as(_dp)):
            raise KeyError("treat: a neighbour of a solid-boundary cell is missing")
        return F

    def marker_correct(self):
        """The markers' direct forcing of the finest level's fcol (one call)."""
        lib().mto_marker_correct(self.h)

    def force_log(self):
        n = lib().mto_force_log(self.h, None, 0)
        out = np.zeros(3 * max(n, 1))
        lib().mto_force_log(self.h, out.ctypes.data_as(_dp), n)
        return out[: 3 * n].reshape(n, 3)


class DenseOracleCase(OracleCase):
    """The oracle's dense mode: the same algorithm and the same bits as OracleCase,
    on sorted arrays with index tables and threads, fast enough for full-size runs.
    Whole steps, field access and the force log only (no single seams).
    from_state=OracleCase starts from that case's current state; otherwise the
    initial state is built from the level sets without allocating the maps."""

    def __init__(self, case, nee_sequential=True, from_state=None):
        if from_state is not None:
            self._mapcase = from_state
            self.case, self.Q = from_state.case, from_state.Q
            self.d = lib().mto_dense_create(from_state.h, 0)
        else:
            self._mapcase = OracleCase(case, nee_sequential, _init=False)
            self.case, self.Q = case, case["Q"]
            self.d = lib().mto_dense_create(self._mapcase.h, 1)
        if not self.d:
            raise RuntimeError("dense oracle: a transfer child or boundary neighbour is missing")
        self.h = None  # the seams of the map mode are not offered

    def __del__(self):
        if getattr(self, "d", None) and _lib is not None:
            _lib.mto_dense_destroy(self.d)
            self.d = None

    def step(self, n=1):
        lib().mto_dense_step(self.d, n)

    def get(self, level, what, keys=None):
        if keys is None:
            keys = self.case["levels"][level]["keys"]
        keys, pk = _k(keys)
        n = keys.size
        rows = {0: self.Q, 1: self.Q, 2: 1, 3: 3, 4: 3}[what]
        out = np.empty(rows * n)
        if lib().mto_dense_get(self.d, level, what, pk, n, out.ctypes.data_as(_dp)):
            raise KeyError("key not in level")
        return out.reshape(rows, n) if rows > 1 else out

    def set(self, level, what, values, keys=None):
        if keys is None:
            keys = self.case["levels"][level]["keys"]
        keys, pk = _k(keys)
        v, pv = _d(values)
        if lib().mto_dense_set(self.d, level, what, pk, keys.size, pv):
            raise KeyError("key not in level")

    def force_log(self):
        n = lib().mto_dense_force_log(self.d, None, 0)
        out = np.zeros(3 * max(n, 1))
        lib().mto_dense_force_log(self.d, out.ctypes.data_as(_dp), n)
        return out[: 3 * n].reshape(n, 3)

    def _no_seam(self, *a):
        raise NotImplementedError("the dense oracle runs whole steps only")

    macros = collide = stream = explode = coalesce = bc_after_stream = ib_treat = marker_correct = _no_seam

    def set_tau(self, taus):
        self._mapcase.set_tau(taus)

    def diagnostics(self):
        """infoPrint's velocity error and per-level max |v|, diagL2MaxF and
        diagL2RhoMin of the current state (the map mode keeps v_old as the
        reference does)."""
        vals = np.zeros(25)
        keys = np.zeros(3, np.uint64)
        qs = np.zeros(2, np.int32)
        lib().mto_diagnostics(self.h, vals.ctypes.data_as(_dp), keys.ctypes.data_as(_u64p),
                              qs.ctypes.data_as(_ip))
        n = self.case["Lmax"] + 1
        return {"velocity_error": vals[0], "max_v": vals[1:1 + n], "n_big": vals[9:9 + n].astype(np.int64),
                "max_f": vals[17], "max_fcol": vals[18], "min_rho_f": vals[19], "min_rho_fcol": vals[20],
                "n_neg_f": int(vals[21]), "n_neg_fcol": int(vals[22]), "n_low_f": int(vals[23]),
                "worst_is_dis": int(vals[24]), "max_f_key": int(keys[0]), "max_fcol_key": int(keys[1]),
                "worst_key": int(keys[2]), "max_f_q": int(qs[0]), "max_fcol_q": int(qs[1])}

    def set_mdf_sweeps(self, n):
        self._mapcase.set_mdf_sweeps(n)


# ---- primitives ----------------------------------------------------------------
def feq(Q, rho, u):
    out = np.zeros(Q)
    uu, pu = _d(u)
    lib().mto_feq(Q, rho, pu, out.ctypes.data_as(_dp))
    return out


def feq_heluo(Q, rho, u, rho0=1.0):
    out = np.zeros(Q)
    uu, pu = _d(u)
    lib().mto_feq_heluo(Q, rho, pu, rho0, out.ctypes.data_as(_dp))
    return out


def macros_cell(f):
    f, pf = _d(f)
    rho = np.zeros(1)
    u = np.zeros(3)
    lib().mto_macros_cell(f.size, pf, rho.ctypes.data_as(_dp), u.ctypes.data_as(_dp))
    return rho[0], u


def cumulant(f, rho, u, tau):
    f, pf = _d(f)
    uu, pu = _d(u)
    out = np.zeros(27)
    lib().mto_cumulant(pf, rho, pu, tau, out.ctypes.data_as(_dp))
    return out


def central_fwd(f, u):
    f, pf = _d(f)
    uu, pu = _d(u)
    k = np.zeros(27)
    lib().mto_central_fwd(pf, pu, k.ctypes.data_as(_dp))
    return k.reshape(3, 3, 3)


def central_bwd(k, u):
    k, pk = _d(k)
    uu, pu = _d(u)
    f = np.zeros(27)
    lib().mto_central_bwd(pk, pu, f.ctypes.data_as(_dp))
    return f


def to_cumulants(k, rho):
    k, pk = _d(k)
    C_ = np.zeros(27)
    lib().mto_to_cumulants(pk, rho, C_.ctypes.data_as(_dp))
    return C_.reshape(3, 3, 3)


def from_cumulants(C_, rho, k_in):
    """Backward cumulant transform in place on a copy of k_in (members the
    reference never reassigns keep k_in's values)."""
    C_, pc = _d(C_)
    k = np.ascontiguousarray(np.array(k_in, dtype=np.float64).reshape(27))
    lib().mto_from_cumulants(pc, rho, k.ctypes.data_as(_dp))
    return k.reshape(3, 3, 3)


def stress_norm(Q, f, fe):
    f, pf = _d(f)
    fe, pe = _d(fe)
    return lib().mto_stress_norm(Q, pf, pe)


def smag_tau(tau, C_, qn):
    return lib().mto_smag_tau(tau, C_, qn)


def find_neighbor(key, level, Lmax, dx, dy, dz):
    return int(lib().mto_find_neighbor(int(key), level, Lmax, dx, dy, dz))
