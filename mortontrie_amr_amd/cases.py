"""Case generators: analytic geometry -> per-level Morton-keyed cell sets.

This is the input side of the hot path, the role the reference's
``Lat_Manager::voxelize`` plays (Lat_Voxelize.cpp:17-33) for the analytic
geometries the workloads name (lid-driven cavity, refined sphere in a channel,
refined boxes).  It produces, per refinement level, the same *set contract* the
reference's level sets carry (header/Lat_Manager.hpp:245-316):

* ``lat_f``            -> KIND_FLUID
* ``lat_overlap_F2C``  -> KIND_F2C   (coarse cells covered by the fine band)
* ``lat_overlap_C2F``  -> KIND_C2F   (fine children of F2C cells)
* ``ghost_overlap_C2F``-> KIND_GHOST_C2F
* ``ghost_overlap_F2C``-> KIND_GHOST_F2C (a flag on lat_f cells)
* ``lat_sf``           -> KIND_SOLID (finest level, populations zero)
* ``lat_bc[26]``       -> KIND_BCRES (level-0 frozen equilibrium reservoir)
* ``dis``              -> KIND_DIS + per-cell link fractions

Keys follow the reference layout (Morton_Assist.cpp:185-193, Lat_Voxelize.cpp:245):
x bit i -> key bit 3i, y -> 3i+1, z -> 3i+2, of the cell's lower corner in
finest-level units, so a level-L cell key is encode(i,j,k) << 3*(Lmax-L).

Everything here is plain numpy on dense per-level boolean grids; the sets it
returns are uploaded unsorted and the device builds its tables from them.
"""
from __future__ import annotations

import math
import numpy as np

# ---- cell kinds (uint16 bit flags), shared with include/mt_amr.h ----------
KIND_FLUID = 0x0001
KIND_F2C = 0x0002
KIND_C2F = 0x0004
KIND_GHOST_C2F = 0x0008
KIND_GHOST_F2C = 0x0010
KIND_SOLID = 0x0020
KIND_BCRES = 0x0040
KIND_DIS = 0x0080
FACE_W, FACE_E, FACE_S, FACE_N, FACE_B, FACE_T = (0x0100 << i for i in range(6))
FACE_BITS = (FACE_W, FACE_E, FACE_S, FACE_N, FACE_B, FACE_T)

# ---- face boundary kinds ---------------------------------------------------
BC_NONE = 0        # missing neighbours keep their previous f (reference stream rule)
BC_RESERVOIR = 1   # frozen feq reservoir cells (compiled reference layout) + NEE on their f
BC_ZOUHE = 2       # wet-node velocity boundary (Hecht-Harting closure, generalised)
BC_OUTFLOW = 3     # zero-gradient copy from the inward neighbour
BC_PERIODIC = 4    # neighbour tables wrap
BC_NEE = 5         # wet-node velocity wall by non-equilibrium extrapolation from x - n

COLLIDE_LBGK = 0
COLLIDE_CUMULANT = 1

IB_NONE = 0
IB_FH = 1
IB_DW = 2
IB_UNIFORM = 3     # UniformBoundary::treat: seam path (mt_ib_treat) only

# D3Q19 / D3Q27 velocity sets in the reference order (header/Constants.h:492-538)
EX27 = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0,
                 1, -1, 1, -1, -1, 1, -1, 1], dtype=np.int64)
EY27 = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1,
                 1, -1, -1, 1, -1, 1, 1, -1], dtype=np.int64)
EZ27 = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1,
                 1, -1, 1, -1, 1, -1, 1, -1], dtype=np.int64)


def reported(u_physical):
    """Velocity in the reference's reported frame.

    The reference streams by pulling f_q(x) <- fcol_q(x + e_q)
    (lbm_kernel/Stream_Manager.cpp:53-62), so population q travels along -e_q and
    the velocity its macros report, sum_q f_q e_q / rho (LBM_Manager.cpp:983-1001),
    is minus the physical velocity.  Boundary and initial velocities of the
    workloads are physical (flow enters at an inflow face); this converts them."""
    return -u_physical


def velocities(Q):
    if Q not in (19, 27):
        raise ValueError("Q must be 19 or 27")
    return EX27[:Q], EY27[:Q], EZ27[:Q]


def part1by2(v):
    """Spread the low 21 bits of v so bit i lands on bit 3i."""
    v = np.asarray(v).astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def compact1by2(v):
    v = np.asarray(v).astype(np.uint64) & np.uint64(0x1249249249249249)
    v = (v ^ (v >> np.uint64(2))) & np.uint64(0x10C30C30C30C30C3)
    v = (v ^ (v >> np.uint64(4))) & np.uint64(0x100F00F00F00F00F)
    v = (v ^ (v >> np.uint64(8))) & np.uint64(0x1F0000FF0000FF)
    v = (v ^ (v >> np.uint64(16))) & np.uint64(0x1F00000000FFFF)
    v = (v ^ (v >> np.uint64(32))) & np.uint64(0x1FFFFF)
    return v


def encode_key(x, y, z, level, Lmax):
    """Reference cell key of the level-`level` cell (x,y,z) (Lat_Voxelize.cpp:245)."""
    s = np.uint64(Lmax - level)
    x = np.asarray(x).astype(np.uint64) << s
    y = np.asarray(y).astype(np.uint64) << s
    z = np.asarray(z).astype(np.uint64) << s
    return part1by2(x) | (part1by2(y) << np.uint64(1)) | (part1by2(z) << np.uint64(2))


def decode_key(key, level, Lmax):
    key = np.asarray(key).astype(np.uint64)
    s = np.uint64(Lmax - level)
    x = compact1by2(key) >> s
    y = compact1by2(key >> np.uint64(1)) >> s
    z = compact1by2(key >> np.uint64(2)) >> s
    return x.astype(np.int64), y.astype(np.int64), z.astype(np.int64)


# ---- dense-grid helpers ----------------------------------------------------
def _shift(mask, dx, dy, dz, periodic=(False, False, False)):
    """out[x,y,z] = mask[x+dx, y+dy, z+dz]; False outside unless periodic."""
    out = mask
    for axis, d in enumerate((dx, dy, dz)):
        if d == 0:
            continue
        if periodic[axis]:
            out = np.roll(out, -d, axis=axis)
        else:
            rolled = np.zeros_like(out)
            n = out.shape[axis]
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if d > 0:
                src[axis] = slice(d, n)
                dst[axis] = slice(0, n - d)
            else:
                src[axis] = slice(0, n + d)
                dst[axis] = slice(-d, n)
            rolled[tuple(dst)] = out[tuple(src)]
            out = rolled
    return out


def _dilate(mask, offsets, periodic=(False, False, False)):
    out = mask.copy()
    for (dx, dy, dz) in offsets:
        if dx == dy == dz == 0:
            continue
        out |= _shift(mask, -dx, -dy, -dz, periodic)
    return out


_OFF26 = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def _qoffsets(Q):
    ex, ey, ez = velocities(Q)
    return [(int(ex[q]), int(ey[q]), int(ez[q])) for q in range(1, Q)]


def _upsample(mask):
    return mask.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def _centers(dims, level):
    h = 1.0 / (1 << level)
    xs = (np.arange(dims[0]) + 0.5) * h
    ys = (np.arange(dims[1]) + 0.5) * h
    zs = (np.arange(dims[2]) + 0.5) * h
    return xs[:, None, None], ys[None, :, None], zs[None, None, :]


def _emit(kind_grid, level, Lmax):
    x, y, z = np.nonzero(kind_grid)
    keys = encode_key(x, y, z, level, Lmax)
    kinds = kind_grid[x, y, z].astype(np.uint16)
    order = np.argsort(keys, kind="stable")
    return {"keys": keys[order], "kinds": kinds[order]}


def _face_bits(kind_grid, dims, faces, cellmask):
    """Mark level-0 cells lying on each non-periodic domain face."""
    for f in range(6):
        if faces[f][0] == BC_PERIODIC:
            continue
        axis, side = f // 2, f % 2
        sl = [slice(None)] * 3
        sl[axis] = 0 if side == 0 else dims[axis] - 1
        sub = kind_grid[tuple(sl)]
        m = cellmask[tuple(sl)]
        sub[m] |= FACE_BITS[f]
        kind_grid[tuple(sl)] = sub


def build_nested(dims0, Lmax, refine, solid=None, faces=None, Q=19, ib_dist=None,
                 reservoir=False):
    """Generic nested-refinement set builder.

    refine(L, cx, cy, cz) -> bool grid: level-L cells to be covered by level L+1.
    solid(cx, cy, cz) -> bool grid evaluated on the finest level.
    ib_dist(cx, cy, cz, ex, ey, ez, h) -> per-cell link fraction or NaN (finest).
    reservoir=True puts a frozen lat_bc shell around the level-0 box (the compiled
    reference layout, Lat_Voxelize.cpp:242-288) instead of wet-node face cells.
    """
    faces = faces or [(BC_NONE, 0.0, 0.0, 0.0)] * 6
    periodic = tuple(faces[2 * a][0] == BC_PERIODIC for a in range(3))
    qoff = _qoffsets(Q)
    levels = []
    D = np.ones(dims0, dtype=bool)
    C2F = np.zeros(dims0, dtype=bool)
    prev = None
    grids = []
    for L in range(Lmax + 1):
        dims = tuple(int(d) << L for d in dims0)
        cx, cy, cz = _centers(dims, L)
        if L > 0:
            D = _upsample(prev["R"])
            C2F = _upsample(prev["F2C"])
        shell = np.zeros(dims, bool)
        if L == 0 and reservoir:
            shell[0, :, :] = shell[-1, :, :] = True
            shell[:, 0, :] = shell[:, -1, :] = True
            shell[:, :, 0] = shell[:, :, -1] = True
        if L < Lmax:
            R = D & refine(L, cx, cy, cz)
            if L == 0 and reservoir:
                if (R & _dilate(shell, _OFF26)).any():
                    raise ValueError("refinement touches the boundary layer")
            notR = ~R
            band = R & _dilate(notR, _OFF26, periodic if L == 0 else (False,) * 3)
            edge = np.zeros(dims, bool)
            edge[0, :, :] = edge[-1, :, :] = edge[:, 0, :] = edge[:, -1, :] = True
            edge[:, :, 0] = edge[:, :, -1] = True
            if (R & _dilate(edge, _OFF26)).any():
                raise ValueError(f"level {L}: refinement touches the domain edge")
            interior = R & ~band
            if (_dilate(R, _OFF26) & C2F).any():
                raise ValueError(f"level {L}: refinement too close to the C2F band")
        else:
            R = np.zeros(dims, bool)
            band = np.zeros(dims, bool)
            interior = R
        sol = np.zeros(dims, bool)
        if L == Lmax and solid is not None:
            sol = D & solid(cx, cy, cz) & ~C2F
        elif solid is not None:
            if (D & ~interior & solid(cx, cy, cz)).any():
                raise ValueError(f"level {L}: solid is not fully inside the finest level")
        bcres = np.zeros(dims, bool)
        if L == 0 and reservoir:
            bcres = shell.copy()
        fluid = D & ~interior & ~band & ~C2F & ~sol & ~bcres
        kind = np.zeros(dims, dtype=np.uint16)
        kind[fluid] |= KIND_FLUID
        kind[band] |= KIND_F2C
        kind[C2F] |= KIND_C2F
        kind[bcres] |= KIND_BCRES
        if L > 0:
            ghost = _dilate(C2F, qoff) & ~D
            # the parents of the ghost layer must be coarse fluid cells
            par = prev["kind"].repeat(2, 0).repeat(2, 1).repeat(2, 2)
            if (ghost & ((par & KIND_FLUID) == 0)).any():
                raise ValueError(f"level {L}: ghost_C2F parent is not a fluid cell")
            kind[ghost] |= KIND_GHOST_C2F
            # ghost_F2C on the coarse level (Lat_Voxelize.cpp:715-731)
            pk = prev["kind"]
            gf = _dilate(prev["F2C"], qoff) & ((pk & KIND_FLUID) != 0)
            pk[gf] |= KIND_GHOST_F2C
        if L == Lmax and solid is not None:
            # keep only solid cells that some fluid cell can pull from
            live = _dilate(fluid | C2F, qoff) & sol
            kind[sol & ~live] = 0
            kind[sol & live] |= KIND_SOLID
        if L == 0:
            if reservoir:
                _face_bits(kind, dims, [(BC_RESERVOIR, 0, 0, 0)] * 6, bcres)
            else:
                _face_bits(kind, dims, faces, fluid)
        rec = {"R": R, "F2C": band, "kind": kind, "dims": dims}
        grids.append(rec)
        prev = rec
    dis_keys = np.zeros(0, np.uint64)
    dis_vals = np.zeros((0, Q - 1))
    if ib_dist is not None:
        top = grids[Lmax]
        dims = top["dims"]
        cx, cy, cz = _centers(dims, Lmax)
        h = 1.0 / (1 << Lmax)
        ex, ey, ez = velocities(Q)
        fl = (top["kind"] & (KIND_FLUID | KIND_C2F)) != 0
        vals = np.full(dims + (Q - 1,), -1.0)
        for q in range(1, Q):
            t = ib_dist(cx, cy, cz, ex[q], ey[q], ez[q], h)
            ok = fl & ~np.isnan(t)
            vals[..., q - 1][ok] = t[ok]
        has = fl & (vals >= 0).any(axis=-1)
        top["kind"][has] |= KIND_DIS
        x, y, z = np.nonzero(has)
        dk = encode_key(x, y, z, Lmax, Lmax)
        order = np.argsort(dk)
        dis_keys = dk[order]
        dis_vals = vals[x, y, z][order]
    for L in range(Lmax + 1):
        levels.append(_emit(grids[L]["kind"], L, Lmax))
    return levels, dis_keys, dis_vals


def _base_case(Q, Lmax, tau0, collision, dims0, faces, levels, dis_keys, dis_vals,
               u0=(0.0, 0.0, 0.0), rho0=1.0, ib=IB_NONE, les_C=0.0, cd=None, name="", options=0):
    return {
        "name": name, "Q": Q, "Lmax": Lmax, "tau0": float(tau0),
        "collision": collision, "les_C": float(les_C), "rho0": float(rho0),
        "u0": tuple(float(v) for v in u0), "dims0": tuple(int(d) for d in dims0),
        "faces": [tuple(f) for f in faces], "levels": levels,
        "dis_keys": dis_keys, "dis_vals": np.ascontiguousarray(dis_vals, dtype=np.float64),
        "ib": ib, "cd": cd or {}, "options": int(options),
    }


def taus(case):
    """Per-level relaxation time tau_L = 2^L (tau_0 - 1/2) + 1/2 (LBM_Manager.cpp:814-818)."""
    return [(1 << L) * (case["tau0"] - 0.5) + 0.5 for L in range(case["Lmax"] + 1)]


# ---- workloads ---------------------------------------------------------------
def cavity(n=64, Q=19, tau=0.6, lid=0.05, collision=COLLIDE_LBGK, refine_box=None, name=None, wall=BC_ZOUHE):
    """Lid-driven cavity: n^3 wet-node cells, no-slip on five faces and a moving
    lid (physical u_x=lid) on z-max, all of face kind `wall` (Zou-He by
    default); f0 = feq(1, 0).  With refine_box = ((x0,y0,z0),(x1,y1,z1)) in
    coarse cells, that box is refined once."""
    faces = [(wall, 0.0, 0.0, 0.0)] * 5 + [(wall, reported(float(lid)), 0.0, 0.0)]
    dims0 = (n, n, n) if np.isscalar(n) else tuple(n)
    Lmax = 0 if refine_box is None else 1

    def refine(L, cx, cy, cz):
        (x0, y0, z0), (x1, y1, z1) = refine_box
        return ((cx >= x0) & (cx < x1) & (cy >= y0) & (cy < y1) & (cz >= z0) & (cz < z1))

    levels, dk, dv = build_nested(dims0, Lmax, refine, faces=faces, Q=Q)
    return _base_case(Q, Lmax, tau, collision, dims0, faces, levels, dk, dv,
                      name=name or f"cavity{dims0[0]}")


def reservoir_box(dims0=(16, 12, 12), Q=19, tau=0.6, u0=(0.05, 0.0, 0.0), collision=COLLIDE_LBGK):
    """The compiled reference layout: a frozen lat_bc shell holding feq(rho0,u0)
    around lat_f (LBM_Manager.cpp:482-505, Lat_Voxelize.cpp:242-288), NEE on all
    faces (LBM_Manager.cpp:52-80)."""
    faces = [(BC_RESERVOIR,) + tuple(u0)] * 6
    levels, dk, dv = build_nested(dims0, 0, None, faces=faces, Q=Q, reservoir=True)
    return _base_case(Q, 0, tau, collision, dims0, faces, levels, dk, dv, u0=u0,
                      name="reservoir_box")


def sphere_ib_dist(center, radius):
    C = np.asarray(center, dtype=np.float64)

    def dist(cx, cy, cz, ex, ey, ez, h):
        dx = cx - C[0]
        dy = cy - C[1]
        dz = cz - C[2]
        a = h * h * float(ex * ex + ey * ey + ez * ez)
        b = 2.0 * h * (ex * dx + ey * dy + ez * dz)
        c = dx * dx + dy * dy + dz * dz - radius * radius
        disc = b * b - 4.0 * a * c
        with np.errstate(invalid="ignore"):
            t = (-b - np.sqrt(disc)) / (2.0 * a)
        ok = (disc >= 0) & (c > 0) & (t >= 0.0) & (t <= 1.0)
        return np.where(ok, t, np.nan)

    return dist


def sphere_amr(coarse=(64, 32, 32), radius=8.0, center=(16.0, 16.0, 16.0), shells=(6.0, 3.0),
               Q=27, collision=COLLIDE_CUMULANT, U=0.04, Re=100.0, ib=IB_FH, les_C=0.0,
               periodic_yz=True, name=None, keep_solid=True, with_sphere=True, options=0, periodic_x=False,
               builder="host"):
    """Refined sphere in a channel: Zou-He inflow u_x=U at x=0, outflow at x-max,
    periodic y and z; level L+1 covers coarse cells whose centre lies within
    radius+shells[L] of the sphere centre; cells inside the sphere are solid on the
    finest level and the cut links carry analytic fractions.  Re = U*2R/nu0.
    builder="gpu" builds the level sets on the device (mt_levelset_build) with
    the same rules and arithmetic as the host builder."""
    Lmax = len(shells)
    nu0 = abs(U) * 2.0 * radius / Re
    tau0 = 3.0 * nu0 + 0.5
    C = np.asarray(center, dtype=np.float64)
    pk = BC_PERIODIC if periodic_yz else BC_ZOUHE
    ur = reported(U)
    faces = [(BC_ZOUHE, ur, 0.0, 0.0), (BC_OUTFLOW, 0.0, 0.0, 0.0)] + [(pk, 0.0, 0.0, 0.0)] * 4
    if periodic_x:
        faces[0] = faces[1] = (BC_PERIODIC, 0.0, 0.0, 0.0)

    def r_of(cx, cy, cz):
        return np.sqrt((cx - C[0]) ** 2 + (cy - C[1]) ** 2 + (cz - C[2]) ** 2)

    def refine(L, cx, cy, cz):
        return r_of(cx, cy, cz) <= radius + shells[L]

    def solid(cx, cy, cz):
        return r_of(cx, cy, cz) < radius

    if builder == "gpu":
        from . import amr
        levels, dk, dv = amr.build_levels(Q, Lmax, coarse, faces, shells, center=C, radius=radius,
                                          solid=with_sphere)
    else:
        levels, dk, dv = build_nested(coarse, Lmax, refine, solid=solid if with_sphere else None, faces=faces,
                                      Q=Q, ib_dist=sphere_ib_dist(C, radius) if with_sphere else None)
    if not keep_solid:
        top = levels[Lmax]
        keep = (top["kinds"] & KIND_SOLID) == 0
        top["keys"], top["kinds"] = top["keys"][keep], top["kinds"][keep]
    if not with_sphere:
        ib = IB_NONE
    d_lb = 2.0 * radius * (1 << Lmax)
    return _base_case(Q, Lmax, tau0, collision, coarse, faces, levels, dk, dv,
                      u0=(ur, 0.0, 0.0), ib=ib, les_C=les_C,
                      cd={"U": ur, "d": d_lb}, name=name or "sphere_amr", options=options)


def level_counts(case):
    out = []
    for lv in case["levels"]:
        k = lv["kinds"]
        out.append({
            "cells": int(k.size),
            "fluid": int(((k & KIND_FLUID) != 0).sum()),
            "f2c": int(((k & KIND_F2C) != 0).sum()),
            "c2f": int(((k & KIND_C2F) != 0).sum()),
            "ghost_c2f": int(((k & KIND_GHOST_C2F) != 0).sum()),
            "solid": int(((k & KIND_SOLID) != 0).sum()),
            "bcres": int(((k & KIND_BCRES) != 0).sum()),
            "dis": int(((k & KIND_DIS) != 0).sum()),
            "c2f_dis": int((((k & KIND_C2F) != 0) & ((k & KIND_DIS) != 0)).sum()),
        })
    return out


def updates_per_coarse_step(case):
    """Cell updates per coarse step, a level-L sub-step counting once: sum_L (lat_f+C2F)_L * 2^L."""
    tot = 0
    for L, c in enumerate(level_counts(case)):
        tot += (c["fluid"] + c["c2f"]) << L
    return tot


def reference_sphere(coarse=(24, 16, 16), radius=3.0, center=(8.0, 8.0, 8.0), shells=(2.5, 1.25), Q=19,
                     collision=COLLIDE_LBGK, U=0.05, Re=100.0, ib=IB_FH, name=None):
    """The configuration the compiled reference runs (LBM_Manager.cpp:38-80,
    header/Constants.h:164-235): LBGK, a frozen lat_bc reservoir holding
    feq(rho0, U) on all six faces with NEE on its f, FH link interpolation plus
    multi-direct forcing on the finest level, every cell initialised to feq(1, U)."""
    Lmax = len(shells)
    nu0 = U * 2.0 * radius / Re
    tau0 = 3.0 * nu0 + 0.5
    C = np.asarray(center, dtype=np.float64)
    faces = [(BC_RESERVOIR, U, 0.0, 0.0)] * 6

    def r_of(cx, cy, cz):
        return np.sqrt((cx - C[0]) ** 2 + (cy - C[1]) ** 2 + (cz - C[2]) ** 2)

    levels, dk, dv = build_nested(coarse, Lmax, lambda L, cx, cy, cz: r_of(cx, cy, cz) <= radius + shells[L],
                                  solid=lambda cx, cy, cz: r_of(cx, cy, cz) < radius, faces=faces, Q=Q,
                                  ib_dist=sphere_ib_dist(C, radius), reservoir=True)
    return _base_case(Q, Lmax, tau0, collision, coarse, faces, levels, dk, dv, u0=(U, 0.0, 0.0), ib=ib,
                      cd={"U": U, "d": 2.0 * radius * (1 << Lmax)}, name=name or "reference_sphere")


# ---- bodies given as triangle meshes -----------------------------------------------
def icosphere(center, radius, subdiv=2):
    """A closed triangle mesh of the sphere: the icosahedron, each face split into
    4 subdiv times, vertices pushed onto the sphere; [n][3][3], outward winding."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(i, j):
            k = (min(i, j), max(i, j))
            if k not in cache:
                m = V[i] + V[j]
                V.append(m / np.linalg.norm(m))
                cache[k] = len(V) - 1
            return cache[k]

        G = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            G += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = G
    P = np.array(V) * float(radius) + np.asarray(center, dtype=np.float64)
    return P[np.array(F)]


def read_stl(path, scale=1.0, offset=(0.0, 0.0, 0.0)):
    """Triangles [n][3][3] of an ASCII STL file (the format Shape::geofile_stl reads,
    Shape.cpp:341-548): 'vertex x y z' lines in threes, then scale and offset."""
    verts = []
    with open(path) as fh:
        for line in fh:
            w = line.split()
            if w and w[0] == "vertex":
                verts.append([float(w[1]), float(w[2]), float(w[3])])
    T = np.array(verts, dtype=np.float64).reshape(-1, 3, 3)
    return T * float(scale) + np.asarray(offset, dtype=np.float64)


def _tri_dist2(T, px, py, pz):
    """Squared distance from the points to the triangle T (closest point by regions,
    Ericson, Real-Time Collision Detection 5.1.5), vectorised over the points."""
    A, B, Cv = T
    ab, ac = B - A, Cv - A
    apx, apy, apz = px - A[0], py - A[1], pz - A[2]
    d1 = ab[0] * apx + ab[1] * apy + ab[2] * apz
    d2 = ac[0] * apx + ac[1] * apy + ac[2] * apz
    bpx, bpy, bpz = px - B[0], py - B[1], pz - B[2]
    d3 = ab[0] * bpx + ab[1] * bpy + ab[2] * bpz
    d4 = ac[0] * bpx + ac[1] * bpy + ac[2] * bpz
    cpx, cpy, cpz = px - Cv[0], py - Cv[1], pz - Cv[2]
    d5 = ab[0] * cpx + ab[1] * cpy + ab[2] * cpz
    d6 = ac[0] * cpx + ac[1] * cpy + ac[2] * cpz
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    with np.errstate(divide="ignore", invalid="ignore"):
        v_ab = d1 / (d1 - d3)
        w_ac = d2 / (d2 - d6)
        w_bc = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        den = 1.0 / (va + vb + vc)
    v_in, w_in = vb * den, vc * den

    def d2of(qx, qy, qz):
        return (px - qx) * (px - qx) + (py - qy) * (py - qy) + (pz - qz) * (pz - qz)

    out = d2of(A[0] + ab[0] * v_in + ac[0] * w_in, A[1] + ab[1] * v_in + ac[1] * w_in, A[2] + ab[2] * v_in + ac[2] * w_in)
    r = (va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0)
    out = np.where(r, d2of(B[0] + w_bc * (Cv[0] - B[0]), B[1] + w_bc * (Cv[1] - B[1]), B[2] + w_bc * (Cv[2] - B[2])), out)
    r = (vb <= 0) & (d2 >= 0) & (d6 <= 0)
    out = np.where(r, d2of(A[0] + w_ac * ac[0], A[1] + w_ac * ac[1], A[2] + w_ac * ac[2]), out)
    r = (d6 >= 0) & (d5 <= d6)
    out = np.where(r, d2of(Cv[0], Cv[1], Cv[2]), out)
    r = (vc <= 0) & (d1 >= 0) & (d3 <= 0)
    out = np.where(r, d2of(A[0] + v_ab * ab[0], A[1] + v_ab * ab[1], A[2] + v_ab * ab[2]), out)
    r = (d3 >= 0) & (d4 <= d3)
    out = np.where(r, d2of(B[0], B[1], B[2]), out)
    r = (d1 <= 0) & (d2 <= 0)
    out = np.where(r, d2of(A[0], A[1], A[2]), out)
    return out


class MeshBody:
    """Host (numpy) geometry of a closed triangle mesh, the rules the device
    builder applies: inside by the generalised winding number (Van Oosterom-
    Strackee solid angles), distance to the surface, and the first crossing along
    a link (Moller-Trumbore, Shape.cpp:809-843)."""

    def __init__(self, triangles):
        self.T = np.asarray(triangles, dtype=np.float64).reshape(-1, 3, 3)
        self.lo = self.T.reshape(-1, 3).min(axis=0)
        self.hi = self.T.reshape(-1, 3).max(axis=0)

    def _near(self, cx, cy, cz, margin):
        """Broadcast point arrays and the mask of points within `margin` of the
        mesh's bounding box (the only ones a triangle test can change)."""
        px, py, pz = np.broadcast_arrays(cx, cy, cz)
        m = np.ones(px.shape, bool)
        for a, p in enumerate((px, py, pz)):
            m &= (p >= self.lo[a] - margin) & (p <= self.hi[a] + margin)
        return px, py, pz, m

    def inside(self, cx, cy, cz):
        px, py, pz, m = self._near(cx, cy, cz, 0.0)
        out = np.zeros(px.shape, bool)
        if m.any():
            out[m] = self._inside(px[m], py[m], pz[m])
        return out

    def _inside(self, px, py, pz):
        w = np.zeros(px.shape)
        for T in self.T:
            a = [T[0][k] - p for k, p in enumerate((px, py, pz))]
            b = [T[1][k] - p for k, p in enumerate((px, py, pz))]
            c = [T[2][k] - p for k, p in enumerate((px, py, pz))]
            la = np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
            lb = np.sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2])
            lc = np.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2])
            det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0])
            den = la * lb * lc + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * lc + \
                (a[0] * c[0] + a[1] * c[1] + a[2] * c[2]) * lb + (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) * la
            w = w + 2.0 * np.arctan2(det, den)
        return np.abs(w) > 2.0 * np.pi

    def dist2(self, cx, cy, cz, margin=np.inf):
        """Squared distance to the surface; points farther than `margin` from the
        bounding box read inf."""
        px, py, pz, m = self._near(cx, cy, cz, margin)
        out = np.full(px.shape, np.inf)
        if m.any():
            sub = np.full(int(m.sum()), np.inf)
            qx, qy, qz = px[m], py[m], pz[m]
            for T in self.T:
                sub = np.minimum(sub, _tri_dist2(T, qx, qy, qz))
            out[m] = sub
        return out

    def refine(self, shell):
        def f(L, cx, cy, cz):
            return self.inside(cx, cy, cz) | (self.dist2(cx, cy, cz, shell[L] * 1.001) <= shell[L] * shell[L])
        return f

    def link(self, cx, cy, cz, ex, ey, ez, h):
        px, py, pz, m = self._near(cx, cy, cz, 2.0 * h)
        best = np.full(px.shape, 2.0)
        if m.any():
            best[m] = self._link(px[m], py[m], pz[m], ex, ey, ez, h)
        return best

    def _link(self, px, py, pz, ex, ey, ez, h):
        d = (h * ex, h * ey, h * ez)
        best = np.full(px.shape, 2.0)
        for T in self.T:
            e1 = T[1] - T[0]
            e2 = T[2] - T[0]
            sx, sy, sz = px - T[0][0], py - T[0][1], pz - T[0][2]
            hx = d[1] * e2[2] - d[2] * e2[1]
            hy = d[2] * e2[0] - d[0] * e2[2]
            hz = d[0] * e2[1] - d[1] * e2[0]
            a = e1[0] * hx + e1[1] * hy + e1[2] * hz
            if abs(a) < 1e-300:
                continue
            f = 1.0 / a
            u = f * (sx * hx + sy * hy + sz * hz)
            qx = sy * e1[2] - sz * e1[1]
            qy = sz * e1[0] - sx * e1[2]
            qz = sx * e1[1] - sy * e1[0]
            v = f * (d[0] * qx + d[1] * qy + d[2] * qz)
            t = f * (e2[0] * qx + e2[1] * qy + e2[2] * qz)
            ok = (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t >= 0)
            best = np.where(ok & (t < best), t, best)
        return best

    def ib_dist(self):
        def dist(cx, cy, cz, ex, ey, ez, h):
            out = self.link(cx, cy, cz, float(ex), float(ey), float(ez), h)
            inside = self.inside(cx, cy, cz)
            return np.where((out <= 1.0) & ~inside, out, np.nan)
        return dist


def mesh_amr(triangles, coarse=(64, 32, 32), shells=(6.0, 3.0), Q=27, collision=COLLIDE_CUMULANT, U=0.04, Re=100.0,
             d_ref=None, ib=IB_FH, les_C=0.0, options=2, name=None, builder="gpu"):
    """A body given as a closed triangle mesh (an STL surface) in the channel of
    sphere_amr: Zou-He inflow at x=0, outflow at x-max, periodic y/z.  Level L+1
    covers the cells inside the body or within shells[L] of its surface; the finest
    level's inside cells are solid and the cut links carry Moller-Trumbore
    fractions.  Re = U d_ref / nu0 with d_ref the body's extent along y by default.
    builder="gpu" builds the level sets on the device (mt_levelset_build), "host"
    with numpy (the same rules; slow, for checking)."""
    T = np.asarray(triangles, dtype=np.float64).reshape(-1, 3, 3)
    Lmax = len(shells)
    if d_ref is None:
        d_ref = float(T[..., 1].max() - T[..., 1].min())
    nu0 = abs(U) * d_ref / Re
    tau0 = 3.0 * nu0 + 0.5
    ur = reported(U)
    faces = [(BC_ZOUHE, ur, 0.0, 0.0), (BC_OUTFLOW, 0.0, 0.0, 0.0)] + [(BC_PERIODIC, 0.0, 0.0, 0.0)] * 4
    if builder == "gpu":
        from . import amr
        levels, dk, dv = amr.build_levels(Q, Lmax, coarse, faces, shells, triangles=T, solid=True)
    else:
        body = MeshBody(T)
        levels, dk, dv = build_nested(coarse, Lmax, body.refine(shells), solid=body.inside, faces=faces, Q=Q,
                                      ib_dist=body.ib_dist())
    return _base_case(Q, Lmax, tau0, collision, coarse, faces, levels, dk, dv, u0=(ur, 0.0, 0.0), ib=ib,
                      les_C=les_C, cd={"U": ur, "d": d_ref * (1 << Lmax)}, name=name or "mesh_amr", options=options)


# ---- packed batching ---------------------------------------------------------
def pack(members, gap=4, origins=None, dims0=None):
    """Pack independent cases into one case of members (mt_case_set_members).

    Member i keeps its own cells, faces, relaxation time and initial state in
    the box [origin0_i, origin0_i + dims0_i) of level-0 cells; the boxes are laid
    side by side along x, 4-aligned, `gap` cells apart.  Keys are re-encoded into
    the packed case's coordinates and Lmax (the largest member Lmax); the solid
    boundary (dis) applies on the packed finest level, so members with dis must
    reach it.  All members share Q, collision and options."""
    if not members:
        raise ValueError("no members")
    Q, coll, opts = members[0]["Q"], members[0]["collision"], members[0].get("options", 0)
    ibs = {m["ib"] for m in members if len(m["dis_keys"])}
    if len(ibs) > 1:
        raise ValueError("members with a solid boundary must share its method")
    ib = ibs.pop() if ibs else IB_NONE
    Lmax = max(m["Lmax"] for m in members)
    for m in members:
        if m["Q"] != Q or m["collision"] != coll or m.get("options", 0) != opts:
            raise ValueError("members must share Q, collision and options")
        if len(m["dis_keys"]) and m["Lmax"] != Lmax:
            raise ValueError("a member with a solid boundary must reach the packed finest level")
    if origins is None:
        origins, ox = [], 0
        for m in members:
            origins.append((ox, 0, 0))
            ox += (m["dims0"][0] + 3) // 4 * 4 + gap
        dims0 = (ox - gap, max(m["dims0"][1] for m in members), max(m["dims0"][2] for m in members))
    origins = [tuple(int(v) for v in o) for o in origins]

    def keys_of(m, o, what, L):
        # re-encoded keys are cached in the member, per placement
        cache = m.setdefault("_packed", {})
        k = (what, L, o, Lmax)
        if k not in cache:
            src = m["dis_keys"] if what == "dis" else m["levels"][L]["keys"]
            cache[k] = member_keys_at(src, L, m["Lmax"], o, Lmax)
        return cache[k]

    # (the device sorts the keys itself: levels are uploaded in member order)
    levels = []
    for L in range(Lmax + 1):
        ks, kd = [], []
        for m, o in zip(members, origins):
            if L > m["Lmax"]:
                continue
            ks.append(keys_of(m, o, "cells", L))
            kd.append(m["levels"][L]["kinds"])
        keys = np.concatenate(ks) if ks else np.zeros(0, np.uint64)
        kinds = np.concatenate(kd) if kd else np.zeros(0, np.uint16)
        levels.append({"keys": keys, "kinds": kinds})
    dks, dvs = [], []
    for m, o in zip(members, origins):
        if len(m["dis_keys"]):
            dks.append(keys_of(m, o, "dis", Lmax))
            dvs.append(m["dis_vals"])
    dis_keys = np.concatenate(dks) if dks else np.zeros(0, np.uint64)
    dis_vals = np.concatenate(dvs) if dvs else np.zeros((0, Q - 1))
    mdesc = [{"origin0": o, "dims0": tuple(m["dims0"]), "faces": m["faces"], "tau0": m["tau0"],
              "les_C": m["les_C"], "rho0": m["rho0"], "u0": m["u0"], "Lmax": m["Lmax"],
              "cd": m.get("cd", {})} for m, o in zip(members, origins)]
    case = _base_case(Q, Lmax, members[0]["tau0"], coll, dims0, [(BC_NONE, 0.0, 0.0, 0.0)] * 6, levels,
                      dis_keys, dis_vals, ib=ib, name=f"packed[{len(members)}]", options=opts)
    case["members"] = mdesc
    return case


def member_keys_at(keys, level, member_Lmax, origin0, packed_Lmax):
    """A member's level-`level` keys in the packed case's coordinates."""
    x, y, z = decode_key(np.asarray(keys, dtype=np.uint64), level, member_Lmax)
    return encode_key(x + (origin0[0] << level), y + (origin0[1] << level), z + (origin0[2] << level),
                      level, packed_Lmax)


def member_keys(packed, i, level, keys):
    """Translate member i's own keys on `level` into the packed case's keys."""
    m = packed["members"][i]
    return member_keys_at(keys, level, m["Lmax"], m["origin0"], packed["Lmax"])


def cavity_batch(n_cases=256, seed=0, Q=19, edges=(16, 24, 32, 48, 64), options=2, wall=BC_NEE):
    """The packed-batching workload: n_cases lid-driven cavities, edge L drawn
    from `edges`, lid speed U ~ U[0.02, 0.08], tau ~ U[0.52, 0.9] (drawn per case
    in the order L, U, tau from numpy.random.default_rng(seed)); every odd-index
    case is two-level, refined in an axis-aligned box whose volume fraction is
    drawn from U[0.1, 0.3] and whose lower corner is drawn uniformly among the
    positions 2 or more cells from every face (box draws follow the case's
    L, U, tau).  The walls are wet-node NEE (BC_NEE): the Zou-He closure with
    extrapolated edges diverges below tau ~ 0.6 in these cavities, NEE walls
    stay stable down to tau = 0.52 at lid 0.08.  Returns the list of member
    cases (pack() joins them)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_cases):
        L = int(rng.choice(np.asarray(edges)))
        U = float(rng.uniform(0.02, 0.08))
        tau = float(rng.uniform(0.52, 0.9))
        box = None
        if i % 2 == 1:
            frac = float(rng.uniform(0.10, 0.30))
            e = max(4, int(round(L * frac ** (1.0 / 3.0))))
            e = min(e, L - 4)
            lo = rng.integers(2, L - e - 2 + 1, size=3)
            box = (tuple(int(v) for v in lo), tuple(int(v) + e for v in lo))
        c = cavity(L, Q=Q, tau=tau, lid=U, refine_box=box, name=f"cavity{L}_{i}", wall=wall)
        c["options"] = int(options)
        out.append(c)
    return out


SERVE_DOMAINS = ((48, 24, 24), (64, 32, 32), (96, 48, 48))


def sphere_request(rng, U=0.04, shell=3.0, options=2):
    """One served request: a two-level D3Q27 cumulant sphere in a channel
    (Zou-He inflow, outflow, periodic y/z), coarse domain drawn from
    SERVE_DOMAINS, sphere radius ~ U[4, 12] coarse cells and Re ~ U[20, 200]
    (draws in that order).  The sphere sits at (nx/3, ny/2, nz/2); its radius is
    capped so that the refined shell keeps 2 coarse cells from every face."""
    dims = SERVE_DOMAINS[int(rng.integers(0, len(SERVE_DOMAINS)))]
    r = float(rng.uniform(4.0, 12.0))
    Re = float(rng.uniform(20.0, 200.0))
    cx, cy, cz = dims[0] / 3.0, dims[1] / 2.0, dims[2] / 2.0
    r = min(r, cy - shell - 2.0, cx - shell - 2.0)
    case = sphere_amr(dims, radius=r, center=(cx, cy, cz), shells=(shell,), U=U, Re=Re, options=options,
                      name=f"sphere{dims[0]}_r{r:.2f}_Re{Re:.0f}")
    case["request"] = {"dims": dims, "radius": r, "Re": Re}
    return case


# ---- Lagrangian-marker sphere ----------------------------------------------------
def roma3(r):
    """Three-point kernel of Roma, Peskin & Berger (J. Comput. Phys. 153, 1999):
    sums to 1 over unit-spaced nodes and has zero first moment."""
    r = np.abs(np.asarray(r, dtype=np.float64))
    out = np.zeros_like(r)
    a = r <= 0.5
    out[a] = (1.0 + np.sqrt(1.0 - 3.0 * r[a] ** 2)) / 3.0
    b = (r > 0.5) & (r < 1.5)
    out[b] = (5.0 - 3.0 * r[b] - np.sqrt(np.maximum(0.0, 1.0 - 3.0 * (1.0 - r[b]) ** 2))) / 6.0
    return out


def fibonacci_sphere(n):
    """n points on the unit sphere: z_i = 1 - 2(i + 1/2)/n, azimuth i times the
    golden angle pi (3 - sqrt 5)."""
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - 2.0 * (i + 0.5) / n
    rr = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * np.pi * (3.0 - np.sqrt(5.0))
    return np.stack([np.cos(phi) * rr, np.sin(phi) * rr, z], axis=1)


def marker_stencils(xyz, Lmax):
    """Per marker (finest-level cell units, cell i centred at i + 1/2) the 27
    cells of its 3-point kernel support and their weights."""
    c = np.asarray(xyz, dtype=np.float64) - 0.5          # node-index coordinates
    base = np.floor(c).astype(np.int64) - 1                # candidates base .. base + 3
    keys, ws = [], []
    offs = np.array([(i, j, k) for k in range(4) for j in range(4) for i in range(4)])
    for m in range(c.shape[0]):
        idx = base[m][None, :] + offs
        w = roma3(idx[:, 0] - c[m, 0]) * roma3(idx[:, 1] - c[m, 1]) * roma3(idx[:, 2] - c[m, 2])
        keep = np.nonzero(w > 0)[0]
        if keep.size > 27:
            raise ValueError("marker support exceeds 27 cells")
        kk = np.zeros(27, np.uint64)
        ww = np.zeros(27)
        kk[:keep.size] = encode_key(idx[keep, 0], idx[keep, 1], idx[keep, 2], Lmax, Lmax)
        kk[keep.size:] = kk[0]
        ww[:keep.size] = w[keep]
        keys.append(kk)
        ws.append(ww)
    return np.array(keys, dtype=np.uint64), np.array(ws)


def sphere_markers(coarse=(64, 32, 32), radius=8.0, center=(16.0, 16.0, 16.0), shells=(6.0, 3.0),
                   n_markers=6000, Q=27, collision=COLLIDE_CUMULANT, U=0.04, Re=100.0, les_C=0.1, options=2,
                   builder="host"):
    """The sphere of sphere_amr represented by n_markers Lagrangian markers on a
    Fibonacci lattice (explicit direct forcing on the finest level, see the
    device's marker kernels) instead of cut links; the grid is the same refined
    channel with no solid cells.  Smagorinsky constant les_C."""
    case = sphere_amr(coarse, radius=radius, center=center, shells=shells, Q=Q, collision=collision, U=U,
                      Re=Re, les_C=les_C, with_sphere=False, options=options, name="sphere_markers", builder=builder)
    Lmax = case["Lmax"]
    s = float(1 << Lmax)
    R = radius * s
    xyz = np.asarray(center, dtype=np.float64) * s + R * fibonacci_sphere(n_markers)
    keys, w = marker_stencils(xyz, Lmax)
    top = case["levels"][Lmax]
    pure = top["keys"][((top["kinds"] & KIND_FLUID) != 0)
                       & ((top["kinds"] & (KIND_C2F | KIND_GHOST_C2F | KIND_SOLID)) == 0)]
    if not np.isin(keys[w != 0], pure).all():
        raise ValueError("marker support reaches beyond the finest level's fluid cells: widen shells[-1]")
    case["markers"] = {"xyz": xyz, "keys": keys, "w": w, "dv": np.full(n_markers, 4.0 * np.pi * R * R / n_markers),
                       "ub": np.zeros(3)}
    case["cd"] = {"U": reported(U), "d": 2.0 * R}
    return case
