"""Samples of the flow fields along long runs, for parity at the stated step counts.

Two workloads at their full size:
  cavity  the 64^3 lid-driven cavity (D3Q19 LBGK, tau 0.6, Zou-He walls, lid 0.05), 1000 steps
  sphere  the 3-level sphere (64x32x32 coarse, r 8, D3Q27 cumulant, FP64, both-sub-step
          explosions), 4000 coarse steps
At each checkpoint the density and velocity of a fixed sample of fluid cells per level
(up to 8192 cells, numpy default_rng(0)) are recorded, with the full-field sums
sum rho and sum |u|^2 per level.

    python tests/golden/parity_curves.py oracle cavity OUT.npz   # dense CPU oracle
    python tests/golden/parity_curves.py device cavity OUT.npz   # the HIP path
    python tests/golden/parity_curves.py compare A.npz B.npz     # rel-L2 per checkpoint

The oracle run of the sphere takes about 40 minutes on 8 cores; every fourth
sampled cell at 100, 1000 and 4000 steps, with a tolerance of 1e-12 on the relative
L2 of rho and u, is kept as tests/golden/sphere_4000.npz (`python
tests/golden/parity_curves.py fixture SPHERE_ORACLE.npz`), which
tests/test_gpu_workloads.py compares the device with.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from mortontrie_amr_amd import cases as C  # noqa: E402

CHECKPOINTS = {
    "cavity": [1, 2, 5, 10, 20, 50, 100, 200, 300, 500, 700, 1000],
    "sphere": [1, 2, 5, 10, 20, 50, 100, 200, 500, 1000, 1500, 2000, 2500, 3000, 3500, 4000],
}
SAMPLE = 8192


def workload(name):
    if name == "cavity":
        return C.cavity(64)
    if name == "sphere":
        return C.sphere_amr(options=2)
    raise ValueError(name)


def sample_keys(case):
    out = []
    rng = np.random.default_rng(0)
    for lv in case["levels"]:
        k = lv["keys"][(lv["kinds"] & (C.KIND_FLUID | C.KIND_F2C)) != 0]
        pick = np.sort(rng.choice(k.size, min(k.size, SAMPLE), replace=False))
        out.append(k[pick])
    return out


def record(sim, case, keys, rec, step):
    for L, k in enumerate(keys):
        allk = case["levels"][L]["keys"]
        fl = (case["levels"][L]["kinds"] & (C.KIND_FLUID | C.KIND_F2C)) != 0
        rho_all = sim.rho(L, allk[fl])
        u_all = sim.u(L, allk[fl])
        rec[f"rho_L{L}_s{step}"] = sim.rho(L, k)
        rec[f"u_L{L}_s{step}"] = sim.u(L, k)
        rec[f"sumrho_L{L}_s{step}"] = np.array(rho_all.sum())
        rec[f"sumu2_L{L}_s{step}"] = np.array((u_all ** 2).sum())


def run(kind, name, out):
    case = workload(name)
    keys = sample_keys(case)
    if kind == "oracle":
        from oracle import pyoracle as O
        sim = O.DenseOracleCase(case)
    else:
        from mortontrie_amr_amd import amr
        sim = amr.AmrCase(case)
    rec = {"checkpoints": np.array(CHECKPOINTS[name])}
    for L, k in enumerate(keys):
        rec[f"keys_L{L}"] = k
    done = 0
    t0 = time.time()
    for s in CHECKPOINTS[name]:
        sim.step(s - done)
        done = s
        record(sim, case, keys, rec, s)
        print(f"{kind} {name} step {s} ({time.time() - t0:.1f} s)", flush=True)
    np.savez_compressed(out, **rec)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def compare(a_path, b_path):
    a, b = np.load(a_path), np.load(b_path)
    levels = sorted({int(k.split("_L")[1]) for k in a.files if k.startswith("keys_L")})
    print("step  " + "  ".join(f"rho_L{L}      u_L{L}        " for L in levels))
    for s in a["checkpoints"]:
        row = []
        for L in levels:
            row.append(f"{rel(a[f'rho_L{L}_s{s}'], b[f'rho_L{L}_s{s}']):.2e}  {rel(a[f'u_L{L}_s{s}'], b[f'u_L{L}_s{s}']):.2e}")
        print(f"{int(s):5d} " + "  ".join(row))


def fixture(oracle_path):
    src = np.load(oracle_path)
    out = {"checkpoints": np.array([100, 1000, 4000])}
    for L in range(3):
        out[f"keys_L{L}"] = src[f"keys_L{L}"][::4]
    for s in (100, 1000, 4000):
        for L in range(3):
            out[f"rho_L{L}_s{s}"] = src[f"rho_L{L}_s{s}"][::4]
            out[f"u_L{L}_s{s}"] = src[f"u_L{L}_s{s}"][:, ::4]
        out[f"tol_s{s}"] = np.array([1e-12, 1e-12])
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere_4000.npz"), **out)


if __name__ == "__main__":
    if sys.argv[1] == "compare":
        compare(sys.argv[2], sys.argv[3])
    elif sys.argv[1] == "fixture":
        fixture(sys.argv[2])
    else:
        run(sys.argv[1], sys.argv[2], sys.argv[3])
