"""Field output from the device: snapshots, the Tecplot flow file and the AMR node dump.

The Tecplot file is checked against the reference's own writer (IO_TECPLOTH::write,
tecplot.cpp:16-177, run through oracle/_ref by tests/golden/make_tecplot_fixture.py):
the same case, per-level relaxation times and perturbed state, two coarse steps on
the device, and the file's
canonical form (each element's 8 corner coordinates and U, V, W, exactly as
printed, element lines sorted) has the same SHA-256 as the reference's.  Bit-equal
fields and the same number formatting are both needed for that.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

from mortontrie_amr_amd import amr
from mortontrie_amr_amd import cases as C
from oracle import refcase as R

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import tecplot_form  # noqa: E402

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(HERE, "golden", "tecplot_ref.json")))


def fixture_case():
    g = FIX["geometry"]
    case = C.reference_sphere(coarse=tuple(g["coarse"]), radius=g["radius"], center=tuple(g["center"]),
                              shells=tuple(g["shells"]), U=FIX["U"])
    case["tau0"] = FIX["tau"][0]
    return case


def device_from(case, state):
    d = amr.AmrCase(case)
    d.set_tau(FIX["tau"])
    for L, s in enumerate(state):
        keys = case["levels"][L]["keys"]
        d.set(L, amr.FIELD_F, s["f"])
        d.set(L, amr.FIELD_FCOL, s["fcol"])
        ok = ~np.isnan(s["rho"])
        d.set(L, amr.FIELD_RHO, s["rho"][ok], keys[ok])
        d.set(L, amr.FIELD_U, s["u"][:, ok], keys[ok])
    return d


def test_tecplot_equals_reference_writer(tmp_path):
    case = fixture_case()
    d = device_from(case, R.perturbed_state(case, FIX["seed"], FIX["amp"]))
    d.step(FIX["steps"])
    path = tmp_path / "flow.dat"
    d.write_tecplot(path, dx0=FIX["dx0"], convert_u=FIX["convert_u"])
    got = tecplot_form.digest(path.read_text())
    d.close()
    print(got, flush=True)
    assert (got["nodes"], got["elements"]) == (FIX["nodes"], FIX["elements"])
    assert got["sha256"] == FIX["sha256"]


def test_snapshot_and_node_dump(tmp_path):
    case = C.reference_sphere(coarse=(24, 20, 20), radius=2.0, center=(10, 10, 10), shells=(4.5, 1.5))
    d = amr.AmrCase(case)
    d.step(3)
    for L in range(case["Lmax"] + 1):
        keys, kinds = case["levels"][L]["keys"], case["levels"][L]["kinds"].astype(np.int64)
        for mask in (0, C.KIND_FLUID, C.KIND_SOLID, 0xFFFF):
            want = (kinds & (C.KIND_FLUID | C.KIND_F2C)) != 0 if mask == 0 else (kinds & mask) != 0
            k_want = np.sort(keys[want])
            k, rho, u = d.snapshot(L, mask)
            assert np.array_equal(k, k_want), (L, mask)
            if len(k) and mask != C.KIND_SOLID:
                assert np.array_equal(rho, d.rho(L, k), equal_nan=True)
                assert np.array_equal(u, d.u(L, k), equal_nan=True)
    path = tmp_path / "nodes.txt"
    d.write_nodes(path)
    lines = path.read_text().splitlines()
    fine = np.sort(case["levels"][case["Lmax"]]["keys"])
    assert len(lines) == len(fine)
    assert all(len(s) == 64 for s in lines)
    assert np.array_equal(np.array([int(s, 2) for s in lines], np.uint64), fine)
    for bad in ([0.6, 0.7], [0.6, 0.5, 0.7], [0.6, 0.7, 0.8, 0.9]):
        with pytest.raises(amr.MtError):
            d.set_tau(bad)
    with pytest.raises(amr.MtError):
        d.set_tau([0.6, 0.7, 0.9], member=1)
    d.close()
