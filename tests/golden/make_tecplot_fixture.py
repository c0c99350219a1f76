"""Writes tests/golden/tecplot_ref.json: the digest of the reference's own Tecplot
flow file (IO_TECPLOTH::write, run through oracle/_ref) for a small case of the
reference's layout after two coarse steps from a perturbed state, with the
parameters a device run needs to write the same file (tests/test_gpu_output.py).

    python tests/golden/make_tecplot_fixture.py     (needs the reference's sources)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from mortontrie_amr_amd import cases as C  # noqa: E402
from oracle import refcase as R  # noqa: E402
import tecplot_form  # noqa: E402

GEOM = dict(coarse=(24, 20, 20), radius=2.0, center=(10, 10, 10), shells=(4.5, 1.5))
SEED, AMP, STEPS = 23, 0.02, 2


def main():
    par, _ = R.run(C.reference_sphere(**GEOM), None, [])
    U = float(par["bc_u"][0][0])
    case = C.reference_sphere(**GEOM, U=U)
    st = R.perturbed_state(case, SEED, AMP)
    par, _ = R.run(case, st, [(R.STEP, STEPS, 0), (R.TECPLOT, 0, 0)])
    text, cf_u = par["tecplot"][0]
    out = {"geometry": {k: list(v) if isinstance(v, tuple) else v for k, v in GEOM.items()}, "U": U,
           "tau": [float(t) for t in par["tau"]], "seed": SEED, "amp": AMP, "steps": STEPS, "dx0": R.C_DX,
           "convert_u": cf_u, **tecplot_form.digest(text)}
    with open(os.path.join(HERE, "tecplot_ref.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
