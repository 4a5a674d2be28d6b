"""Loader of the golden fixtures in tests/golden (made by make_golden.py
from the compiled reference)."""
import glob
import json
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Case:
    def __init__(self, path, index, params, nd, outputs):
        self.path = path
        self.name = f"{os.path.basename(path)[:-4]}[{params['numTs']}]"
        self.index = index
        self.params = params
        self.nd = nd
        self.out = outputs  # field name -> array (ux, uy, uz, density, Fx, Fy, Fz)
        self.precision = "f32" if path.endswith("_f32.npz") else "f64"

    @property
    def q(self):
        return int(self.params["latticeType"][3:])

    @property
    def dims(self):
        return self.params["nx"], self.params["ny"], self.params["nz"]

    def kwargs(self):
        p = self.params
        return dict(omega=p["omega"], u_lbm=p["uLBM"], rho_lbm=p["rhoLBM"], cs=p["cs"],
                    dynamics=p["dynamics"])


def load_cases(pattern="*.npz"):
    cases = []
    for path in sorted(glob.glob(os.path.join(GOLDEN_DIR, pattern))):
        z = np.load(path)
        meta = json.loads(bytes(z["meta"]).decode())
        for i, p in enumerate(meta["params"]):
            out = {k[: -len(f"_{i}")]: z[k] for k in z.files if k.endswith(f"_{i}")}
            cases.append(Case(path, i, p, z["ndType"], out))
    return cases
