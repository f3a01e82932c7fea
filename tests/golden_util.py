"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_names():
    """fixtures holding states (the digest-only code_* fixtures excepted)"""
    return sorted(n for n in (os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))
                  if not n.startswith("code_"))


def code_names():
    """the reference's shipped parfiles run by the reference: digests only"""
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN_DIR, "code_*.npz")))


class Golden:
    def __init__(self, name):
        self.name = name
        z = np.load(os.path.join(GOLDEN_DIR, name + ".npz"), allow_pickle=False)
        self.d = {k: z[k] for k in z.files}
        self.params = json.loads(str(self.d["params"]))
        self.N_p = self.d["N_p"].astype(np.int32)
        self.r_faces = self.d["r_faces"]
        self.z_faces = self.d["z_faces"]
        self.NR, self.NZ = len(self.N_p), len(self.z_faces) - 1
        self.NQ = self.params["NUM_C"] + self.params["NUM_N"]
        self.Nz = int(self.N_p.sum()) * self.NZ
        self.tiph0 = self.d.get("tiph0")
        self.dphi0 = self.d.get("dphi0")
        self.prim0 = self.d["prim0"].reshape(self.NQ, self.Nz) if "prim0" in self.d else None

    def has(self, key):
        return key in self.d

    def get(self, key):
        v = self.d[key]
        if v.dtype == np.float64 and v.size == self.NQ * self.Nz and v.size != self.Nz:
            return v.reshape(self.NQ, self.Nz)
        return v

    def steps(self):
        out = []
        for k in self.d:
            if k.startswith("run/step") and k.endswith("/prim"):
                out.append(int(k[len("run/step"):-len("/prim")]))
        return sorted(out)
