"""Loads the golden vectors written by tests/golden/make_golden.py."""
import os

import numpy as np

from lsm_amd.inputs import Grid

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Fixture:
    def __init__(self, name):
        self.name = name
        self.d = dict(np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False))
        d = self.d
        self.grid = Grid(d["in_x"], d["in_y"], d["in_z"],
                         csi=np.stack([d["in_csi0"], d["in_csi1"], d["in_csi2"]]),
                         eta=np.stack([d["in_eta0"], d["in_eta1"], d["in_eta2"]]),
                         zet=np.stack([d["in_zet0"], d["in_zet1"], d["in_zet2"]]),
                         aj=d["in_aj"])

    def inp(self, k):
        return self.d[f"in_{k}"]

    def par(self, k):
        v = self.d[f"p_{k}"]
        return v.item() if v.ndim == 0 else tuple(int(t) for t in v)

    def kw(self, key, k, default=None):
        v = self.d.get(f"kw_{key}_{k}")
        if v is None:
            return default
        return v.item() if v.ndim == 0 else tuple(int(t) for t in v)

    def out(self, key, n=0):
        return self.d[f"out_{key}_{n}"]

    def has(self, key):
        return f"out_{key}_0" in self.d

    @property
    def velocity(self):
        return self.inp("u"), self.inp("v"), self.inp("w")


def load(name):
    return Fixture(name)
