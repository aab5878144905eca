"""Readers for the golden fixtures in tests/golden/ (converted by make_fixtures.py)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class RefFile:
    def __init__(self, z, name):
        self.name = name
        self.data = z[name + "/data"]
        self.columns = [str(c) for c in z[name + "/columns"]]
        self.header = json.loads(str(z[name + "/header"]))

    def col(self, name):
        return self.columns.index(name)

    def values(self, key):
        return [float(v) for v in self.header.get(key, [])]

    def param(self, key):
        p = self.header["params"]
        return float(p[p.index(key) + 1])


def load(fixture):
    z = np.load(os.path.join(GOLDEN, fixture))
    names = sorted({k.rsplit("/", 1)[0] for k in z.files})
    return z, names


def files(fixture):
    z, names = load(fixture)
    return [RefFile(z, n) for n in names]
