"""Shared test helpers: fixtures, systems and error norms."""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from swinmd_amd import systems as S  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def build_system(spec):
    kind = spec["builder"]
    args = dict(spec["args"])
    if kind == "water":
        if "types" in args:
            args["types"] = {"spce": S.SPCE_TYPES, "amoeba": S.AMOEBA_TYPES, "polar": S.POLAR_TYPES}[args["types"]]
        if "lattice" in args:
            args["lattice"] = tuple(args["lattice"])
        return S.water(**args)
    if kind == "lj_fcc":
        return S.lj_fcc(**args)
    if kind == "co2":
        return S.co2(**args)
    raise ValueError(kind)


class Fixture:
    def __init__(self, name):
        d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        self.name = name
        self.spec = json.loads(str(d["spec"]))
        self.cfg = json.loads(str(d["cfg"]))
        self.meta = json.loads(str(d["meta"]))
        self.mode = str(d["mode"])
        self.extra = str(d["extra"])
        self.system = build_system(self.spec)
        self.system.pos = np.array(d["pos"], dtype=np.float64)
        self.ref = {k[4:]: np.array(d[k]) for k in d.files if k.startswith("ref_")}

    def vec(self, key):
        return self.ref[key].reshape(-1, 3)

    def pairs(self):
        off = self.ref["pl_off"].astype(np.int32)
        nb = self.ref["pl_nb"].astype(np.int32)
        return off, nb


def rel_rms(a, b):
    """sqrt(sum |a-b|^2 / sum |b|^2) over all sites."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = float(np.sum(b * b))
    if den == 0.0:
        return float(np.sqrt(np.sum(a * a)))
    return float(np.sqrt(np.sum((a - b) ** 2) / den))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def pair_set(off, nb):
    """Unordered pair set {(min, max)} from CSR rows."""
    rows = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off))
    j = nb.astype(np.int64)
    lo = np.minimum(rows, j)
    hi = np.maximum(rows, j)
    return set(zip(lo.tolist(), hi.tolist()))


def ensure_oracle():
    from oracle import oracle as O
    O.build()
    return O
