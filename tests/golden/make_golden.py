"""Generates the golden fixtures in tests/golden/*.npz from the compiled SwinMD
reference (oracle/_ref/ref_harness, built by `make -f oracle/Makefile.ref`).

Run in the build container only (the reference sources are needed to build the
harness):  python tests/golden/make_golden.py

Each fixture stores the inputs (builder name and arguments, positions, run
settings) and the reference's outputs: pair lists, per-term forces,
energies, Lustig derivatives and virials, or the state after N leap-frog
steps.  Inputs are the synthetic systems of swinmd_amd.systems written in the
reference's own input formats.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle.refinput import write_inputs  # noqa: E402
from swinmd_amd import systems as S  # noqa: E402
from tests.helpers import build_system  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.dirname(os.path.abspath(__file__))


def run(spec, mode, extra, cfg):
    s = build_system(spec)
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, s, **cfg)
        cmd = [HARNESS, d, mode] + ([extra] if extra else [])
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"harness failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        meta = {}
        for line in open(os.path.join(d, "meta.txt")):
            k, v = line.split()
            meta[k] = float(v)
        arrays = read_outputs(d)
    return s, meta, arrays


def read_outputs(d):
    """The harness's *.bin arrays (int32 for names starting pl_ or i_, else
    float64) and, when present, names.txt as an array of "key value" rows."""
    arrays = {}
    for f in os.listdir(d):
        if f.endswith(".bin"):
            name = f[:-4]
            dt = np.int32 if name.startswith(("pl_", "i_")) else np.float64
            arrays[name] = np.fromfile(os.path.join(d, f), dtype=dt)
    names = os.path.join(d, "names.txt")
    if os.path.exists(names):
        arrays["names"] = np.array([ln.rstrip("\n") for ln in open(names)])
    # text the reference's own writers produced, kept whole
    for f, key in (("state.out", "state_text"), ("press_res.out", "press_res_text"), ("pressac.out", "pressac_text"),
                   ("pressure_tensor.out", "pressure_tensor_text")):
        path = os.path.join(d, f)
        if os.path.exists(path):
            arrays[key] = np.array(open(path).read())
    return arrays


# The reference's own example inputs (its test/ directory), kept as data under
# tests/golden/ref_test_inputs/.
REF_TEST_INPUTS = os.path.join(OUT, "ref_test_inputs")


def run_ref_test_inputs(mode, extra=None, edits=None):
    """Runs the harness on a copy of the reference's example inputs; edits
    maps a config.txt line index to its replacement."""
    with tempfile.TemporaryDirectory() as d:
        for f in os.listdir(REF_TEST_INPUTS):
            shutil.copy(os.path.join(REF_TEST_INPUTS, f), d)
        if edits:
            cfg = os.path.join(d, "config.txt")
            lines = open(cfg).read().split("\n")
            for k, v in edits.items():
                lines[k] = v
            open(cfg, "w").write("\n".join(lines))
        cmd = [HARNESS, d, mode] + ([extra] if extra else [])
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=1200)
        if r.returncode != 0:
            raise RuntimeError(f"harness failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        meta = {}
        for line in open(os.path.join(d, "meta.txt")):
            k, v = line.split()
            meta[k] = float(v)
        return meta, read_outputs(d)


def save_ref_test(name, mode, extra, edits, meta, arrays):
    path = os.path.join(OUT, name + ".npz")
    payload = {"mode": mode, "extra": extra or "", "meta": json.dumps(meta),
               "edits": json.dumps({str(k): v for k, v in (edits or {}).items()})}
    for k, v in arrays.items():
        payload["ref_" + k] = v
    np.savez_compressed(path, **payload)
    print(f"{path}: {os.path.getsize(path) / 1024:.0f} KiB, {len(arrays)} arrays")


# (fixture name, harness mode, extra argument, config.txt line edits) on the
# reference's own inputs
HARMONIC_BENDING = {14: "0\t0.00001 1\t0.01\t// constrainMode; tolerance; bendType 1 (harmonic)"}
REF_TEST_CASES = [
    ("ref_test_inputs_parsed", "inputs", None, None),
    ("ref_test_bonded_amoeba", "forces", "bond,angle,ub", None),
    ("ref_test_bonded_harmonic", "forces", "bond,angle,ub", HARMONIC_BENDING),
    ("ref_test_lf_forces", "lf_forces", None, None),
]


def save(name, spec, cfg, mode, extra, s, meta, arrays):
    path = os.path.join(OUT, name + ".npz")
    payload = {"spec": json.dumps(spec), "cfg": json.dumps(cfg), "mode": mode, "extra": extra or "",
               "meta": json.dumps(meta), "pos": s.pos}
    for k, v in arrays.items():
        payload["ref_" + k] = v
    np.savez_compressed(path, **payload)
    print(f"{path}: {os.path.getsize(path) / 1024:.0f} KiB, {len(arrays)} arrays")


CASES = [
    # (name, spec, mode, extra, cfg)
    ("water216_brute_forces",
     {"builder": "water", "args": {"n_mol": 216, "box": 1.9, "seed": 3, "lattice": [6, 6, 6]}},
     "forces", "lj,ewald",
     {"cutoff": 0.6, "cut_buff": 0.05, "alpha": 3.5, "kcut": S.kcut_for(27, 1.9), "surface": True, "lustig": True}),
    ("water512_cells_forces",
     {"builder": "water", "args": {"n_mol": 512, "density_mol": 33.4, "seed": 5, "lattice": [8, 8, 8]}},
     "forces", "lj,ewald",
     {"cutoff": 0.7, "cut_buff": 0.1, "alpha": 3.2, "kcut": 20.0, "surface": False, "lustig": True}),
    ("water512_cells_pairlist",
     {"builder": "water", "args": {"n_mol": 512, "density_mol": 33.4, "seed": 5, "lattice": [8, 8, 8]}},
     "pairlist", None, {"cutoff": 0.7, "cut_buff": 0.1}),
    ("water216_brute_pairlist",
     {"builder": "water", "args": {"n_mol": 216, "box": 1.9, "seed": 3, "lattice": [6, 6, 6]}},
     "pairlist", None, {"cutoff": 0.6, "cut_buff": 0.05}),
    ("lj864_cells_pairlist",
     {"builder": "lj_fcc", "args": {"cells": 6, "rho": 0.8442, "jitter": 0.05, "seed": 7}},
     "pairlist", None, {"cutoff": 2.5, "cut_buff": 0.3, "constraint": False, "ewald": False}),
    ("lj864_forces",
     {"builder": "lj_fcc", "args": {"cells": 6, "rho": 0.8442, "jitter": 0.05, "seed": 7}},
     "forces", "lj,mie",
     {"cutoff": 2.5, "cut_buff": 0.3, "constraint": False, "ewald": False, "switch_on": True,
      "switch_dist": 2.3, "lustig": True, "mie": True, "n_mie": 12.0, "m_mie": 6.0}),
    ("lj864_mie147_forces",
     {"builder": "lj_fcc", "args": {"cells": 6, "rho": 0.8442, "jitter": 0.05, "seed": 7}},
     "forces", "mie",
     {"cutoff": 2.5, "cut_buff": 0.3, "constraint": False, "ewald": False, "switch_on": True,
      "switch_dist": 2.2, "lustig": True, "mie": True, "n_mie": 14.0, "m_mie": 7.0}),
    ("amoeba216_buff_forces",
     {"builder": "water", "args": {"n_mol": 216, "density_mol": 33.4, "seed": 11, "lattice": [6, 6, 6],
                                   "r_oh": 0.09572, "hoh_deg": 108.5, "types": "amoeba"}},
     "forces", "buff",
     {"cutoff": 0.55, "cut_buff": 0.05, "switch_on": True, "switch_dist": 0.5, "buff": True, "lustig": True,
      "ewald": False}),
    ("co2_216_buck_forces",
     {"builder": "co2", "args": {"n_mol": 216, "density_mol": 10.0, "seed": 13}},
     "forces", "buck",
     {"cutoff": 0.8, "cut_buff": 0.05, "switch_on": True, "switch_dist": 0.7, "lustig": True,
      "ewald": False, "coulomb": True}),
    ("amoeba216_induction_cg",
     {"builder": "water", "args": {"n_mol": 216, "density_mol": 33.4, "seed": 19, "lattice": [6, 6, 6],
                                   "r_oh": 0.09572, "hoh_deg": 108.5, "types": "amoeba"}},
     "induction", "",
     {"cutoff": 0.6, "cut_buff": 0.01, "induction": True, "ind_damp": True, "ind_min": True, "cutoff_ind": 0.6,
      "a_damp": 0.39, "ewald": False}),
    ("amoeba216_induction_direct",
     {"builder": "water", "args": {"n_mol": 216, "density_mol": 33.4, "seed": 23, "lattice": [6, 6, 6],
                                   "r_oh": 0.09572, "hoh_deg": 108.5, "types": "amoeba"}},
     "induction", "",
     {"cutoff": 0.6, "cut_buff": 0.01, "induction": True, "ind_damp": False, "ind_min": False,
      "cutoff_ind": 0.55, "a_damp": 0.39, "ewald": False}),
    ("amoeba216_multipole",
     {"builder": "water", "args": {"n_mol": 216, "density_mol": 33.4, "seed": 41, "lattice": [6, 6, 6],
                                   "r_oh": 0.09572, "hoh_deg": 108.5, "types": "amoeba"}},
     "multipole", "",
     {"cutoff": 0.7, "cut_buff": 0.05, "cutoff_ew": 0.65, "alpha": 3.0, "kcut": 14.0, "multipole": True,
      "torques": True, "surface": True, "constraint": True, "ewald": False}),
    ("amoeba216_multipole_flexible",
     {"builder": "water", "args": {"n_mol": 216, "density_mol": 33.4, "seed": 43, "lattice": [6, 6, 6],
                                   "r_oh": 0.09572, "hoh_deg": 108.5, "types": "amoeba"}},
     "multipole", "",
     {"cutoff": 0.6, "cut_buff": 0.1, "cutoff_ew": 0.7, "alpha": 3.4, "kcut": 12.0, "multipole": True,
      "torques": True, "surface": False, "constraint": False, "ewald": False}),
    ("amoeba216_indforces_cg",
     {"builder": "water", "args": {"n_mol": 216, "density_mol": 33.4, "seed": 47, "lattice": [6, 6, 6],
                                   "r_oh": 0.09572, "hoh_deg": 108.5, "types": "amoeba"}},
     "induction_forces", "",
     {"cutoff": 0.7, "cut_buff": 0.05, "cutoff_ew": 0.7, "alpha": 3.0, "kcut": 14.0, "induction": True,
      "torques": True, "surface": True, "constraint": True, "ewald": False, "ind_damp": True, "ind_min": True,
      "cutoff_ind": 0.7, "a_damp": 0.39, "damp_amoeba": False}),
    ("amoeba216_indforces_tinker_flexible",
     {"builder": "water", "args": {"n_mol": 216, "density_mol": 33.4, "seed": 53, "lattice": [6, 6, 6],
                                   "r_oh": 0.09572, "hoh_deg": 108.5, "types": "amoeba"}},
     "induction_forces", "",
     {"cutoff": 0.65, "cut_buff": 0.05, "cutoff_ew": 0.68, "alpha": 3.3, "kcut": 13.0, "induction": True,
      "torques": True, "surface": False, "constraint": False, "ewald": False, "ind_damp": True, "ind_min": True,
      "cutoff_ind": 0.65, "a_damp": 0.39, "damp_amoeba": True}),
    ("amoeba216_indforces_tinker_direct",
     {"builder": "water", "args": {"n_mol": 216, "density_mol": 33.4, "seed": 59, "lattice": [6, 6, 6],
                                   "r_oh": 0.09572, "hoh_deg": 108.5, "types": "amoeba"}},
     "induction_forces", "",
     {"cutoff": 0.7, "cut_buff": 0.05, "cutoff_ew": 0.7, "alpha": 3.0, "kcut": 14.0, "induction": True,
      "torques": True, "surface": True, "constraint": True, "ewald": False, "ind_damp": False, "ind_min": False,
      "cutoff_ind": 0.6, "a_damp": 0.39, "damp_amoeba": True}),
    ("water512_md10",
     {"builder": "water", "args": {"n_mol": 512, "density_mol": 33.4, "seed": 5, "lattice": [8, 8, 8]}},
     "md", "10",
     {"cutoff": 0.7, "cut_buff": 0.1, "alpha": 3.2, "kcut": 20.0, "dt": 0.002, "temperature": 300.0,
      "coupl": 0.1, "shake_tol": 1e-8, "max_count": 10000, "seed": 17, "atom_thermo": True}),
    ("water216_md20_moltherm",
     {"builder": "water", "args": {"n_mol": 216, "box": 1.9, "seed": 3, "lattice": [6, 6, 6]}},
     "md", "20",
     {"cutoff": 0.6, "cut_buff": 0.05, "alpha": 3.5, "kcut": S.kcut_for(27, 1.9), "dt": 0.002,
      "temperature": 280.0, "coupl": 0.05, "shake_tol": 1e-8, "max_count": 10000, "seed": 29,
      "atom_thermo": False}),
    # rigid polarizable water (the AMOEBA-like site parameters, H eps = 0 as the
    # polarizable configs set it): buffered 14-7, multipole Ewald with torques,
    # induced dipoles by the reference's CG, then SHAKE and Nose-Hoover, 10
    # leap-frog steps (LFIntegrator.cpp:442-473, 561-586)
    ("amoeba216_md10_polar",
     {"builder": "water", "args": {"n_mol": 216, "density_mol": 33.4, "seed": 67, "lattice": [6, 6, 6],
                                   "types": "polar"}},
     "md", "10",
     {"cutoff": 0.7, "cut_buff": 0.05, "cutoff_ew": 0.7, "alpha": 3.0, "kcut": 14.0, "buff": True,
      "switch_on": False, "multipole": True, "induction": True, "torques": True, "surface": False,
      "constraint": True, "ewald": False, "ind_damp": True, "ind_min": True, "cutoff_ind": 0.7, "a_damp": 0.39,
      "damp_amoeba": False, "dt": 0.002, "temperature": 300.0, "coupl": 0.1, "shake_tol": 1e-8,
      "max_count": 10000, "seed": 71, "atom_thermo": True, "lustig": True}),
    # the 1,024-molecule polarizable water of the headline workload's small
    # form (a seeded subset of an 11^3 lattice in L = 3.1289 nm, polarizable
    # site parameters): 3 steps from the lattice start with its parameters
    # (cutoff 0.9 nm, Ewald alpha 3.5 /nm, |n|^2 <= 81, Thole a = 0.39)
    ("polar1024_md3",
     {"builder": "water", "args": {"n_mol": 1024, "box": 3.1289, "seed": 2025, "lattice": [11, 11, 11],
                                   "subset": True, "types": "polar"}},
     "md", "3",
     {"cutoff": 0.9, "cut_buff": 0.1, "cutoff_ew": 0.9, "alpha": 3.5, "kcut": 18.128713661985863, "buff": True,
      "switch_on": False, "multipole": True, "induction": True, "torques": True, "surface": False,
      "constraint": True, "ewald": False, "ind_damp": True, "ind_min": True, "cutoff_ind": 0.9, "a_damp": 0.39,
      "damp_amoeba": False, "dt": 0.002, "temperature": 300.0, "coupl": 0.1, "shake_tol": 1e-8,
      "max_count": 10000, "seed": 73, "atom_thermo": True, "lustig": True}),
    # 864 LJ atoms on the fcc lattice at rho* = 0.8442 (sigma = eps = m = 1),
    # energy-shifted cutoff 2.5 sigma, unconstrained leap-frog with the
    # self-consistent Nose-Hoover velocity loop (LFIntegrator.cpp:587-643):
    # T* = 0.722 is 0.722 / BOLTZMAN K, Q = 10 is tau = sqrt(Q / (Ndf kB T))
    # with Ndf = 3N - 3; dt* = 0.005.  A 0.1 sigma buffer makes the Verlet
    # trigger fire within the run.
    ("lj864_md20_nh",
     {"builder": "lj_fcc", "args": {"cells": 6, "rho": 0.8442, "jitter": 0.05, "seed": 7}},
     "md", "20",
     {"cutoff": 2.5, "cut_buff": 0.1, "constraint": False, "ewald": False, "switch_on": True,
      "switch_dist": 2.5, "lustig": True, "dt": 0.005, "temperature": 0.722 / 0.0083145112,
      "coupl": (10.0 / (2589 * 0.722)) ** 0.5, "seed": 17, "atom_thermo": True}),
]


def main(only=None):
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: make -f oracle/Makefile.ref")
    for name, spec, mode, extra, cfg in CASES:
        if only and name not in only:
            continue
        s, meta, arrays = run(spec, mode, extra, cfg)
        save(name, spec, cfg, mode, extra, s, meta, arrays)
    for name, mode, extra, edits in REF_TEST_CASES:
        if only and name not in only:
            continue
        meta, arrays = run_ref_test_inputs(mode, extra, edits)
        save_ref_test(name, mode, extra, edits, meta, arrays)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
