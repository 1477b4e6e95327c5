"""Host-side state set-up that the reference's Ensemble performs before the
first step: Maxwell-Boltzmann velocities from its own random stream.

Ensemble::drand (src/Ensemble.cpp:1853-1864) is the linear congruential
generator s <- (1029 s + 221591) mod 2^20 started from the config seed;
Ensemble::gauss (:1831-1850) maps twelve uniforms to a normal deviate with an
odd polynomial; Ensemble::init_velocity (:1752-1827) removes the mean momentum
and scales to the target temperature over ndf degrees of freedom.
"""
from __future__ import annotations

import numpy as np

BOLTZ = 0.0083145112          # src/NEMD_defs.h:108, kJ/(mol K)
_LCG_A, _LCG_C, _LCG_M = 1029, 221591, 1048576
# Polynomial coefficients of Ensemble::gauss (src/Ensemble.cpp:1837-1841).
_G1, _G3, _G5, _G7, _G9 = 3.949846138, 0.252408784, 0.076542912, 0.008355968, 0.029899776


def uniform_stream(seed: int, count: int) -> np.ndarray:
    """The first `count` values of Ensemble::drand from `seed`."""
    out = np.empty(count, dtype=np.float64)
    s = int(seed) & 0xFFFFFFFF
    for t in range(count):
        s = (s * _LCG_A + _LCG_C) % _LCG_M
        out[t] = s
    return out / _LCG_M


def gauss_stream(seed: int, count: int) -> np.ndarray:
    """`count` deviates of Ensemble::gauss; each consumes twelve uniforms in order."""
    u = uniform_stream(seed, 12 * count).reshape(count, 12)
    total = np.zeros(count)
    for k in range(12):          # left-to-right sum, as the reference's loop adds
        total = total + u[:, k]
    r = (total - 6.0) / 4.0
    r2 = r * r
    rst = (((_G9 * r2 + _G7) * r2 + _G5) * r2 + _G3) * r2 + _G1
    return rst * r


def degrees_of_freedom(system, per_atom: bool = True, constraint_on: bool = True) -> int:
    """ndf of Ensemble::init_velocity (:1760-1771): per atom, less one per
    bond and angle when the bonds are constrained; per molecule otherwise."""
    if per_atom:
        if constraint_on:
            return 3 * system.n_atoms - len(system.bonds) - len(system.angles) - 3
        return 3 * system.n_atoms - 3
    return 3 * system.n_mols - 3


def init_velocity(system, temperature: float, seed: int, per_atom: bool = True, ndf: int | None = None) -> np.ndarray:
    """Atom velocities [n, 3] (nm/ps) as Ensemble::init_velocity sets them."""
    if ndf is None:
        ndf = degrees_of_freedom(system, per_atom)
    n = system.n_atoms if per_atom else system.n_mols
    mom = gauss_stream(seed, 3 * n).reshape(n, 3)
    mom = mom - np.cumsum(mom, axis=0)[-1] / n      # running sums, in the reference's order
    if per_atom:
        mass = system.mass
    else:
        mass = np.zeros(system.n_mols)
        np.add.at(mass, system.mol, system.mass)
    l2 = (mom[:, 0] * mom[:, 0] + mom[:, 1] * mom[:, 1]) + mom[:, 2] * mom[:, 2]
    tavg = float(np.cumsum(l2 / mass)[-1]) / (ndf * BOLTZ)
    scale = np.sqrt(temperature / tavg)
    v = (scale * mom) / mass[:, None]
    if per_atom:
        return v
    return v[system.mol]
