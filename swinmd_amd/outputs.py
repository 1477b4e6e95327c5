"""The reference's restart and pressure output files.

Writers of Integrator::write_state (src/Integrator.cpp:275-302),
NHIntegrator::write_vtMinus (src/integrators/NHIntegrator.cpp:1782-1803) and
Ensemble::write_press_res / write_pressac / write_pressureTensor
(src/Ensemble.cpp:1153-1172, :1887-1912).  The formatting is native
(csrc/outputs.cpp, std::ostream with the reference's manipulators); these
functions pass arrays through the C ABI.  ``precision`` 0 keeps the stream
default of 6 significant digits, what the reference writes; 17 makes a
restart lossless.  The readers are in :mod:`swinmd_amd.inputs`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .inputs import InputError

__all__ = ["write_state", "write_vt_minus", "write_press_res", "write_pressac", "write_pressure_tensor"]


def _call(fn, *args):
    err = C.create_string_buffer(512)
    rc = fn(*args, err, C.sizeof(err))
    if rc != _lib.SMD_OK:
        raise InputError(rc, err.value.decode(errors="replace") or _lib.load().smd_strerror(rc).decode())


def _v(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise InputError(_lib.SMD_EINVAL, f"expected {n} values, got {a.size}")
    return a


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def write_state(path, step, pos, real_pos, vel, precision=0):
    pos, real_pos, vel = (_v(x) for x in (pos, real_pos, vel))
    n = pos.size // 3
    if real_pos.size != 3 * n or vel.size != 3 * n:
        raise InputError(_lib.SMD_EINVAL, "positions, real positions and velocities differ in size")
    _call(_lib.load().smd_write_state, os.fsencode(os.fspath(path)), int(step), n, _p(pos), _p(real_pos), _p(vel),
          int(precision))


def write_vt_minus(path, step, v, precision=0):
    v = _v(v)
    _call(_lib.load().smd_write_vt_minus, os.fsencode(os.fspath(path)), int(step), v.size // 3, _p(v),
          int(precision))


def write_press_res(path, step, atom_pressure, mol_pressure, append=True):
    _call(_lib.load().smd_write_press_res, os.fsencode(os.fspath(path)), int(bool(append)), int(step),
          _p(_v(atom_pressure, 9)), _p(_v(mol_pressure, 9)))


def write_pressac(path, step, atom_stress, mol_stress, samples, append=True):
    _call(_lib.load().smd_write_pressac, os.fsencode(os.fspath(path)), int(bool(append)), int(step),
          _p(_v(atom_stress, 5)), _p(_v(mol_stress, 5)), int(samples))


def write_pressure_tensor(path, counter1, atom_pressure, mol_pressure, atom_shear_sum, mol_shear_sum, counter,
                          append=True):
    _call(_lib.load().smd_write_pressure_tensor, os.fsencode(os.fspath(path)), int(bool(append)), int(counter1),
          _p(_v(atom_pressure, 9)), _p(_v(mol_pressure, 9)), float(atom_shear_sum), float(mol_shear_sum),
          int(counter))
