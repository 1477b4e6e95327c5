"""The reference's input files: config.txt, sysData.in, coordinate, velocity
and restart files.

``SimConfiguration`` keeps the reference's getter names
(src/SimConfiguration.h:209-317); ``Parameters`` keeps its member names
(src/Parameters.h:150-190).  The parsing itself is native: the readers in
libswinmd_hip.so (csrc/inputs.cpp), called through the C ABI.  Where the
reference would carry on after a failed extraction with a zero or stale value
(or exit on a missing file, ERRMSG), these raise ``InputError`` naming the
file, line and field.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .ensemble import degrees_of_freedom, init_velocity
from .systems import AtomType, System

__all__ = ["InputError", "SimConfiguration", "Parameters", "read_coordinate", "read_velocity", "read_real_pos",
           "read_state", "build_system", "load_inputs"]


class InputError(_lib.SmdError):
    pass


def _call(fn, *args):
    err = C.create_string_buffer(512)
    rc = fn(*args, err, C.sizeof(err))
    if rc != _lib.SMD_OK:
        raise InputError(rc, err.value.decode(errors="replace") or _lib.load().smd_strerror(rc).decode())


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


# getter name -> field of smd_config (src/SimConfiguration.h:209-317)
_GETTERS = {
    "get_timestep": "time_step", "get_n_steps": "n_steps", "get_start_steps": "start_steps",
    "get_short_ts_freq": "short_ts_freq", "get_long_ts_freq": "long_ts_freq", "get_avg_ts_freq": "avg_ts_freq",
    "get_print_ts_freq": "print_ts_freq", "get_backup_ts_freq": "backup_ts_freq",
    "get_start_sampling": "start_sampling", "get_sampling_ts_freq": "sampling_ts_freq",
    "get_start_TTCF": "start_ttcf", "get_ttcf_ts_freq": "ttcf_ts_freq", "get_integrator_type": "integrator_type",
    "get_ensemble_status": "ensemble_status", "get_rdf_nbins": "rdf_nbins", "get_seed": "seed",
    "get_cutoff": "cutoff", "get_cutoffEw": "cutoff_ew", "get_cutoffInd": "cutoff_ind", "get_adamping": "a_damp",
    "get_cutBuff": "cut_buff", "get_k_cutoff": "k_cutoff", "get_alpha": "alpha",
    "get_cell_side_len": "cell_side_len", "get_density": "density", "get_temperature": "temperature",
    "get_coupl_const": "coupl_thermo", "get_pressure": "pressure", "get_baros_const": "coupl_baros",
    "get_CFB": "cfb", "get_DFB": "dfb", "get_TFB": "tfb", "get_tolerance": "tolerance",
    "get_max_count": "max_count", "get_switch_distance": "switch_dist", "get_kl_ub": "kl_ub",
    "get_l0_ub": "l0_ub", "get_nMie_param": "n_mie", "get_mMie_param": "m_mie", "get_nBuff_param": "n_buff",
    "get_mBuff_param": "m_buff", "get_deltaBuff_param": "delta_buff", "get_gammaBuff_param": "gamma_buff",
    "get_hfactor_reduction": "h_factor", "get_volume": "volume", "get_energy": "energy", "get_rate": "rate",
    "get_run_type": "run_type", "get_constraint_mode": "constraint_mode", "get_bending_type": "bending_type",
    # file names
    "get_coordinate_file": "coordinate_file", "get_velocity_file": "velocity_file",
    "get_sysData_file": "sys_data_file", "get_result_file": "result_file", "get_restart_file": "restart_file",
    "get_trajectory_file": "trajectory_file", "get_dump_file": "dump_file",
    "get_mol_traject_file": "mol_trajectory_file", "get_mol_velocity_file": "mol_velocity_file",
    "get_thermo_type": "thermo_type",
    # flags
    "is_read_velocity": "read_velocity", "is_new_start": "is_new_start", "is_constraint_on": "constraint_on",
    "is_switch_on": "switch_on", "write_trajectory": "write_trajectory", "use_long_correct": "long_correct",
    "use_cell_list": "use_cell_list", "use_TTCF": "ttcf", "use_harmonic_potential": "harmonic_bend",
    "use_amoeba_bending": "amoeba_bend", "use_atom_thermo": "use_atom_thermo",
    "compute_coulomb": "coulomb", "compute_mie": "mie", "compute_buffamoeba": "buff_vdw",
    "compute_harmonic_ub": "harmonic_ub", "compute_ewald_force": "ewald", "compute_surf_correct": "surf_correct",
    "compute_wolf": "wolf", "compute_multipol": "multipole", "compute_induction": "induction",
    "compute_torques": "torques", "use_damping_induction": "damp_induction",
    "do_induction_minimization": "induction_min", "use_damping_indAmoeba": "damp_amoeba",
    "has_ext_forces": "ext_forces", "has_fixed_atom": "fixed_atom", "compute_msd": "msd", "compute_rdf": "rdf",
    "compute_lustig": "lustig", "compute_vacf": "vacf",
    "is_constant_volume": "constant_volume", "is_constant_pressure": "constant_pressure",
    "is_constant_temperature": "constant_temperature", "is_constant_energy": "constant_energy",
    "do_equilibrium": "do_equilibrium", "do_compression": "do_compression", "do_shear": "do_shear",
    "do_elongation": "do_elongation",
}
_FLAGS = set(_lib._CONFIG_FLAGS)
_NAMES = set(_lib._CONFIG_NAMES)


def _getter(field):
    def get(self):
        v = getattr(self._require(), field)
        if field in _NAMES:
            return v.decode()
        if field in _FLAGS:
            return bool(v)
        return v
    get.__name__ = field
    return get


class SimConfiguration:
    """SimConfiguration (src/SimConfiguration.h): construct with the config
    file name, then ``read_config_file()``; the getters keep the reference's
    names.  ``as_dict()`` gives every field."""

    def __init__(self, configFile):
        if configFile is None:
            raise InputError(_lib.SMD_EINVAL, "Missing simulation configuration file")
        self.sysConfigFile = os.fspath(configFile)
        self._c = None

    def read_config_file(self):
        c = _lib.Config()
        _call(_lib.load().smd_config_read, _path(self.sysConfigFile), C.byref(c))
        self._c = c
        return self

    def _require(self):
        if self._c is None:
            raise InputError(_lib.SMD_EINVAL, "read_config_file() not called")
        return self._c

    def get_box_dimension(self):
        c = self._require()
        return np.array([c.lx_box, c.ly_box, c.lz_box])

    def get_box_volume(self):
        c = self._require()
        return c.lx_box * c.ly_box * c.lz_box

    def as_dict(self):
        c = self._require()
        out = {}
        for name, _ in _lib.Config._fields_:
            v = getattr(c, name)
            out[name] = v.decode() if name in _NAMES else v
        return out


for _g, _f in _GETTERS.items():
    setattr(SimConfiguration, _g, _getter(_f))


def _rows(ptr, n, width):
    if n == 0:
        return np.zeros((0, width), dtype=np.int32)
    return np.ctypeslib.as_array(ptr, shape=(n * width,)).reshape(n, width).copy()


class Parameters:
    """Parameters::read_sys_data (src/Parameters.cpp:76-382): sizes, force
    field tables and topology of sysData.in, with the reference's member
    names.  Parameter rows are dicts; topology arrays are numpy int32."""

    def __init__(self, sysDataFile):
        self.sysDataFile = os.fspath(sysDataFile)
        lib = _lib.load()
        h = C.c_void_p()
        _call(lib.smd_sysdata_read, _path(self.sysDataFile), C.byref(h))
        try:
            v = _lib.SysdataView()
            rc = lib.smd_sysdata_view_get(h, C.byref(v))
            if rc != _lib.SMD_OK:
                raise InputError(rc, "smd_sysdata_view_get failed")
            self._copy(v)
        finally:
            lib.smd_sysdata_free(h)

    def _copy(self, v):
        z = v.sizes
        self.nMolecules, self.nMolTypes = z.n_mols, z.n_mol_types
        self.nAtoms, self.nAtomTypes = z.n_atoms, z.n_atom_types
        self.nBonds, self.nBondTypes = z.n_bonds, z.n_bond_types
        self.nAngles, self.nAngleTypes = z.n_angles, z.n_angle_types
        self.nDihedrals, self.nDihedralTypes = z.n_dihedrals, z.n_dihedral_types
        self.nImpropers, self.nImproperTypes = z.n_impropers, z.n_improper_types
        self.atomParams = [dict(atomType=a.type, mass=a.mass, charge=a.charge, eps=a.eps, sigma=a.sigma,
                                dip=tuple(a.dip), quad=tuple(a.quad), polar=a.polar,
                                typeName=a.type_name.decode(), atomName=a.atom_name.decode())
                           for a in (v.atom_params[k] for k in range(self.nAtomTypes))]
        self.bondParams = [dict(bondType=b.type, atomType1=b.t1, atomType2=b.t2, r0=b.r0, k=b.k)
                           for b in (v.bond_params[k] for k in range(self.nBondTypes))]
        self.angleParams = [dict(angleType=a.type, atomType1=a.t1, atomType2=a.t2, atomType3=a.t3,
                                 theta0=a.theta0, k=a.k, theta0_deg=a.theta0_deg)
                            for a in (v.angle_params[k] for k in range(self.nAngleTypes))]

        def tors(ptr, n):
            return [dict(torsionType=t.type, atomType1=t.t1, atomType2=t.t2, atomType3=t.t3, atomType4=t.t4,
                         k=t.k, d=t.d, n=t.n) for t in (ptr[k] for k in range(n))]

        self.dihedralParams = tors(v.dihedral_params, self.nDihedralTypes)
        self.improperParams = tors(v.improper_params, self.nImproperTypes)
        self.molTypes = [dict(molType=m.type, numOfAtoms=m.n_atoms, name=m.name.decode())
                         for m in (v.mol_types[k] for k in range(self.nMolTypes))]
        self.atomType = _rows(v.atom_type, self.nAtoms, 1).ravel()
        self.atomMol = _rows(v.atom_mol, self.nAtoms, 1).ravel()
        self.molType = _rows(v.mol_type, self.nMolecules, 1).ravel()
        self.bondArray = _rows(v.bonds, self.nBonds, 3)
        self.angleArray = _rows(v.angles, self.nAngles, 4)
        self.dihedralArray = _rows(v.dihedrals, self.nDihedrals, 5)
        self.improperArray = _rows(v.impropers, self.nImpropers, 5)


def read_coordinate(path, n_atoms: int, skip_lines: int = 0) -> np.ndarray:
    """Ensemble::read_coordinate (src/Ensemble.cpp:630-670): count line, then
    "x y z" per atom."""
    out = np.zeros((int(n_atoms), 3))
    _call(_lib.load().smd_read_xyz, _path(path), int(n_atoms), int(skip_lines),
          out.ctypes.data_as(C.c_void_p))
    return out


read_velocity = read_coordinate       # Ensemble::read_velocity (:712-758), same layout
read_real_pos = read_coordinate       # Ensemble::read_realPos (:672-710)


def read_state(path, n_atoms: int):
    """The restart file (src/Ensemble.cpp:299-317): positions, unwrapped
    positions and velocities, each block after one skipped line."""
    n = int(n_atoms)
    pos, real, vel = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    _call(_lib.load().smd_read_state, _path(path), n, p(pos), p(real), p(vel))
    return pos, real, vel


def build_system(params: Parameters, positions: np.ndarray, density: float, name: str = "inputs") -> System:
    """A System from sysData.in and coordinates, as Ensemble::set_mol_system
    assembles atoms and molecules (src/Ensemble.cpp:140-227).  The box is the
    one Ensemble::compute_bound_box sets (src/Ensemble.cpp:369-442): a cube of
    volume nAtoms/density when density > 0."""
    if params.nMolTypes != 1:
        raise InputError(_lib.SMD_ENOTSUP, f"{params.nMolTypes} molecule types; one is supported")
    if not density > 0:
        raise InputError(_lib.SMD_ENOTSUP, "density <= 0 (box from the atoms' bounding box) is not supported")
    mol = np.asarray(params.atomMol, dtype=np.int32)
    if np.any(np.diff(mol) < 0) or np.any(mol < 0):
        raise InputError(_lib.SMD_EINVAL, "atoms of one molecule must be contiguous, molecules in order")
    types = [AtomType(mass=a["mass"], charge=a["charge"], eps=a["eps"], sigma=a["sigma"], dip=a["dip"],
                      quad=a["quad"], polar=a["polar"], type_name=a["typeName"], atom_name=a["atomName"])
             for a in params.atomParams]
    mt = params.molTypes[0]
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    if pos.shape != (params.nAtoms, 3):
        raise InputError(_lib.SMD_EINVAL, f"positions shape {pos.shape}, expected ({params.nAtoms}, 3)")
    return System(name=name, types=types, mol_type_name=mt["name"], atoms_per_mol=mt["numOfAtoms"],
                  pos=pos.copy(), type=np.asarray(params.atomType, dtype=np.int32), mol=mol, density=density,
                  bonds=[tuple(int(x) for x in r) for r in params.bondArray],
                  angles=[tuple(int(x) for x in r) for r in params.angleArray],
                  bond_params=[(b["atomType1"], b["atomType2"], b["r0"], b["k"]) for b in params.bondParams],
                  angle_params=[(a["atomType1"], a["atomType2"], a["atomType3"], a["theta0_deg"], a["k"])
                                for a in params.angleParams])


def load_inputs(directory, config_file: str = "config.txt"):
    """Reads config.txt and the files it names (relative to ``directory``,
    as the reference resolves them against its working directory).  Returns
    (config, parameters, system, velocities) as the Ensemble constructor sets
    them up (src/Ensemble.cpp:283-318): velocities from the velocity file
    when readVelocity is on, else drawn by Ensemble::init_velocity with the
    config's seed and temperature; positions and velocities from state.in on
    a restart."""
    d = os.fspath(directory)
    cfg = SimConfiguration(os.path.join(d, config_file)).read_config_file()
    params = Parameters(os.path.join(d, cfg.get_sysData_file()))
    vel = None
    if cfg.is_new_start():
        pos = read_coordinate(os.path.join(d, cfg.get_coordinate_file()), params.nAtoms)
        if cfg.is_read_velocity():
            vel = read_velocity(os.path.join(d, cfg.get_velocity_file()), params.nAtoms)
    else:
        pos, _, vel = read_state(os.path.join(d, "state.in"), params.nAtoms)
    system = build_system(params, pos, cfg.get_density(), name=os.path.basename(os.path.abspath(d)))
    if vel is None:
        per_atom = cfg.use_atom_thermo()
        ndf = degrees_of_freedom(system, per_atom, cfg.is_constraint_on())
        vel = init_velocity(system, cfg.get_temperature(), cfg.get_seed(), per_atom=per_atom, ndf=ndf)
    return cfg, params, system, vel
