"""ctypes declarations for libswinmd_hip.so (include/swinmd_hip.h).

The library is built in-tree by ``python -m swinmd_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswinmd_hip.so")

SMD_OK = 0
SMD_EINVAL, SMD_ENOMEM, SMD_EDEVICE, SMD_ESTATE = -1, -2, -3, -4
SMD_ECAPACITY, SMD_ESHAKE, SMD_ENOCONV, SMD_ENOTSUP = -5, -6, -7, -8
SMD_EIO = -9
PREC_FP64, PREC_MIXED = 0, 1
VDW_LJ, VDW_MIE, VDW_BUCK, VDW_BUFF147 = 0, 1, 2, 3


class Topology(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("n_mols", C.c_int32), ("n_types", C.c_int32),
                ("atom_type", C.c_void_p), ("mol_id", C.c_void_p), ("scaled_charge", C.c_void_p),
                ("mass", C.c_void_p), ("excl_offsets", C.c_void_p), ("excl_atoms", C.c_void_p)]


class VdwParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("cutoff", C.c_double), ("switch_dist", C.c_double),
                ("switch_on", C.c_int32), ("use_constraint", C.c_int32), ("lustig", C.c_int32),
                ("coulomb", C.c_int32), ("sigma2", C.c_void_p), ("eps4", C.c_void_p),
                ("n_mie", C.c_double), ("m_mie", C.c_double), ("buck_a", C.c_void_p),
                ("buck_b", C.c_void_p), ("buck_c", C.c_void_p), ("co2_bond", C.c_double),
                ("type_sigma", C.c_void_p), ("type_eps", C.c_void_p), ("n_buff", C.c_double),
                ("m_buff", C.c_double), ("delta", C.c_double), ("gamma", C.c_double),
                ("hfactor", C.c_double), ("e_lrc", C.c_double), ("v_lrc", C.c_double),
                ("du_lrc", C.c_double), ("d2u_lrc", C.c_double)]


class VdwOut(C.Structure):
    _fields_ = [("energy", C.c_double), ("lj_energy", C.c_double), ("virial", C.c_double * 9),
                ("ulus", C.c_double), ("dudv", C.c_double), ("d2udv", C.c_double)]


class EwaldParams(C.Structure):
    _fields_ = [("alpha", C.c_double), ("real_cut", C.c_double), ("kcut", C.c_double),
                ("use_constraint", C.c_int32), ("surface", C.c_int32)]


class EwaldOut(C.Structure):
    _fields_ = [("real_energy", C.c_double), ("real_deriv", C.c_double), ("real_deriv2", C.c_double),
                ("self_energy", C.c_double), ("charged_energy", C.c_double), ("intra_energy", C.c_double),
                ("intra_vir_correct", C.c_double), ("mol_deriv", C.c_double), ("mol_deriv2", C.c_double),
                ("surf_energy", C.c_double), ("surf_deriv", C.c_double), ("surf_deriv2", C.c_double),
                ("virial", C.c_double * 9)]


class EwaldLongOut(C.Structure):
    _fields_ = [("energy", C.c_double), ("deriv", C.c_double), ("deriv2", C.c_double),
                ("virial", C.c_double * 9)]


class IntegParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("temperature", C.c_double), ("coupl", C.c_double),
                ("tol", C.c_double), ("max_count", C.c_int32), ("nose_hoover", C.c_int32),
                ("atom_thermo", C.c_int32)]


class MdParams(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("cut_buff", C.c_double), ("ewald", EwaldParams),
                ("integ", IntegParams), ("use_cells", C.c_int32), ("long_steps", C.c_int32),
                ("lustig", C.c_int32), ("prec", C.c_int32), ("sigma2", C.c_void_p), ("eps4", C.c_void_p),
                ("use_graph", C.c_int32), ("total_steps", C.c_int64), ("constraint_mode", C.c_int32),
                ("ewald_on", C.c_int32), ("switch_on", C.c_int32)]


class MdPressure(C.Structure):
    _fields_ = [("virial", C.c_double * 9), ("mol_virial", C.c_double * 9), ("atom_pressure", C.c_double * 9),
                ("mol_pressure", C.c_double * 9), ("mol_pt", C.c_double), ("counter", C.c_int64),
                ("samples", C.c_int64), ("atom_stress", C.c_double * 5), ("mol_stress", C.c_double * 5),
                ("lj_dudv", C.c_double), ("real_deriv", C.c_double), ("long_deriv", C.c_double),
                ("mol_deriv", C.c_double), ("surf_deriv", C.c_double)]


class MdReport(C.Structure):
    _fields_ = [("steps", C.c_int64), ("list_updates", C.c_int64), ("e_lj", C.c_double),
                ("e_real", C.c_double), ("e_long", C.c_double), ("e_self", C.c_double),
                ("e_intra", C.c_double), ("e_surf", C.c_double), ("atom_ke2", C.c_double),
                ("mol_ke2", C.c_double), ("fnh", C.c_double), ("virial", C.c_double * 9)]


class InduceParams(C.Structure):
    _fields_ = [("cutoff_ind", C.c_double), ("a_damp", C.c_double), ("damp_tij", C.c_int32),
                ("do_min", C.c_int32), ("tol", C.c_double), ("max_iter", C.c_int32), ("guess", C.c_int32),
                ("prec", C.c_int32), ("precond", C.c_int32), ("tol_rms", C.c_int32)]

DEBYE_E_NM = 0.020819434   # 1 Debye in e nm


class InduceReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("residual", C.c_double)]


MLP_FEATURES, MLP_HIDDEN, MLP_OUTPUTS, MLP_NPARAMS, MLP_NNORM = 32, 128, 3, 37635, 70


class MlpTrainParams(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("lr_final", C.c_float),
                ("seed", C.c_uint64), ("renormalize", C.c_int32)]


class MlpTrainReport(C.Structure):
    _fields_ = [("samples", C.c_int64), ("steps", C.c_int32), ("loss_first", C.c_double),
                ("loss_last", C.c_double), ("ms", C.c_double)]


class MpParams(C.Structure):
    _fields_ = [("alpha", C.c_double), ("real_cut", C.c_double), ("use_constraint", C.c_int32),
                ("surface", C.c_int32), ("torques", C.c_int32)]


class MpOut(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("self_energy", "charged_energy", "real_energy", "intra_energy",
                                           "surf_energy", "real_deriv", "mol_deriv", "surf_deriv", "surf_deriv2",
                                           "tot_torque")] + [("virial", C.c_double * 9)]


class MpLongOut(C.Structure):
    _fields_ = [("energy", C.c_double), ("deriv", C.c_double), ("deriv2", C.c_double), ("tot_torque", C.c_double),
                ("virial", C.c_double * 9)]


class IndfParams(C.Structure):
    _fields_ = [("alpha", C.c_double), ("real_cut", C.c_double), ("a_damp", C.c_double),
                ("use_constraint", C.c_int32), ("surface", C.c_int32), ("torques", C.c_int32),
                ("damp_amoeba", C.c_int32), ("do_min", C.c_int32)]


class IndfOut(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("self_energy", "real_energy", "intra_energy", "surf_energy",
                                           "self_deriv", "real_deriv", "mol_deriv", "surf_deriv", "surf_deriv2",
                                           "tot_torque")] + [("virial", C.c_double * 9)]


class PolarMdParams(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("cut_buff", C.c_double), ("vdw", VdwParams), ("mp", MpParams),
                ("mp_kcut", C.c_double), ("indf", IndfParams), ("ind_kcut", C.c_double),
                ("induce", InduceParams), ("integ", IntegParams), ("use_cells", C.c_int32), ("prec", C.c_int32),
                ("use_graph", C.c_int32), ("total_steps", C.c_int64)]


class PolarMdReport(C.Structure):
    _fields_ = [("steps", C.c_int64), ("list_updates", C.c_int64), ("cg_iterations", C.c_int64),
                ("last_iterations", C.c_int32), ("last_residual", C.c_double)] + \
               [(k, C.c_double) for k in ("e_vdw", "e_mp_real", "e_mp_self", "e_mp_intra", "e_mp_long",
                                           "e_ind_real", "e_ind_self", "e_ind_intra", "e_ind_long", "atom_ke2",
                                           "mol_ke2", "fnh")]


class BatchParams(C.Structure):
    _fields_ = [("n_replicas", C.c_int32), ("n_mols", C.c_void_p), ("box", C.c_void_p), ("shear_rate", C.c_void_p),
                ("kcut", C.c_void_p)] + \
               [(k, C.c_double) for k in ("mass_h", "mass_o", "q_h", "q_o", "sigma_o", "eps_o", "r_oh", "r_hh",
                                           "cutoff", "cut_buff", "alpha", "dt", "temperature", "coupl",
                                           "shake_tol")] + \
               [(k, C.c_int32) for k in ("max_count", "nose_hoover", "prec", "nb_cap")]


class BatchReport(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("tilt", "fnh", "e_lj", "e_real", "e_long", "e_intra", "e_self",
                                           "atom_ke2", "mol_ke2")] + \
               [("virial", C.c_double * 9), ("mol_kin", C.c_double * 9), ("mol_correction", C.c_double * 9),
                ("steps", C.c_int64), ("list_updates", C.c_int64), ("n_kvec", C.c_int32), ("max_row", C.c_int32)]


class BatchTiles(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("pair_lanes", "sfac_chunk", "kforce_tile", "overlap", "use_graph")]


_NAME80 = C.c_char * 80
_CONFIG_NAMES = ("sys_data_file", "coordinate_file", "velocity_file", "result_file", "restart_file",
                 "trajectory_file", "dump_file", "mol_trajectory_file", "mol_velocity_file", "thermo_type")
_CONFIG_I32 = ("n_steps", "start_steps", "short_ts_freq", "long_ts_freq", "avg_ts_freq", "print_ts_freq",
               "backup_ts_freq", "start_sampling", "sampling_ts_freq", "start_ttcf", "ttcf_ts_freq", "rdf_nbins",
               "integrator_type", "ensemble_status", "run_type", "constraint_mode", "max_count", "bending_type")
_CONFIG_F64 = ("temperature", "pressure", "volume", "energy", "density", "coupl_thermo", "coupl_baros", "rate",
               "cutoff", "cutoff_ew", "cutoff_ind", "k_cutoff", "alpha", "cut_buff", "cell_side_len",
               "lx_box", "ly_box", "lz_box", "cfb", "dfb", "tfb", "tolerance", "kl_ub", "l0_ub", "switch_dist",
               "a_damp", "n_mie", "m_mie", "n_buff", "m_buff", "delta_buff", "gamma_buff", "h_factor")
_CONFIG_FLAGS = ("write_trajectory", "is_new_start", "read_velocity", "use_atom_thermo", "use_cell_list",
                 "constraint_on", "harmonic_bend", "amoeba_bend", "harmonic_ub", "long_correct", "switch_on",
                 "coulomb", "ewald", "surf_correct", "wolf", "multipole", "induction", "torques",
                 "damp_induction", "induction_min", "damp_amoeba", "buff_vdw", "mie", "ext_forces", "fixed_atom",
                 "ttcf", "rdf", "msd", "vacf", "lustig", "constant_volume", "constant_pressure",
                 "constant_temperature", "constant_energy", "do_equilibrium", "do_compression", "do_shear",
                 "do_elongation")


class Config(C.Structure):
    _fields_ = ([(k, _NAME80) for k in _CONFIG_NAMES] + [("time_step", C.c_double)]
                + [(k, C.c_int32) for k in _CONFIG_I32] + [("seed", C.c_uint32)]
                + [(k, C.c_double) for k in _CONFIG_F64] + [(k, C.c_int32) for k in _CONFIG_FLAGS])


class SysdataSizes(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("n_mols", "n_mol_types", "n_atoms", "n_atom_types", "n_bonds",
                                         "n_bond_types", "n_angles", "n_angle_types", "n_dihedrals",
                                         "n_dihedral_types", "n_impropers", "n_improper_types")]


class AtomParam(C.Structure):
    _fields_ = [("type", C.c_int32), ("mass", C.c_double), ("charge", C.c_double), ("eps", C.c_double),
                ("sigma", C.c_double), ("dip", C.c_double * 3), ("quad", C.c_double * 6), ("polar", C.c_double),
                ("type_name", C.c_char * 16), ("atom_name", C.c_char * 16)]


class BondParam(C.Structure):
    _fields_ = [("type", C.c_int32), ("t1", C.c_int32), ("t2", C.c_int32), ("r0", C.c_double), ("k", C.c_double)]


class AngleParam(C.Structure):
    _fields_ = [("type", C.c_int32), ("t1", C.c_int32), ("t2", C.c_int32), ("t3", C.c_int32),
                ("theta0", C.c_double), ("k", C.c_double), ("theta0_deg", C.c_double)]


class TorsionParam(C.Structure):
    _fields_ = [("type", C.c_int32), ("t1", C.c_int32), ("t2", C.c_int32), ("t3", C.c_int32), ("t4", C.c_int32),
                ("k", C.c_double), ("d", C.c_int32), ("n", C.c_int32)]


class MolType(C.Structure):
    _fields_ = [("type", C.c_int32), ("n_atoms", C.c_int32), ("name", C.c_char * 16)]


class SysdataView(C.Structure):
    _fields_ = [("sizes", SysdataSizes), ("atom_params", C.POINTER(AtomParam)),
                ("bond_params", C.POINTER(BondParam)), ("angle_params", C.POINTER(AngleParam)),
                ("dihedral_params", C.POINTER(TorsionParam)), ("improper_params", C.POINTER(TorsionParam)),
                ("mol_types", C.POINTER(MolType))] + [
        (k, C.POINTER(C.c_int32)) for k in ("atom_type", "atom_mol", "mol_type", "bonds", "angles", "dihedrals",
                                            "impropers")]


class BondedParams(C.Structure):
    _fields_ = [("potential", C.c_int32), ("n_bond_types", C.c_int32), ("n_angle_types", C.c_int32),
                ("bond_r0", C.c_void_p), ("bond_k", C.c_void_p), ("angle_theta0", C.c_void_p),
                ("angle_k", C.c_void_p), ("n_bonds", C.c_int64), ("n_angles", C.c_int64), ("bonds", C.c_void_p),
                ("angles", C.c_void_p), ("kl_ub", C.c_double), ("l0_ub", C.c_double)]


class BondedOut(C.Structure):
    _fields_ = [("energy", C.c_double), ("deriv", C.c_double), ("deriv2", C.c_double), ("virial", C.c_double * 9)]


# name -> (restype, argtypes); every entry is a symbol declared in include/swinmd_hip.h
_P = C.c_void_p
SIGNATURES = {
    "smd_abi_version": (C.c_int, []),
    "smd_strerror": (C.c_char_p, [C.c_int]),
    "smd_ctx_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(Topology), C.c_int]),
    "smd_ctx_destroy": (None, [_P]),
    "smd_last_error": (C.c_char_p, [_P]),
    "smd_sync": (C.c_int, [_P]),
    "smd_set_box": (C.c_int, [_P, _P]),
    "smd_set_positions": (C.c_int, [_P, _P, C.c_int64]),
    "smd_get_positions": (C.c_int, [_P, _P, C.c_int64]),
    "smd_set_velocities": (C.c_int, [_P, _P, C.c_int64]),
    "smd_get_velocities": (C.c_int, [_P, _P, C.c_int64]),
    "smd_set_mol_centers": (C.c_int, [_P, _P, C.c_int64]),
    "smd_get_mol_centers": (C.c_int, [_P, _P, C.c_int64]),
    "smd_compute_mol_centers": (C.c_int, [_P]),
    "smd_zero_forces": (C.c_int, [_P]),
    "smd_get_forces": (C.c_int, [_P, _P, C.c_int64]),
    "smd_build_pairlist": (C.c_int, [_P, C.c_double, C.c_double, C.c_int]),
    "smd_pairlist_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "smd_get_pairlist": (C.c_int, [_P, _P, _P, C.c_int64]),
    "smd_set_pairlist_capacity": (C.c_int, [_P, C.c_int32]),
    "smd_vdw_compute": (C.c_int, [_P, C.POINTER(VdwParams), C.c_int, _P, C.POINTER(VdwOut)]),
    "smd_lj_long_range_correct": (C.c_int, [_P, _P, _P, C.c_double, _P]),
    "smd_ewald_setup": (C.c_int, [_P, C.POINTER(EwaldParams), C.POINTER(C.c_int32)]),
    "smd_ewald_compute": (C.c_int, [_P, C.c_int, _P, C.POINTER(EwaldOut)]),
    "smd_ewald_compute_long": (C.c_int, [_P, C.c_int, _P, C.POINTER(EwaldLongOut)]),
    "smd_ewald_compute_long_static": (C.c_int, [_P, C.c_int, C.POINTER(EwaldLongOut)]),
    "smd_set_constraints": (C.c_int, [_P, _P, C.c_int64]),
    "smd_set_vt_minus": (C.c_int, [_P, _P, C.c_int64]),
    "smd_get_vt_minus": (C.c_int, [_P, _P, C.c_int64]),
    "smd_set_mol_momenta": (C.c_int, [_P, _P, C.c_int64]),
    "smd_shake": (C.c_int, [_P, C.POINTER(IntegParams), C.c_double, _P]),
    "smd_lf_atoms": (C.c_int, [_P, C.POINTER(IntegParams), C.c_double]),
    "smd_mol_kinetics": (C.c_int, [_P, _P]),
    "smd_nh_chain": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                               _P, C.POINTER(C.c_double)]),
    "smd_md_init": (C.c_int, [_P, C.POINTER(MdParams)]),
    "smd_polar_md_init": (C.c_int, [_P, C.POINTER(PolarMdParams)]),
    "smd_polar_md_run": (C.c_int, [_P, C.c_int64, C.POINTER(PolarMdReport)]),
    "smd_polar_time_kernel": (C.c_int, [_P, C.c_char_p, C.c_int, C.POINTER(C.c_double)]),
    "smd_batch_create": (C.c_int, [C.POINTER(_P), C.POINTER(BatchParams), C.c_int]),
    "smd_batch_destroy": (None, [_P]),
    "smd_batch_last_error": (C.c_char_p, [_P]),
    "smd_batch_set_state": (C.c_int, [_P, _P, _P, _P]),
    "smd_batch_start": (C.c_int, [_P]),
    "smd_batch_run": (C.c_int, [_P, C.c_int64]),
    "smd_batch_get_state": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "smd_batch_get_reports": (C.c_int, [_P, _P]),
    "smd_batch_set_tiles": (C.c_int, [_P, C.POINTER(BatchTiles)]),
    "smd_batch_get_tiles": (C.c_int, [_P, C.POINTER(BatchTiles)]),
    "smd_batch_autotune": (C.c_int, [_P, C.c_int, _P, C.POINTER(C.c_int32)]),
    "smd_batch_time_kernel": (C.c_int, [_P, C.c_char_p, C.c_int, C.POINTER(C.c_double)]),
    "smd_md_run": (C.c_int, [_P, C.c_int64, C.POINTER(MdReport)]),
    "smd_md_pressure": (C.c_int, [_P, C.POINTER(MdPressure)]),
    "smd_set_real_positions": (C.c_int, [_P, _P, C.c_int64]),
    "smd_get_real_positions": (C.c_int, [_P, _P, C.c_int64]),
    "smd_md_get_thermostat": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "smd_md_set_thermostat": (C.c_int, [_P, C.c_double, C.c_double]),
    "smd_write_state": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, _P, _P, _P, C.c_int32, C.c_char_p, C.c_size_t]),
    "smd_write_vt_minus": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, _P, C.c_int32, C.c_char_p, C.c_size_t]),
    "smd_write_press_res": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, _P, _P, C.c_char_p, C.c_size_t]),
    "smd_write_pressac": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, _P, _P, C.c_int64, C.c_char_p, C.c_size_t]),
    "smd_write_pressure_tensor": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, _P, _P, C.c_double, C.c_double,
                                            C.c_int64, C.c_char_p, C.c_size_t]),
    "smd_profile_enable": (C.c_int, [_P, C.c_int]),
    "smd_profile_get": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "smd_md_time_family": (C.c_int, [_P, C.c_char_p, C.c_int, C.POINTER(C.c_double)]),
    "smd_set_tuning": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "smd_set_multipoles": (C.c_int, [_P, _P, _P, _P, C.c_int64]),
    "smd_induce": (C.c_int, [_P, C.POINTER(InduceParams), C.POINTER(InduceReport)]),
    "smd_get_induction": (C.c_int, [_P, _P, _P, _P, _P, C.c_int64]),
    "smd_induce_time_apply": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    "smd_mlp_init": (C.c_int, [_P, C.c_uint64]),
    "smd_mlp_set_params": (C.c_int, [_P, _P, C.c_int64]),
    "smd_mlp_get_params": (C.c_int, [_P, _P, C.c_int64]),
    "smd_mlp_set_norm": (C.c_int, [_P, _P, C.c_int64]),
    "smd_mlp_get_norm": (C.c_int, [_P, _P, C.c_int64]),
    "smd_mlp_forward": (C.c_int, [_P, _P, C.c_int64, _P]),
    "smd_mlp_loss_grad": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(C.c_double), _P]),
    "smd_mlp_descriptors": (C.c_int, [_P, _P, C.c_int64]),
    "smd_mlp_add_frame": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "smd_mlp_clear_frames": (C.c_int, [_P]),
    "smd_mlp_train": (C.c_int, [_P, C.POINTER(MlpTrainParams), C.POINTER(MlpTrainReport)]),
    "smd_multipole_setup": (C.c_int, [_P, C.POINTER(MpParams), C.c_double, C.POINTER(C.c_int32)]),
    "smd_multipole_compute": (C.c_int, [_P, C.c_int, _P, C.POINTER(MpOut)]),
    "smd_multipole_compute_long": (C.c_int, [_P, C.c_int, _P, C.POINTER(MpLongOut)]),
    "smd_induction_setup": (C.c_int, [_P, C.POINTER(IndfParams), C.c_double, C.POINTER(C.c_int32)]),
    "smd_induction_compute": (C.c_int, [_P, C.c_int, _P, C.POINTER(IndfOut)]),
    "smd_induction_compute_long": (C.c_int, [_P, C.c_int, _P, C.POINTER(MpLongOut)]),
    "smd_bonded_setup": (C.c_int, [_P, C.POINTER(BondedParams)]),
    "smd_bond_compute": (C.c_int, [_P, _P, C.POINTER(BondedOut)]),
    "smd_angle_compute": (C.c_int, [_P, _P, C.POINTER(BondedOut)]),
    "smd_ub_compute": (C.c_int, [_P, _P, C.POINTER(BondedOut)]),
    "smd_config_read": (C.c_int, [C.c_char_p, C.POINTER(Config), C.c_char_p, C.c_size_t]),
    "smd_sysdata_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]),
    "smd_sysdata_view_get": (C.c_int, [_P, C.POINTER(SysdataView)]),
    "smd_sysdata_free": (None, [_P]),
    "smd_read_xyz": (C.c_int, [C.c_char_p, C.c_int64, C.c_int32, _P, C.c_char_p, C.c_size_t]),
    "smd_read_state": (C.c_int, [C.c_char_p, C.c_int64, _P, _P, _P, C.c_char_p, C.c_size_t]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing; build it with `python -m swinmd_amd.build` "
                          "(the package has no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code
