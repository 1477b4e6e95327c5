"""Input readers (config.txt, sysData.in, coordinate/velocity/state files)
against what the compiled reference's own readers make of the reference's
example inputs (test/config.txt, test/sysData.in, test/coordinate.in, kept as
data under tests/golden/ref_test_inputs/; the expected values are in
tests/golden/ref_test_inputs_parsed.npz, written by the reference harness's
"inputs" mode).  CPU only: the readers are host code."""
import json
import os

import numpy as np
import pytest

from tests.helpers import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")
INPUTS = os.path.join(GOLD, "ref_test_inputs")


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLD, "ref_test_inputs_parsed.npz"))


@pytest.fixture(scope="module")
def loaded():
    from swinmd_amd import inputs
    return inputs.load_inputs(INPUTS)


# the order of the harness's cfg_num record (oracle/ref_harness.cpp, "inputs")
CFG_NUM = ["get_timestep", "get_n_steps", "get_start_steps", "get_short_ts_freq", "get_long_ts_freq",
           "get_avg_ts_freq", "get_print_ts_freq", "get_backup_ts_freq", "get_start_sampling",
           "get_sampling_ts_freq", "get_start_TTCF", "get_ttcf_ts_freq", "get_integrator_type",
           "get_ensemble_status", "get_rdf_nbins", "get_seed", "get_cutoff", "get_cutoffEw", "get_cutoffInd",
           "get_adamping", "get_cutBuff", "get_k_cutoff", "get_alpha", "get_cell_side_len", "get_temperature",
           "get_coupl_const", "get_pressure", "get_baros_const", "get_tolerance", "get_max_count",
           "get_switch_distance", "get_kl_ub", "get_l0_ub", "get_nMie_param", "get_mMie_param", "get_nBuff_param",
           "get_mBuff_param", "get_deltaBuff_param", "get_gammaBuff_param", "get_hfactor_reduction",
           "is_read_velocity", "is_new_start", "is_constraint_on", "is_switch_on", "write_trajectory",
           "use_long_correct", "use_cell_list", "use_TTCF", "use_harmonic_potential", "use_amoeba_bending",
           "use_atom_thermo", "compute_coulomb", "compute_mie", "compute_buffamoeba", "compute_harmonic_ub",
           "compute_ewald_force", "compute_surf_correct", "compute_wolf", "compute_multipol", "compute_induction",
           "compute_torques", "use_damping_induction", "do_induction_minimization", "use_damping_indAmoeba",
           "has_ext_forces", "has_fixed_atom", "compute_msd", "compute_rdf", "compute_lustig", "compute_vacf"]


def test_config_matches_the_reference_getters(ref, loaded):
    cfg = loaded[0]
    want = ref["ref_cfg_num"]
    assert len(want) == len(CFG_NUM)
    got = np.array([float(getattr(cfg, g)()) for g in CFG_NUM])
    bad = [(g, w, v) for g, w, v in zip(CFG_NUM, want, got) if w != v]
    assert not bad, bad
    names = dict(line.split(" ", 1) for line in ref["ref_names"])
    assert cfg.get_sysData_file() == names["sysData"]
    assert cfg.get_coordinate_file() == names["coordinate"]
    assert cfg.get_velocity_file() == names["velocity"]
    assert cfg.get_result_file() == names["result"]
    assert cfg.get_restart_file() == names["restart"]
    assert cfg.get_dump_file() == names["dump"]
    assert cfg.get_thermo_type() == names["thermo"]
    assert cfg.get_mol_traject_file() == names["moltraj"]
    assert cfg.get_mol_velocity_file() == names["molvel"]
    # the second kCutoff on the Ewald line is the one kept
    assert cfg.get_k_cutoff() == 18.6


def test_sysdata_matches_the_reference(ref, loaded):
    p = loaded[1]
    sizes = [p.nMolecules, p.nMolTypes, p.nAtoms, p.nAtomTypes, p.nBonds, p.nBondTypes, p.nAngles, p.nAngleTypes,
             p.nDihedrals, p.nDihedralTypes, p.nImpropers, p.nImproperTypes]
    assert sizes == ref["ref_i_sizes"].tolist()
    ap = ref["ref_atom_params"].reshape(p.nAtomTypes, 15)
    for t, a in enumerate(p.atomParams):
        row = [a["atomType"], a["mass"], a["charge"], a["eps"], a["sigma"], *a["dip"], *a["quad"], a["polar"]]
        assert row == ap[t].tolist()
    bp = ref["ref_bond_params"].reshape(p.nBondTypes, 5)
    for t, b in enumerate(p.bondParams):
        assert [b["bondType"], b["atomType1"], b["atomType2"], b["r0"], b["k"]] == bp[t].tolist()
    gp = ref["ref_angle_params"].reshape(p.nAngleTypes, 6)
    for t, a in enumerate(p.angleParams):
        # theta0 in radians, bit for bit (tmp*2*PI/360.0 in the reference)
        assert [a["angleType"], a["atomType1"], a["atomType2"], a["atomType3"], a["theta0"], a["k"]] == gp[t].tolist()
    mt = ref["ref_i_mol_types"].reshape(p.nMolTypes, 2)
    assert [[m["molType"], m["numOfAtoms"]] for m in p.molTypes] == mt.tolist()
    at = ref["ref_i_atoms"].reshape(p.nAtoms, 2)
    np.testing.assert_array_equal(p.atomType, at[:, 0])
    np.testing.assert_array_equal(p.atomMol, at[:, 1])
    np.testing.assert_array_equal(p.molType, ref["ref_i_mols"])
    np.testing.assert_array_equal(p.bondArray, ref["ref_i_bonds"].reshape(p.nBonds, 3))
    np.testing.assert_array_equal(p.angleArray, ref["ref_i_angles"].reshape(p.nAngles, 4))


def test_system_from_inputs_matches_the_reference_ensemble(ref, loaded):
    cfg, p, s, vel = loaded
    meta = json.loads(str(ref["meta"]))
    assert s.box[0] == meta["box"] and s.volume == meta["volume"]
    np.testing.assert_array_equal(s.pos.ravel(), ref["ref_init_pos"])
    mq = ref["ref_mass_scaled_charge"].reshape(s.n_atoms, 2)
    np.testing.assert_array_equal(s.mass, mq[:, 0])
    np.testing.assert_array_equal(s.scaled_charge, mq[:, 1])
    off, ex = s.exclusions()
    np.testing.assert_array_equal(off, ref["ref_i_excl_off"])
    np.testing.assert_array_equal(ex, ref["ref_i_excl"])
    # isNewStart with readVelocity off: Ensemble::init_velocity with the seed,
    # temperature and ndf = 3N - 3 (unconstrained, atom thermostat)
    np.testing.assert_allclose(vel.ravel(), ref["ref_init_vel"], rtol=1e-13, atol=1e-15)


def test_reader_errors_name_the_file_and_line(tmp_path):
    from swinmd_amd import _lib, inputs
    with pytest.raises(inputs.InputError) as e:
        inputs.SimConfiguration(tmp_path / "missing.txt").read_config_file()
    assert e.value.code == _lib.SMD_EIO and "can't open" in str(e.value)
    lines = open(os.path.join(INPUTS, "config.txt")).read().splitlines()
    lines[4] = "0.001\tmany\t0\t// timeStep; numSteps; startSteps"
    bad = tmp_path / "config.txt"
    bad.write_text("\n".join(lines) + "\n")
    with pytest.raises(inputs.InputError) as e:
        inputs.SimConfiguration(bad).read_config_file()
    assert ":5:" in str(e.value) and "nSteps" in str(e.value)
    # a heading out of place in sysData.in
    text = open(os.path.join(INPUTS, "sysData.in")).read().replace("!BondParams", "!BondParameters", 1)
    (tmp_path / "sys.in").write_text(text)
    with pytest.raises(inputs.InputError) as e:
        inputs.Parameters(tmp_path / "sys.in")
    assert "!BondParams" in str(e.value)
    # the coordinate count must equal the number of atoms (read_coordinate)
    (tmp_path / "c.in").write_text("2\n0 0 0\n1 1 1\n")
    with pytest.raises(inputs.InputError) as e:
        inputs.read_coordinate(tmp_path / "c.in", 3)
    assert "count line" in str(e.value)
    (tmp_path / "c2.in").write_text("2\n0 0 0\n1 1\n")
    with pytest.raises(inputs.InputError) as e:
        inputs.read_coordinate(tmp_path / "c2.in", 2)
    assert ":3:" in str(e.value)


def test_istream_integer_extraction_stops_at_the_first_non_digit(tmp_path):
    """'>> int' on "20.5" reads 20 and leaves ".5" for the next extraction,
    as the reference's stream reads do."""
    from swinmd_amd import inputs
    lines = open(os.path.join(INPUTS, "config.txt")).read().splitlines()
    lines[4] = "0.001\t20.5\t// timeStep; numSteps; startSteps (startSteps reads .5 -> 0, as an int fails)"
    p = tmp_path / "config.txt"
    p.write_text("\n".join(lines) + "\n")
    with pytest.raises(inputs.InputError):
        inputs.SimConfiguration(p).read_config_file()   # ".5" is not an int: the reference would read 0
    lines[4] = "0.001\t20\t3\t// timeStep; numSteps; startSteps"
    lines[12] = "1.10\t1.10\t14\t.1\t// cutoff; cutoffEw; kCutoff; cutBuffer"
    p.write_text("\n".join(lines) + "\n")
    cfg = inputs.SimConfiguration(p).read_config_file()
    assert cfg.get_n_steps() == 20 and cfg.get_start_steps() == 3 and cfg.get_cutBuff() == 0.1


def test_state_file_round_trip(tmp_path):
    """state.in: three coordinate blocks, each after one skipped line
    (src/Ensemble.cpp:299-317)."""
    from swinmd_amd import inputs
    rng = np.random.default_rng(4)
    blocks = [rng.normal(size=(5, 3)) for _ in range(3)]
    with open(tmp_path / "state.in", "w") as f:
        for title, b in zip(("coordinates", "real positions", "velocities"), blocks):
            f.write(f"# {title}\n{len(b)}\n")
            for x, y, z in b:
                f.write(f"{float(x)!r} {float(y)!r} {float(z)!r}\n")
    pos, real, vel = inputs.read_state(tmp_path / "state.in", 5)
    for got, want in zip((pos, real, vel), blocks):
        np.testing.assert_array_equal(got, want)
