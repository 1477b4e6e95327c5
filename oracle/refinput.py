"""Writers for the SwinMD reference's input formats (config.txt, sysData.in,
coordinate.in).  TEST INFRASTRUCTURE: used to feed oracle/_ref/ref_harness
when generating golden vectors and for the CPU baseline.

Layouts follow SimConfiguration::read_config_file
(src/SimConfiguration.cpp:695-932) and Parameters::read_sys_data
(src/Parameters.cpp:76-382); the coordinate file is read by
Ensemble::read_coordinate(FILE*) (src/Ensemble.cpp:630-673).
"""
from __future__ import annotations

import os


def _b(x):
    return "t" if x else "f"


def _g(x):
    return repr(float(x))


def write_sysdata(path, s):
    n_at, n_mol = s.n_atoms, s.n_mols
    lines = ["!sysData file", "!Size_of_system"]
    lines.append(f"{n_mol}\t1\t! number of molecules and number of molecule types")
    lines.append(f"{n_at}\t{len(s.types)}\t! number of atoms and number of atom types")
    lines.append(f"{len(s.bonds)}\t{len(s.bond_params)}\t! bonds")
    lines.append(f"{len(s.angles)}\t{len(s.angle_params)}\t! angles")
    lines.append("0\t0\t! dihedrals")
    lines.append("0\t0\t! impropers")
    lines.append("")
    lines.append("!AtomParams")
    lines.append("!typeID Mass Charge Epsilon Sigma dipx dipy dipz quadxx quadxy quadxz quadyy quadyz quadzz polar typeName atomName")
    for i, t in enumerate(s.types):
        vals = [t.mass, t.charge, t.eps, t.sigma, *t.dip, *t.quad, t.polar]
        lines.append(f"{i}\t" + "\t".join(_g(v) for v in vals) + f"\t{t.type_name}\t{t.atom_name}")
    lines.append("")
    lines.append("!BondParams")
    lines.append("!bondTypeID atomType1 atomType2 length forceConstant")
    for i, (t1, t2, r0, k) in enumerate(s.bond_params):
        lines.append(f"{i}\t{t1}\t{t2}\t{_g(r0)}\t{_g(k)}")
    lines.append("")
    lines.append("!AngleParams")
    lines.append("!angTypeID aType1 aType2 aType3 degree K")
    for i, (t1, t2, t3, deg, k) in enumerate(s.angle_params):
        lines.append(f"{i}\t{t1}\t{t2}\t{t3}\t{_g(deg)}\t{_g(k)}")
    lines.append("")
    lines.append("!DihedralParams")
    lines.append("!TypeID aType1 aType2 aType3 aType4 K d n")
    lines.append("")
    lines.append("!ImproperParams")
    lines.append("!TypeID aType1 aType2 aType3 aType4 K d n")
    lines.append("")
    lines.append("!MolTypes")
    lines.append("!molTypeID numOfAtoms molNames")
    lines.append(f"0\t{s.atoms_per_mol}\t{s.mol_type_name}")
    lines.append("")
    lines.append("!Atoms")
    lines.append("!atomID atomType molID molType")
    for i in range(n_at):
        lines.append(f"{i}\t{int(s.type[i])}\t{int(s.mol[i])}\t0")
    lines.append("")
    lines.append("!Bonds")
    lines.append("!bondID atom1 atom2 bondType")
    for i, (a1, a2, bt) in enumerate(s.bonds):
        lines.append(f"{i}\t{a1}\t{a2}\t{bt}")
    lines.append("")
    lines.append("!Angles")
    lines.append("!angleID atom1 atom2 atom3 angleType")
    for i, (a1, a2, a3, at) in enumerate(s.angles):
        lines.append(f"{i}\t{a1}\t{a2}\t{a3}\t{at}")
    lines.append("")
    lines.append("!Dihedrals")
    lines.append("!dihedralID atom1 atom2 atom3 atom4 dihedralType")
    lines.append("")
    lines.append("!Impropers")
    lines.append("!ImproperID atom1 atom2 atom3 atom4 dihedralType")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_coordinates(path, pos):
    with open(path, "w") as f:
        f.write(f"{len(pos)}\n")
        for p in pos:
            f.write(f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n")


def write_config(path, s, *, cutoff, cut_buff, alpha=3.5, kcut=10.0, dt=0.002, nsteps=1,
                 long_steps=1, temperature=300.0, atom_thermo=True, coupl=0.1, thermostat="NoseHoover",
                 seed=1, use_cells=True, constraint=True, shake_tol=1e-8, max_count=10000,
                 lcorr=False, switch_on=False, switch_dist=0.0, coulomb=False, ewald=True,
                 surface=False, buff=False, n_buff=14, m_buff=7, delta=0.07, gamma=0.12,
                 hfactor=0.91, mie=False, n_mie=12, m_mie=6, lustig=False, cutoff_ew=None,
                 induction=False, ind_damp=False, ind_min=False, cutoff_ind=None, a_damp=0.39,
                 multipole=False, torques=False, damp_amoeba=False):
    cutoff_ew = cutoff if cutoff_ew is None else cutoff_ew
    cutoff_ind = cutoff if cutoff_ind is None else cutoff_ind
    L = [
        "sysData.in\tcoordinate.in\t// input file names",
        "result.out\trestart.out\tdump.out\t// output file names",
        "f\ttrajectory.out\t// writeTrajectory",
        f"t\t{seed}\tf\tvelocity.in\t// isNewStart; seed; readVelocity; velocityFile",
        f"{_g(dt)}\t{nsteps}\t0\t// timeStep; numSteps; startSteps",
        f"1\t{long_steps}\t// nonbondTsFrequency; longElectTsFrequency",
        "10\t10\t10\t// computeAvgTsFreq; printTsFreq; backupTsFreq",
        "2\t// integrator type: leap frog",
        f"0\t{_g(temperature)}\t0.101\t{1 if atom_thermo else 2}\t{_g(coupl)}\t{thermostat}\t0.5\t// NVT",
        "0\t0.0\t// testMode; rate",
        f"{_g(cutoff)}\t{_g(cutoff_ew)}\t{_g(kcut)}\t{_g(cut_buff)}\t// cutoff; cutoffEw; kCutoff; cutBuffer",
        f"{_g(s.density)}\t// density",
        "1.0\t1.0\t1.0\t// box (overridden by density)",
        f"{_b(use_cells)}\t1.5\t// useCellPairList; cellSideLen",
        (f"2\t{_g(shake_tol)}\t{max_count}\t// SHAKE" if constraint else "0\t0.00001\t1\t// no constraint"),
        "f\t0.0\t0.0\t// computeharmoUB",
        f"{_b(lcorr)}\t{_b(switch_on)}\t{_g(switch_dist)}\t// LRC; switchOn; switchDistance",
        f"{_b(coulomb)}\t{_b(ewald)}\t{_b(surface)}\tf\t{_g(alpha)}\t{_g(kcut)}\t// coulomb ewald surface wolf alpha kcut",
        (f"{_b(multipole)}\t{_b(induction)}\t{_b(torques)}\t{_b(ind_damp)}\t{_b(ind_min)}\t{_b(damp_amoeba)}"
         f"\t{_g(cutoff_ind)}\t{_g(a_damp)}"
         "\t// multipole; induction; torques; dampInduction; inductionMin; dampAmoeba; cutoffInd; a"),
        f"{_b(buff)}\t{_g(n_buff)}\t{_g(m_buff)}\t{_g(delta)}\t{_g(gamma)}\t{_g(hfactor)}\t// buffered 14-7",
        f"{_b(mie)}\t{_g(n_mie)}\t{_g(m_mie)}\t// Mie",
        "f\tf\t// hasExtForces; hasFixedAtom",
        "30000\t20\t// sampling",
        "f\t0\t0\t// TTCF",
        "f\t400\t// RDF",
        "f\tmsd.xyz\t// MSD",
        "f\tvacf.out\t// VACF",
        f"{_b(lustig)}\t// Lustig averages",
    ]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_inputs(d, s, pos=None, **cfg):
    os.makedirs(d, exist_ok=True)
    write_sysdata(os.path.join(d, "sysData.in"), s)
    write_coordinates(os.path.join(d, "coordinate.in"), s.pos if pos is None else pos)
    write_config(os.path.join(d, "config.txt"), s, **cfg)
