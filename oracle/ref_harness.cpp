// ref_harness.cpp -- drives the compiled SwinMD reference (objects built from
// the reference's own sources by oracle/Makefile.ref) to produce golden
// vectors and CPU-baseline timings.  TEST INFRASTRUCTURE ONLY: it is run by
// tests/golden/make_golden.py and by bench.py's cpu_baseline leg, never by the
// product path.
//
// Usage:  ref_harness <dir> forces <force-list>   (lj,mie,buck,buff,ewald)
//         ref_harness <dir> induction
//         ref_harness <dir> multipole
//         ref_harness <dir> induction_forces
//         ref_harness <dir> pairlist
//         ref_harness <dir> inputs
//         ref_harness <dir> md <nsteps>
//         ref_harness <dir> time <nsteps>
//         ref_harness <dir> time_polar <reps>
//         ref_harness <dir> lf_forces
// <dir> holds config.txt, sysData.in and coordinate.in in the reference's
// formats; outputs (*.bin raw little-endian arrays, meta.txt) are written there.
//
// Reference defects are neutralised so that outputs are well defined:
//  * Heap objects are read before they are written in several places:
//    Parameters::read_sys_data never stores the atom types it reads
//    (src/Parameters.cpp:112-134) and then indexes the LJ table with them
//    (:371-380); LJForce reads eLrc/vLrc/dUlrc/d2Ulrc when the long-range
//    correction is off (src/forces/LJForce.cpp:232-247).  The harness replaces
//    the global operator new with a zero-filling one, so these reads see 0.
//  * The AtomParams section is then parsed here and the LJ table rebuilt with
//    Parameters::build_lj_param_table().
#include <mpi.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <iostream>
#include <list>
#include <map>
#include <set>

// LFIntegrator.h pulls in every force header in the order the reference's own
// translation units see them (several headers define clashing macros).  The
// induction mode reads Induction's private state (Pi, Ei0, the Local2Global
// frames); access specifiers do not change the class layout, so the
// reference's headers are read here with every member public.
#define private public
#define protected public
#include "integrators/LFIntegrator.h"
#undef private
#undef protected
// forces/NccForce.h defines a macro named like AtomParam's polarizability
// member; the harness reads that member, so the macro is dropped here.
#undef polar

// Zero-filled storage for every new-expression in the process.
void* operator new(std::size_t n) {
    void* p = std::calloc(1, n ? n : 1);
    if (!p) throw std::bad_alloc();
    return p;
}
void* operator new[](std::size_t n) { return ::operator new(n); }
void operator delete(void* p) noexcept { std::free(p); }
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete(void* p, std::size_t) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

// The reference defines these two globals next to its main() (NEMDMain.cpp),
// which the harness replaces.
int mpi_size_g;
int mpi_rank_g;

namespace {

std::string g_dir;

void write_bin(const std::string& name, const void* data, size_t bytes) {
    std::string path = g_dir + "/" + name + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    if (bytes) std::fwrite(data, 1, bytes, f);
    std::fclose(f);
}

struct Meta {
    std::ostringstream os;
    void put(const std::string& k, double v) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.17g", v);
        os << k << " " << buf << "\n";
    }
    void save() {
        std::ofstream f((g_dir + "/meta.txt").c_str());
        f << os.str();
    }
};

// Parse the AtomParams section the way the reader intended and store it.
void repair_atom_params(Parameters* p, const char* sysdata) {
    std::ifstream in(sysdata);
    std::string line;
    while (std::getline(in, line))
        if (line.compare(0, 11, "!AtomParams") == 0) break;
    std::getline(in, line);  // column comment
    for (unsigned i = 0; i < p->nAtomTypes; ++i) {
        AtomParam ap;
        std::memset(&ap, 0, sizeof ap);
        std::string tname, aname;
        in >> ap.atomType >> ap.mass >> ap.charge >> ap.eps >> ap.sigma >> ap.dipx >> ap.dipy >> ap.dipz >>
            ap.quadxx >> ap.quadxy >> ap.quadxz >> ap.quadyy >> ap.quadyz >> ap.quadzz >> ap.polar >> tname >>
            aname;
        std::strncpy(ap.typeName, tname.c_str(), sizeof ap.typeName - 1);
        std::strncpy(ap.atomName, aname.c_str(), sizeof ap.atomName - 1);
        std::getline(in, line);
        p->atomParams[ap.atomType] = ap;
    }
    p->build_lj_param_table();
}

template <class T>
T* make_zeroed(Ensemble* e) {
    return new T(e);   // storage is zero-filled by the operator new above
}

std::vector<double> forces_of(Ensemble* e) {
    std::vector<double> f(3 * e->nAtoms);
    for (int i = 0; i < e->nAtoms; ++i) {
        f[3 * i] = e->atoms[i].force.x;
        f[3 * i + 1] = e->atoms[i].force.y;
        f[3 * i + 2] = e->atoms[i].force.z;
    }
    return f;
}

void zero_forces(Ensemble* e) {
    for (int i = 0; i < e->nAtoms; ++i) e->atoms[i].force = 0.0;
    for (int k = 0; k < 9; ++k) e->virial[k] = 0.0, e->longVirial[k] = 0.0;
}

void dump_state(Ensemble* e, const std::string& tag) {
    std::vector<double> p(3 * e->nAtoms), v(3 * e->nAtoms), d(3 * e->nAtoms);
    for (int i = 0; i < e->nAtoms; ++i) {
        const Atom& a = e->atoms[i];
        p[3 * i] = a.position.x; p[3 * i + 1] = a.position.y; p[3 * i + 2] = a.position.z;
        v[3 * i] = a.velocity.x; v[3 * i + 1] = a.velocity.y; v[3 * i + 2] = a.velocity.z;
        d[3 * i] = a.displacement.x; d[3 * i + 1] = a.displacement.y; d[3 * i + 2] = a.displacement.z;
    }
    write_bin(tag + "pos", p.data(), p.size() * 8);
    write_bin(tag + "vel", v.data(), v.size() * 8);
    write_bin(tag + "disp", d.data(), d.size() * 8);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <dir> forces|pairlist|md|time [...]\n", argv[0]);
        return 2;
    }
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &mpi_rank_g);
    MPI_Comm_size(MPI_COMM_WORLD, &mpi_size_g);
    g_dir = argv[1];
    const std::string mode = argv[2];
    if (chdir(g_dir.c_str()) != 0) { std::perror("chdir"); return 2; }

    SimConfiguration* cfg = new SimConfiguration("config.txt");
    cfg->read_config_file();
    Parameters* params = new Parameters(cfg->get_sysData_file());
    repair_atom_params(params, cfg->get_sysData_file());

    Ensemble* ens = new Ensemble();
    ens->set_mol_system(params);
    ens->set_configuration(cfg);
    ens->compute_bound_box();
    ens->set_pairlist();
    std::ofstream* sysout = new std::ofstream("sysData.out");
    ens->set_sysdata_file(sysout);
    ens->set_result_file(sysout);

    Meta meta;
    meta.put("box", ens->boxLx);
    meta.put("volume", ens->volume);
    meta.put("used_cells", ens->useCellPairList ? 1 : 0);
    {
        std::vector<double> v(3 * ens->nAtoms);
        for (int i = 0; i < ens->nAtoms; ++i) {
            v[3 * i] = ens->atoms[i].velocity.x;
            v[3 * i + 1] = ens->atoms[i].velocity.y;
            v[3 * i + 2] = ens->atoms[i].velocity.z;
        }
        write_bin("init_vel", v.data(), v.size() * 8);
        for (int i = 0; i < ens->nAtoms; ++i) {
            v[3 * i] = ens->atoms[i].position.x;
            v[3 * i + 1] = ens->atoms[i].position.y;
            v[3 * i + 2] = ens->atoms[i].position.z;
        }
        write_bin("init_pos", v.data(), v.size() * 8);
        meta.put("alpha", cfg->get_alpha());
        meta.put("kcut", cfg->get_k_cutoff());
        meta.put("cutoff", cfg->get_cutoff());
        meta.put("cut_buff", cfg->get_cutBuff());
        meta.put("surface", cfg->compute_surf_correct() ? 1 : 0);
    }

    if (mode == "inputs") {
        // What the reference's readers made of the files: SimConfiguration
        // (getters, src/SimConfiguration.h:209-317), Parameters (sizes,
        // tables and topology, src/Parameters.h:100-190) and the Ensemble's
        // per-atom data and exclusion table (src/Ensemble.cpp:140-367).
        SimConfiguration* c = cfg;
        const double num[] = {
            c->get_timestep(), (double)c->get_n_steps(), (double)c->get_start_steps(), (double)c->get_short_ts_freq(),
            (double)c->get_long_ts_freq(), (double)c->get_avg_ts_freq(), (double)c->get_print_ts_freq(),
            (double)c->get_backup_ts_freq(), (double)c->get_start_sampling(), (double)c->get_sampling_ts_freq(),
            (double)c->get_start_TTCF(), (double)c->get_ttcf_ts_freq(), (double)c->get_integrator_type(),
            (double)c->get_ensemble_status(), (double)c->get_rdf_nbins(), (double)c->get_seed(), c->get_cutoff(),
            c->get_cutoffEw(), c->get_cutoffInd(), c->get_adamping(), c->get_cutBuff(), c->get_k_cutoff(),
            c->get_alpha(), c->get_cell_side_len(), c->get_temperature(), c->get_coupl_const(), c->get_pressure(),
            c->get_baros_const(), c->get_tolerance(), c->get_max_count(), c->get_switch_distance(), c->get_kl_ub(),
            c->get_l0_ub(), c->get_nMie_param(), c->get_mMie_param(), c->get_nBuff_param(), c->get_mBuff_param(),
            c->get_deltaBuff_param(), c->get_gammaBuff_param(), c->get_hfactor_reduction(),
            (double)c->is_read_velocity(), (double)c->is_new_start(), (double)c->is_constraint_on(),
            (double)c->is_switch_on(), (double)c->write_trajectory(), (double)c->use_long_correct(),
            (double)c->use_cell_list(), (double)c->use_TTCF(), (double)c->use_harmonic_potential(),
            (double)c->use_amoeba_bending(), (double)c->use_atom_thermo(), (double)c->compute_coulomb(),
            (double)c->compute_mie(), (double)c->compute_buffamoeba(), (double)c->compute_harmonic_ub(),
            (double)c->compute_ewald_force(), (double)c->compute_surf_correct(), (double)c->compute_wolf(),
            (double)c->compute_multipol(), (double)c->compute_induction(), (double)c->compute_torques(),
            (double)c->use_damping_induction(), (double)c->do_induction_minimization(),
            (double)c->use_damping_indAmoeba(), (double)c->has_ext_forces(), (double)c->has_fixed_atom(),
            (double)c->compute_msd(), (double)c->compute_rdf(), (double)c->compute_lustig(), (double)c->compute_vacf()};
        write_bin("cfg_num", num, sizeof num);
        {
            std::ofstream f("names.txt");
            f << "sysData " << c->get_sysData_file() << "\ncoordinate " << c->get_coordinate_file()
              << "\nvelocity " << c->get_velocity_file() << "\nresult " << c->get_result_file() << "\nrestart "
              << c->get_restart_file() << "\ndump " << c->get_dump_file() << "\nthermo " << c->get_thermo_type()
              << "\nmoltraj " << c->get_mol_traject_file() << "\nmolvel " << c->get_mol_velocity_file() << "\n";
        }
        Parameters* P = params;
        const int sizes[] = {(int)P->nMolecules, (int)P->nMolTypes, (int)P->nAtoms, (int)P->nAtomTypes,
                             (int)P->nBonds, (int)P->nBondTypes, (int)P->nAngles, (int)P->nAngleTypes,
                             (int)P->nDihedrals, (int)P->nDihedralTypes, (int)P->nImpropers, (int)P->nImproperTypes};
        write_bin("i_sizes", sizes, sizeof sizes);
        std::vector<double> ap;
        for (unsigned t = 0; t < P->nAtomTypes; ++t) {
            const AtomParam& a = P->atomParams[t];
            for (double v : {(double)a.atomType, a.mass, a.charge, a.eps, a.sigma, a.dipx, a.dipy, a.dipz, a.quadxx,
                             a.quadxy, a.quadxz, a.quadyy, a.quadyz, a.quadzz, a.polar})
                ap.push_back(v);
        }
        write_bin("atom_params", ap.data(), ap.size() * 8);
        std::vector<double> bp, gp;
        for (unsigned t = 0; t < P->nBondTypes; ++t) {
            const BondParam& b = P->bondParams[t];
            for (double v : {(double)b.bondType, (double)b.atomType1, (double)b.atomType2, b.r0, b.k}) bp.push_back(v);
        }
        for (unsigned t = 0; t < P->nAngleTypes; ++t) {
            const AngleParam& a = P->angleParams[t];
            for (double v : {(double)a.angleType, (double)a.atomType1, (double)a.atomType2, (double)a.atomType3,
                             a.theta0, a.k})
                gp.push_back(v);
        }
        write_bin("bond_params", bp.data(), bp.size() * 8);
        write_bin("angle_params", gp.data(), gp.size() * 8);
        std::vector<int> topo, mols, bonds, angles, mtypes;
        for (unsigned t = 0; t < P->nMolTypes; ++t) {
            mtypes.push_back(P->molTypes[t].molType);
            mtypes.push_back(P->molTypes[t].numOfAtoms);
        }
        for (unsigned i = 0; i < P->nAtoms; ++i) {
            topo.push_back(P->atomArray[i].atomType);
            topo.push_back((int)P->atomArray[i].molID);
        }
        for (unsigned m = 0; m < P->nMolecules; ++m) mols.push_back(P->molArray[m].molType);
        for (unsigned b = 0; b < P->nBonds; ++b)
            for (int v : {P->bondArray[b].atom1, P->bondArray[b].atom2, P->bondArray[b].bondType}) bonds.push_back(v);
        for (unsigned a = 0; a < P->nAngles; ++a)
            for (int v : {P->angleArray[a].atom1, P->angleArray[a].atom2, P->angleArray[a].atom3,
                          P->angleArray[a].angleType})
                angles.push_back(v);
        write_bin("i_mol_types", mtypes.data(), mtypes.size() * 4);
        write_bin("i_atoms", topo.data(), topo.size() * 4);
        write_bin("i_mols", mols.data(), mols.size() * 4);
        write_bin("i_bonds", bonds.data(), bonds.size() * 4);
        write_bin("i_angles", angles.data(), angles.size() * 4);
        std::vector<double> mq(2 * ens->nAtoms);
        for (int i = 0; i < ens->nAtoms; ++i) {
            mq[2 * i] = ens->atoms[i].mass;
            mq[2 * i + 1] = ens->atoms[i].scaledCharge;
        }
        write_bin("mass_scaled_charge", mq.data(), mq.size() * 8);
        std::vector<int> eoff(ens->nAtoms + 1, 0), excl;
        for (int i = 0; i < ens->nAtoms; ++i) {
            for (int k = 0; ens->exclusionTable[i][k] != -1; ++k) excl.push_back(ens->exclusionTable[i][k]);
            eoff[i + 1] = (int)excl.size();
        }
        write_bin("i_excl_off", eoff.data(), eoff.size() * 4);
        write_bin("i_excl", excl.data(), excl.size() * 4);
    } else if (mode == "pairlist") {
        std::vector<int> off(ens->nAtoms + 1, 0), nb;
        for (int i = 0; i < ens->nAtoms; ++i) {
            Atom& a = ens->atoms[i];
            for (int k = 0; k < a.get_list_size(); ++k) nb.push_back(a.myPairList[k]->atomID);
            off[i + 1] = (int)nb.size();
        }
        write_bin("pl_off", off.data(), off.size() * 4);
        write_bin("pl_nb", nb.data(), nb.size() * 4);
    } else if (mode == "forces") {
        const std::string list = argc > 3 ? argv[3] : "lj,ewald";
        for (int m = 0; m < ens->nMols; ++m) ens->molecules[m].mass_center();
        std::vector<double> com(3 * ens->nMols);
        for (int m = 0; m < ens->nMols; ++m) {
            com[3 * m] = ens->molecules[m].massCenter.x;
            com[3 * m + 1] = ens->molecules[m].massCenter.y;
            com[3 * m + 2] = ens->molecules[m].massCenter.z;
        }
        write_bin("com", com.data(), com.size() * 8);
        std::stringstream ss(list);
        std::string item;
        while (std::getline(ss, item, ',')) {
            zero_forces(ens);
            std::memset(&ens->myLustig, 0, sizeof ens->myLustig);
            std::memset(&ens->myEnergy, 0, sizeof ens->myEnergy);
            double vir[9];
            if (item == "lj" || item == "mie" || item == "buck" || item == "buff") {
                Force* f = nullptr;
                if (item == "lj") f = make_zeroed<LJForce>(ens);
                if (item == "mie") f = make_zeroed<MieForce>(ens);
                if (item == "buck") f = make_zeroed<BuckinghamForce>(ens);
                if (item == "buff") f = make_zeroed<BuffvdWForce>(ens);
                f->compute();
                f->get_virial(vir);
                std::vector<double> F = forces_of(ens);
                write_bin(item + "_force", F.data(), F.size() * 8);
                write_bin(item + "_virial", vir, sizeof vir);
                meta.put(item + "_energy", f->get_energy());
                meta.put(item + "_ljEnergy", ens->myEnergy.ljEnergy);
                meta.put(item + "_ulus", ens->myLustig.ljUlus);
                meta.put(item + "_dudv", ens->myLustig.ljdUlus);
                meta.put(item + "_d2udv", ens->myLustig.ljd2Ulus);
            } else if (item == "bond" || item == "angle" || item == "ub") {
                // BondForce / AngleForce / UreyBradleyForce::compute
                Force* f = nullptr;
                if (item == "bond") f = make_zeroed<BondForce>(ens);
                if (item == "angle") f = make_zeroed<AngleForce>(ens);
                if (item == "ub") f = make_zeroed<UreyBradleyForce>(ens);
                f->compute();
                f->get_virial(vir);
                std::vector<double> F = forces_of(ens);
                write_bin(item + "_force", F.data(), F.size() * 8);
                write_bin(item + "_virial", vir, sizeof vir);
                meta.put(item + "_energy", f->get_energy());
                if (item == "bond") {
                    meta.put("bond_deriv", ens->myLustig.BondDeriv);
                    meta.put("bond_deriv2", ens->myLustig.BondSecDeriv);
                    meta.put("bond_ensEnergy", ens->myEnergy.bondEnergy);
                } else if (item == "angle") {
                    meta.put("angle_ensEnergy", ens->myEnergy.angleEnergy);
                } else {
                    meta.put("ub_deriv", ens->myLustig.ubDeriv);
                    meta.put("ub_deriv2", ens->myLustig.ubSecDeriv);
                    meta.put("ub_ensEnergy", ens->myEnergy.ubEnergy);
                }
            } else if (item == "ewald") {
                EwaldForce* ew = make_zeroed<EwaldForce>(ens);
                ew->compute();
                ew->get_virial(vir);
                std::vector<double> F = forces_of(ens);
                write_bin("ewald_short_force", F.data(), F.size() * 8);
                write_bin("ewald_short_virial", vir, sizeof vir);
                meta.put("ewald_real", ens->myEnergy.realEnergy);
                meta.put("ewald_correct", ens->myEnergy.correctEnergy);
                meta.put("ewald_molCorrect", ens->myEnergy.molCorrectEnergy);
                meta.put("ewald_surfCorrect", ens->myEnergy.surfCorrectEnergy);
                meta.put("ewald_intravirCorrect", ens->myLustig.intravirCorrect);
                meta.put("ewald_realDeriv", ens->myLustig.realDeriv);
                meta.put("ewald_realSecDeriv", ens->myLustig.realSecDeriv);
                meta.put("ewald_molDeriv", ens->myLustig.molDeriv);
                meta.put("ewald_molSecDeriv", ens->myLustig.molSecDeriv);
                meta.put("ewald_surfDeriv", ens->myLustig.surfDeriv);
                meta.put("ewald_surfSecDeriv", ens->myLustig.surfSecDeriv);
                zero_forces(ens);
                ew->compute_long();
                ew->get_long_virial(vir);
                F = forces_of(ens);
                write_bin("ewald_long_force", F.data(), F.size() * 8);
                write_bin("ewald_long_virial", vir, sizeof vir);
                meta.put("ewald_long", ens->myEnergy.longEnergy);
                meta.put("ewald_longDeriv", ens->myLustig.longDeriv);
                meta.put("ewald_longSecDeriv", ens->myLustig.longSecDeriv);
            } else {
                std::fprintf(stderr, "unknown force '%s'\n", item.c_str());
                return 2;
            }
        }
    } else if (mode == "induction") {
        // Induction::compute's first steps (src/forces/Induction.cpp:503-507):
        // frames of reference, then the permanent field and induced dipoles.
        for (int m = 0; m < ens->nMols; ++m) ens->molecules[m].mass_center();
        Induction* ind = new Induction(ens);
        ind->myL2G->FrameofReference();
        ind->InducedDipoles();
        const int n = ens->nAtoms;
        std::vector<double> dip(3 * n), quad(6 * n), e0(ind->Ei0, ind->Ei0 + 3 * n), pi(ind->Pi, ind->Pi + 3 * n);
        for (int i = 0; i < n; ++i) {
            dip[3 * i] = ind->myL2G->dip[i].x;
            dip[3 * i + 1] = ind->myL2G->dip[i].y;
            dip[3 * i + 2] = ind->myL2G->dip[i].z;
            quad[6 * i] = ind->myL2G->Qxx[i];
            quad[6 * i + 1] = ind->myL2G->Qxy[i];
            quad[6 * i + 2] = ind->myL2G->Qxz[i];
            quad[6 * i + 3] = ind->myL2G->Qyy[i];
            quad[6 * i + 4] = ind->myL2G->Qyz[i];
            quad[6 * i + 5] = ind->myL2G->Qzz[i];
        }
        write_bin("ind_dip", dip.data(), dip.size() * 8);
        write_bin("ind_quad", quad.data(), quad.size() * 8);
        write_bin("ind_ei0", e0.data(), e0.size() * 8);
        write_bin("ind_pi", pi.data(), pi.size() * 8);
        meta.put("cutoff_ind", ens->myConfig->get_cutoffInd());
        meta.put("a_damp", ens->myConfig->get_adamping());
    } else if (mode == "multipole") {
        // Multipole::compute() then compute_long() on the same coordinates
        // (src/forces/Multipole.cpp:163-240), forces of each written apart.
        for (int m = 0; m < ens->nMols; ++m) ens->molecules[m].mass_center();
        Multipole* mp = make_zeroed<Multipole>(ens);
        zero_forces(ens);
        std::memset(&ens->myLustig, 0, sizeof ens->myLustig);
        std::memset(&ens->myEnergy, 0, sizeof ens->myEnergy);
        double vir[9];
        mp->compute();
        mp->get_virial(vir);
        std::vector<double> F = forces_of(ens);
        write_bin("mp_short_force", F.data(), F.size() * 8);
        write_bin("mp_short_virial", vir, sizeof vir);
        meta.put("mp_real", ens->myEnergy.realEnergy);
        meta.put("mp_correct", ens->myEnergy.correctEnergy);
        meta.put("mp_molCorrect", ens->myEnergy.molCorrectEnergy);
        meta.put("mp_surfCorrect", ens->myEnergy.surfCorrectEnergy);
        meta.put("mp_realDeriv", ens->myLustig.realDeriv);
        meta.put("mp_molDeriv", ens->myLustig.molDeriv);
        meta.put("mp_surfDeriv", ens->myLustig.surfDeriv);
        meta.put("mp_surfSecDeriv", ens->myLustig.surfSecDeriv);
        const int n = ens->nAtoms;
        std::vector<double> dip(3 * n), quad(6 * n);
        for (int i = 0; i < n; ++i) {
            dip[3 * i] = mp->myL2G->dip[i].x;
            dip[3 * i + 1] = mp->myL2G->dip[i].y;
            dip[3 * i + 2] = mp->myL2G->dip[i].z;
            quad[6 * i] = mp->myL2G->Qxx[i];
            quad[6 * i + 1] = mp->myL2G->Qxy[i];
            quad[6 * i + 2] = mp->myL2G->Qxz[i];
            quad[6 * i + 3] = mp->myL2G->Qyy[i];
            quad[6 * i + 4] = mp->myL2G->Qyz[i];
            quad[6 * i + 5] = mp->myL2G->Qzz[i];
        }
        write_bin("mp_dip", dip.data(), dip.size() * 8);
        write_bin("mp_quad", quad.data(), quad.size() * 8);
        zero_forces(ens);
        mp->compute_long();
        mp->get_long_virial(vir);
        F = forces_of(ens);
        write_bin("mp_long_force", F.data(), F.size() * 8);
        write_bin("mp_long_virial", vir, sizeof vir);
        meta.put("mp_long", ens->myEnergy.longEnergy);
        meta.put("mp_longDeriv", ens->myLustig.longDeriv);
        meta.put("mp_longSecDeriv", ens->myLustig.longSecDeriv);
        meta.put("mp_totTorque", ens->myEnergy.TotTorq);
        meta.put("cutoff_ew", ens->myConfig->get_cutoffEw());
    } else if (mode == "induction_forces") {
        // Induction::compute() (dipole solve, then the energy, force and torque
        // terms) and compute_long() (src/forces/Induction.cpp:486-590).
        for (int m = 0; m < ens->nMols; ++m) ens->molecules[m].mass_center();
        Induction* ind = make_zeroed<Induction>(ens);
        zero_forces(ens);
        std::memset(&ens->myLustig, 0, sizeof ens->myLustig);
        std::memset(&ens->myEnergy, 0, sizeof ens->myEnergy);
        double vir[9];
        ind->compute();
        ind->get_virial(vir);
        std::vector<double> F = forces_of(ens);
        write_bin("indf_short_force", F.data(), F.size() * 8);
        write_bin("indf_short_virial", vir, sizeof vir);
        const int n = ens->nAtoms;
        std::vector<double> pi(ind->Pi, ind->Pi + 3 * n);
        write_bin("indf_pi", pi.data(), pi.size() * 8);
        meta.put("indf_real", ens->myEnergy.realPolEnergy);
        meta.put("indf_self", ens->myEnergy.correctPolEnergy);
        meta.put("indf_molCorrect", ens->myEnergy.molCorrectPolEnergy);
        meta.put("indf_surf", ens->myEnergy.surfacePolEnergy);
        meta.put("indf_realDeriv", ens->myLustig.realPolDeriv);
        meta.put("indf_molDeriv", ens->myLustig.molPolDeriv);
        meta.put("indf_surfDeriv", ens->myLustig.surfPolDeriv);
        meta.put("indf_surfSecDeriv", ens->myLustig.surfSecPolDeriv);
        meta.put("indf_selfDeriv", ens->myLustig.selfPolDeriv);
        zero_forces(ens);
        ind->compute_long();
        ind->get_long_virial(vir);
        F = forces_of(ens);
        write_bin("indf_long_force", F.data(), F.size() * 8);
        write_bin("indf_long_virial", vir, sizeof vir);
        meta.put("indf_long", ens->myEnergy.longPolEnergy);
        meta.put("indf_longDeriv", ens->myLustig.longPolDeriv);
        meta.put("indf_longSecDeriv", ens->myLustig.longSecPolDeriv);
        meta.put("indf_totTorque", ind->totTrqI);
        meta.put("cutoff_ew", ens->myConfig->get_cutoffEw());
        meta.put("cutoff_ind", ens->myConfig->get_cutoffInd());
    } else if (mode == "time_polar") {
        // The polarizable water's force evaluation, as LFIntegrator::run calls
        // it (src/integrators/LFIntegrator.cpp:382-470): buffered 14-7, then
        // Multipole and Induction compute() and compute_long(), reps times.
        const int reps = argc > 3 ? std::atoi(argv[3]) : 1;
        for (int m = 0; m < ens->nMols; ++m) ens->molecules[m].mass_center();
        BuffvdWForce* vdw = make_zeroed<BuffvdWForce>(ens);
        Multipole* mp = make_zeroed<Multipole>(ens);
        Induction* ind = make_zeroed<Induction>(ens);
        auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; ++r) {
            zero_forces(ens);
            vdw->compute();
            mp->compute();
            ind->compute();
            mp->compute_long();
            ind->compute_long();
        }
        auto t1 = std::chrono::steady_clock::now();
        meta.put("seconds", std::chrono::duration<double>(t1 - t0).count());
        meta.put("reps", reps);
        meta.put("e_buff", vdw->get_energy());
        meta.put("e_mp_real", ens->myEnergy.realEnergy);
        meta.put("e_pol_real", ens->myEnergy.realPolEnergy);
        meta.put("e_mp_long", ens->myEnergy.longEnergy);
        meta.put("e_pol_long", ens->myEnergy.longPolEnergy);
        std::vector<double> F = forces_of(ens);
        write_bin("polar_force", F.data(), F.size() * 8);
        std::printf("{\"reps\": %d, \"seconds\": %.6f}\n", reps,
                    std::chrono::duration<double>(t1 - t0).count());
    } else if (mode == "lf_forces") {
        // One force evaluation as LFIntegrator::run performs it
        // (src/integrators/LFIntegrator.cpp:382-557), with the terms
        // Integrator::initialise creates for this config
        // (src/Integrator.cpp:111-197): vdW, then Ewald or multipoles (+
        // induction), long-range virial added, then bond, angle and
        // Urey-Bradley.  Wolf and constrained-bond configs are not handled.
        for (int m = 0; m < ens->nMols; ++m) ens->molecules[m].mass_center();
        zero_forces(ens);
        Force* vdw = cfg->compute_buffamoeba() ? (Force*)make_zeroed<BuffvdWForce>(ens)
                                               : (Force*)make_zeroed<LJForce>(ens);
        vdw->compute();
        if (cfg->compute_ewald_force()) {
            EwaldForce* ew = make_zeroed<EwaldForce>(ens);
            ew->compute();
            ew->compute_long();
        } else if (cfg->compute_multipol()) {
            Multipole* mp = make_zeroed<Multipole>(ens);
            mp->compute();
            mp->compute_long();
            if (cfg->compute_induction()) {
                Induction* ind = make_zeroed<Induction>(ens);
                ind->compute();
                ind->compute_long();
            }
        }
        for (int j = 0; j < 9; ++j) ens->virial[j] += ens->longVirial[j];
        const bool flexible = !cfg->is_constraint_on();
        if (flexible && ens->nBonds > 0) make_zeroed<BondForce>(ens)->compute();
        if (flexible && ens->nAngles > 0) make_zeroed<AngleForce>(ens)->compute();
        if (cfg->compute_harmonic_ub()) make_zeroed<UreyBradleyForce>(ens)->compute();
        Energy& E = ens->myEnergy;
        const double induc = E.realPolEnergy + E.longPolEnergy + E.correctPolEnergy + E.molCorrectPolEnergy +
                             E.surfacePolEnergy;
        const double tot = E.ljEnergy + E.realEnergy + E.longEnergy + E.molCorrectEnergy + E.surfCorrectEnergy +
                           E.correctEnergy + E.angleEnergy + E.bondEnergy + E.ubEnergy + induc + E.TBEnergy;
        meta.put("ljEnergy", E.ljEnergy);
        meta.put("realEnergy", E.realEnergy);
        meta.put("longEnergy", E.longEnergy);
        meta.put("correctEnergy", E.correctEnergy);
        meta.put("molCorrectEnergy", E.molCorrectEnergy);
        meta.put("surfCorrectEnergy", E.surfCorrectEnergy);
        meta.put("realPolEnergy", E.realPolEnergy);
        meta.put("longPolEnergy", E.longPolEnergy);
        meta.put("correctPolEnergy", E.correctPolEnergy);
        meta.put("molCorrectPolEnergy", E.molCorrectPolEnergy);
        meta.put("surfacePolEnergy", E.surfacePolEnergy);
        meta.put("bondEnergy", E.bondEnergy);
        meta.put("angleEnergy", E.angleEnergy);
        meta.put("ubEnergy", E.ubEnergy);
        meta.put("inducEnergy", induc);
        meta.put("totEnergy", tot);
        std::vector<double> F = forces_of(ens);
        write_bin("lf_force", F.data(), F.size() * 8);
        write_bin("lf_virial", ens->virial, 9 * sizeof(double));
    } else if (mode == "md" || mode == "time") {
        const int nsteps = argc > 3 ? std::atoi(argv[3]) : 1;
        // The reference never initialises Molecule::massCenter before the
        // first force call; set it, as the oracle and the HIP path do.
        for (int m = 0; m < ens->nMols; ++m) ens->molecules[m].mass_center();
        LFIntegrator* lf = new LFIntegrator(ens, cfg);   // initialise() -> start()
        if (mode == "md") dump_state(ens, "start_");
        auto t0 = std::chrono::steady_clock::now();
        lf->run(nsteps);
        auto t1 = std::chrono::steady_clock::now();
        const double secs = std::chrono::duration<double>(t1 - t0).count();
        meta.put("seconds", secs);
        meta.put("steps", nsteps);
        meta.put("e_lj", ens->myEnergy.ljEnergy);
        meta.put("e_real", ens->myEnergy.realEnergy);
        meta.put("e_long", ens->myEnergy.longEnergy);
        meta.put("e_correct", ens->myEnergy.correctEnergy);
        meta.put("e_molCorrect", ens->myEnergy.molCorrectEnergy);
        meta.put("e_atomKin", ens->myEnergy.atomKinEnergy);
        meta.put("e_molKin", ens->myEnergy.molKinEnergy);
        if (mode == "md") {
            dump_state(ens, "end_");
            std::vector<double> F = forces_of(ens);
            write_bin("end_force", F.data(), F.size() * 8);
            // the last step's computePressure (src/Integrator.cpp:205-270) with
            // the shake-virial correction (LFIntegrator.cpp:805-814), and the
            // stress sums of Ensemble::accumulate (src/Ensemble.cpp:814-894)
            write_bin("end_virial", ens->virial, 9 * sizeof(double));
            write_bin("end_mol_virial", ens->molVirial, 9 * sizeof(double));
            write_bin("end_atom_pressure", ens->atomPressure, 9 * sizeof(double));
            write_bin("end_mol_pressure", ens->molPressure, 9 * sizeof(double));
            const double st[10] = {ens->atomStress.pressure, ens->atomStress.shearStress,
                                   ens->atomStress.antiShearStress, ens->atomStress.firstNormStress,
                                   ens->atomStress.secondNormStress, ens->molStress.pressure,
                                   ens->molStress.shearStress, ens->molStress.antiShearStress,
                                   ens->molStress.firstNormStress, ens->molStress.secondNormStress};
            write_bin("stress_sums", st, sizeof st);
            meta.put("acc_counter", ens->counter);
            meta.put("acc_luscount", ens->luscount);
            meta.put("totsteps", ens->totsteps);
            meta.put("molPt", ens->myAvLus.P);
            meta.put("ljdUlus", ens->myLustig.ljdUlus);
            meta.put("realDeriv", ens->myLustig.realDeriv);
            meta.put("longDeriv", ens->myLustig.longDeriv);
            meta.put("molDeriv", ens->myLustig.molDeriv);
            meta.put("surfDeriv", ens->myLustig.surfDeriv);
            // the reference's own writers on this end state: the restart file
            // and the pressure lines (text kept as fixtures)
            lf->write_state();   // state.out in the working directory
            {
                std::ofstream pr("press_res.out");
                ens->write_press_res(pr, nsteps);
                std::ofstream pa("pressac.out");
                ens->write_pressac(pa, nsteps);
                std::ofstream pt("pressure_tensor.out");
                ens->write_pressureTensor(pt, nsteps);
            }
            std::vector<double> rp(3 * ens->nAtoms);
            for (int i = 0; i < ens->nAtoms; ++i) {
                rp[3 * i] = ens->atoms[i].realPos.x;
                rp[3 * i + 1] = ens->atoms[i].realPos.y;
                rp[3 * i + 2] = ens->atoms[i].realPos.z;
            }
            write_bin("end_real_pos", rp.data(), rp.size() * 8);
            meta.put("current_steps", lf->currentSteps);
        }
        std::printf("{\"steps\": %d, \"seconds\": %.6f}\n", nsteps, secs);
    } else {
        std::fprintf(stderr, "unknown mode '%s'\n", mode.c_str());
        return 2;
    }
    meta.save();
    std::fflush(stdout);
    MPI_Finalize();
    return 0;
}
