// inputs.cpp -- readers of the reference's input files (host code, no GPU).
//
//   config.txt   SimConfiguration::init / read_config_file
//                src/SimConfiguration.cpp:29-84, :695-932, :933-978
//   sysData.in   Parameters::read_sys_data, check_heading
//                src/Parameters.cpp:76-382, :1309-1324
//   coordinates  Ensemble::read_coordinate / read_realPos / read_velocity
//                src/Ensemble.cpp:630-758; restart file :299-317
//
// The reference reads config.txt and sysData.in with std::istream
// extraction: whitespace-separated tokens, an int stops at the first
// character that cannot continue it (so "1.10" read as an int is 1 and
// leaves ".10"), and "skip the rest" discards up to the next newline.  The
// token reader below follows those rules; where the reference would carry on
// after a failed extraction with a zero or stale value, it reports the file,
// line and field instead.
#include <cctype>
#include <cerrno>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/swinmd_hip.h"

namespace {

struct ReadError {
    std::string msg;
};

[[noreturn]] void raise(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw ReadError{buf};
}

void put_err(char* err, size_t errlen, const std::string& m) {
    if (err && errlen) {
        std::strncpy(err, m.c_str(), errlen - 1);
        err[errlen - 1] = '\0';
    }
}

std::string slurp(const char* path) {
    std::ifstream in(path, std::ios::in | std::ios::binary);
    if (!in.is_open()) raise("can't open %s", path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

// Extraction in the manner of std::istream >> on a whole file.
class Tokens {
public:
    Tokens(std::string text, const char* path) : t_(std::move(text)), path_(path) {}

    int line() const { return line_; }

    // SKIP_REST: ignore(max, '\n')
    void skip_rest() {
        while (p_ < t_.size() && t_[p_] != '\n') ++p_;
        if (p_ < t_.size()) { ++p_; ++line_; }
    }

    std::string word(const char* what) {
        skip_ws();
        const size_t b = p_;
        while (p_ < t_.size() && !std::isspace((unsigned char)t_[p_])) ++p_;
        if (p_ == b) fail(what);
        return t_.substr(b, p_ - b);
    }

    // the token as a flag: "t" is true, anything else false (boolFlag == "t")
    bool flag(const char* what) { return word(what) == "t"; }

    long long integer(const char* what) {
        skip_ws();
        size_t q = p_;
        if (q < t_.size() && (t_[q] == '+' || t_[q] == '-')) ++q;
        const size_t d0 = q;
        while (q < t_.size() && std::isdigit((unsigned char)t_[q])) ++q;
        if (q == d0) fail(what);
        const long long v = std::strtoll(t_.c_str() + p_, nullptr, 10);
        p_ = q;
        return v;
    }

    double real(const char* what) {
        skip_ws();
        size_t q = p_;
        if (q < t_.size() && (t_[q] == '+' || t_[q] == '-')) ++q;
        size_t digits = 0;
        while (q < t_.size() && std::isdigit((unsigned char)t_[q])) { ++q; ++digits; }
        if (q < t_.size() && t_[q] == '.') {
            ++q;
            while (q < t_.size() && std::isdigit((unsigned char)t_[q])) { ++q; ++digits; }
        }
        if (!digits) fail(what);
        if (q < t_.size() && (t_[q] == 'e' || t_[q] == 'E')) {
            size_t e = q + 1;
            if (e < t_.size() && (t_[e] == '+' || t_[e] == '-')) ++e;
            const size_t e0 = e;
            while (e < t_.size() && std::isdigit((unsigned char)t_[e])) ++e;
            if (e == e0) fail(what);
            q = e;
        }
        const std::string num = t_.substr(p_, q - p_);
        p_ = q;
        return std::strtod(num.c_str(), nullptr);
    }

    // check_heading: skip a line, read the heading word, skip the rest of its
    // line and (when it matches) the column-comment line after it.
    bool heading(const char* expect) {
        skip_rest();
        skip_ws();
        size_t b = p_;
        while (p_ < t_.size() && !std::isspace((unsigned char)t_[p_])) ++p_;
        const std::string h = t_.substr(b, p_ - b);
        skip_rest();
        if (h != expect) return false;
        skip_rest();
        return true;
    }

    [[noreturn]] void fail(const char* what) const {
        raise("%s:%d: cannot read %s", path_, line_, what);
    }

private:
    void skip_ws() {
        while (p_ < t_.size() && std::isspace((unsigned char)t_[p_])) {
            if (t_[p_] == '\n') ++line_;
            ++p_;
        }
    }

    std::string t_;
    const char* path_;
    size_t p_ = 0;
    int line_ = 1;
};

void copy_name(char (&dst)[80], const std::string& s, const char* what, const Tokens& tk) {
    if (s.size() >= sizeof dst) raise("line %d: %s longer than %zu characters", tk.line(), what, sizeof dst - 1);
    std::memcpy(dst, s.c_str(), s.size() + 1);
}

template <size_t N>
void copy_short(char (&dst)[N], const std::string& s) {
    const size_t k = s.size() < N - 1 ? s.size() : N - 1;
    std::memcpy(dst, s.c_str(), k);
    dst[k] = '\0';
}

int32_t as_i32(long long v, const Tokens& tk, const char* what) {
    if (v < INT32_MIN || v > INT32_MAX) raise("line %d: %s out of range", tk.line(), what);
    return (int32_t)v;
}

// SimConfiguration::init (src/SimConfiguration.cpp:29-84)
void config_defaults(smd_config& c) {
    std::memset(&c, 0, sizeof c);
    c.n_steps = 1;
    c.start_steps = 0;
    c.time_step = 1.0;
    c.short_ts_freq = 1;
    c.long_ts_freq = 4;
    c.avg_ts_freq = c.print_ts_freq = c.backup_ts_freq = c.n_steps;
    c.start_sampling = c.start_steps;
    c.sampling_ts_freq = c.n_steps;
    c.start_ttcf = c.start_steps;
    c.ttcf_ts_freq = c.n_steps;
    c.integrator_type = 0;
    c.rdf_nbins = 100;
    c.coupl_thermo = 1.0;
    c.coupl_baros = 1.0;
    c.lx_box = c.ly_box = c.lz_box = 1.0;
    c.cutoff = 1.2;
    c.k_cutoff = 5;
    c.cut_buff = 0.0;
    c.cutoff_ew = 1.0;
    c.rate = 0.0;
    c.cutoff_ind = 1.0;
    c.a_damp = 200.0;
    c.density = 1.0;
    c.switch_dist = 1.0;
    c.n_mie = 12.0;
    c.m_mie = 6.0;
    c.n_buff = 14.0;
    c.m_buff = 7.0;
    c.h_factor = 0.91;
    c.delta_buff = 0.07;
    c.gamma_buff = 0.12;
    c.cell_side_len = 1.0;
    c.temperature = 1.0;
    c.tolerance = 0.00001;
    c.max_count = 10000;
    c.seed = 1;
    c.bending_type = 1;
    c.use_atom_thermo = 1;
}

void read_config(const char* path, smd_config& c) {
    Tokens tk(slurp(path), path);
    config_defaults(c);
    copy_name(c.sys_data_file, tk.word("sysDataFile"), "sysDataFile", tk);
    copy_name(c.coordinate_file, tk.word("coordinateFile"), "coordinateFile", tk);
    tk.skip_rest();
    copy_name(c.result_file, tk.word("resultFile"), "resultFile", tk);
    copy_name(c.restart_file, tk.word("restartFile"), "restartFile", tk);
    copy_name(c.dump_file, tk.word("dumpFile"), "dumpFile", tk);
    tk.skip_rest();
    if (tk.flag("writeTrajectory flag")) {
        c.write_trajectory = 1;
        copy_name(c.trajectory_file, tk.word("trajectoryFile"), "trajectoryFile", tk);
    }
    tk.skip_rest();
    c.is_new_start = tk.flag("isNewStart flag");
    c.seed = (uint32_t)tk.integer("seed");
    c.read_velocity = tk.flag("readVelocity flag");
    copy_name(c.velocity_file, tk.word("velocityFile"), "velocityFile", tk);
    tk.skip_rest();
    c.time_step = tk.real("timeStep");
    c.n_steps = as_i32(tk.integer("nSteps"), tk, "nSteps");
    c.start_steps = as_i32(tk.integer("startSteps"), tk, "startSteps");
    tk.skip_rest();
    c.short_ts_freq = as_i32(tk.integer("shortTsFreq"), tk, "shortTsFreq");
    c.long_ts_freq = as_i32(tk.integer("longTsFreq"), tk, "longTsFreq");
    tk.skip_rest();
    c.avg_ts_freq = as_i32(tk.integer("averageTsFreq"), tk, "averageTsFreq");
    c.print_ts_freq = as_i32(tk.integer("printTsFreq"), tk, "printTsFreq");
    c.backup_ts_freq = as_i32(tk.integer("backupTsFreq"), tk, "backupTsFreq");
    tk.skip_rest();
    c.integrator_type = as_i32(tk.integer("integratorType"), tk, "integratorType");
    tk.skip_rest();
    const int ens = as_i32(tk.integer("ensembleStatus"), tk, "ensembleStatus");
    const double p1 = tk.real("ensemble parameter 1");
    const double p2 = tk.real("ensemble parameter 2");
    const long long atom_flag = tk.integer("atom/molecule thermostat flag");
    // set_ensemble_status (:951-978)
    c.ensemble_status = ens;
    switch (ens) {
    case 0: c.constant_temperature = c.constant_volume = 1; c.temperature = p1; c.volume = p2; break;
    case 1: c.constant_temperature = c.constant_pressure = 1; c.temperature = p1; c.pressure = p2; break;
    case 2: c.constant_volume = c.constant_energy = 1; c.energy = p1; c.volume = p2; break;
    case 3: c.constant_pressure = c.constant_energy = 1; c.energy = p1; c.pressure = p2; break;
    default: break;
    }
    if (atom_flag != 1) c.use_atom_thermo = 0;
    c.coupl_thermo = tk.real("couplthermo");
    copy_name(c.thermo_type, tk.word("thermotype"), "thermotype", tk);
    c.coupl_baros = tk.real("couplBaros");
    tk.skip_rest();
    // set_run_type (:933-949)
    c.run_type = as_i32(tk.integer("runType"), tk, "runType");
    c.rate = tk.real("rate");
    c.do_equilibrium = 1;   // init() default, kept whatever the run type
    if (c.run_type == 1) c.do_compression = 1;
    if (c.run_type == 2) c.do_shear = 1;
    if (c.run_type == 3) c.do_elongation = 1;
    tk.skip_rest();
    c.cutoff = tk.real("cutoff");
    c.cutoff_ew = tk.real("cutoffEw");
    c.k_cutoff = tk.real("kCutoff");
    c.cut_buff = tk.real("cutBuff");
    tk.skip_rest();
    c.density = tk.real("density");
    tk.skip_rest();
    c.lx_box = tk.real("lxBox");
    c.ly_box = tk.real("lyBox");
    c.lz_box = tk.real("lzBox");
    tk.skip_rest();
    c.use_cell_list = tk.flag("useCellPairList flag");
    c.cell_side_len = tk.real("cellSideLen");
    tk.skip_rest();
    c.constraint_mode = as_i32(tk.integer("constraintMode"), tk, "constraintMode");
    if (c.constraint_mode == 1) {          // GAUSSIAN_CONSTRAINT
        c.constraint_on = 1;
        c.cfb = tk.real("CFB");
        c.dfb = tk.real("DFB");
        c.tfb = tk.real("TFB");
        tk.skip_rest();
    } else if (c.constraint_mode == 2) {   // SHAKE_CONSTRAINT
        c.constraint_on = 1;
        c.tolerance = tk.real("tolerance");
        c.max_count = as_i32(tk.integer("maxCount"), tk, "maxCount");
        tk.skip_rest();
    } else if (c.constraint_mode == 0) {   // NO_CONSTRAINT
        c.tolerance = tk.real("tolerance");
        c.bending_type = as_i32(tk.integer("bendingType"), tk, "bendingType");
        if (c.bending_type == 1) c.harmonic_bend = 1;
        else if (c.bending_type == 2) c.amoeba_bend = 1;
        tk.skip_rest();
    }
    // (any other mode reads nothing more from this line, as the reference)
    c.harmonic_ub = tk.flag("computeharmoUB flag");
    c.kl_ub = tk.real("klUB");
    c.l0_ub = tk.real("l0UB");
    tk.skip_rest();
    c.long_correct = tk.flag("long-range correction flag");
    c.switch_on = tk.flag("switchOn flag");
    c.switch_dist = tk.real("switchDist");
    tk.skip_rest();
    c.coulomb = tk.flag("computeCoulomb flag");
    c.ewald = tk.flag("computeEwald flag");
    c.surf_correct = tk.flag("computeSurfCorrect flag");
    c.wolf = tk.flag("computeWolf flag");
    c.alpha = tk.real("alpha");
    c.k_cutoff = tk.real("kCutoff");
    tk.skip_rest();
    c.multipole = tk.flag("computeMultiPol flag");
    c.induction = tk.flag("computeInduction flag");
    c.torques = tk.flag("computeTorq flag");
    c.damp_induction = tk.flag("dampInduction flag");
    c.induction_min = tk.flag("doInductionMin flag");
    c.damp_amoeba = tk.flag("dampAmoeba flag");
    c.cutoff_ind = tk.real("cutoffInd");
    c.a_damp = tk.real("adamp");
    tk.skip_rest();
    c.buff_vdw = tk.flag("computeBuffvdW flag");
    c.n_buff = tk.real("nBuff");
    c.m_buff = tk.real("mBuff");
    c.delta_buff = tk.real("deltaBuff");
    c.gamma_buff = tk.real("gammaBuff");
    c.h_factor = tk.real("Hfactor");
    tk.skip_rest();
    c.mie = tk.flag("computeMie flag");
    c.n_mie = tk.real("nMie");
    c.m_mie = tk.real("mMie");
    tk.skip_rest();
    c.ext_forces = tk.flag("hasExtForces flag");
    c.fixed_atom = tk.flag("hasFixedAtom flag");
    tk.skip_rest();
    c.start_sampling = as_i32(tk.integer("startSampling"), tk, "startSampling");
    c.sampling_ts_freq = as_i32(tk.integer("samplingTsFreq"), tk, "samplingTsFreq");
    tk.skip_rest();
    c.ttcf = tk.flag("useTTCF flag");
    c.start_ttcf = as_i32(tk.integer("startTTCF"), tk, "startTTCF");
    c.ttcf_ts_freq = as_i32(tk.integer("ttcfTsFreq"), tk, "ttcfTsFreq");
    tk.skip_rest();
    c.rdf = tk.flag("computeRDF flag");
    c.rdf_nbins = as_i32(tk.integer("nBins"), tk, "nBins");
    tk.skip_rest();
    c.msd = tk.flag("computeMSD flag");
    copy_name(c.mol_trajectory_file, tk.word("molTrajectoryFile"), "molTrajectoryFile", tk);
    tk.skip_rest();
    c.vacf = tk.flag("computeVACF flag");
    copy_name(c.mol_velocity_file, tk.word("molVelocityFile"), "molVelocityFile", tk);
    tk.skip_rest();
    c.lustig = tk.flag("computeLustig flag");
}

struct SysData {
    smd_sysdata_sizes sz{};
    std::vector<smd_atom_param> atom_params;
    std::vector<smd_bond_param> bond_params;
    std::vector<smd_angle_param> angle_params;
    std::vector<smd_torsion_param> dihedral_params, improper_params;
    std::vector<smd_mol_type> mol_types;
    std::vector<int32_t> atom_type, atom_mol, mol_type;
    std::vector<int32_t> bonds, angles, dihedrals, impropers;
};

int32_t index_in(long long v, int32_t n, const Tokens& tk, const char* what) {
    if (v < 0 || v >= n) raise("line %d: %s %lld outside [0, %d)", tk.line(), what, v, n);
    return (int32_t)v;
}

void read_sysdata(const char* path, SysData& d) {
    Tokens tk(slurp(path), path);
    tk.skip_rest();   // two title lines
    tk.skip_rest();
    auto count = [&](const char* what) { return as_i32(tk.integer(what), tk, what); };
    d.sz.n_mols = count("number of molecules");
    d.sz.n_mol_types = count("number of molecule types");
    tk.skip_rest();
    d.sz.n_atoms = count("number of atoms");
    d.sz.n_atom_types = count("number of atom types");
    tk.skip_rest();
    d.sz.n_bonds = count("number of bonds");
    d.sz.n_bond_types = count("number of bond types");
    tk.skip_rest();
    d.sz.n_angles = count("number of angles");
    d.sz.n_angle_types = count("number of angle types");
    tk.skip_rest();
    d.sz.n_dihedrals = count("number of dihedrals");
    d.sz.n_dihedral_types = count("number of dihedral types");
    tk.skip_rest();
    d.sz.n_impropers = count("number of impropers");
    d.sz.n_improper_types = count("number of improper types");
    tk.skip_rest();
    const smd_sysdata_sizes& z = d.sz;
    for (int32_t v : {z.n_mols, z.n_mol_types, z.n_atoms, z.n_atom_types, z.n_bonds, z.n_bond_types, z.n_angles,
                      z.n_angle_types, z.n_dihedrals, z.n_dihedral_types, z.n_impropers, z.n_improper_types})
        if (v < 0) raise("%s: negative count in the size section", path);

    auto need = [&](const char* h) {
        if (!tk.heading(h)) raise("%s:%d: section heading %s expected", path, tk.line(), h);
    };
    // AtomParams (:101-142); stored at the type the line names
    d.atom_params.assign(z.n_atom_types, smd_atom_param{});
    if (z.n_atom_types) {
        need("!AtomParams");
        for (int k = 0; k < z.n_atom_types; ++k) {
            smd_atom_param ap{};
            ap.type = index_in(tk.integer("atomType"), z.n_atom_types, tk, "atom type");
            ap.mass = tk.real("mass");
            ap.charge = tk.real("charge");
            ap.eps = tk.real("eps");
            ap.sigma = tk.real("sigma");
            for (double& v : ap.dip) v = tk.real("dipole");
            for (double& v : ap.quad) v = tk.real("quadrupole");
            ap.polar = tk.real("polar");
            copy_short(ap.type_name, tk.word("typeName"));
            copy_short(ap.atom_name, tk.word("atomName"));
            tk.skip_rest();
            d.atom_params[ap.type] = ap;
        }
    } else {
        tk.heading("!AtomParams");
    }
    d.bond_params.assign(z.n_bond_types, smd_bond_param{});
    if (z.n_bond_types) {
        need("!BondParams");
        for (int k = 0; k < z.n_bond_types; ++k) {
            smd_bond_param& b = d.bond_params[k];
            b.type = as_i32(tk.integer("bondType"), tk, "bondType");
            b.t1 = as_i32(tk.integer("atomType1"), tk, "atomType1");
            b.t2 = as_i32(tk.integer("atomType2"), tk, "atomType2");
            b.r0 = tk.real("r0");
            b.k = tk.real("k");
            tk.skip_rest();
        }
    } else {
        tk.heading("!BondParams");
    }
    d.angle_params.assign(z.n_angle_types, smd_angle_param{});
    if (z.n_angle_types) {
        need("!AngleParams");
        for (int k = 0; k < z.n_angle_types; ++k) {
            smd_angle_param& a = d.angle_params[k];
            a.type = as_i32(tk.integer("angleType"), tk, "angleType");
            a.t1 = as_i32(tk.integer("atomType1"), tk, "atomType1");
            a.t2 = as_i32(tk.integer("atomType2"), tk, "atomType2");
            a.t3 = as_i32(tk.integer("atomType3"), tk, "atomType3");
            a.theta0_deg = tk.real("theta0");   // degrees in the file
            a.theta0 = a.theta0_deg * 2 * 3.14159265358979323846 / 360.0;
            a.k = tk.real("k");
            tk.skip_rest();
        }
    } else {
        tk.heading("!AngleParams");
    }
    auto torsions = [&](std::vector<smd_torsion_param>& v, int32_t n, const char* h) {
        v.assign(n, smd_torsion_param{});
        if (!n) { tk.heading(h); return; }
        need(h);
        for (int k = 0; k < n; ++k) {
            smd_torsion_param& t = v[k];
            t.type = as_i32(tk.integer("torsionType"), tk, "torsionType");
            t.t1 = as_i32(tk.integer("atomType1"), tk, "atomType1");
            t.t2 = as_i32(tk.integer("atomType2"), tk, "atomType2");
            t.t3 = as_i32(tk.integer("atomType3"), tk, "atomType3");
            t.t4 = as_i32(tk.integer("atomType4"), tk, "atomType4");
            t.k = tk.real("k");
            t.d = as_i32(tk.integer("d"), tk, "d");
            t.n = as_i32(tk.integer("n"), tk, "n");
            tk.skip_rest();
        }
    };
    torsions(d.dihedral_params, z.n_dihedral_types, "!DihedralParams");
    torsions(d.improper_params, z.n_improper_types, "!ImproperParams");
    d.mol_types.assign(z.n_mol_types, smd_mol_type{});
    if (z.n_mol_types) {
        need("!MolTypes");
        for (int k = 0; k < z.n_mol_types; ++k) {
            smd_mol_type& m = d.mol_types[k];
            m.type = as_i32(tk.integer("molType"), tk, "molType");
            m.n_atoms = as_i32(tk.integer("numOfAtoms"), tk, "numOfAtoms");
            copy_short(m.name, tk.word("molecule name"));
            tk.skip_rest();
        }
    }
    d.atom_type.assign(z.n_atoms, -1);
    d.atom_mol.assign(z.n_atoms, -1);
    d.mol_type.assign(z.n_mols, -1);
    if (z.n_atoms) {
        need("!Atoms");
        for (int k = 0; k < z.n_atoms; ++k) {
            const int32_t id = index_in(tk.integer("atomID"), z.n_atoms, tk, "atom ID");
            d.atom_type[id] = index_in(tk.integer("atomType"), z.n_atom_types, tk, "atom type");
            const int32_t mol = index_in(tk.integer("molID"), z.n_mols, tk, "molecule ID");
            d.atom_mol[id] = mol;
            d.mol_type[mol] = as_i32(tk.integer("molType"), tk, "molType");
            tk.skip_rest();
        }
    }
    auto rows = [&](std::vector<int32_t>& v, int32_t n, int width, const char* h, const char* what) {
        v.assign((size_t)n * width, -1);
        if (!n) return;
        need(h);
        for (int k = 0; k < n; ++k) {
            const int32_t id = index_in(tk.integer(what), n, tk, what);
            for (int c = 0; c + 1 < width; ++c)
                v[(size_t)id * width + c] = index_in(tk.integer("atom index"), z.n_atoms, tk, "atom index");
            v[(size_t)id * width + width - 1] = as_i32(tk.integer("type"), tk, "type");
            tk.skip_rest();
        }
    };
    rows(d.bonds, z.n_bonds, 3, "!Bonds", "bond ID");
    rows(d.angles, z.n_angles, 4, "!Angles", "angle ID");
    rows(d.dihedrals, z.n_dihedrals, 5, "!Dihedrals", "dihedral ID");
    rows(d.impropers, z.n_impropers, 5, "!Impropers", "improper ID");
}

// read_coordinate: fgets a count line, sscanf "%d", then one fgets +
// sscanf "%lf%lf%lf" per atom (src/Ensemble.cpp:630-670).
struct Lines {
    explicit Lines(const char* path) : path_(path) {
        f_ = std::fopen(path, "r");
        if (!f_) raise("can't open %s", path);
    }
    ~Lines() { if (f_) std::fclose(f_); }
    std::string next() {
        std::string s;
        int ch;
        bool any = false;
        while ((ch = std::fgetc(f_)) != EOF) {
            any = true;
            if (ch == '\n') break;
            s.push_back((char)ch);
        }
        if (!any) raise("%s:%d: unexpected end of file", path_, line_ + 1);
        ++line_;
        return s;
    }
    int line() const { return line_; }
    const char* path_;

private:
    std::FILE* f_ = nullptr;
    int line_ = 0;
};

void read_block(Lines& in, int64_t n, double* out) {
    const std::string head = in.next();
    long long cnt = 0;
    if (std::sscanf(head.c_str(), "%lld", &cnt) != 1 || cnt != n)
        raise("%s:%d: count line must give %lld atoms", in.path_, in.line(), (long long)n);
    for (int64_t i = 0; i < n; ++i) {
        const std::string l = in.next();
        double x, y, z;
        if (std::sscanf(l.c_str(), "%lf%lf%lf", &x, &y, &z) != 3)
            raise("%s:%d: three numbers expected", in.path_, in.line());
        out[3 * i] = x;
        out[3 * i + 1] = y;
        out[3 * i + 2] = z;
    }
}

template <class F>
int guarded(char* err, size_t errlen, F&& f) {
    try {
        f();
        put_err(err, errlen, "");
        return SMD_OK;
    } catch (const ReadError& e) {
        put_err(err, errlen, e.msg);
        return SMD_EIO;
    } catch (const std::bad_alloc&) {
        put_err(err, errlen, "host allocation failed");
        return SMD_ENOMEM;
    }
}

}  // namespace

struct smd_sysdata {
    SysData d;
};

extern "C" int smd_config_read(const char* path, smd_config* out, char* err, size_t errlen) {
    if (!path || !out) return SMD_EINVAL;
    return guarded(err, errlen, [&] { read_config(path, *out); });
}

extern "C" int smd_sysdata_read(const char* path, smd_sysdata** out, char* err, size_t errlen) {
    if (!path || !out) return SMD_EINVAL;
    *out = nullptr;
    smd_sysdata* h = nullptr;
    const int rc = guarded(err, errlen, [&] {
        h = new smd_sysdata;
        read_sysdata(path, h->d);
    });
    if (rc != SMD_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return SMD_OK;
}

extern "C" int smd_sysdata_view_get(const smd_sysdata* h, smd_sysdata_view* v) {
    if (!h || !v) return SMD_EINVAL;
    const SysData& d = h->d;
    v->sizes = d.sz;
    v->atom_params = d.atom_params.data();
    v->bond_params = d.bond_params.data();
    v->angle_params = d.angle_params.data();
    v->dihedral_params = d.dihedral_params.data();
    v->improper_params = d.improper_params.data();
    v->mol_types = d.mol_types.data();
    v->atom_type = d.atom_type.data();
    v->atom_mol = d.atom_mol.data();
    v->mol_type = d.mol_type.data();
    v->bonds = d.bonds.data();
    v->angles = d.angles.data();
    v->dihedrals = d.dihedrals.data();
    v->impropers = d.impropers.data();
    return SMD_OK;
}

extern "C" void smd_sysdata_free(smd_sysdata* h) { delete h; }

extern "C" int smd_read_xyz(const char* path, int64_t n_atoms, int32_t skip_lines, double* out, char* err,
                            size_t errlen) {
    if (!path || !out || n_atoms < 0 || skip_lines < 0) return SMD_EINVAL;
    return guarded(err, errlen, [&] {
        Lines in(path);
        for (int k = 0; k < skip_lines; ++k) in.next();
        read_block(in, n_atoms, out);
    });
}

extern "C" int smd_read_state(const char* path, int64_t n_atoms, double* pos, double* real_pos, double* vel,
                              char* err, size_t errlen) {
    if (!path || !pos || !real_pos || !vel || n_atoms < 0) return SMD_EINVAL;
    return guarded(err, errlen, [&] {
        Lines in(path);
        in.next();
        read_block(in, n_atoms, pos);
        in.next();
        read_block(in, n_atoms, real_pos);
        in.next();
        read_block(in, n_atoms, vel);
    });
}
