/*
 * oracle.h -- CPU restatement of SwinMD's nonbonded force path, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in swinmd_amd/: it is loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and by
 * nothing else.  The product path never calls it and never falls back to it.
 *
 * Every function restates one reference routine, following its loop order and
 * floating-point expression order so that results agree with the compiled
 * reference to rounding (checked against golden vectors in tests/golden/).
 * Reference paths below are relative to the SwinMD source root.
 *
 * Units are the reference's: nm, ps, amu, kJ/mol, charges pre-scaled by
 * sqrt(138.93548501) (src/Ensemble.cpp:191).  All arithmetic is FP64.
 */
#ifndef SWINMD_ORACLE_H
#define SWINMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Virial index order XX XY XZ YX YY YZ ZX ZY ZZ (src/NEMD_defs.h:45-53). */

/* System view.  Atoms of molecule m are mol_start[m] .. mol_start[m+1]-1, in
 * the order Molecule::myAtoms holds them (atom-index order,
 * src/Ensemble.cpp:214-221). */
typedef struct {
    int32_t n_atoms, n_mols;
    const double *pos;        /* [3*n_atoms] AoS */
    const int32_t *type;      /* [n_atoms] */
    const int32_t *mol;       /* [n_atoms] molecule id */
    const double *q;          /* [n_atoms] scaled charge */
    const double *mass;       /* [n_atoms] */
    const int32_t *mol_start; /* [n_mols+1] */
    const double *com;        /* [3*n_mols] Molecule::massCenter */
    const int32_t *excl_off;  /* [n_atoms+1] exclusion table (CSR) */
    const int32_t *excl;      /* ascending per row, as std::set orders it */
    double box[3];
    double tilt;              /* Lees-Edwards xy tilt of the sheared replicas (0: the reference's box) */
} orc_sys;

/* Cell grid of CellManager::set_cellManager (src/utils/CellManager.cpp:43-84).
 * Returns 1 when the grid is usable, 0 when len > box/3 (reference falls back
 * to the brute-force list). */
int orc_cell_setup(double len, const double box[3], int32_t n[3],
                   double cell_len[3], double rcell_len[3]);

/* Half pair list (Ensemble::set_pairlist / update_pairlist,
 * src/Ensemble.cpp:454-490, CellManager::update_pair_list
 * src/utils/CellManager.cpp:216-314, Ensemble::build_pairlist :492-525).
 * r_list2 = (cutoff+cutBuff)^2 evaluated as the reference does.
 * Rows are in reference order.  off has n_atoms+1 entries.  Returns the
 * number of pairs, or -(needed) when cap is too small (nothing written past
 * cap).  *used_cells tells which path ran. */
int64_t orc_pairlist(const orc_sys *s, double cutoff, double cut_buff,
                     int use_cells, int32_t *off, int32_t *nb, int64_t cap,
                     int *used_cells);

/* Kinds of short-range van der Waals term. */
enum { ORC_LJ = 0, ORC_MIE = 1, ORC_BUCK = 2, ORC_BUFF147 = 3 };

typedef struct {
    int32_t kind;
    int32_t n_types;
    double cutoff;            /* SimConfiguration cutoff (vdW) */
    double switch_dist;       /* get_switch_distance() (not squared) */
    int32_t switch_on, use_constraint, lustig, coulomb;
    /* LJ / Mie: LJPairParam table, sigma^2 and 4*eps, [n_types^2]
     * (Parameters::set_lj_param, src/Parameters.cpp:1360-1372). */
    const double *sigma2, *eps4;
    /* Mie exponents (src/forces/MieForce.cpp:32-35). */
    double n_mie, m_mie;
    /* Buckingham A,B,C [n_types^2] and CO2 geometry
     * (src/forces/BuckinghamForce.h:19-37, .cpp:372-410). */
    const double *buck_a, *buck_b, *buck_c;
    double co2_bond;          /* bondParams[0].r0 */
    /* Buffered 14-7: per-type R and eps (AtomParam.sigma, .eps) and shape. */
    const double *type_sigma, *type_eps;
    double n_buff, m_buff, delta, gamma, hfactor;
    /* Long-range correction values added the way compute() adds them. */
    double e_lrc, v_lrc, du_lrc, d2u_lrc;
} orc_vdw_params;

typedef struct {
    double energy;       /* Force::energy after compute() */
    double lj_energy;    /* what compute() stores in myEnergy.ljEnergy */
    double virial[9];    /* Force::virial (compute() then does ensemble.virial -= virial) */
    double ulus, dudv, d2udv;  /* myLustig.ljUlus / ljdUlus / ljd2Ulus */
} orc_vdw_out;

/* LJForce::compute (src/forces/LJForce.cpp:58-255), MieForce::compute
 * (src/forces/MieForce.cpp:62-241), BuckinghamForce::compute
 * (src/forces/BuckinghamForce.cpp:79-287), BuffvdWForce::compute
 * (src/forces/BuffvdWForce.cpp:93-343).  force is accumulated (+=). */
int orc_vdw(const orc_sys *s, const int32_t *off, const int32_t *nb,
            const orc_vdw_params *p, double *force, orc_vdw_out *out);

/* Buffered 14-7 pair table (BuffvdWForce::set_buff_param,
 * src/forces/BuffvdWForce.cpp:392-433): R_ij and eps_ij, [n^2]. */
void orc_buff_table(int n_types, const double *type_sigma,
                    const double *type_eps, double *r_ij, double *eps_ij);

/* LJ long-range correction (LJForce::long_range_correct,
 * src/forces/LJForce.cpp:260-370), with the accumulator csix started at 0. */
void orc_lj_lrc(const orc_sys *s, int n_types, const double *sigma2,
                const double *eps4, double cutoff, double *e_lrc,
                double *v_lrc, double *du_lrc, double *d2u_lrc);

/* Ewald pieces of EwaldForce (src/forces/EwaldForce.cpp). */
typedef struct {
    double energy, deriv, deriv2;   /* term energy and its Lustig volume derivatives */
    double extra;                   /* intraMolCorrect: intraVircorrect */
    double virial[9];
} orc_ewald_out;

/* Self and charged-system energies (ctor, :69-79). */
void orc_ewald_self(const orc_sys *s, double alpha, double volume,
                    double *self_energy, double *charged_energy);
/* real_term (:296-527). */
void orc_ewald_real(const orc_sys *s, const int32_t *off, const int32_t *nb,
                    double alpha, double real_cut, int use_constraint,
                    double *force, orc_ewald_out *out);
/* intraMolCorrect (:709-813). */
void orc_ewald_intra(const orc_sys *s, double alpha, double volume,
                     int use_constraint, double *force, orc_ewald_out *out);
/* surfDipoleCorrect (:816-937). */
void orc_ewald_surface(const orc_sys *s, double volume, int use_constraint,
                       double *force, orc_ewald_out *out);
/* set_up (:198-257): half-space k set.  Returns K, or -(needed) if cap is too
 * small.  kvec [3K], kmod2 [K], expk [K]. */
int orc_ewald_kvectors(double kcut, const double box[3], double alpha,
                       double *kvec, double *kmod2, double *expk, int cap);
/* reciprocal_term (:529-698) with compute_long's zeroing (:178-196). */
void orc_ewald_recip(const orc_sys *s, int K, const double *kvec,
                     const double *kmod2, const double *expk, double alpha,
                     double volume, int use_constraint, double *force,
                     orc_ewald_out *out);

/* ---- integrator pieces (src/integrators/LFIntegrator.cpp) ---------------- */

/* Per-molecule SHAKE with the Nose-Hoover molecular friction, NVT branch of
 * LFIntegrator::shake (:1678-1828, :1996-1997).  bond2 holds the 3 squared
 * constraint lengths per molecule ([3*n_mols]) as initialise() builds them
 * (:143-167).  Water molecules only (3 sites, cyclic).  pos, vt_minus,
 * velocity, displacement are updated in place; mol_mom is Molecule::momenta
 * from the previous step.  Returns 0, or -1 on "Shake constraint failure",
 * -2 when a molecule does not converge in max_count sweeps (the reference
 * exits in both cases). */
int orc_shake(const orc_sys *s, double *pos, double *vt_minus,
              double *velocity, double *displacement, const double *force,
              const double *bond2, const double *mol_mom, double fnh,
              int nose_hoover, double dt, double tol, int max_count,
              double virial[9]);

/* Molecular centres, momenta and kinetic energies after the update
 * (:650-686).  Writes com [3*n_mols] and mol_mom [3*n_mols]. */
void orc_mol_kinetics(const orc_sys *s, const double *pos,
                      const double *velocity, double *com, double *mol_mom,
                      double *atom_ke2, double *mol_ke2);

/* Leap-frog Nose-Hoover friction update (:688-722).  *fnh_minus is fnhmi. */
void orc_nh_friction(double ke2, double ndf, double temperature,
                     double coupl, double dt, double *fnh, double *fnh_minus);

/* Unconstrained leap-frog step with the self-consistent NH velocity loop
 * (:587-643, NVT).  Returns 0, or -1 for "NH iteration Error" (NHcount==10). */
int orc_lf_atoms(int n, double *pos, double *vt_minus, double *velocity,
                 double *displacement, const double *force,
                 const double *mass, double fnh, int nose_hoover, double dt);

/* Martyna-Tuckerman-Klein chain (NHIntegrator::chain,
 * src/integrators/NHIntegrator.cpp:1814-1912) for the given kinetic energy.
 * vnh,gnh,qnh have 4 entries; vlog/glog the barostat.  Returns the velocity
 * scale factor. */
double orc_nh_chain(double ekin, double temperature, double dt, double ndf,
                    double press_diff, double volume, double wbar,
                    double *vnh, double *gnh, const double *qnh, double *vlog,
                    double *glog);

/* Seeded Maxwell-Boltzmann velocities (Ensemble::init_velocity/gauss/drand,
 * src/Ensemble.cpp:1752-1864).  per_atom selects useAtomThermo.  The LCG
 * state is *seed and is advanced. */
void orc_init_velocity(const orc_sys *s, int per_atom, int ndf,
                       double temperature, unsigned int *seed,
                       double *velocity);

/* ---- one full leap-frog NVT step for rigid point-charge water ------------ */

typedef struct {
    double cutoff, cut_buff, alpha, kcut, dt, temperature, coupl, tol;
    int32_t max_count, use_cells, long_steps, surface;
    int32_t nose_hoover, atom_thermo;   /* thermostat flags (useAtomThermo) */
    int32_t n_types;
    const double *sigma2, *eps4;  /* LJ table */
    const double *bond2;          /* [3*n_mols] squared constraint lengths */
    /* 1: SHAKE of rigid 3-site water; 0: the unconstrained per-atom update
     * with the self-consistent Nose-Hoover loop (LFIntegrator.cpp:587-643) */
    int32_t constraint;
    int32_t ewald;                /* point-charge Ewald terms on */
    int32_t switch_on;            /* LJ energy shifted by E(rc) (LJForce.cpp:152-168) */
} orc_md_params;

typedef struct {
    /* state owned by the caller, updated in place */
    double *pos, *vt_minus, *velocity, *displacement, *force, *com, *mol_mom;
    double fnh, fnh_minus;
    int64_t step;
    int64_t list_updates;
    /* scratch owned by the caller */
    int32_t *off, *nb; int64_t cap;
    double *kvec, *kmod2, *expk; int32_t K;
    /* last step's energies */
    double e_lj, e_real, e_long, e_self, e_intra, e_surf, mol_ke2, atom_ke2;
    double virial[9];
} orc_md_state;

/* LFIntegrator::run body (:278-941) for LJ (+ point-charge Ewald) with SHAKE
 * or the unconstrained update, NH(NVT), minus output and pressure averaging.  Returns 0 or a negative
 * error (pair list capacity, SHAKE). */
int orc_md_step(const orc_sys *s_in, const orc_md_params *p, orc_md_state *st);

/* LFIntegrator::start (:991-1174): first force evaluation and
 * vtMinus = v - dt/2 F/m.  The pair list in st must be built already. */
int orc_md_start(const orc_sys *s_in, const orc_md_params *p, orc_md_state *st);

/* ---- sheared replicas (molecular SLLOD, Lees-Edwards) ---------------------
 * No counterpart in the reference: the project's definition, reducing to
 * orc_md_step (brute-force list) at zero shear rate.  orc_sys.tilt is the
 * x offset of the image across +y. */

/* Reciprocal vectors of the sheared cell with |k| < kcut, one per +-k pair.
 * Returns K or -(needed) when cap is too small. */
int orc_ewald_kvectors_le(double kcut, const double box[3], double tilt, double alpha,
                          double *kvec, double *kmod2, double *expk, int cap);
/* Builds the brute-force list and the k set (capacity kcap) at tilt, then
 * does orc_md_start. */
int orc_sllod_start(const orc_sys *s_in, const orc_md_params *p, orc_md_state *st, double tilt, int kcap);
/* One step at shear rate gamma (1/ps).  *tilt and *since_build (steps since
 * the list was built) are advanced.  Rigid water only (p->constraint = 1). */
int orc_sllod_step(const orc_sys *s_in, const orc_md_params *p, orc_md_state *st, double gamma,
                   double *tilt, int64_t *since_build, int kcap);


/* ---- polarizable water ----------------------------------------------------- */

/* Local2Global::FrameofReference (src/utils/Local2Global.cpp:90-225): local
 * site dipoles dip_loc[3n] and quadrupoles quad_loc[6n] (xx xy xz yy yz zz,
 * AtomParam dipx.. quadzz) rotated into the lab frame of each water's sites
 * (H1, O, H2); quad receives the traced form divided by 3, as the reference
 * stores it.  Returns -1 unless every molecule has 3 sites. */
int orc_local2global(const orc_sys *s, const double *dip_loc, const double *quad_loc,
                     double *dip, double *quad);

typedef struct {
    double cutoff_ind;   /* get_cutoffInd() */
    double a_damp;       /* get_adamping(): Thole a */
    int32_t damp_tij;    /* use_damping_induction() */
    int32_t do_min;      /* do_induction_minimization(): CG, else Pi = alpha E0 */
} orc_ind_params;

/* Induction::InducedDipoles (src/forces/Induction.cpp:181-455) and
 * ConjugateGradient (:2401-2467): permanent field ei0[3n] of the lab-frame
 * multipoles (Thole-damped, non-excluded pairs within cutoff_ind, minimum
 * image, no Ewald) and the induced dipoles pi[3n] solving A pi = ei0 with
 * A = diag(1/alpha) - T over all pairs within cutoff_ind.  iters and
 * final_res report the CG iterations and the last residual norm.  Returns
 * -2 when a non-excluded pair lies inside one molecule (the reference stops),
 * -3 on allocation failure. */
int orc_induced_dipoles(const orc_sys *s, const double *dip, const double *quad, const double *polar,
                        const orc_ind_params *p, double *ei0, double *pi, int *iters, double *final_res);

/* Permanent multipoles (charge, dipole, quadrupole) under Ewald:
 * Multipole::compute and compute_long (src/forces/Multipole.cpp). */
typedef struct {
    double alpha;             /* get_alpha() */
    double real_cut;          /* get_cutoffEw() */
    int32_t use_constraint;   /* is_constraint_on() */
    int32_t surface;          /* compute_surf_correct() */
    int32_t torques;          /* compute_torques() */
} orc_mp_params;
typedef struct {
    double self_energy, charged_energy;   /* myEnergy.correctEnergy = their sum */
    double real_energy, intra_energy, surf_energy, long_energy;
    double real_deriv, mol_deriv, surf_deriv, surf_deriv2, long_deriv, long_deriv2;
    double tot_torque;        /* sum of |torque| over sites and parts (myEnergy.TotTorq) */
    double virial[9];         /* Force::virial after compute(): real space + surface */
    double long_virial[9];    /* the class's longVirial after compute_long() */
} orc_mp_out;
/* compute() (:163-216): self_correct (:338-395), real_term over the half list
 * (:398-949), intraMolCorrect (:1290-1750), surfDipoleCorrect (:1753-1945),
 * each part's torques turned into forces by torques2forces (:1946-2136).
 * dip[3n]/quad[6n] are the lab-frame multipoles orc_local2global returns.
 * force is accumulated.  Returns -1 unless every molecule has 3 sites, -3 on
 * allocation failure. */
int orc_multipole_short(const orc_sys *s, const int32_t *off, const int32_t *nb, const double *dip,
                        const double *quad, const orc_mp_params *p, double *force, orc_mp_out *out);
/* compute_long() / reciprocal_term (:218-240, :951-1279) over the k set of
 * orc_ewald_kvectors (Multipole::set_up builds the same set, :242-302). */
int orc_multipole_long(const orc_sys *s, int K, const double *kvec, const double *kmod2, const double *expk,
                       const double *dip, const double *quad, const orc_mp_params *p, double *force,
                       orc_mp_out *out);

/* Energies, forces and torques of the induced dipoles Pi, treated as fixed,
 * against the permanent multipoles under Ewald: Induction::compute after
 * InducedDipoles, and compute_long (src/forces/Induction.cpp). */
typedef struct {
    double alpha;             /* get_alpha() */
    double real_cut;          /* get_cutoffEw() */
    double a_damp;            /* Thole a (get_adamping()) */
    int32_t use_constraint;   /* is_constraint_on() */
    int32_t surface;          /* compute_surf_correct() */
    int32_t torques;          /* compute_torques() */
    int32_t damp_amoeba;      /* use_damping_indAmoeba(): Tinker-style damping of the real space */
    int32_t do_min;           /* do_induction_minimization(): adds the induced-induced force terms */
} orc_indf_params;
typedef struct {
    double self_energy, real_energy, intra_energy, surf_energy, long_energy;   /* myEnergy.*PolEnergy */
    double self_deriv, real_deriv, mol_deriv, surf_deriv, surf_deriv2, long_deriv, long_deriv2;
    double tot_torque;
    double virial[9];         /* Force::virial after compute() */
    double long_virial[9];    /* the class's longVirial after compute_long() */
} orc_indf_out;
/* compute() minus the dipole solve (:486-565): self_correct (:688-760),
 * real_term over every pair j > i within cutoffEw, minimum image (:762-1393),
 * intraMolCorrect (:1760-2186), surfDipoleCorrect (:2188-2399), each with
 * torques2forces (:2557-2744).  pind [3n] are the induced dipoles, polar [n]
 * the site polarizabilities.  force is accumulated. */
int orc_induction_short(const orc_sys *s, const double *dip, const double *quad, const double *polar,
                        const double *pind, const orc_indf_params *p, double *force, orc_indf_out *out);
/* compute_long / reciprocal_term (:567-590, :1395-1758). */
int orc_induction_long(const orc_sys *s, int K, const double *kvec, const double *kmod2, const double *expk,
                       const double *dip, const double *quad, const double *pind, const orc_indf_params *p,
                       double *force, orc_indf_out *out);

/* ---- bonded terms of flexible molecules ---------------------------------
 * out[12] = energy, Lustig first and second derivative (0 for angles), the
 * term's virial (virial[XY] = f.x r.y here, as these three routines index
 * it).  potential: 1 harmonic, 2 AMOEBA anharmonic.  force is accumulated. */
/* BondForce::compute (src/forces/BondForce.cpp:46-122); bonds [nb][3]. */
int orc_bond(const orc_sys *s, int64_t nb, const int32_t *bonds, const double *r0, const double *k,
             int potential, double *force, double out[12]);
/* AngleForce::compute (src/forces/AngleForce.cpp:35-126); angles [na][4],
 * theta0 in radians. */
int orc_angle(const orc_sys *s, int64_t na, const int32_t *angles, const double *theta0, const double *k,
              int potential, double *force, double out[12]);
/* UreyBradleyForce::compute (src/forces/UreyBradleyForce.cpp:41-101):
 * between atoms 0 and 2 of every molecule. */
int orc_ub(const orc_sys *s, double kl, double l0, double *force, double out[12]);

/* ---- pressure tensor ------------------------------------------------------
 * The sums of Integrator::computePressure (src/Integrator.cpp:205-270) for
 * one state: mc = sum f (x) (r - R_mol) with mc[XY] = f.y (r.x - R.x); ka =
 * sum over atoms of p p / m and km = sum over molecules of P P / M, full 3x3
 * with the symmetric components copied as the reference does; p = v m, P the
 * molecule's summed momentum, R its mass centre (com [3*n_mols]). */
void orc_pressure_sums(const orc_sys *s, const double *pos, const double *vel, const double *force,
                       const double *com, double mc[9], double ka[9], double km[9]);

#ifdef __cplusplus
}
#endif
#endif
