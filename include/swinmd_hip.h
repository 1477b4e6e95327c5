/*
 * swinmd_hip.h -- C ABI of libswinmd_hip.so, the MI355X (gfx950) implementation
 * of SwinMD's nonbonded force path.
 *
 * The reference reaches its nonbonded terms through the Force plugin interface
 * (src/Force.h:30-72: compute() accumulates into atoms[i].force and writes
 * energies, Lustig derivatives and the virial) and through
 * EwaldForce::compute_long (src/forces/EwaldForce.h:129); the integrator
 * (src/integrators/LFIntegrator.cpp) owns SHAKE and the Nose-Hoover update.
 * Each entry point below replaces one of those routines; a reference-side
 * Force subclass marshals its AoS Atom array into these calls.
 *
 * Conventions
 *  - Every function returns SMD_OK (0) or a negative SMD_E* code; the library
 *    never exits.  smd_last_error() describes the last failure of a context.
 *  - Host buffers are caller-owned; device buffers are owned by the context.
 *  - Arrays of 3-vectors are AoS doubles [3*n] (x0 y0 z0 x1 ...), the layout of
 *    the reference's forceArray (src/integrators/LFIntegrator.cpp:480-499).
 *  - Virials are 9 doubles in the order XX XY XZ YX YY YZ ZX ZY ZZ
 *    (src/NEMD_defs.h:45-53).
 *  - Force entry points take `double* f_accum`: when non-NULL the term's forces
 *    are added (+=) into that host array, as compute() adds into
 *    atoms[i].force; when NULL they are added into the context's device force
 *    accumulator (see smd_zero_forces / smd_get_forces) and nothing crosses
 *    PCIe.
 *  - A context is single-threaded; all work is ordered on one HIP stream that
 *    the context owns.  Units: nm, ps, amu, kJ/mol, charges pre-scaled by
 *    sqrt(138.93548501) as the reference's Atom::scaledCharge.
 */
#ifndef SWINMD_HIP_H
#define SWINMD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMD_ABI_VERSION 1

enum smd_status {
    SMD_OK = 0,
    SMD_EINVAL = -1,      /* bad argument or inconsistent topology */
    SMD_ENOMEM = -2,      /* device allocation failed */
    SMD_EDEVICE = -3,     /* HIP runtime error */
    SMD_ESTATE = -4,      /* call out of order (e.g. no pair list yet) */
    SMD_ECAPACITY = -5,   /* pair list row capacity exceeded; rebuild with more */
    SMD_ESHAKE = -6,      /* SHAKE constraint failure (reference: ERRORMSG) */
    SMD_ENOCONV = -7,     /* SHAKE did not converge in max_count sweeps */
    SMD_ENOTSUP = -8,     /* configuration the library does not implement */
    SMD_EIO = -9          /* input file missing or malformed (readers, state files) */
};

enum smd_prec { SMD_PREC_FP64 = 0, SMD_PREC_MIXED = 1 };
enum smd_vdw_kind { SMD_VDW_LJ = 0, SMD_VDW_MIE = 1, SMD_VDW_BUCK = 2, SMD_VDW_BUFF147 = 3 };
enum smd_list_mode { SMD_LIST_CELLS = 1, SMD_LIST_BRUTE = 0 };

typedef struct smd_ctx smd_ctx;

/* Topology: what Parameters/Ensemble::set_mol_system hold
 * (src/Ensemble.cpp:140-227).  Atoms of one molecule must be contiguous and
 * in Molecule::myAtoms order; mol_id must be non-decreasing.  The exclusion
 * table is Ensemble::exclusionTable (src/Ensemble.cpp:323-367) as CSR. */
typedef struct {
    int32_t n_atoms, n_mols, n_types;
    const int32_t* atom_type;      /* [n_atoms] */
    const int32_t* mol_id;         /* [n_atoms] */
    const double* scaled_charge;   /* [n_atoms] */
    const double* mass;            /* [n_atoms] */
    const int32_t* excl_offsets;   /* [n_atoms+1] */
    const int32_t* excl_atoms;     /* [excl_offsets[n_atoms]] */
} smd_topology;

/* ---- lifecycle ---------------------------------------------------------- */
int smd_abi_version(void);
const char* smd_strerror(int code);
/* Creates a context on `device` and uploads the topology (replaces the
 * allocation half of Ensemble::set_mol_system). */
int smd_ctx_create(smd_ctx** out, const smd_topology* topo, int device);
void smd_ctx_destroy(smd_ctx* ctx);
const char* smd_last_error(const smd_ctx* ctx);
int smd_sync(smd_ctx* ctx);

/* ---- state (Atom::position / velocity / Molecule::massCenter) ----------- */
/* Ensemble::compute_bound_box result (src/Ensemble.cpp:370-445). */
int smd_set_box(smd_ctx* ctx, const double box[3]);
int smd_set_positions(smd_ctx* ctx, const double* xyz, int64_t n_atoms);
int smd_get_positions(smd_ctx* ctx, double* xyz, int64_t n_atoms);
int smd_set_velocities(smd_ctx* ctx, const double* v, int64_t n_atoms);
int smd_get_velocities(smd_ctx* ctx, double* v, int64_t n_atoms);
/* Molecule::massCenter, read by the Lustig terms (LJForce.cpp:96,112). */
int smd_set_mol_centers(smd_ctx* ctx, const double* com, int64_t n_mols);
int smd_get_mol_centers(smd_ctx* ctx, double* com, int64_t n_mols);
/* Molecule::mass_center for every molecule (src/Molecule.h:61-67). */
int smd_compute_mol_centers(smd_ctx* ctx);
int smd_zero_forces(smd_ctx* ctx);
int smd_get_forces(smd_ctx* ctx, double* f, int64_t n_atoms);

/* ---- (a) cell grid and half pair list ----------------------------------- */
/* Ensemble::set_pairlist + update_pairlist (src/Ensemble.cpp:454-490):
 * CellManager grid and 13-cell stencil when use_cells and the box allows it
 * (src/utils/CellManager.cpp:43-314), else the brute-force i<j list
 * (src/Ensemble.cpp:492-525).  Pairs satisfy r^2 < (cutoff+cut_buff)^2 under
 * the reference's arithmetic, so the pair set is the reference's bit for bit. */
int smd_build_pairlist(smd_ctx* ctx, double cutoff, double cut_buff, int use_cells);
/* Sizes of the current list: total pairs, longest row, whether cells ran. */
int smd_pairlist_info(smd_ctx* ctx, int64_t* n_pairs, int32_t* max_row, int32_t* used_cells);
/* Per-atom rows (Atom::myPairList, src/Atom.h:61) in atom order and in the
 * reference's visiting order within a row: offsets [n_atoms+1], nb [cap]. */
int smd_get_pairlist(smd_ctx* ctx, int32_t* offsets, int32_t* nb, int64_t cap);
/* Row capacity of the device list (entries per atom); 0 picks automatically. */
int smd_set_pairlist_capacity(smd_ctx* ctx, int32_t row_capacity);

/* ---- (b) van der Waals ------------------------------------------------- */
typedef struct {
    int32_t kind;              /* smd_vdw_kind */
    double cutoff;             /* SimConfiguration::get_cutoff() */
    double switch_dist;        /* get_switch_distance() (not squared) */
    int32_t switch_on;         /* is_switch_on() */
    int32_t use_constraint;    /* is_constraint_on(): rij.rijm in Lustig terms */
    int32_t lustig;            /* compute_lustig() */
    int32_t coulomb;           /* compute_coulomb(): GROMACS-shifted Coulomb */
    /* LJ/Mie: LJPairParam sigma^2 and 4*eps per type pair [n_types^2]
     * (Parameters::set_lj_param, src/Parameters.cpp:1360-1372). */
    const double* sigma2;
    const double* eps4;
    double n_mie, m_mie;       /* MieForce exponents */
    /* Buckingham A, B, C [n_types^2] (BuckinghamParam) and CO2 O-O length. */
    const double* buck_a;
    const double* buck_b;
    const double* buck_c;
    double co2_bond;
    /* Buffered 14-7: per-type R and eps (AtomParam.sigma, .eps) and shape. */
    const double* type_sigma;
    const double* type_eps;
    double n_buff, m_buff, delta, gamma, hfactor;
    /* Long-range correction values, added the way compute() adds them. */
    double e_lrc, v_lrc, du_lrc, d2u_lrc;
} smd_vdw_params;

typedef struct {
    double energy;      /* Force::energy */
    double lj_energy;   /* value compute() stores in myEnergy.ljEnergy */
    double virial[9];   /* Force::virial; compute() does ensemble.virial -= virial */
    double ulus, dudv, d2udv;   /* myLustig.ljUlus / ljdUlus / ljd2Ulus */
} smd_vdw_out;

/* LJForce::compute (src/forces/LJForce.cpp:58-255), MieForce::compute
 * (src/forces/MieForce.cpp:62-241), BuckinghamForce::compute
 * (src/forces/BuckinghamForce.cpp:79-287), BuffvdWForce::compute
 * (src/forces/BuffvdWForce.cpp:93-343), over the current pair list. */
int smd_vdw_compute(smd_ctx* ctx, const smd_vdw_params* p, int prec, double* f_accum, smd_vdw_out* out);

/* LJForce::long_range_correct (src/forces/LJForce.cpp:260-370): eLrc, vLrc,
 * dUlrc, d2Ulrc from the LJ table and the exclusions (host arithmetic). */
int smd_lj_long_range_correct(smd_ctx* ctx, const double* sigma2, const double* eps4, double cutoff,
                              double out4[4]);

/* ---- (b)(c) Ewald for point charges ------------------------------------ */
typedef struct {
    double alpha;          /* get_alpha() */
    double real_cut;       /* get_cutoff() (EwaldForce.cpp:37) */
    double kcut;           /* get_k_cutoff() */
    int32_t use_constraint;
    int32_t surface;       /* compute_surf_correct() */
} smd_ewald_params;

typedef struct {
    double real_energy, real_deriv, real_deriv2;            /* real_term */
    double self_energy, charged_energy;                     /* ctor :69-79 */
    double intra_energy, intra_vir_correct, mol_deriv, mol_deriv2;   /* intraMolCorrect */
    double surf_energy, surf_deriv, surf_deriv2;            /* surfDipoleCorrect */
    double virial[9];                                       /* EwaldForce::virial */
} smd_ewald_out;

typedef struct {
    double energy, deriv, deriv2;   /* longEnergy, longderiv, long2ndDeriv */
    double virial[9];               /* longVirial (all nine entries) */
} smd_ewald_long_out;

/* EwaldForce ctor + set_up (src/forces/EwaldForce.cpp:21-89, :198-257):
 * k-vector set, exp(-k^2/4a^2)/k^2 and the self energies.  Returns K via
 * n_kvec when non-NULL. */
int smd_ewald_setup(smd_ctx* ctx, const smd_ewald_params* p, int32_t* n_kvec);
/* EwaldForce::compute (:128-176): real_term, intraMolCorrect and, when
 * enabled, surfDipoleCorrect. */
int smd_ewald_compute(smd_ctx* ctx, int prec, double* f_accum, smd_ewald_out* out);
/* EwaldForce::compute_long (:178-196) -> reciprocal_term (:529-698). */
int smd_ewald_compute_long(smd_ctx* ctx, int prec, double* f_accum, smd_ewald_long_out* out);
/* smd_ewald_compute_long without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The reciprocal-space scratch (per-atom phases and moments, per-k
 * coefficients, structure-factor partials) lives in one static device array
 * of SMD_STATIC_SCRATCH_BYTES per GPU, held until the process ends; a system
 * that needs more returns SMD_ECAPACITY.  Forces are added to the context's
 * device force accumulator (smd_get_forces).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, threads or
 * contexts) race. */
#define SMD_STATIC_SCRATCH_BYTES ((size_t)16 << 20)
int smd_ewald_compute_long_static(smd_ctx* ctx, int prec, smd_ewald_long_out* out);

/* ---- (e) SHAKE and Nose-Hoover ----------------------------------------- */
typedef struct {
    double dt;              /* get_timestep() */
    double temperature;     /* get_temperature() */
    double coupl;           /* get_coupl_const() (NH tau) */
    double tol;             /* get_tolerance() */
    int32_t max_count;      /* get_max_count() */
    int32_t nose_hoover;    /* thermotype == "NoseHoover" */
    int32_t atom_thermo;    /* use_atom_thermo() */
} smd_integrator_params;

/* Squared constraint lengths per molecule, 3 per water (H1-O, O-H2, H-H), as
 * LFIntegrator::initialise builds them (src/integrators/LFIntegrator.cpp:143-167). */
int smd_set_constraints(smd_ctx* ctx, const double* bond2, int64_t n_mols);
/* Leap-frog velocities: vtMinus (LFIntegrator::vtMinus). */
int smd_set_vt_minus(smd_ctx* ctx, const double* v, int64_t n_atoms);
int smd_get_vt_minus(smd_ctx* ctx, double* v, int64_t n_atoms);
/* Molecule::momenta read by SHAKE's NH friction term. */
int smd_set_mol_momenta(smd_ctx* ctx, const double* p, int64_t n_mols);
/* LFIntegrator::shake NVT branch (:1678-1828) with friction fnh, using the
 * device forces.  Updates positions, displacements, velocities and vtMinus;
 * writes the constraint virial (already divided by dt^2). */
int smd_shake(smd_ctx* ctx, const smd_integrator_params* p, double fnh, double virial_out[9]);
/* Unconstrained leap-frog with the self-consistent NH velocity loop
 * (:587-643). */
int smd_lf_atoms(smd_ctx* ctx, const smd_integrator_params* p, double fnh);
/* Molecular centres, momenta and kinetic energies (:650-686); out2 receives
 * {2*atomKinEnergy, 2*molKinEnergy} as the reference accumulates them. */
int smd_mol_kinetics(smd_ctx* ctx, double out2[2]);
/* MTK Nose-Hoover chain (NHIntegrator::chain,
 * src/integrators/NHIntegrator.cpp:1814-1912) on the device velocities:
 * state4 = {vnh[4], gnh[4], qnh[4], vlog, glog, wbar}; returns the scale. */
int smd_nh_chain(smd_ctx* ctx, double ekin, double temperature, double dt, double ndf, double press_diff,
                 double volume, double* chain_state, double* scale_out);

/* ---- (d) induced dipoles of polarizable water ------------------------------ */
/* Local-frame permanent dipoles [3n], quadrupoles [6n] (xx xy xz yy yz zz)
 * and polarizabilities [n] of every atom: AtomParam dipx..quadzz, polar as
 * Ensemble::set_mol_system copies them onto the atoms (src/Ensemble.cpp:190-201).
 * Molecules must be 3-site waters in the order H1, O, H2. */
int smd_set_multipoles(smd_ctx* ctx, const double* dip_local, const double* quad_local, const double* polar,
                       int64_t n_atoms);

typedef struct {
    double cutoff_ind;     /* get_cutoffInd(); must not exceed the pair-list cutoff */
    double a_damp;         /* get_adamping(): Thole a */
    int32_t damp_tij;      /* use_damping_induction(): damp T in A */
    int32_t do_min;        /* do_induction_minimization(): CG, else Pi = alpha E0 */
    double tol;            /* CG stops when |r| < tol; 0 -> 1e-14 (ConjugateGradient) */
    int32_t max_iter;      /* 0 -> 3n, the reference's bound */
    int32_t guess;         /* 0: Pi = 1 (the reference); 1: the previous solution; 2: alpha E0;
                              3: the learned guess (smd_mlp_*); 4: the previous solution plus
                              alpha (E0 - E0 of the previous guess-4 solve) */
    int32_t prec;          /* smd_prec: SMD_PREC_FP64, or SMD_PREC_MIXED (FP32 coupling table and
                              A v, FP64 vectors, scalars and residual; preconditioned solve only) */
    int32_t precond;       /* 0: the reference's CG; 1: CG preconditioned by M^-1 = diag(alpha);
                              2: CG preconditioned by the inverse of each 3-site molecule's 9x9
                              diagonal block of A (block Jacobi).  1 and 2 have no counterpart in
                              the reference, whose solve is unpreconditioned */
    int32_t tol_rms;       /* 0: stop when |r|_2 < tol (the reference, :2431-2457); 1: stop when
                              the RMS over polarizable sites of |alpha_i r_i| -- the dipole step
                              of a Jacobi iteration, in e nm -- is < tol (1e-6 D = 2.0819e-8
                              e nm); 1 requires precond 1 or 2 and 0 requires precond 0 */
} smd_induce_params;

/* 1 Debye in e nm: the unit of the tol_rms tolerance's configs. */
#define SMD_DEBYE_E_NM 0.020819434

typedef struct {
    int32_t iterations;    /* CG updates performed */
    int32_t converged;     /* a residual test ended the loop */
    double residual;       /* |r| (tol_rms 0) or the RMS dipole residual (tol_rms 1) after the last update */
} smd_induce_report;

/* Induction::compute's first steps (src/forces/Induction.cpp:503-507):
 * Local2Global::FrameofReference (src/utils/Local2Global.cpp:90-225), the
 * permanent field and A of Induction::InducedDipoles (:181-455) and the
 * ConjugateGradient solve (:2401-2467), on the current positions and pair
 * list.  A is applied without being stored. */
int smd_induce(smd_ctx* ctx, const smd_induce_params* p, smd_induce_report* report);
/* Lab-frame dipoles [3n] and quadrupoles [6n] (Local2Global::dip, Qxx..Qzz),
 * the permanent field Ei0 [3n] and the induced dipoles Pi [3n] of the last
 * smd_induce; any pointer may be NULL. */
int smd_get_induction(smd_ctx* ctx, double* dip, double* quad, double* e0, double* mu, int64_t n_atoms);
/* Average microseconds of one launch of the CG's A v (the coupling-table
 * product of the last minimised smd_induce or polarizable step, on the
 * current dipoles), reps launches back to back; the solution is left
 * unchanged.  SMD_ESTATE before any minimised solve.  No counterpart in the
 * reference. */
int smd_induce_time_apply(smd_ctx* ctx, int reps, double* avg_us);

/* ---- learned initial guess of the induced dipoles --------------------------- */
/* No counterpart in the reference, whose ConjugateGradient starts from Pi = 1
 * (src/forces/Induction.cpp:2404-2410).  smd_induce with guess 3 starts CG
 * from Pi_i = alpha_i E0_i + F_i^T (sd * y_i + mean), y_i the output of a
 * per-site network on a 32-value descriptor of site i's environment in its
 * molecular frame F_i, then refines to the same tolerance: the solve decides
 * the result, the network only the number of iterations.
 *   descriptor  alpha E0, alpha T alpha E0, alpha sum_j q_j r/r^3 of the
 *               intermolecular partners in four shells to 0.6 nm split by the
 *               sign of q_j, alpha, positive-site flag (needs a pair list of
 *               cutoff and a CG cutoffInd >= 0.6 nm; 3-site waters H1, O, H2)
 *   network     32 -> 128 -> 128 -> 128 -> 3, SiLU, BF16 weights and
 *               activations on the matrix cores, FP32 accumulation
 *   training    on the GPU: mean squared error of the standardised residual
 *               F_i (Pi_i - alpha_i E0_i) of converged solves, Adam with a
 *               cosine learning-rate decay; deterministic for a given seed
 * Parameters in the order W1[128][32] b1 W2[128][128] b2 W3[128][128] b3
 * W4[3][128] b4[3]; weights [out][in].  Normalisation: feature mean[32],
 * feature sd[32], output mean[3], output sd[3]. */
#define SMD_MLP_FEATURES 32
#define SMD_MLP_HIDDEN 128
#define SMD_MLP_OUTPUTS 3
#define SMD_MLP_NPARAMS 37635
#define SMD_MLP_NNORM 70

/* Allocates the network on the context (once), draws uniform weights
 * (+-sqrt(6/fan_in), the last layer x0.1) from the seed, zero biases, an
 * identity normalisation that smd_induce does not accept until
 * smd_mlp_train or smd_mlp_set_norm has run. */
int smd_mlp_init(smd_ctx* ctx, uint64_t seed);
int smd_mlp_set_params(smd_ctx* ctx, const float* params, int64_t count);
int smd_mlp_get_params(smd_ctx* ctx, float* params, int64_t count);
int smd_mlp_set_norm(smd_ctx* ctx, const float* norm, int64_t count);
int smd_mlp_get_norm(smd_ctx* ctx, float* norm, int64_t count);
/* The network on host rows: x [rows][32] standardised descriptors (rounded to
 * BF16), y [rows][3] standardised outputs. */
int smd_mlp_forward(smd_ctx* ctx, const float* x, int64_t rows, float* y);
/* Mean squared error over rows x 3 of the network's outputs against target
 * [rows][3], and its gradient [SMD_MLP_NPARAMS] in the parameter order
 * (the training kernels' forward and backward pass); either output may be NULL. */
int smd_mlp_loss_grad(smd_ctx* ctx, const float* x, const float* target, int64_t rows, double* loss, float* grad);
/* Raw descriptors [n][32] of the state of the last converged smd_induce. */
int smd_mlp_descriptors(smd_ctx* ctx, float* desc, int64_t n_atoms);
/* Appends the descriptors and targets of the last converged smd_induce to the
 * training set on the device; total_rows (may be NULL) gets its size. */
int smd_mlp_add_frame(smd_ctx* ctx, int64_t* total_rows);
int smd_mlp_clear_frames(smd_ctx* ctx);

typedef struct {
    int32_t epochs;
    int32_t batch;         /* rows per Adam step, >= 16 */
    float lr;              /* initial learning rate, decayed by a cosine to lr_final */
    float lr_final;
    uint64_t seed;         /* order of the rows in each epoch */
    int32_t renormalize;   /* recompute the normalisation from the training set (always on the first run) */
} smd_mlp_train_params;

typedef struct {
    int64_t samples;       /* rows of the training set */
    int32_t steps;         /* Adam steps */
    double loss_first;     /* mean squared error over the first epoch, standardised units */
    double loss_last;      /* over the last epoch */
    double ms;             /* device time of the training */
} smd_mlp_train_report;

int smd_mlp_train(smd_ctx* ctx, const smd_mlp_train_params* p, smd_mlp_train_report* report);

/* ---- permanent multipoles under Ewald: class Multipole ------------------ */
/* The sites' local multipoles come from smd_set_multipoles; each call below
 * first rotates them into the lab frame (Local2Global::FrameofReference, as
 * compute() and compute_long() do, src/forces/Multipole.cpp:176,222).  3-site
 * water (H1, O, H2) only; FP64 only. */
typedef struct {
    double alpha;             /* SimConfiguration::get_alpha() */
    double real_cut;          /* get_cutoffEw(); at most the pair list's cutoff + buffer */
    int32_t use_constraint;   /* is_constraint_on() */
    int32_t surface;          /* compute_surf_correct() */
    int32_t torques;          /* compute_torques() */
} smd_mp_params;

/* Multipole::compute() results (:163-216). */
typedef struct {
    double self_energy, charged_energy;   /* myEnergy.correctEnergy = their sum */
    double real_energy;                   /* myEnergy.realEnergy */
    double intra_energy;                  /* myEnergy.molCorrectEnergy */
    double surf_energy;                   /* myEnergy.surfCorrectEnergy */
    double real_deriv, mol_deriv;         /* myLustig.realDeriv, molDeriv */
    double surf_deriv, surf_deriv2;       /* myLustig.surfDeriv, surfSecDeriv */
    double tot_torque;                    /* sum of |torque| over sites and parts */
    double virial[9];                     /* Force::virial (caller applies ensemble.virial -= virial) */
} smd_mp_out;

/* Multipole::compute_long() results (:218-240). */
typedef struct {
    double energy;                        /* myEnergy.longEnergy */
    double deriv, deriv2;                 /* myLustig.longDeriv, longSecDeriv */
    double tot_torque;
    double virial[9];                     /* the class's longVirial (all nine entries) */
} smd_mp_long_out;

/* Induction's energy, force and torque terms for the induced dipoles of the
 * last smd_induce (src/forces/Induction.cpp). */
typedef struct {
    double alpha;             /* get_alpha() */
    double real_cut;          /* get_cutoffEw(); at most the pair list's cutoff + buffer */
    double a_damp;            /* Thole a (get_adamping()) */
    int32_t use_constraint;   /* is_constraint_on() */
    int32_t surface;          /* compute_surf_correct() */
    int32_t torques;          /* compute_torques() */
    int32_t damp_amoeba;      /* use_damping_indAmoeba(): Tinker-style real-space damping */
    int32_t do_min;           /* do_induction_minimization(): induced-induced force terms */
} smd_indf_params;

/* Induction::compute() results after the dipole solve (:486-565). */
typedef struct {
    double self_energy;                   /* myEnergy.correctPolEnergy */
    double real_energy;                   /* myEnergy.realPolEnergy */
    double intra_energy;                  /* myEnergy.molCorrectPolEnergy */
    double surf_energy;                   /* myEnergy.surfacePolEnergy */
    double self_deriv, real_deriv, mol_deriv;   /* myLustig.selfPolDeriv, realPolDeriv, molPolDeriv */
    double surf_deriv, surf_deriv2;       /* myLustig.surfPolDeriv, surfSecPolDeriv */
    double tot_torque;
    double virial[9];                     /* Force::virial */
} smd_indf_out;

/* Multipole ctor and set_up (:21-115, :242-302): parameters and the
 * half-space k set k^2 < kcut^2; *n_kvec receives its size. */
int smd_multipole_setup(smd_ctx* ctx, const smd_mp_params* p, double kcut, int32_t* n_kvec);
/* Multipole::compute (:163-216): self_correct, real_term over the pair list,
 * intraMolCorrect, surfDipoleCorrect, each with its torques turned into
 * forces (torques2forces :1946-2136).  Forces are added to f_accum [3n] when
 * it is not NULL.  prec: SMD_PREC_FP64. */
int smd_multipole_compute(smd_ctx* ctx, int prec, double* f_accum, smd_mp_out* out);
/* Multipole::compute_long / reciprocal_term (:218-240, :951-1279). */
int smd_multipole_compute_long(smd_ctx* ctx, int prec, double* f_accum, smd_mp_long_out* out);
/* Induction ctor and set_up (src/forces/Induction.cpp:19-158, :592-652). */
int smd_induction_setup(smd_ctx* ctx, const smd_indf_params* p, double kcut, int32_t* n_kvec);
/* Induction::compute after InducedDipoles (:486-565): self_correct
 * (:688-760), real_term over every pair within cutoffEw including the
 * excluded ones (:762-1393), intraMolCorrect (:1760-2186), surfDipoleCorrect
 * (:2188-2399), each with torques2forces (:2557-2744).  Uses the induced
 * dipoles of the last smd_induce. */
int smd_induction_compute(smd_ctx* ctx, int prec, double* f_accum, smd_indf_out* out);
/* Induction::compute_long / reciprocal_term (:567-590, :1395-1758). */
int smd_induction_compute_long(smd_ctx* ctx, int prec, double* f_accum, smd_mp_long_out* out);

/* ---- whole step: LFIntegrator::run ---------------------------------------
 * Rigid point-charge water (LJ + Ewald, SHAKE) or unconstrained LJ atoms
 * (the per-atom update with the self-consistent Nose-Hoover loop,
 * src/integrators/LFIntegrator.cpp:587-643), NVT. */
typedef struct {
    double cutoff, cut_buff;
    smd_ewald_params ewald;
    smd_integrator_params integ;
    int32_t use_cells;
    int32_t long_steps;     /* longElectTsFrequency */
    int32_t lustig;         /* config computeLustig: the LJ volume derivative enters
                               myAvLus.P (the Ewald terms' derivatives are always
                               evaluated, as the reference does) */
    int32_t prec;           /* smd_prec of the pair and k-space math */
    const double* sigma2;   /* LJ table [n_types^2] */
    const double* eps4;
    int32_t use_graph;      /* replay each step as one captured hipGraph */
    int64_t total_steps;    /* config nSteps (Ensemble::totsteps): the stress sums
                               take the steps after the first 15% of it */
    int32_t constraint_mode;   /* config constraintMode: 2 SHAKE of rigid 3-site
                                  water, 0 none (atoms updated one by one) */
    int32_t ewald_on;       /* computeEwald: point-charge Ewald terms (else LJ only) */
    int32_t switch_on;      /* switchOn: LJ energy shifted by E(rc) (LJForce.cpp:152-168) */
} smd_md_params;

typedef struct {
    int64_t steps, list_updates;
    double e_lj, e_real, e_long, e_self, e_intra, e_surf;
    double atom_ke2, mol_ke2, fnh;
    double virial[9];
} smd_md_report;

/* LFIntegrator::initialise/start (:79-225, :991-1174): builds the pair list
 * if none, evaluates forces and sets vtMinus = v - dt/2 F/m. */
int smd_md_init(smd_ctx* ctx, const smd_md_params* p);
/* nsteps iterations of LFIntegrator::run's body (:278-941) without output:
 * Verlet trigger and rebuild, drift removal, LJ (+ Ewald), SHAKE or the
 * unconstrained update, kinetics, NH friction, molecule rewrap.  Work stays
 * on the device. */
int smd_md_run(smd_ctx* ctx, int64_t nsteps, smd_md_report* out);

/* The last step's pressure tensors and the stress sums since smd_md_init:
 * Integrator::computePressure (src/Integrator.cpp:205-270) with the
 * shake-virial correction (src/integrators/LFIntegrator.cpp:805-814),
 * myAvLus.P (:730-743) and the atomStress/molStress sums of
 * Ensemble::accumulate (src/Ensemble.cpp:814-894), each 5 values: pressure,
 * shearStress, antiShearStress, firstNormStress, secondNormStress.  virial is
 * Ensemble::virial as the reference holds it (its long-range part without
 * the ZZ component, which EwaldForce::compute_long never copies,
 * src/forces/EwaldForce.cpp:194). */
typedef struct {
    double virial[9], mol_virial[9];
    double atom_pressure[9], mol_pressure[9];
    double mol_pt;
    int64_t counter, samples;
    double atom_stress[5], mol_stress[5];
    /* the last step's myLustig terms that enter mol_pt: ljdUlus, realDeriv,
     * longDeriv, molDeriv, surfDeriv */
    double lj_dudv, real_deriv, long_deriv, mol_deriv, surf_deriv;
} smd_md_pressure_out;

int smd_md_pressure(smd_ctx* ctx, smd_md_pressure_out* out);

/* ---- whole step: LFIntegrator::run for rigid polarizable water ----------
 * The force block of src/integrators/LFIntegrator.cpp:385-470 for a
 * polarizable config -- BuffvdWForce (or LJForce) compute, Multipole compute
 * and compute_long, Induction compute (the induced-dipole solve) and
 * compute_long -- then SHAKE, kinetics and the Nose-Hoover update
 * (:561-722) and the per-step molecule rewrap (:897-910), on the device.
 * Each step is three parts: the forces up to the solve's first iteration and
 * the tail after it are captured hipGraphs; the solve runs in captured
 * batches of cg_batch iterations until the device's stop flag is set (the
 * host reads it between batches).  Rigid 3-site water (SHAKE) only.  The
 * context must hold the multipoles (smd_set_multipoles) and constraints
 * (smd_set_constraints); the multipole and induction k sets are built here. */
typedef struct {
    double cutoff, cut_buff;    /* vdW cutoff; the list is built at cutoff + cut_buff */
    smd_vdw_params vdw;         /* SMD_VDW_BUFF147 or SMD_VDW_LJ (tables as smd_vdw_compute takes) */
    smd_mp_params mp;
    double mp_kcut;
    smd_indf_params indf;
    double ind_kcut;
    smd_induce_params induce;   /* the solve: start, preconditioner, tolerance, precision */
    smd_integrator_params integ;
    int32_t use_cells;
    int32_t prec;               /* smd_prec of the vdW, multipole and induction real-space pair
                                   math (the solve's is induce.prec) */
    int32_t use_graph;          /* replay the two force parts as captured hipGraphs */
    int64_t total_steps;        /* config nSteps (the stress sums' 15% rule) */
} smd_polar_md_params;

typedef struct {
    int64_t steps, list_updates;
    int64_t cg_iterations;      /* solve iterations summed over the steps of this call */
    int32_t last_iterations;    /* the last step's */
    double last_residual;       /* the last step's final residual (the solve's own measure) */
    double e_vdw;               /* myEnergy.ljEnergy of BuffvdWForce / LJForce */
    double e_mp_real, e_mp_self, e_mp_intra, e_mp_long;      /* Multipole's realEnergy, correctEnergy
                                                                (self part), molCorrectEnergy, longEnergy */
    double e_ind_real, e_ind_self, e_ind_intra, e_ind_long;  /* Induction's realPolEnergy, correctPolEnergy,
                                                                molCorrectPolEnergy, longPolEnergy */
    double atom_ke2, mol_ke2, fnh;
} smd_polar_md_report;

/* LFIntegrator::start for the polarizable model: list, k sets, one force
 * evaluation (with a complete solve) and vtMinus = v - dt/2 F/m. */
int smd_polar_md_init(smd_ctx* ctx, const smd_polar_md_params* p);
int smd_polar_md_run(smd_ctx* ctx, int64_t nsteps, smd_polar_md_report* out);
/* Average microseconds of one launch of the polarizable step's real-space
 * kernel "mp_real" (Multipole::real_term) or "ind_real" (Induction::
 * real_term), in the step's precision and form (the mixed "ind_real" over
 * the solve's coupling table when the step uses it), reps launches back to
 * back on the current state; the step's forces and sums are left
 * unchanged.  No counterpart in the reference. */
int smd_polar_time_kernel(smd_ctx* ctx, const char* name, int reps, double* avg_us);

/* ---- checkpoint state ---------------------------------------------------
 * Atom::realPos, the unwrapped positions of the restart file: set with the
 * positions by smd_set_positions (as Ensemble::read_coordinate does), then
 * advanced by every integration step and untouched by the molecule rewrap;
 * smd_set_real_positions overrides them (Ensemble::read_realPos,
 * src/Ensemble.cpp:672-710). */
int smd_set_real_positions(smd_ctx* ctx, const double* r, int64_t n_atoms);
int smd_get_real_positions(smd_ctx* ctx, double* r, int64_t n_atoms);
/* The Nose-Hoover friction of the MD step (fnh, fnhmi of LFIntegrator::run
 * :688-709).  The reference's restart files do not keep it; a lossless
 * continuation restores it with the set call.  No counterpart in the
 * reference. */
int smd_md_get_thermostat(smd_ctx* ctx, double* fnh, double* fnh_minus);
int smd_md_set_thermostat(smd_ctx* ctx, double fnh, double fnh_minus);

/* ---- measurement -------------------------------------------------------- */
/* When enabled, every launch of the named kernel families is bracketed by HIP
 * events on the context stream. */
int smd_profile_enable(smd_ctx* ctx, int on);
/* Kernel families: "pairlist", "pairs", "ewald_intra", "recip_sfac",
 * "recip_reduce", "recip_force", "shake", "kinetics".  total_ms and launches
 * are cumulative since enabling. */
int smd_profile_get(smd_ctx* ctx, const char* family, double* total_ms, int64_t* launches);
/* Average duration of one launch of an MD-step kernel family ("pairs",
 * "recip_sfac", "recip_reduce", "recip_force"), launched reps times back to
 * back between two events with the current step's inputs, so launch gaps
 * overlap the previous launch as in a profiler's kernel trace.  "pairs" times
 * the pair kernel alone (the mixed path's site packing runs once, untimed).
 * Forces and energy sums are restored afterwards.  No counterpart in the
 * reference. */
int smd_md_time_family(smd_ctx* ctx, const char* family, int reps, double* avg_us);

/* Launch shapes (no counterpart in the reference; results do not depend on
 * them beyond FP rounding order).  "pair_packed" (0/1, default 1): the MD
 * step's mixed LJ + Ewald pairs read packed FP32 hi/lo site records.
 * "recip_overlap" (0/1, default 1): the MD step's k-space on a second stream.
 * "polar_joint_long" (0/1, default 1): the polarizable step runs one k-space
 * pass for Multipole and Induction after the dipole solve when both use the
 * same alpha and kcut (0: a pass per class).
 * "polar_m_sums" (0/1, default 1): with the joint pass and the second
 * stream, the permanent multipoles' structure factors are summed beside the
 * dipole solve with vdW and Multipole's short range (0: after the solve, in
 * the joint pass).
 * "md_fused_shake" (0/1, default 1): the MD step runs SHAKE and the kinetic
 * and pressure sums as one kernel (0: two kernels).
 * "cg_lanes" (8, 16 or 32, default 16): lanes sharing one row of the CG's A v.
 * "cg_graph" (0/1, default 1): smd_induce replays each batch of CG
 * iterations from a captured hipGraph (0: direct launches, for profilers).
 * "cg_batch" (1..64, default 16): CG iterations per batch, i.e. between the
 * host's checks of the stop flag.
 * "pair_lanes": lanes sharing one atom's row
 * in the pair kernels (both the packed mixed kernel and the general one),
 * 4, 8, 16, 32 or 64, or 0 (default: 16 for FP32 pair math, 32 for FP64).
 * Drops a captured MD
 * graph.  SMD_EINVAL for an unknown name or value. */
int smd_set_tuning(smd_ctx* ctx, const char* name, int value);

/* ---- bonded terms of flexible molecules ---------------------------------
 * BondForce (src/forces/BondForce.cpp:46-122), AngleForce
 * (src/forces/AngleForce.cpp:35-126) and UreyBradleyForce
 * (src/forces/UreyBradleyForce.cpp:41-101), the terms Integrator::initialise
 * creates for unconstrained bonds and for computeharmoUB
 * (src/Integrator.cpp:162-192).  potential is the config's bendingType: 1
 * harmonic, 2 AMOEBA anharmonic (bonds and angles alike).  Tables are indexed
 * by type: BondParam r0 and k, AngleParam theta0 (radians) and k.  Every bond
 * and angle must join atoms of one molecule (SMD_ENOTSUP otherwise); the
 * Urey-Bradley term acts between the first and third atom of every molecule
 * (Molecule::myAtoms[0] and [2]). */
typedef struct {
    int32_t potential;
    int32_t n_bond_types, n_angle_types;
    const double* bond_r0;        /* [n_bond_types] */
    const double* bond_k;
    const double* angle_theta0;   /* [n_angle_types] */
    const double* angle_k;
    int64_t n_bonds, n_angles;
    const int32_t* bonds;         /* [n_bonds][3]: atom1 atom2 type (Parameters::bondArray) */
    const int32_t* angles;        /* [n_angles][4]: atom1 atom2 atom3 type (vertex atom2) */
    double kl_ub, l0_ub;          /* config klUB, l0UB */
} smd_bonded_params;

/* energy; the Lustig first and second volume derivatives the term stores
 * (BondDeriv/BondSecDeriv, ubDeriv/ubSecDeriv; 0 for angles); the term's own
 * virial, which the reference subtracts from Ensemble::virial. */
typedef struct {
    double energy, deriv, deriv2;
    double virial[9];
} smd_bonded_out;

int smd_bonded_setup(smd_ctx* ctx, const smd_bonded_params* p);
int smd_bond_compute(smd_ctx* ctx, double* f_accum, smd_bonded_out* out);
int smd_angle_compute(smd_ctx* ctx, double* f_accum, smd_bonded_out* out);
int smd_ub_compute(smd_ctx* ctx, double* f_accum, smd_bonded_out* out);

/* ---- input formats ------------------------------------------------------
 * Host-only readers of the reference's input files.  They need no context
 * and no GPU.  On failure they return SMD_EIO (SMD_EINVAL for bad arguments)
 * and, when err is not NULL, write a message naming the file, line and field
 * into err[0..errlen).  The reference reads the same files with unchecked
 * stream extraction and exits on a missing file (ERRMSG). */

/* SimConfiguration after SimConfiguration::init() (src/SimConfiguration.cpp:
 * 29-84) and read_config_file() (:695-932): the positional config.txt, one
 * record per line, the rest of each line ignored (SKIP_REST), booleans as
 * "t", set_ensemble_status (:951-978) and set_run_type (:933-949) applied.
 * Member names follow the reference's fields; flags are 0/1.  kCutoff is read
 * twice by the reference (cutoff line, then Ewald line); the second value is
 * the one kept. */
typedef struct {
    char sys_data_file[80], coordinate_file[80], velocity_file[80];
    char result_file[80], restart_file[80], trajectory_file[80], dump_file[80];
    char mol_trajectory_file[80], mol_velocity_file[80], thermo_type[80];
    double time_step;
    int32_t n_steps, start_steps, short_ts_freq, long_ts_freq;
    int32_t avg_ts_freq, print_ts_freq, backup_ts_freq;
    int32_t start_sampling, sampling_ts_freq, start_ttcf, ttcf_ts_freq, rdf_nbins;
    int32_t integrator_type;      /* 0 Gear, 1 velocity Verlet, 2 leap-frog */
    int32_t ensemble_status;      /* 0 NVT, 1 NPT, 2 NVE, 3 NPE */
    int32_t run_type;             /* 0 equilibrium, 1 compression, 2 shear, 3 elongation */
    int32_t constraint_mode;      /* 0 none, 1 Gaussian, 2 SHAKE */
    int32_t max_count, bending_type;
    uint32_t seed;
    double temperature, pressure, volume, energy, density;
    double coupl_thermo, coupl_baros, rate;
    double cutoff, cutoff_ew, cutoff_ind, k_cutoff, alpha, cut_buff, cell_side_len;
    double lx_box, ly_box, lz_box;
    double cfb, dfb, tfb, tolerance;
    double kl_ub, l0_ub, switch_dist, a_damp;
    double n_mie, m_mie, n_buff, m_buff, delta_buff, gamma_buff, h_factor;
    int32_t write_trajectory, is_new_start, read_velocity, use_atom_thermo, use_cell_list;
    int32_t constraint_on, harmonic_bend, amoeba_bend, harmonic_ub, long_correct, switch_on;
    int32_t coulomb, ewald, surf_correct, wolf, multipole, induction, torques;
    int32_t damp_induction, induction_min, damp_amoeba, buff_vdw, mie;
    int32_t ext_forces, fixed_atom, ttcf, rdf, msd, vacf, lustig;
    int32_t constant_volume, constant_pressure, constant_temperature, constant_energy;
    int32_t do_equilibrium, do_compression, do_shear, do_elongation;
} smd_config;

int smd_config_read(const char* path, smd_config* out, char* err, size_t errlen);

/* Parameters::read_sys_data (src/Parameters.cpp:76-382) with the section
 * headings checked as check_heading does (:1309-1324).  Atom parameters are
 * stored at their type index as the reader intends (the reference reads them
 * into a copy, :112, and keeps none).  Angle rest values are converted from
 * degrees to radians as the reference does (:186). */
typedef struct {
    int32_t n_mols, n_mol_types, n_atoms, n_atom_types;
    int32_t n_bonds, n_bond_types, n_angles, n_angle_types;
    int32_t n_dihedrals, n_dihedral_types, n_impropers, n_improper_types;
} smd_sysdata_sizes;

typedef struct {
    int32_t type;
    double mass, charge, eps, sigma;
    double dip[3];                /* x y z */
    double quad[6];               /* xx xy xz yy yz zz */
    double polar;
    char type_name[16], atom_name[16];
} smd_atom_param;

typedef struct { int32_t type, t1, t2; double r0, k; } smd_bond_param;
typedef struct {
    int32_t type, t1, t2, t3;
    double theta0, k;             /* theta0 in radians, as the reference stores it */
    double theta0_deg;            /* the file's value */
} smd_angle_param;
typedef struct { int32_t type, t1, t2, t3, t4; double k; int32_t d, n; } smd_torsion_param;
typedef struct { int32_t type, n_atoms; char name[16]; } smd_mol_type;

/* Arrays owned by the sysdata handle (valid until smd_sysdata_free).  Rows
 * are stored at the ID the file gives them, as the reference's arrays are. */
typedef struct {
    smd_sysdata_sizes sizes;
    const smd_atom_param* atom_params;         /* [n_atom_types] */
    const smd_bond_param* bond_params;         /* [n_bond_types] */
    const smd_angle_param* angle_params;       /* [n_angle_types] */
    const smd_torsion_param* dihedral_params;  /* [n_dihedral_types] */
    const smd_torsion_param* improper_params;  /* [n_improper_types] */
    const smd_mol_type* mol_types;             /* [n_mol_types] */
    const int32_t* atom_type;                  /* [n_atoms] */
    const int32_t* atom_mol;                   /* [n_atoms] */
    const int32_t* mol_type;                   /* [n_mols] */
    const int32_t* bonds;                      /* [n_bonds][3]: atom1 atom2 type */
    const int32_t* angles;                     /* [n_angles][4]: atom1 atom2 atom3 type */
    const int32_t* dihedrals;                  /* [n_dihedrals][5] */
    const int32_t* impropers;                  /* [n_impropers][5] */
} smd_sysdata_view;

typedef struct smd_sysdata smd_sysdata;
int smd_sysdata_read(const char* path, smd_sysdata** out, char* err, size_t errlen);
int smd_sysdata_view_get(const smd_sysdata* sd, smd_sysdata_view* out);
void smd_sysdata_free(smd_sysdata* sd);

/* A coordinate or velocity block as Ensemble::read_coordinate / read_velocity
 * / read_realPos read it (src/Ensemble.cpp:630-758): a count line that must
 * equal n_atoms, then one "x y z" line per atom.  skip_lines lines are
 * skipped first (the restart branch skips one line before each block,
 * :303-317). */
int smd_read_xyz(const char* path, int64_t n_atoms, int32_t skip_lines, double* out, char* err, size_t errlen);

/* The restart file as the Ensemble constructor reads it when isNewStart is
 * off (src/Ensemble.cpp:299-317): a skipped line, the coordinate block, a
 * skipped line, the unwrapped ("real") positions, a skipped line, the
 * velocities.  Each output is [n_atoms][3]. */
int smd_read_state(const char* path, int64_t n_atoms, double* pos, double* real_pos, double* vel, char* err,
                   size_t errlen);

/* ---- output files -------------------------------------------------------
 * Host-only writers of the reference's restart and pressure files, through
 * std::ostream with the reference's manipulators.  precision 0 keeps the
 * stream default (6 significant digits, what the reference writes); 17 makes
 * the restart lossless.  append 0 truncates the file first.  SMD_EIO when the
 * file cannot be written. */
/* Integrator::write_state (src/Integrator.cpp:275-302): the file
 * smd_read_state reads. */
int smd_write_state(const char* path, int64_t step, int64_t n_atoms, const double* pos, const double* real_pos,
                    const double* vel, int32_t precision, char* err, size_t errlen);
/* NHIntegrator::write_vtMinus (src/integrators/NHIntegrator.cpp:1782-1803):
 * a title line, then a coordinate block (read back by smd_read_xyz with
 * skip_lines 1, as read_vtMinus does, :1960-1994).  The reference writes the
 * atoms' velocities there; pass whichever array is to be kept. */
int smd_write_vt_minus(const char* path, int64_t step, int64_t n_atoms, const double* v, int32_t precision,
                       char* err, size_t errlen);
/* Ensemble::write_press_res (src/Ensemble.cpp:1153-1160): one line of the
 * atomic and molecular pressure, shear, antisymmetric and normal stresses. */
int smd_write_press_res(const char* path, int append, int64_t step, const double atom_pressure[9],
                        const double mol_pressure[9], char* err, size_t errlen);
/* Ensemble::write_pressac (:1161-1171): the stress sums over samples (no
 * line when samples is 0). */
int smd_write_pressac(const char* path, int append, int64_t step, const double atom_stress[5],
                      const double mol_stress[5], int64_t samples, char* err, size_t errlen);
/* Ensemble::write_pressureTensor (:1887-1912): both tensors and the mean
 * shear stresses (atomStress.shearStress / counter, molStress's likewise). */
int smd_write_pressure_tensor(const char* path, int append, int64_t counter1, const double atom_pressure[9],
                              const double mol_pressure[9], double atom_shear_sum, double mol_shear_sum,
                              int64_t counter, char* err, size_t errlen);

/* ---- batched sheared replicas (planar Couette flow) ----------------------
 * No counterpart in the reference, which has neither shear flow nor replica
 * batching: independent replicas of rigid 3-site water (site order H, O, H;
 * LJ on O-O, point-charge Ewald, SHAKE, Nose-Hoover on the molecular
 * peculiar kinetic energy), each at its own shear rate, stepped together.
 * The step of one replica is LFIntegrator::run (src/integrators/
 * LFIntegrator.cpp:278-941) for this model with molecular SLLOD: the
 * streaming velocity gamma*y along x at each molecule's centre, the term
 * -gamma P_y/M along x in the peculiar update, Lees-Edwards images (the
 * image across +y is offset by the tilt along x), the tilt advanced by
 * gamma*L*dt and kept in [-L/2, L/2), the reciprocal lattice of the sheared
 * cell with |k| < kcut, a brute-force pair list rebuilt when
 * 2|d| + |gamma| rlist t > cut_buff (d the non-affine displacement, t the
 * time since the build).  Zero shear rate gives the reference's step.
 * Units are the reference's (nm, ps, amu, kJ/mol, charges scaled by
 * sqrt(138.93548501)).  One step is one hipGraph launch for all replicas. */
typedef struct smd_batch smd_batch;
typedef struct {
    int32_t n_replicas;
    const int32_t* n_mols;        /* [R] molecules per replica */
    const double* box;            /* [R] cubic box edge (nm) */
    const double* shear_rate;     /* [R] gamma (1/ps) */
    const double* kcut;           /* [R] reciprocal-space cutoff (1/nm) */
    double mass_h, mass_o;        /* site masses (amu) */
    double q_h, q_o;              /* scaled site charges */
    double sigma_o, eps_o;        /* LJ between O sites (nm, kJ/mol); H sites have none */
    double r_oh, r_hh;            /* constraint lengths (nm) */
    double cutoff, cut_buff, alpha;   /* LJ and real-space cutoff, list buffer, Ewald alpha */
    double dt, temperature, coupl, shake_tol;
    int32_t max_count;            /* SHAKE sweeps */
    int32_t nose_hoover;          /* 1: NH on the molecular kinetic energy; 0: none */
    int32_t prec;                 /* smd_prec: FP64, or MIXED (FP32 pair and k-space math, FP64 sums) */
    int32_t nb_cap;               /* pair-list row capacity (0: from the density) */
} smd_batch_params;
typedef struct {
    double tilt, fnh;             /* after the last step */
    double e_lj, e_real, e_long, e_intra, e_self;   /* at the start of the last step */
    double atom_ke2, mol_ke2;     /* sum m v.v and sum P.P/M after the last step */
    double virial[9];             /* ensemble virial: -(LJ + real) + reciprocal + constraint */
    double mol_kin[9];            /* sum_m P_a P_b / M (peculiar) */
    double mol_correction[9];     /* sum_i F_i,b (r_i - R_m(i))_a at the start of the last step */
    int64_t steps, list_updates;
    int32_t n_kvec, max_row;      /* k-vectors of the last step; longest list row at the last build */
} smd_batch_report;
typedef struct {
    int32_t pair_lanes;           /* lanes per atom row of the pair kernel */
    int32_t sfac_chunk;           /* atoms per block of the structure-factor pass */
    int32_t kforce_tile;          /* k-vectors per LDS tile of the reciprocal force pass */
    int32_t overlap;              /* 1: pair and reciprocal passes on two graph branches */
    int32_t use_graph;            /* 1: one captured graph launch per step; 0: direct launches */
} smd_batch_tiles;
int smd_batch_create(smd_batch** out, const smd_batch_params* p, int device);
void smd_batch_destroy(smd_batch* b);
const char* smd_batch_last_error(const smd_batch* b);
/* Positions, velocities [3 * total atoms], replica after replica, and the
 * replicas' tilts [R]; molecules must be whole (centres are taken from
 * them).  Resets the thermostat and momenta as Ensemble's start does. */
int smd_batch_set_state(smd_batch* b, const double* pos, const double* vel, const double* tilt);
/* LFIntegrator::start for every replica: list, k set, forces, vtMinus. */
int smd_batch_start(smd_batch* b);
/* nsteps steps (one graph launch each); checks the device status at the end. */
int smd_batch_run(smd_batch* b, int64_t nsteps);
int smd_batch_get_state(smd_batch* b, double* pos, double* vel, double* vt_minus, double* force, double* tilt);
int smd_batch_get_reports(smd_batch* b, smd_batch_report* rep);   /* [R] */
int smd_batch_set_tiles(smd_batch* b, const smd_batch_tiles* t);
int smd_batch_get_tiles(const smd_batch* b, smd_batch_tiles* t);
/* Times every candidate tile shape on the current state (each kernel alone,
 * HIP events, reps launches) and keeps the fastest.  times_us receives
 * n_cand rows of (pair_lanes, sfac_chunk, kforce_tile, pair us, sfac us,
 * kforce us) when not NULL; *n_cand is its capacity in rows on entry. */
int smd_batch_autotune(smd_batch* b, int reps, double* times_us, int32_t* n_cand);
/* Average microseconds of one launch of "list", "pairs", "kvec", "sfac",
 * "kfin", "kforce" or "shake" on the current state (the list pass is timed
 * as a full rebuild; the state is left as it was). */
int smd_batch_time_kernel(smd_batch* b, const char* name, int reps, double* avg_us);

#ifdef __cplusplus
}
#endif
#endif
