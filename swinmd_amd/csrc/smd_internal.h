// smd_internal.h -- context layout and device helpers shared by the gfx950
// kernels of libswinmd_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/swinmd_hip.h"

namespace smd {

constexpr int WAVE = 64;
constexpr int KIN_PARTS = 32;   // SHAKE virial 9, kinetic sums 2, pressure sums 21
// Pair-list entries of the reference's half list carry this bit.
constexpr uint32_t OWNED = 0x80000000u;
constexpr int32_t INDEX_MASK = 0x7fffffff;
constexpr double PI = 3.14159265358979323846;       // src/NEMD_defs.h:77
constexpr double SQRT_PI = 1.7724538509055160273;   // src/NEMD_defs.h:78
constexpr double BOLTZ = 0.0083145112;              // src/NEMD_defs.h:108

// Reduction sites of block_part (see there).
enum RedSite : int {
    RS_BOND = 0, RS_ANGLE, RS_UB, RS_EW_INTRA, RS_EW_SURF, RS_RECIP, RS_MP_REAL, RS_MP_INTRA, RS_MP_SURF,
    RS_MP_LONG, RS_IND_REAL, RS_IND_INTRA, RS_IND_SURF, RS_IND_LONG, RS_PAIR_LJ, RS_PAIR_EW, RS_PACKED,
    // the torque kernel's dU/dV and torque sums, one pair per launch of a step
    RS_MP_DUDV, RS_MP_TRQ, RS_MP_DUDV_L, RS_MP_TRQ_L, RS_IND_DUDV, RS_IND_TRQ, RS_IND_DUDV_L, RS_IND_TRQ_L,
    RS_COUNT
};
constexpr int RED_NV = 25;
struct Red {
    double* base = nullptr;
    int cap = 0;                  // blocks per site
};
inline size_t red_doubles(int cap) { return (size_t)RS_COUNT * RED_NV * cap; }
// A finish that waits for red_flush (inside a step, where one kernel then
// adds every site's partials).
struct RedPending {
    int site;
    unsigned nb;
    int nv;
    double* dst;
};

// Device status bits (set by kernels, read by the host after a sync).
enum : unsigned {
    ST_LIST_OVERFLOW = 1u,
    ST_SHAKE_FAIL = 2u,
    ST_SHAKE_NOCONV = 4u,
    ST_NH_ITER = 8u,
};

// Slots of the device accumulator for scalar results (FP64, zeroed per use).
// The pair kernel writes its 13 vdW values (E, ELEC, DUDV, D2UDV, VIR[9]) and
// its 12 Ewald values (E, D1, D2, VIR[9]) contiguously from T_LJ_E / T_EW_E.
enum TermSlot {
    T_LJ_E = 0, T_LJ_ELEC, T_LJ_DUDV, T_LJ_D2UDV, T_LJ_VIR,                     // VIR: 9
    T_EW_E = T_LJ_VIR + 9, T_EW_D1, T_EW_D2, T_EW_VIR,                           // VIR: 9
    T_IN_E = T_EW_VIR + 9, T_IN_U, T_IN_D1, T_IN_D2,
    T_SU_QR, T_SU_QRR = T_SU_QR + 3, T_SU_VIR = T_SU_QRR + 3,                     // 3, 3, 9
    T_RC_E = T_SU_VIR + 9, T_RC_D1, T_RC_D2, T_RC_VIR,                            // VIR: 9
    T_RC_FF = T_RC_VIR + 9, T_RC_FFK2, T_RC_D2UV1, T_RC_D2URR,
    T_SH_VIR, T_KE_ATOM = T_SH_VIR + 9, T_KE_MOL, T_VOM, T_MTOT = T_VOM + 3,
    // MD step pressure (Integrator::computePressure): per-molecule sums, then
    // the step's tensors.  MC: sum f (x) (r - R_mol); KA, KM: atom and
    // molecule p p / m (xx xy xz yy yz zz)
    T_PR_MC, T_PR_KA = T_PR_MC + 9, T_PR_KM = T_PR_KA + 6, T_PR_VIR = T_PR_KM + 6,
    T_PR_ATOM = T_PR_VIR + 9, T_PR_MOL = T_PR_ATOM + 9, T_PR_MOLVIR = T_PR_MOL + 9, T_PR_MOLPT = T_PR_MOLVIR + 9,
    T_PR_LUS,   // [5] ljdUlus, realDeriv, longDeriv, molDeriv, surfDeriv of the step
    T_COUNT = T_PR_LUS + 5
};

// Stream-ordered event pairs for the profiling families.
struct Prof {
    bool on = false;
    struct Rec { std::string fam; hipEvent_t a, b; };
    std::vector<Rec> pending;
    std::map<std::string, std::pair<double, int64_t>> acc;
};

struct Box {
    double L[3];
    double h[3];     // L/2, as Ensemble::compute_bound_box stores halfL
    double inv[3];   // 1/L
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int n = 0, nmol = 0, ntypes = 0;
    Box box{};
    bool box_set = false;

    // topology
    int32_t* d_type = nullptr;
    int32_t* d_mol = nullptr;
    int32_t* d_mol_start = nullptr;   // [nmol+1]
    double* d_q = nullptr;
    double* d_mass = nullptr;
    int32_t* d_excl_off = nullptr;
    int32_t* d_excl = nullptr;
    std::vector<int32_t> h_excl_off, h_excl, h_type, h_mol_start;
    std::vector<double> h_mass, h_q;

    // state
    double* d_pos = nullptr;      // [3n]
    double* d_vel = nullptr;      // [3n] Atom::velocity
    double* d_vtm = nullptr;      // [3n] vtMinus
    double* d_disp = nullptr;     // [3n] Atom::displacement
    double* d_force = nullptr;    // [3n] device force accumulator
    double* d_com = nullptr;      // [3nmol]
    double* d_mom = nullptr;      // [3nmol] Molecule::momenta
    double* d_bond2 = nullptr;    // [3nmol]
    double* d_terms = nullptr;    // [T_COUNT]
    unsigned* d_status = nullptr;
    int* d_flag = nullptr;        // [8]: 0 rebuild-needed, 1 drift-apply, 2 kinetics ticket,
                                  // 3 SHAKE ticket, 4 surface-sum ticket, 5 multipole/induction site-sum ticket
    // Fixed-order block partials of the sums that feed the trajectory (drift
    // momentum, SHAKE virial, kinetic and pressure sums, surface dipoles): each
    // block stores its sums, and a fixed-order total replaces FP64 atomics, so
    // the trajectory does not depend on block scheduling.
    double* d_part = nullptr;
    Red red;                      // reduction arena (block_part, red_finish)
    bool red_defer = false;       // red_finish queues instead of launching (red_flush launches)
    std::vector<RedPending> red_pending;
    double* static_scratch = nullptr;   // this device's address of ewald.hip's static scratch array
    double* part_vom = nullptr;   // [4][nb_atoms]: k_step_head
    double* part_kin = nullptr;   // [KIN_PARTS][nb_mols64]: kinetics kernels
    double* part_sh = nullptr;    // [9][nb_mols]: k_shake
    double* part_su = nullptr;    // [16][nb_atoms]: surface sums
    int nb_atoms = 0;             // blocks of 256 over the atoms
    int nb_mols = 0;              // blocks of 256 over the molecules
    int nb_mols64 = 0;            // blocks of 64 over the molecules
    double* d_scalars = nullptr;  // [8]: fnh, fnh_minus, ...
    double* h_pinned = nullptr;   // pinned staging for small results
    int* h_cg_done = nullptr;     // mapped host memory the solve's stopping test writes: [0] stopped,
                                  // [1] iterations, [2] reason, [4..5] the residual (double)
    int* d_cg_done = nullptr;     // its device address

    // pair list (ELL rows, reference order within a row)
    int32_t* d_cell_of = nullptr;
    int32_t* d_cell_cnt = nullptr;    // [ncell+1]
    int32_t* d_cell_start = nullptr;  // [ncell+1]
    int32_t* d_cell_atoms = nullptr;  // [n]
    int32_t* d_list_seg = nullptr;    // [28][n] per-atom hit counts of the cell-pair segments
    uint32_t* d_list_mask = nullptr;  // [27][n][list_mask_words] hit bits of the cell-pair pass
    int list_mask_words = 0;
    int32_t* d_nb = nullptr;          // [n * stride]
    int32_t* d_nb_cnt = nullptr;      // [n] full row length
    int32_t* d_nb_own = nullptr;      // [n] entries tagged OWNED (the reference row)
    int32_t stride = 0;               // allocated row capacity
    int32_t want_stride = 0;          // user-requested capacity (0 = auto)
    bool list_valid = false;
    bool list_cells = false;
    int ncell[3] = {0, 0, 0};
    double cell_rlen[3] = {0, 0, 0};
    double list_cutoff = 0, list_buff = 0;
    int cell_alloc = 0;

    // Ewald setup
    bool ewald_ready = false;
    smd_ewald_params ew{};
    int K = 0;
    int nmax[3] = {0, 0, 0};
    double self_energy = 0, charged_energy = 0;
    double* d_kvec = nullptr;     // [3K]
    double* d_kmod2 = nullptr;    // [K]
    double* d_expk = nullptr;     // [K]
    int32_t* d_kn = nullptr;      // [3K] integer (nx, ny, nz)
    int4* d_kgrp = nullptr;       // [n_kgrp] runs of consecutive nz in one (nx, ny) row
    int32_t* d_ksplit = nullptr;  // [RF_WAVES+1] group ranges of the reciprocal force pass
    int n_kgrp = 0;
    double* d_sfac_part = nullptr;  // [chunks][K][10] per-chunk structure-factor sums
    double* d_kf = nullptr;         // [K] (b C, b S) in the force pass's precision
    double* d_e1 = nullptr;         // [n][3] complex exp(i 2 pi r_p / L_p), then [n] (q, r - R_mol)
    int sfac_chunks = 0;
    size_t sfac_lds[2] = {0, 0};    // dynamic LDS of the structure-factor pass (FP64, FP32)
    size_t force_lds[2] = {0, 0};   // dynamic LDS of the force pass (FP64, FP32)

    // vdW tables (device)
    double* d_vdw_a = nullptr;    // [ntypes^2] per kind meaning
    double* d_vdw_b = nullptr;
    double* d_vdw_c = nullptr;
    double* d_epos = nullptr;     // [3n] effective (virtual-site) positions
    double* d_fa = nullptr;       // [n] force fraction kept on the atom
    double* d_fb = nullptr;       // [n] fraction moved to the host atom
    int32_t* d_host = nullptr;    // [n]
    double* d_fvirt = nullptr;    // [3n] force on each (virtual) site before redistribution

    // induced dipoles
    bool multipoles_set = false;
    int npol = 0;                  // sites with a polarizability above POLMIN
    float4* d_cg32 = nullptr;      // mixed solve: FP32 copy of the CG direction [n]
    float4* d_ind_rec32 = nullptr; // mixed solve: hi/lo FP32 site records [2n]
    int4* d_ind_fix = nullptr;     // mixed A v: fixed-point positions and 1/sqrt(alpha) [n]
    double* d_e0_prev = nullptr;   // E0 of the previous solve with guess 4 [3n]
    int4* d_vrec = nullptr;        // fixed-point virtual-site records of the mixed buffered 14-7 [n]
    double* d_pc_blk = nullptr;    // block preconditioner: inverse blocks [nmol][45]
    float* d_shell32 = nullptr;    // the learned guess's charge shells from the mixed field pass [n][24]
    bool ind_shells_valid = false; // d_shell32 holds the shells of the current solve's state
    int mp_prec = 0;               // smd_prec of the multipole / induction real-space pair math
    float4* d_site32 = nullptr;    // mixed real space: FP32 site records [3n]
    float4* d_site32_ind = nullptr;  // the mixed field pass's copy of them [3n]
    float4* d_pind32 = nullptr;    // mixed real space: FP32 induced dipoles [n]
    double* d_dipl = nullptr;     // [3n] local-frame dipoles
    double* d_quadl = nullptr;    // [6n] local-frame quadrupoles
    double* d_polar = nullptr;    // [n]
    double* d_dipg = nullptr;     // [3n] lab-frame dipoles
    double* d_quadg = nullptr;    // [6n] lab-frame quadrupoles (traced / 3)
    double* d_e0 = nullptr;       // [3n] permanent field
    double* d_mu = nullptr;       // [3n] induced dipoles
    double* d_cg = nullptr;       // [9n] CG r, v, z
    double* d_cg_part = nullptr;  // [4][cg_parts] block partials
    int* d_cg_state = nullptr;    // [4] done, iterations, stop reason
    double* d_cg_scal = nullptr;  // [4] last residual norm
    int cg_graph = 1;                   // replay CG batches from a captured graph
    int cg_lanes = 16;                  // lanes per row of the CG's A v
    int cg_batch = 16;                  // CG iterations per captured batch
    hipGraphExec_t cg_exec = nullptr;   // a captured batch of CG iterations
    std::string cg_key;                 // the kernel arguments it was captured with
    int it_stride = 0;            // coupling table of the solve (induction.hip)
    int64_t it_cap = 0;
    int32_t* d_it_cnt = nullptr;
    int32_t* d_it_j = nullptr;
    double* d_it_a = nullptr;     // (a, b) pairs [n][it_stride]

    // permanent multipoles and induced dipoles under Ewald
    struct KSet {                   // half-space k set of one Ewald class
        int K = 0, chunks = 1;
        double* kvec = nullptr;     // [3K]
        double* kmod2 = nullptr;    // [K]
        double* expk = nullptr;     // [K]
        double* part = nullptr;     // [chunks][K][sums]
        double* kf = nullptr;       // [K][4] per-k force-pass coefficients
        int* kstep = nullptr;       // [K] 1 when k is the previous k with nz + 1 (same nx, ny)
        int ksplit = 1;             // k-ranges of the force pass
        double* fpart = nullptr;    // [ksplit][n][6] force and torque per k-range
        double alpha = 0, kcut = 0;   // the set is fixed by these and the box
    };
    bool mp_ready = false, indf_ready = false;
    smd_mp_params mp{};
    smd_indf_params indf{};
    KSet mp_ks, ind_ks;
    double* d_mp_trq = nullptr;     // [5][3n] torques: real space, intramolecular, surface, k-space, self
    double* d_mp_terms = nullptr;   // accumulator slots
    int cg_parts = 0;

    // MD driver
    bool md_ready = false;
    smd_md_params md{};
    std::vector<double> md_sigma2, md_eps4;
    int64_t md_step = 0, md_list_updates = 0;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    // Reciprocal space of the MD step runs on a second stream, forked after
    // the list rebuild's rewrap and joined before SHAKE; its forces go to
    // d_frecip, which SHAKE adds after the short-range terms.
    hipStream_t stream2 = nullptr;
    hipStream_t polar_main = nullptr;   // the step's stream while a polarizable k-space fork is open
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_cg = nullptr;       // after the polarizable step's solve batches
    hipEvent_t ev_fork2 = nullptr;    // the polarizable step's second fork (after the field pass)
    hipEvent_t list_fork = nullptr;   // recorded by the list build after its preamble when set
    double* d_frecip = nullptr;       // [3n]
    bool recip_store = false;         // k_recip_force stores into d_frecip instead of adding
    bool shake_frecip = false;        // k_shake adds d_frecip to the forces it reads
    int recip_overlap = 1;
    int polar_joint = 1;            // one k-space pass for Multipole and Induction in the polarizable step
    int polar_m_sums = 1;           // the polarizable step's part M also sums the permanent structure factors
    int md_fused_shake = 1;           // the MD step's SHAKE and kinetics in one kernel

    Prof prof;
    int pair_lanes = 0;           // lanes per atom row in the pair kernels (0: by precision)
    int pair_packed = 1;
    bool pack_reuse = false;
    // Ensemble::accumulate stress sums of the MD step: [0] counter,
    // [1] samples taken, [2..6] atomStress, [7..11] molStress
    double* d_stress = nullptr;
    // Atom::realPos, the unwrapped positions the restart file keeps
    // (advanced by every integration step, untouched by the rewrap)
    double* d_real = nullptr;
    // bonded terms of flexible molecules (bonded.hip)
    bool bonded_ready = false, bd_ub_ok = false;
    int bd_potential = 1, bd_n_bond_types = 0;
    double bd_kl = 0.0, bd_l0 = 0.0;
    void* d_bd_bonds = nullptr;     // int4 [n_bonds] grouped by molecule
    void* d_bd_angles = nullptr;    // int4 [n_angles]
    int32_t* d_bd_bond_off = nullptr;
    int32_t* d_bd_angle_off = nullptr;
    double* d_bd_par = nullptr;
    double* d_bd_out = nullptr;
    // mixed LJ + Ewald of the MD step from packed site records
    void* d_pack = nullptr;       // float4 [3n]: packed records, then offsets from the molecule centres
    // learned initial guess of the induced dipoles (learned_guess.hip)
    struct Mlp* mlp = nullptr;
    // polarizable MD step (polar_md.hip)
    smd_polar_md_params pmd{};
    smd_vdw_params pmd_vdw{};           // pmd.vdw with the step's own table copies
    std::vector<double> pmd_tab[4];     // sigma2, eps4, type sigma, type eps
    double* d_polar_terms = nullptr;    // [2][POLAR_TERMS]: Multipole's, Induction's accumulators
    hipGraphExec_t pmd_a = nullptr;     // captured forces up to the solve
    hipGraphExec_t pmd_a2 = nullptr;    // captured field pass (when parts M1 and M2 run beside the solve)
    hipGraphExec_t pmd_a3 = nullptr;    // captured solve start (idem)
    hipGraphExec_t pmd_m = nullptr;     // captured vdW on the second stream (M1)
    hipGraphExec_t pmd_m2 = nullptr;    // captured Multipole short range on the second stream (M2)
    hipGraphExec_t pmd_c = nullptr;     // captured induction terms and integration tail
    bool pmd_ready = false;
    int pmd_last_iters = 0;             // the previous step's solve iterations
    smd_induce_params ind_last{};   // parameters of the last smd_induce
    bool ind_last_converged = false;
    bool ind_table_built = false;   // a minimised solve has filled the coupling table
    bool ind_real_table = false;    // Induction's mixed real space may walk the current solve's coupling table
    bool mp_real_table = false;     // Multipole's mixed real space may walk the current solve's coupling table
};

}  // namespace smd

struct smd_ctx : public smd::Ctx {};

namespace smd {

int fail(Ctx* c, int code, const char* fmt, ...);
// learned guess (learned_guess.hip) and the induction helpers it uses (induction.hip)
void mlp_release(Ctx* c);
int mlp_guess(Ctx* c, double* x, const double* z, float* x32);
int induce_begin(Ctx* c, const smd_induce_params* p, bool frames);
int induce_field(Ctx* c, const smd_induce_params* p, bool frames);
int induce_start(Ctx* c, const smd_induce_params* p);
int induce_batches(Ctx* c, const smd_induce_params* p, int batches);
int induce_poll(Ctx* c, int* done, int* iters, int* reason, double* res);
int frames_run(Ctx* c);
constexpr int POLAR_TERMS = 128;   // doubles of one multipole/induction accumulator (>= M_COUNT)
int multipole_enqueue(Ctx* c, double* terms);
int induction_enqueue(Ctx* c, double* terms);
bool polar_joint_long(Ctx* c);
int polar_joint_enqueue(Ctx* c, double* mp_terms, double* ind_terms, bool with_mp_short, bool mp_sums_done);
int polar_mp_short_enqueue(Ctx* c, double* mp_terms, bool with_sums);
bool polar_fork_begin(Ctx* c);
void polar_fork_end(Ctx* c);
void polar_join(Ctx* c);
int polar_classes_check(Ctx* c);
int polar_terms_energies(Ctx* c, const double* mp_t, const double* ind_t, double* e);
int induce_direct_apply(Ctx* c, const smd_induce_params* p, double* x, double* z);
int hip_fail(Ctx* c, hipError_t e, const char* what);

#define SMD_HIP(c, expr)                                        \
    do {                                                        \
        hipError_t _e = (expr);                                 \
        if (_e != hipSuccess) return smd::hip_fail((c), _e, #expr); \
    } while (0)

// Profiling bracket around a launch.
// Enqueues on another stream for the scope (every launch helper uses c->stream).
struct StreamSwap {
    Ctx* c;
    hipStream_t keep;
    StreamSwap(Ctx* c_, hipStream_t s) : c(c_), keep(c_->stream) { c->stream = s; }
    ~StreamSwap() { c->stream = keep; }
};

struct ProfScope {
    Ctx* c;
    const char* fam;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(Ctx* c_, const char* f) : c(c_), fam(f) {
        if (c->prof.on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~ProfScope() {
        if (c->prof.on) {
            (void)hipEventRecord(b, c->stream);
            c->prof.pending.push_back({fam, a, b});
        }
    }
};

// ---- device helpers -------------------------------------------------------

// Ensemble::apply_pbc (src/Ensemble.h:319-334), half-open [-L/2, L/2): the
// reference's loops subtract or add L until the value is inside.  The first
// step is taken with selects (it is the only one for |x| < 1.5 L, which covers
// displacements between atoms of a wrapped box); further steps, bounded so a
// corrupt coordinate cannot hang a wave, run only when still outside.  Any
// |x| < 8 L gives the reference's result, operation for operation.
__device__ __forceinline__ double pbc1(double x, double L, double h) {
    x = x >= h ? x - L : x;
    x = x < -h ? x + L : x;
    if (x >= h || x < -h) {
        for (int it = 0; it < 7 && x >= h; ++it) x -= L;
        for (int it = 0; it < 7 && x < -h; ++it) x += L;
    }
    return x;
}

__device__ __forceinline__ void pbc(double& x, double& y, double& z, const Box& b) {
    x = pbc1(x, b.L[0], b.h[0]);
    y = pbc1(y, b.L[1], b.h[1]);
    z = pbc1(z, b.L[2], b.h[2]);
}

// Minimum image for a displacement known to lie within 1.5 L of the image
// it wraps to: one shift by the nearest multiple of L.  Equals pbc() except
// for |x| = L/2 exactly (where pbc() keeps -L/2); the pair kernels use it
// only on pairs of a valid list, whose separations inside the cutoff are far
// below L/2.
__device__ __forceinline__ void pbc_near(double& x, double& y, double& z, const Box& b) {
    x -= b.L[0] * rint(x * b.inv[0]);
    y -= b.L[1] * rint(y * b.inv[1]);
    z -= b.L[2] * rint(z * b.inv[2]);
}

// Vector3::length2 (src/utils/Vector3.h:190-192) with each product and sum
// rounded separately, as the reference's non-FMA x86 build evaluates it.
__device__ __forceinline__ double len2_exact(double x, double y, double z) {
#pragma clang fp contract(off)
    double xx = x * x;
    double yy = y * y;
    double zz = z * z;
    return (xx + yy) + zz;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// ---- FP32 erfc of the mixed-precision real-space kernels ----------------------
// erfc(z) = exp(-z^2) t P(t), t = 1 / (1 + p z): the form of Abramowitz &
// Stegun 7.1.26 with a degree-7 P whose coefficients were fitted (weighted
// least squares toward minimax on the relative error of erfc(z) exp(z^2) / t
// over z in [0, 5.5]; max relative error 2.3e-8 in exact arithmetic, 3.5e-7
// evaluated in FP32).  exp(-z^2) is the factor the force needs anyway, so the
// pair costs one exp2, one rcp and one rsq; taking it from the FP32 r^2 adds
// the 2 z^2 2^-24 relative error of that rounding.
constexpr float ERFC_P = 0.415f;
constexpr float ERFC_C[8] = {0.0771595775739555f, -0.36372916731130467f, 0.5426081378824027f,
                             -0.2276783459634979f, 0.3230029930806543f, 0.17506060444731647f,
                             0.23978895752318216f, 0.23378723961927758f};
constexpr float LOG2E_F = 1.4426950408889634f;

// erfc(alpha r) / r and exp(-alpha^2 r^2) from r^2 and 1/r.
__device__ __forceinline__ void ewald_erfc_f32(float alpha, float alpha2, float r2, float r, float r_1, float& e,
                                               float& ex) {
    ex = __builtin_amdgcn_exp2f(-alpha2 * r2 * LOG2E_F);
    const float t = __builtin_amdgcn_rcpf(1.0f + ERFC_P * (alpha * r));
    float p = ERFC_C[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) p = p * t + ERFC_C[k];
    e = t * p * ex * r_1;
}

// Stores the block's NV sums at part[k * stride + blockIdx.x] (no atomics).
// A fixed-order total of the blocks' partials then gives the same bits
// whatever order the blocks ran in.
template <int NV>
__device__ __forceinline__ void block_store(double (&v)[NV], double* part, int stride) {
    __shared__ double red[NV][16];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += red[threadIdx.x][w];
        part[(size_t)threadIdx.x * stride + blockIdx.x] = s;
    }
    __syncthreads();   // red may be reused by the next call
}

// Fixed-order sum of p[0..n) by one wave: lane l adds p[l], p[l + 64], ...
// in order, then a shuffle tree; every lane gets the same bits.
__device__ __forceinline__ double wave_fixed_sum(const volatile double* p, int n) {
    double s = 0.0;
    for (int k = threadIdx.x % WAVE; k < n; k += WAVE) s += p[k];
    return wave_sum(s);
}

// ---- deterministic block sums ---------------------------------------------
// Each kernel that reduces energies or virials over its grid owns a site of
// the context's reduction arena (at most RED_NV values for at most cap
// blocks).  block_part stores the block's sums there with plain stores; the
// host then enqueues red_finish on the same stream, one small kernel that
// adds the partials to the destination in block order.  The totals do not
// depend on block scheduling, and no block waits on another: a ticket and
// the device-wide release fence it needs cost more than the extra launch on
// a multi-XCD part.  Sites that can run at the same time are distinct.
__device__ __forceinline__ double* red_part(const Red& r, int site) { return r.base + (size_t)site * RED_NV * r.cap; }

// Fixed-order totals of block partials, one wave per row: dst[k] += the sum
// over b < nb of part[k * stride + b], lane l adding b = l, l + 64, ... in
// order (the loads unrolled, so several are in flight), then a shuffle tree.
// Rows from `split` on go to dst2[k - split] instead.
__device__ __forceinline__ void rows_total(const double* __restrict__ part, int stride, int nb, int nrows,
                                           double* dst, int split = 1 << 30, double* dst2 = nullptr) {
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int k = w; k < nrows; k += blockDim.x / WAVE) {
        const double* p = part + (size_t)k * stride;
        double s = 0.0;
#pragma unroll 8
        for (int i = lane; i < nb; i += WAVE) s += p[i];
        s = wave_sum(s);
        if (lane == 0) {
            if (k < split) dst[k] += s;
            else dst2[k - split] += s;
        }
    }
}

template <int NV>
__device__ __forceinline__ void block_part(double (&v)[NV], const Red& r, int site) {
    static_assert(NV <= RED_NV, "block_part: too many values for a reduction site");
    if ((int)gridDim.x > r.cap) return;   // grid-uniform; red_finish reports it
    block_store<NV>(v, red_part(r, site), r.cap);
}

__device__ __forceinline__ void atomic_add3(double* f, double x, double y, double z) {
    unsafeAtomicAdd(f + 0, x);
    unsafeAtomicAdd(f + 1, y);
    unsafeAtomicAdd(f + 2, z);
}

// ---- host-side module entry points (implemented per .hip file) ----------
int pairlist_build(Ctx* c, double cutoff, double cut_buff, int use_cells, bool on_device_flag);
int pairlist_alloc(Ctx* c, int32_t stride);
int vdw_run(Ctx* c, const smd_vdw_params* p, int prec, smd_vdw_out* out, bool read_back);
int vdw_packed_alloc(Ctx* c);
int pairs_lj_ewald(Ctx* c, const double* sigma2, const double* eps4, double cutoff, double alpha,
                   double real_cut, int use_constraint, int lustig, int prec, bool store_forces = false);
int pairs_lj(Ctx* c, double cutoff, int switch_on, int use_constraint, int lustig, int prec, bool store_forces);
int vdw_enqueue(Ctx* c, const smd_vdw_params* p, int prec);
int ewald_setup(Ctx* c, const smd_ewald_params* p);
int ewald_real(Ctx* c, int prec, int lustig);
int ewald_intra(Ctx* c, int lustig);
int ewald_surface(Ctx* c);
// Reciprocal-space passes: phase + structure factors, per-k reduction, forces.
enum : unsigned { RECIP_SFAC = 1u, RECIP_REDUCE = 2u, RECIP_FORCE = 4u, RECIP_ALL = 7u };
int ewald_recip(Ctx* c, int prec, int lustig, unsigned passes = RECIP_ALL);
int static_scratch_init(Ctx* c);
// Adds the nv block partials of a site's last launch (nblocks blocks) to dst,
// in block order, on the context's stream.
int red_finish(Ctx* c, int site, unsigned nblocks, int nv, double* dst);
// Adds rows [0, nrows) of block partials part[k * stride + b], b < nb, to
// dst in block order: one block of FINISH_THREADS on stream s, after the
// kernel that stored them (in place of a last-block ticket, whose
// device-wide release fence costs more than a launch on a multi-XCD part).
constexpr int FINISH_THREADS = 1024;
// Rows from `split` on go to dst2[k - split].
int rows_finish(Ctx* c, hipStream_t s, const double* part, int stride, int nb, int nrows, double* dst,
                int split = 1 << 30, double* dst2 = nullptr);
// Launches one kernel that adds every queued finish (in queue order, each in
// block order) and clears the queue; no-op when the queue is empty.
int red_flush(Ctx* c);
// Queues finishes for the scope (a step's kernels), dropping any left over.
struct RedDefer {
    Ctx* c;
    explicit RedDefer(Ctx* ctx) : c(ctx) { c->red_defer = true; c->red_pending.clear(); }
    ~RedDefer() { c->red_defer = false; c->red_pending.clear(); }
};
int shake_run(Ctx* c, const smd_integrator_params* p, bool fnh_on_device, double fnh, bool md_step = false);
int kinetics_run(Ctx* c, bool rewrap, const smd_integrator_params* md = nullptr);
int lf_atoms_run(Ctx* c, const smd_integrator_params* p, double fnh, bool md_step);
int md_step_enqueue(Ctx* c, bool do_long);
int md_head_enqueue(Ctx* c);
int md_tail_enqueue(Ctx* c);
int md_start_vtm(Ctx* c);
int zero_terms(Ctx* c);
int read_terms(Ctx* c, double* dst);
int check_status(Ctx* c);
int upload_vdw_tables(Ctx* c, const smd_vdw_params* p);

}  // namespace smd
