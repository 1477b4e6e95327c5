// integrate.hip -- SHAKE, Nose-Hoover and the leap-frog step driver.
//
//   LFIntegrator::shake (NVT)       src/integrators/LFIntegrator.cpp:1678-1828
//   unconstrained LF + NH loop      :587-643
//   centres, momenta, kinetic E     :650-686
//   NH friction                     :688-722
//   drift removal                   :354-379
//   Verlet trigger + rewrap         :306-341, :897-910
//   NHIntegrator::chain             src/integrators/NHIntegrator.cpp:1814-1912
//
// One thread per molecule for SHAKE and the molecular sums (the constraint
// sweeps are sequential per molecule in the reference and stay so here);
// whole-system sums are block-reduced FP64 atomics.  The step is a fixed
// sequence of kernels whose list rebuild is gated on a device flag, so it can
// be captured once into a hipGraph and replayed.
#include <type_traits>

#include "smd_internal.h"

namespace smd {

struct IntegArgs {
    int n, nmol;
    Box box;
    const int32_t* mol_start;
    const double* mass;
    double* pos;
    double* vtm;
    double* vel;
    double* disp;
    double* real;      // Atom::realPos
    double* force;
    const double* frecip;   // MD step with reciprocal space on the second stream: its forces, else null
    double* com;
    double* mom;
    const double* bond2;
    double* terms;
    double* scalars;   // 0 fnh, 1 fnh_minus, 2 list updates, 3 steps
    int* flag;
    unsigned* status;
    double dt, tol, temperature, coupl, drmax;
    int max_count, nose_hoover, atom_thermo;
    double fnh_host;
    int fnh_on_device;
    int md_step;      // inside the captured MD step: drift removal and flag reset in SHAKE
    double ndf;       // thermostat degrees of freedom (MD step)
    // MD step pressure and stress sums
    int surface, use_constraint, lustig;
    double* stress;          // [12] (Ctx::d_stress)
    double stress_after;     // samples count once the call counter exceeds 0.15 totsteps
    // fixed-order block partials (Ctx::part_*)
    double* part_vom;        // [4][nb_vom] drift momentum and mass of k_step_head
    double* part_kin;        // [KIN_PARTS][gridDim] kinetics kernels
    double* part_sh;         // [9][gridDim] k_shake
    int nb_vom;
    int zero_force;          // k_step_head also clears the forces (the polarizable step)
};

// The drift velocity of LFIntegrator::run (:354-379) from k_step_head's
// block partials: every block forms the same fixed-order totals, and the
// velocity is zero unless a component exceeds 1e-11.
__device__ __forceinline__ void drift_velocity(const IntegArgs& a, double (&vc)[3]) {
    __shared__ double tot[4];
    if (threadIdx.x < WAVE) {
        // the four sums of wave_fixed_sum's order, their loads issued together
        // (the partials come from an earlier kernel)
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = threadIdx.x; k < a.nb_vom; k += WAVE)
#pragma unroll
            for (int b = 0; b < 4; ++b) s[b] += a.part_vom[(size_t)b * a.nb_vom + k];
#pragma unroll
        for (int b = 0; b < 4; ++b) s[b] = wave_sum(s[b]);
        if (threadIdx.x == 0)
#pragma unroll
            for (int b = 0; b < 4; ++b) tot[b] = s[b];
    }
    __syncthreads();
    const double mt = tot[3];
    const double vx = tot[0] / mt, vy = tot[1] / mt, vz = tot[2] / mt;
    const bool on = fabs(vx) > 1e-11 || fabs(vy) > 1e-11 || fabs(vz) > 1e-11;
    vc[0] = on ? vx : 0.0;
    vc[1] = on ? vy : 0.0;
    vc[2] = on ? vz : 0.0;
}

// Head of the MD step: the Verlet trigger (any |displacement|^2 > cutBuff^2/4,
// LFIntegrator.cpp:306-341) and the block partials of the drift removal's
// momentum sums.  Block 0 clears the step's accumulators.  The rebuild flag
// was cleared by the previous step's integration kernel.
__global__ __launch_bounds__(256) void k_step_head(IntegArgs a) {
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < T_COUNT; k += blockDim.x) a.terms[k] = 0.0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool need = false;
    double v[4] = {0, 0, 0, 0};
    if (i < a.n) {
        if (a.zero_force) { a.force[3 * i] = 0.0; a.force[3 * i + 1] = 0.0; a.force[3 * i + 2] = 0.0; }
        const double x = a.disp[3 * i], y = a.disp[3 * i + 1], z = a.disp[3 * i + 2];
        need = len2_exact(x, y, z) > a.drmax;
        const double m = a.mass[i];
        v[0] = a.vtm[3 * i] * m; v[1] = a.vtm[3 * i + 1] * m; v[2] = a.vtm[3 * i + 2] * m; v[3] = m;
    }
    if (__ballot(need) && (threadIdx.x % WAVE) == 0) a.flag[0] = 1;
    block_store<4>(v, a.part_vom, a.nb_vom);
}

// SHAKE of one water (a.mol_start[m], three sites).  Stores the new
// positions, displacements, unwrapped positions, velocities and vtMinus, adds
// its constraint virial to v, and leaves the new positions and velocities,
// the forces it used and the masses in pn, vel, F and ms.
// vcm: the MD step's drift velocity (LFIntegrator::run :354-379), removed
// from vtMinus (the velocity is rewritten below from vtMinus, so only vtMinus
// needs it); zero outside the MD step.
__device__ __forceinline__ void shake_mol(const IntegArgs& a, int m, const double (&vcm)[3], double (&v)[9],
                                          double (&pn)[3][3], double (&vel)[3][3], double (&F)[3][3],
                                          double (&ms)[3]) {
    {
        const int a0 = a.mol_start[m];
        const double fnh = a.fnh_on_device ? a.scalars[0] : a.fnh_host;
        const double Px = a.mom[3 * m], Py = a.mom[3 * m + 1], Pz = a.mom[3 * m + 2];
        const double tol_2 = 2 * a.tol;
        double ro[3][3], rn[3][3], mr[3], vt[3][3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const int id = a0 + s;
            ms[s] = a.mass[id];
            mr[s] = 1.0 / ms[s];
            const double h = a.dt * mr[s];
            for (int d = 0; d < 3; ++d) {
                ro[s][d] = a.pos[3 * id + d];
                vt[s][d] = a.vtm[3 * id + d] - vcm[d];
                double Fd = a.force[3 * id + d];
                if (a.frecip) {   // the sum the reference forms: short-range terms, then k-space
                    Fd += a.frecip[3 * id + d];
                    a.force[3 * id + d] = Fd;
                }
                F[s][d] = Fd;
                double vp;
                if (a.nose_hoover) {
                    const double P = d == 0 ? Px : (d == 1 ? Py : Pz);
                    vp = vt[s][d] + h * (Fd - (fnh * P) / 3.0);
                } else {
                    vp = vt[s][d] + h * Fd;
                }
                rn[s][d] = ro[s][d] + a.dt * vp;
            }
        }
        // Squared lengths and old-bond vectors are loop invariants; the three
        // constraints are unrolled so every array index is a constant and the
        // per-molecule state stays in registers.
        double bl[3], od[3][3], msum[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int nx = b == 2 ? 0 : b + 1;
            bl[b] = a.bond2[3 * m + b];
            od[b][0] = ro[b][0] - ro[nx][0];
            od[b][1] = ro[b][1] - ro[nx][1];
            od[b][2] = ro[b][2] - ro[nx][2];
            pbc(od[b][0], od[b][1], od[b][2], a.box);
            msum[b] = mr[b] + mr[nx];
        }
        // Minimum image inside the sweeps only for a molecule that spans a
        // quarter box; otherwise apply_pbc is the identity on its bonds.
        bool wrap = false;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int nx = b == 2 ? 0 : b + 1;
#pragma unroll
            for (int d = 0; d < 3; ++d)
                wrap = wrap || fabs(od[b][d]) > 0.25 * a.box.L[d] || fabs(rn[b][d] - rn[nx][d]) > 0.25 * a.box.L[d];
        }
        // Gauss-Seidel sweeps (LFIntegrator.cpp:1499-1540).  The update is
        // predicated instead of branched: an inactive constraint moves its
        // atoms by exactly 0.0, and its quotient is discarded by a select.
        // The sweep loop is compiled twice, with and without the minimum
        // image, and a wave takes the plain one unless one of its molecules
        // needs wrapping.  The constraint virial sum_s f_s n_s (x) o is kept
        // as sum_s f_s n_s per constraint and multiplied by o once.
        int count = 0;
        bool done = false, failed = false;
        double wv[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        auto sweeps = [&](auto wrap_tag) {
            constexpr bool WRAP = decltype(wrap_tag)::value;
            while (!done && count < a.max_count && !failed) {
                done = true;
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const int nx = b == 2 ? 0 : b + 1;
                    double nX = rn[b][0] - rn[nx][0], nY = rn[b][1] - rn[nx][1], nZ = rn[b][2] - rn[nx][2];
                    if (WRAP && wrap) pbc(nX, nY, nZ, a.box);
                    const double diff = bl[b] - len2_exact(nX, nY, nZ);
                    const double oX = od[b][0], oY = od[b][1], oZ = od[b][2];
                    const double rno = oX * nX + oY * nY + oZ * nZ;
                    const bool act = fabs(diff) > bl[b] * tol_2;
                    failed = failed || (act && rno < bl[b] * a.tol);
                    const double q = diff / (2 * rno * msum[b]);
                    const double factor = act ? q : 0.0;
                    wv[b][0] += factor * nX; wv[b][1] += factor * nY; wv[b][2] += factor * nZ;
                    const double fX = factor * oX, fY = factor * oY, fZ = factor * oZ;
                    rn[b][0] += mr[b] * fX; rn[b][1] += mr[b] * fY; rn[b][2] += mr[b] * fZ;
                    rn[nx][0] -= mr[nx] * fX; rn[nx][1] -= mr[nx] * fY; rn[nx][2] -= mr[nx] * fZ;
                    done = done && !act;
                }
                count++;
            }
        };
        if (__any(wrap)) sweeps(std::true_type{});
        else sweeps(std::false_type{});
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int y = 0; y < 3; ++y) v[3 * x + y] += wv[b][x] * od[b][y];
        if (a.md_step && m == 0) a.flag[0] = 0;   // every consumer of the rebuild flag has run
        if (failed) atomicOr(a.status, ST_SHAKE_FAIL);
        else if (!done) atomicOr(a.status, ST_SHAKE_NOCONV);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const int id = a0 + s;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double p0 = a.pos[3 * id + d];
                const double dr = rn[s][d] - p0;
                pn[s][d] = p0 + dr;
                a.pos[3 * id + d] = pn[s][d];
                a.disp[3 * id + d] += dr;
                a.real[3 * id + d] += dr;
                const double vp = dr / a.dt;
                vel[s][d] = (vp + vt[s][d]) * 0.5;
                a.vel[3 * id + d] = vel[s][d];
                a.vtm[3 * id + d] = vp;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_shake(IntegArgs a) {
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double vcm[3] = {0.0, 0.0, 0.0};
    if (a.md_step) drift_velocity(a, vcm);
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < a.nmol) {
        double pn[3][3], vel[3][3], F[3][3], ms[3];
        shake_mol(a, m, vcm, v, pn, vel, F, ms);
    }
    // the constraint virial: block partials, added in block order by
    // k_rows_finish
    block_store<9>(v, a.part_sh, gridDim.x);
}

__global__ __launch_bounds__(256) void k_lf_atoms(IntegArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double vc[3] = {0.0, 0.0, 0.0};
    if (a.md_step) {
        drift_velocity(a, vc);
        if (i == 0) a.flag[0] = 0;   // every consumer of the rebuild flag has run
    }
    if (i >= a.n) return;
    const double fnh = a.fnh_on_device ? a.scalars[0] : a.fnh_host;
    const double mr = 1.0 / a.mass[i];
    // the drift removal of LFIntegrator::run (:354-379) on vtMinus and the
    // velocity (zero outside the MD step)
    for (int d = 0; d < 3; ++d) {
        a.vtm[3 * i + d] -= vc[d];
        a.vel[3 * i + d] -= vc[d];
    }
    double vp[3];
    if (a.nose_hoover) {
        double vel[3] = {a.vel[3 * i], a.vel[3 * i + 1], a.vel[3 * i + 2]};
        double velcheck = 1.0;
        int cnt = 0;
        while (velcheck > 1E-10 && cnt < 11) {
            double vv = 0.0, half[3];
            for (int d = 0; d < 3; ++d) {
                vp[d] = a.vtm[3 * i + d] + a.dt * (mr * a.force[3 * i + d] - fnh * vel[d]);
                half[d] = (vp[d] + a.vtm[3 * i + d]) * 0.5;
                const double dv = half[d] - vel[d];
                vv += dv * dv;
            }
            velcheck = sqrt(vv);
            for (int d = 0; d < 3; ++d) vel[d] = half[d];
            cnt++;
        }
        if (cnt == 10) atomicOr(a.status, ST_NH_ITER);
    } else {
        for (int d = 0; d < 3; ++d) vp[d] = a.vtm[3 * i + d] + a.dt * mr * a.force[3 * i + d];
    }
    for (int d = 0; d < 3; ++d) {
        const double dr = a.dt * vp[d];
        a.pos[3 * i + d] += dr;
        a.disp[3 * i + d] += dr;
        a.real[3 * i + d] += dr;
        a.vel[3 * i + d] = (vp[d] + a.vtm[3 * i + d]) * 0.5;
        a.vtm[3 * i + d] = vp[d];
    }
}

// LF Nose-Hoover friction update (LFIntegrator::run :688-709): fnh, fnhmi.
__device__ __forceinline__ void nh_friction(const IntegArgs& a, double ke_atom, double ke_mol) {
    a.scalars[3] += 1.0;
    if (!a.nose_hoover) { a.scalars[0] = 0.0; return; }
    const double ke = a.atom_thermo ? ke_atom : ke_mol;
    const double qq = a.ndf * BOLTZ * a.temperature * a.coupl * a.coupl;
    const double invqq = 1 / qq;
    const double fm = a.scalars[1];
    const double fp = fm + a.dt * invqq * (ke - a.ndf * BOLTZ * a.temperature);
    a.scalars[0] = 0.5 * (fp + fm);
    a.scalars[1] = fp;
}

// Integrator::computePressure (src/Integrator.cpp:205-270), the shake-virial
// correction of LFIntegrator::run (:805-814), myAvLus.P (:730-743) and the
// stress sums of Ensemble::accumulate (src/Ensemble.cpp:814-894) for the MD
// step, from the step's sums; one thread.  Ensemble::virial is the force
// terms' virials subtracted, the long-range virial added without its ZZ
// component (EwaldForce::compute_long copies i < ZZ, EwaldForce.cpp:194).
__device__ void md_pressure(const IntegArgs& a, const double* t) {
    const double V = a.box.L[0] * a.box.L[1] * a.box.L[2];
    const int sym[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};   // (xx xy xz yy yz zz) -> 3x3
    double vir[9], atomP[9], molP[9];
    const double dt2 = a.dt * a.dt;
    for (int k = 0; k < 9; ++k) {
        vir[k] = -t[T_LJ_VIR + k] - t[T_EW_VIR + k] - (a.surface ? t[T_SU_VIR + k] : 0.0) +
                 (k < 8 ? t[T_RC_VIR + k] : 0.0);
        const double molv = vir[k] - t[T_PR_MC + k];
        atomP[k] = (t[T_PR_KA + sym[k]] + vir[k]) / V;
        molP[k] = (t[T_PR_KM + sym[k]] + molv) / V;
        a.terms[T_PR_MOLVIR + k] = molv;
    }
    for (int k = 0; k < 9; ++k) {
        const double sh = t[T_SH_VIR + k] / dt2;
        vir[k] += sh;
        atomP[k] += sh / V;
        a.terms[T_PR_VIR + k] = vir[k];
        a.terms[T_PR_ATOM + k] = atomP[k];
        a.terms[T_PR_MOL + k] = molP[k];
    }
    // molVirt with the terms this step has: LJ, Ewald real, intramolecular,
    // surface and reciprocal volume derivatives (LFIntegrator.cpp:730-733)
    // LJForce sets ljdUlus only with computeLustig (src/forces/LJForce.cpp:
    // 238-247); EwaldForce's derivatives are always formed, its molDeriv only
    // for unconstrained molecules (EwaldForce.cpp:780-784)
    const double ljdU = a.lustig ? -t[T_LJ_DUDV] / (3.0 * V) : 0.0;
    double surf = 0.0;
    if (a.surface) {   // EwaldForce::surfDipoleCorrect (:854-864)
        const double vr = 1.0 / V, coeff = 2.0 * PI * vr / 3.0;
        const double rr = t[T_SU_QR] * t[T_SU_QR] + t[T_SU_QR + 1] * t[T_SU_QR + 1] + t[T_SU_QR + 2] * t[T_SU_QR + 2];
        const double rR = t[T_SU_QR] * t[T_SU_QRR] + t[T_SU_QR + 1] * t[T_SU_QRR + 1] + t[T_SU_QR + 2] * t[T_SU_QRR + 2];
        surf = a.use_constraint ? -coeff * vr * rr + 2.0 * coeff * vr * rR / 3.0 : -coeff * vr * rr / 3.0;
    }
    const double molDeriv = a.use_constraint ? 0.0 : t[T_IN_D1];
    // reciprocal_term's longderiv, with the force loop's molecular part for
    // constrained molecules (EwaldForce.cpp:671-676)
    const double longDeriv = t[T_RC_D1] + (a.use_constraint ? t[T_RC_FF] : 0.0);
    const double molVirt = ljdU + surf + molDeriv + (longDeriv + t[T_EW_D1]) / (3.0 * V);
    const double molPt = (a.use_constraint ? t[T_KE_MOL] : t[T_KE_ATOM]) / (3.0 * V) - molVirt;
    a.terms[T_PR_MOLPT] = molPt;
    a.terms[T_PR_LUS] = ljdU;
    a.terms[T_PR_LUS + 1] = t[T_EW_D1];
    a.terms[T_PR_LUS + 2] = longDeriv;
    a.terms[T_PR_LUS + 3] = molDeriv;
    a.terms[T_PR_LUS + 4] = surf;
    double* st = a.stress;
    st[0] += 1.0;
    if (st[0] > a.stress_after) {
        st[1] += 1.0;
        st[2] += molPt;
        st[3] += (atomP[1] + atomP[3]) * 0.5;
        st[4] += (atomP[3] - atomP[1]) * 0.5;
        st[5] += (atomP[0] - atomP[4]);
        st[6] += (atomP[4] - atomP[8]);
        st[7] += (molP[0] + molP[4] + molP[8]) / 3.0;
        st[8] += (molP[1] + molP[3]) * 0.5;
        st[9] += (molP[3] - molP[1]) * 0.5;
        st[10] += (molP[0] - molP[4]);
        st[11] += (molP[4] - molP[8]);
    }
}

// One wave per block: the kinetic and pressure sums are wave reductions
// spread over as many CUs as there are waves of molecules.
constexpr int KIN_THREADS = 64;

// Rows of Ctx::part_kin: the kinetic sums (2), the pressure sums (21), the
// SHAKE virial (9).
constexpr int KP_KE = 0, KP_PR = 2, KP_SH = 23;

// The kinetics kernels' totals (kinetic sums, in the MD step the pressure
// sums and, after SHAKE, its virial) and, in the MD step, the Nose-Hoover
// update, the pressure tensors and the stress sums.
// One block of FINISH_THREADS after the kinetics kernel, on its stream.
__global__ __launch_bounds__(FINISH_THREADS) void k_kin_finish(IntegArgs a, int nb, int shake_rows) {
    rows_total(a.part_kin + (size_t)KP_KE * nb, nb, nb, 2, a.terms + T_KE_ATOM);
    if (a.md_step) rows_total(a.part_kin + (size_t)KP_PR * nb, nb, nb, 21, a.terms + T_PR_MC);
    if (shake_rows) rows_total(a.part_kin + (size_t)KP_SH * nb, nb, nb, 9, a.terms + T_SH_VIR);
    if (!a.md_step) return;
    // the step's sums, read once by the whole block (the epilogue's thread
    // then works from LDS, not from a chain of dependent global loads)
    __shared__ double tl[T_COUNT];
    __syncthreads();
    for (int k = threadIdx.x; k < T_COUNT; k += blockDim.x) tl[k] = a.terms[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        nh_friction(a, tl[T_KE_ATOM], tl[T_KE_MOL]);
        md_pressure(a, tl);
    }
}

// Molecular centres, momenta and kinetic sums (:646-686) as block partials;
// k_kin_finish adds them and, in the MD step, applies the thermostat update.
__global__ __launch_bounds__(KIN_THREADS) void k_kinetics(IntegArgs a, int rewrap) {
    double v[2] = {0.0, 0.0};
    double pr[21];   // MD step: sum f (x) (r - R) [9], atom p p / m [6], molecule P P / M [6]
#pragma unroll
    for (int k = 0; k < 21; ++k) pr[k] = 0.0;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < a.nmol) {
        const int s = a.mol_start[m], e = a.mol_start[m + 1];
        double M = 0.0;
        for (int i = s; i < e; ++i) M += a.mass[i];
        double P[3] = {0, 0, 0}, C[3] = {0, 0, 0};
        for (int i = s; i < e; ++i) {
            const double mi = a.mass[i];
            double vv = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double vd = a.vel[3 * i + d];
                P[d] += mi * vd;
                C[d] += mi * a.pos[3 * i + d];
                vv += vd * vd;
            }
            v[0] += vv * mi;
        }
        for (int d = 0; d < 3; ++d) C[d] = C[d] / M;
        v[1] = (P[0] * P[0] + P[1] * P[1] + P[2] * P[2]) / M;
        if (a.md_step) {
            // computePressure's per-molecule sums with the forces of this
            // step and the new positions, centres and momenta
            for (int i = s; i < e; ++i) {
                const double mi = a.mass[i];
                double p[3], f[3], dr[3];
                for (int d = 0; d < 3; ++d) {
                    p[d] = a.vel[3 * i + d] * mi;
                    f[d] = a.force[3 * i + d];
                    dr[d] = a.pos[3 * i + d] - C[d];
                }
                for (int x = 0; x < 3; ++x)
                    for (int y = 0; y < 3; ++y) pr[3 * x + y] += f[y] * dr[x];
                // p p / m as (p / m) p, one reciprocal per atom
                const double rm = 1.0 / mi;
                pr[9] += p[0] * rm * p[0]; pr[10] += p[0] * rm * p[1]; pr[11] += p[0] * rm * p[2];
                pr[12] += p[1] * rm * p[1]; pr[13] += p[1] * rm * p[2]; pr[14] += p[2] * rm * p[2];
            }
            const double rM = 1.0 / M;
            pr[15] += P[0] * rM * P[0]; pr[16] += P[0] * rM * P[1]; pr[17] += P[0] * rM * P[2];
            pr[18] += P[1] * rM * P[1]; pr[19] += P[1] * rM * P[2]; pr[20] += P[2] * rM * P[2];
        }
        for (int d = 0; d < 3; ++d) a.mom[3 * m + d] = P[d];
        if (rewrap) {
            double cx = C[0], cy = C[1], cz = C[2];
            pbc(cx, cy, cz, a.box);
            const double dx = cx - C[0], dy = cy - C[1], dz = cz - C[2];
            for (int i = s; i < e; ++i) { a.pos[3 * i] += dx; a.pos[3 * i + 1] += dy; a.pos[3 * i + 2] += dz; }
            C[0] = cx; C[1] = cy; C[2] = cz;
        }
        for (int d = 0; d < 3; ++d) a.com[3 * m + d] = C[d];
    }
    const int nb = gridDim.x;
    block_store<2>(v, a.part_kin + (size_t)KP_KE * nb, nb);
    if (a.md_step) block_store<21>(pr, a.part_kin + (size_t)KP_PR * nb, nb);
}

// MD step, rigid water: SHAKE of each molecule and, from the positions and
// velocities still in registers, its kinetic and pressure sums (the
// arithmetic of k_kinetics), then the step's epilogue in the last block.
__global__ __launch_bounds__(KIN_THREADS) void k_shake_kinetics(IntegArgs a, int rewrap) {
    double sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double v[2] = {0.0, 0.0};
    double pr[21];
#pragma unroll
    for (int k = 0; k < 21; ++k) pr[k] = 0.0;
    double vcm[3];
    drift_velocity(a, vcm);
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < a.nmol) {
        double pn[3][3], vel[3][3], F[3][3], ms[3];
        shake_mol(a, m, vcm, sv, pn, vel, F, ms);
        const int s0 = a.mol_start[m];
        const double M = (0.0 + ms[0] + ms[1]) + ms[2];
        double P[3] = {0, 0, 0}, C[3] = {0, 0, 0};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            double vv = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                P[d] += ms[s] * vel[s][d];
                C[d] += ms[s] * pn[s][d];
                vv += vel[s][d] * vel[s][d];
            }
            v[0] += vv * ms[s];
        }
        for (int d = 0; d < 3; ++d) C[d] = C[d] / M;
        v[1] = (P[0] * P[0] + P[1] * P[1] + P[2] * P[2]) / M;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            double p[3], dr[3];
            for (int d = 0; d < 3; ++d) {
                p[d] = vel[s][d] * ms[s];
                dr[d] = pn[s][d] - C[d];
            }
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y) pr[3 * x + y] += F[s][y] * dr[x];
            const double rm = 1.0 / ms[s];
            pr[9] += p[0] * rm * p[0]; pr[10] += p[0] * rm * p[1]; pr[11] += p[0] * rm * p[2];
            pr[12] += p[1] * rm * p[1]; pr[13] += p[1] * rm * p[2]; pr[14] += p[2] * rm * p[2];
        }
        const double rM = 1.0 / M;
        pr[15] += P[0] * rM * P[0]; pr[16] += P[0] * rM * P[1]; pr[17] += P[0] * rM * P[2];
        pr[18] += P[1] * rM * P[1]; pr[19] += P[1] * rM * P[2]; pr[20] += P[2] * rM * P[2];
        for (int d = 0; d < 3; ++d) a.mom[3 * m + d] = P[d];
        if (rewrap) {
            double cx = C[0], cy = C[1], cz = C[2];
            pbc(cx, cy, cz, a.box);
            const double dx = cx - C[0], dy = cy - C[1], dz = cz - C[2];
            for (int i = s0; i < s0 + 3; ++i) { a.pos[3 * i] += dx; a.pos[3 * i + 1] += dy; a.pos[3 * i + 2] += dz; }
            C[0] = cx; C[1] = cy; C[2] = cz;
        }
        for (int d = 0; d < 3; ++d) a.com[3 * m + d] = C[d];
    }
    const int nb = gridDim.x;
    block_store<9>(sv, a.part_kin + (size_t)KP_SH * nb, nb);
    block_store<2>(v, a.part_kin + (size_t)KP_KE * nb, nb);
    block_store<21>(pr, a.part_kin + (size_t)KP_PR * nb, nb);
}

__global__ __launch_bounds__(256) void k_mol_centers(int nmol, const int32_t* __restrict__ ms,
                                                     const double* __restrict__ mass, const double* __restrict__ pos,
                                                     double* __restrict__ com) {
    // Molecule::mass_center (src/Molecule.h:61-67)
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    double M = 0.0;
    for (int i = ms[m]; i < ms[m + 1]; ++i) M += mass[i];
    double C[3] = {0, 0, 0};
    for (int i = ms[m]; i < ms[m + 1]; ++i)
        for (int d = 0; d < 3; ++d) C[d] += mass[i] * pos[3 * i + d];
    for (int d = 0; d < 3; ++d) com[3 * m + d] = C[d] / M;
}

__global__ __launch_bounds__(256) void k_start_vtm(int n, const double* __restrict__ vel,
                                                   const double* __restrict__ force, const double* __restrict__ mass,
                                                   double* __restrict__ vtm, double dt) {
    // LFIntegrator::start (:1172-1173): vtMinus = v - 0.5 dt F/m
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mr = 1.0 / mass[i];
    for (int d = 0; d < 3; ++d) vtm[3 * i + d] = vel[3 * i + d] - 0.5 * dt * mr * force[3 * i + d];
}

// MTK chain for one call (NHIntegrator::chain, :1814-1912), one thread.
// st: vnh[4] gnh[4] qnh[4] vlog glog wbar ekin temp dt ndf press_diff volume
__global__ void k_nh_chain(double* st, double* scale_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double* vnh = st;
    double* gnh = st + 4;
    const double* qnh = st + 8;
    double vlog = st[12], glog = st[13];
    const double wbar = st[14];
    double ekin = st[15];
    const double temperature = st[16], dt = st[17], Nf = st[18], diff = st[19], volume = st[20];
    double scale = 1.0;
    double w[3];
    w[0] = 1.0 / (2.0 - pow(2.0, (1.0 / 3.0)));
    w[1] = 1.0 - 2.0 * w[0];
    w[2] = w[0];
    const double kT = BOLTZ * temperature;
    const double dtnc = dt / 5.0;
    const double gn1kt = (Nf + 1.0) * kT;
    const double odnf = 1.0 + 3.0 / Nf;
    double ex;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 3; j++) {
            const double dts = w[j] * dtnc;
            const double dts2 = 0.5 * dts, dts4 = 0.25 * dts, dts8 = 0.125 * dts;
            gnh[3] = (qnh[2] * vnh[2] * vnh[2] - kT) / qnh[3];
            vnh[3] = vnh[3] + gnh[3] * dts4;
            gnh[2] = (qnh[1] * vnh[1] * vnh[1] - kT) / qnh[2];
            ex = exp(-vnh[3] * dts8);
            vnh[2] = ex * (vnh[2] * ex + gnh[2] * dts4);
            gnh[1] = (qnh[0] * vnh[0] * vnh[0] - kT) / qnh[1];
            ex = exp(-vnh[2] * dts8);
            vnh[1] = ex * (vnh[1] * ex + gnh[1] * dts4);
            gnh[0] = (2.0 * ekin + wbar * vlog * vlog - gn1kt) / qnh[0];
            ex = exp(-vnh[1] * dts8);
            vnh[0] = ex * (vnh[0] * ex + gnh[0] * dts4);
            glog = (2.0 * ekin * odnf + 3.0 * volume * diff) / wbar;
            ex = exp(-vnh[0] * dts8);
            vlog = ex * (vlog * ex + glog * dts4);
            ex = exp(-(vnh[0] + vlog * odnf) * dts2);
            scale *= ex;
            ekin = ekin * ex * ex;
            glog = (2.0 * ekin * odnf + 3.0 * volume * diff) / wbar;
            ex = exp(-vnh[0] * dts8);
            vlog = ex * (vlog * ex + glog * dts4);
            gnh[0] = (2.0 * ekin + wbar * vlog * vlog - gn1kt) / qnh[0];
            ex = exp(-vnh[1] * dts8);
            vnh[0] = ex * (vnh[0] * ex + gnh[0] * dts4);
            gnh[1] = (qnh[0] * vnh[0] * vnh[0] - kT) / qnh[1];
            ex = exp(-vnh[2] * dts8);
            vnh[1] = ex * (vnh[1] * ex + gnh[1] * dts4);
            gnh[2] = (qnh[1] * vnh[1] * vnh[1] - kT) / qnh[2];
            ex = exp(-vnh[3] * dts8);
            vnh[2] = ex * (vnh[2] * ex + gnh[2] * dts4);
            gnh[3] = (qnh[2] * vnh[2] * vnh[2] - kT) / qnh[3];
            vnh[3] = vnh[3] + gnh[3] * dts4;
        }
    st[12] = vlog;
    st[13] = glog;
    st[15] = ekin;
    *scale_out = scale;
}

__global__ __launch_bounds__(256) void k_scale(int n3, double* __restrict__ v, const double* __restrict__ s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) v[i] *= *s;
}

// ---------------------------------------------------------------------------

static IntegArgs integ_args(Ctx* c, const smd_integrator_params* p) {
    IntegArgs a{};
    a.n = c->n;
    a.nmol = c->nmol;
    a.box = c->box;
    a.mol_start = c->d_mol_start;
    a.mass = c->d_mass;
    a.pos = c->d_pos;
    a.vtm = c->d_vtm;
    a.vel = c->d_vel;
    a.disp = c->d_disp;
    a.real = c->d_real;
    a.force = c->d_force;
    a.com = c->d_com;
    a.mom = c->d_mom;
    a.bond2 = c->d_bond2;
    a.terms = c->d_terms;
    a.scalars = c->d_scalars;
    a.flag = c->d_flag;
    a.status = c->d_status;
    a.part_vom = c->part_vom;
    a.part_kin = c->part_kin;
    a.part_sh = c->part_sh;
    a.nb_vom = c->nb_atoms;
    if (p) {
        a.dt = p->dt;
        a.tol = p->tol;
        a.temperature = p->temperature;
        a.coupl = p->coupl;
        a.max_count = p->max_count;
        a.nose_hoover = p->nose_hoover;
        a.atom_thermo = p->atom_thermo;
    }
    return a;
}

static double md_ndf(Ctx* c) {
    // Integrator ctor (src/Integrator.cpp:25-29): atomNDF is 6 Nmol - 3 with
    // constraints, else 3 N - 3
    if (c->md.integ.atom_thermo) return c->md.constraint_mode ? (double)(6 * c->nmol - 3) : (double)(3 * c->n - 3);
    return (double)(3 * c->nmol - 3);
}

int shake_run(Ctx* c, const smd_integrator_params* p, bool fnh_on_device, double fnh, bool md_step) {
    IntegArgs a = integ_args(c, p);
    a.fnh_on_device = fnh_on_device ? 1 : 0;
    a.fnh_host = fnh;
    a.md_step = md_step ? 1 : 0;
    a.frecip = c->shake_frecip ? c->d_frecip : nullptr;
    ProfScope ps(c, "shake");
    const int nb = (c->nmol + 255) / 256;
    hipLaunchKernelGGL(k_shake, dim3(nb), dim3(256), 0, c->stream, a);
    SMD_HIP(c, hipGetLastError());
    return rows_finish(c, c->stream, a.part_sh, nb, nb, 9, a.terms + T_SH_VIR);
}

int kinetics_run(Ctx* c, bool rewrap, const smd_integrator_params* md) {
    IntegArgs a = integ_args(c, md);
    if (md) {
        a.md_step = 1;
        a.ndf = md_ndf(c);
        a.surface = c->md.ewald_on ? c->md.ewald.surface : 0;
        a.use_constraint = c->md.ewald_on ? c->md.ewald.use_constraint : (c->md.constraint_mode ? 1 : 0);
        a.lustig = c->md.lustig;
        a.stress = c->d_stress;
        a.stress_after = 0.15 * (double)c->md.total_steps;
    }
    ProfScope ps(c, "kinetics");
    const int nb = (c->nmol + KIN_THREADS - 1) / KIN_THREADS;
    hipLaunchKernelGGL(k_kinetics, dim3(nb), dim3(KIN_THREADS), 0, c->stream, a, rewrap ? 1 : 0);
    hipLaunchKernelGGL(k_kin_finish, dim3(1), dim3(FINISH_THREADS), 0, c->stream, a, nb, 0);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// The MD step's SHAKE and kinetics in one kernel (rigid 3-site water).
static int shake_kinetics_run(Ctx* c, const smd_integrator_params* md) {
    IntegArgs a = integ_args(c, md);
    a.fnh_on_device = 1;
    a.md_step = 1;
    a.frecip = c->shake_frecip ? c->d_frecip : nullptr;
    a.ndf = md_ndf(c);
    a.surface = c->md.ewald_on ? c->md.ewald.surface : 0;
    a.use_constraint = c->md.ewald_on ? c->md.ewald.use_constraint : 1;
    a.lustig = c->md.lustig;
    a.stress = c->d_stress;
    a.stress_after = 0.15 * (double)c->md.total_steps;
    ProfScope ps(c, "shake_kinetics");
    const int nb = (c->nmol + KIN_THREADS - 1) / KIN_THREADS;
    hipLaunchKernelGGL(k_shake_kinetics, dim3(nb), dim3(KIN_THREADS), 0, c->stream, a, 1);
    hipLaunchKernelGGL(k_kin_finish, dim3(1), dim3(FINISH_THREADS), 0, c->stream, a, nb, 1);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

int lf_atoms_run(Ctx* c, const smd_integrator_params* p, double fnh, bool md_step) {
    IntegArgs a = integ_args(c, p);
    a.fnh_host = fnh;
    a.fnh_on_device = md_step ? 1 : 0;
    a.md_step = md_step ? 1 : 0;
    ProfScope ps(c, "shake");
    hipLaunchKernelGGL(k_lf_atoms, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, a);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}


// Pieces of the polarizable MD step (polar_md.hip): the head (trigger,
// drift partials, cleared terms and forces) and the SHAKE + kinetics +
// Nose-Hoover tail, with c->md describing the step.
int md_head_enqueue(Ctx* c) {
    IntegArgs a = integ_args(c, &c->md.integ);
    a.fnh_on_device = 1;
    a.drmax = c->md.cut_buff * c->md.cut_buff / 4.0;   // Integrator.cpp:36
    a.zero_force = 1;
    hipLaunchKernelGGL(k_step_head, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, a);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

int md_tail_enqueue(Ctx* c) {
    c->shake_frecip = false;
    return shake_kinetics_run(c, &c->md.integ);
}

int md_start_vtm(Ctx* c) {
    hipLaunchKernelGGL(k_start_vtm, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n, c->d_vel, c->d_force,
                       c->d_mass, c->d_vtm, c->md.integ.dt);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// One LFIntegrator::run iteration enqueued on the context stream.  Contains
// no host synchronisation and no memset, so it can be captured.
int md_step_enqueue(Ctx* c, bool do_long) {
    const smd_md_params& p = c->md;
    // the force terms' reductions are queued and added by one kernel before
    // the tail reads them
    RedDefer defer(c);
    IntegArgs a = integ_args(c, &p.integ);
    a.fnh_on_device = 1;
    a.drmax = p.cut_buff * p.cut_buff / 4.0;   // Integrator.cpp:36
    const dim3 gA((c->n + 255) / 256), b256(256);
    // trigger + drift sums; the list rebuild (with the molecule rewrap) runs
    // only when the trigger fired, gated on the device flag
    hipLaunchKernelGGL(k_step_head, gA, b256, 0, c->stream, a);
    // Reciprocal space needs the positions after the rebuild's rewrap and
    // nothing of the list: it runs on the second stream beside the list
    // build and the short-range kernels.
    if (!p.ewald_on) {
        // LJ atoms: list, LJ, the unconstrained update (drift removal, NH
        // velocity loop, clears the rebuild flag), then the kinetic and
        // pressure sums, the NH update and the per-step rewrap
        int rc = pairlist_build(c, p.cutoff, p.cut_buff, p.use_cells, true);
        if (!rc) rc = pairs_lj(c, p.cutoff, p.switch_on, p.constraint_mode ? 1 : 0, p.lustig, p.prec, true);
        if (!rc) rc = red_flush(c);
        if (!rc) rc = p.constraint_mode ? shake_run(c, &p.integ, true, 0.0, true) : lf_atoms_run(c, &p.integ, 0.0, true);
        if (!rc) rc = kinetics_run(c, true, &p.integ);
        return rc;
    }
    const bool fork = do_long && c->recip_overlap;
    c->list_fork = fork ? c->ev_fork : nullptr;
    int rc = pairlist_build(c, p.cutoff, p.cut_buff, p.use_cells, true);
    c->list_fork = nullptr;
    if (rc) return rc;
    if (fork) {
        SMD_HIP(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
        {
            StreamSwap sw(c, c->stream2);
            c->recip_store = true;
            rc = ewald_recip(c, p.prec, 1);
            c->recip_store = false;
        }
        if (rc) return rc;
        SMD_HIP(c, hipEventRecord(c->ev_join, c->stream2));
    }
    // the pair kernel is the first force term of the step: it stores forces
    rc = pairs_lj_ewald(c, nullptr, nullptr, p.cutoff, p.ewald.alpha, p.ewald.real_cut, p.ewald.use_constraint,
                        1, p.prec, true);
    if (rc) return rc;
    if ((rc = ewald_intra(c, 1))) return rc;
    if (p.ewald.surface && (rc = ewald_surface(c))) return rc;
    if (fork) SMD_HIP(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    else if (do_long && (rc = ewald_recip(c, p.prec, 1))) return rc;
    if ((rc = red_flush(c))) return rc;
    // drift removal, SHAKE (adds d_frecip), clears the rebuild flag; then the
    // kinetic and pressure sums, the NH update and the per-step rewrap
    c->shake_frecip = fork;
    if (c->md_fused_shake) {
        rc = shake_kinetics_run(c, &p.integ);
    } else {
        rc = shake_run(c, &p.integ, true, 0.0, true);
        if (!rc) rc = kinetics_run(c, true, &p.integ);
    }
    c->shake_frecip = false;
    if (rc) return rc;
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

}  // namespace smd

using namespace smd;

extern "C" int smd_set_constraints(smd_ctx* c, const double* bond2, int64_t n_mols) {
    if (!c || !bond2 || n_mols != c->nmol) return c ? fail(c, SMD_EINVAL, "constraint count mismatch") : SMD_EINVAL;
    for (int m = 0; m < c->nmol; ++m)
        if (c->h_mol_start[m + 1] - c->h_mol_start[m] != 3)
            return fail(c, SMD_ENOTSUP, "SHAKE is implemented for 3-site (water) molecules");
    SMD_HIP(c, hipMemcpy(c->d_bond2, bond2, sizeof(double) * 3 * c->nmol, hipMemcpyHostToDevice));
    return SMD_OK;
}

extern "C" int smd_set_vt_minus(smd_ctx* c, const double* v, int64_t n) {
    if (!c || !v || n != c->n) return SMD_EINVAL;
    SMD_HIP(c, hipMemcpy(c->d_vtm, v, sizeof(double) * 3 * c->n, hipMemcpyHostToDevice));
    return SMD_OK;
}

extern "C" int smd_get_vt_minus(smd_ctx* c, double* v, int64_t n) {
    if (!c || !v || n != c->n) return SMD_EINVAL;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemcpy(v, c->d_vtm, sizeof(double) * 3 * c->n, hipMemcpyDeviceToHost));
    return SMD_OK;
}

extern "C" int smd_set_mol_momenta(smd_ctx* c, const double* p, int64_t n) {
    if (!c || !p || n != c->nmol) return SMD_EINVAL;
    SMD_HIP(c, hipMemcpy(c->d_mom, p, sizeof(double) * 3 * c->nmol, hipMemcpyHostToDevice));
    return SMD_OK;
}

extern "C" int smd_shake(smd_ctx* c, const smd_integrator_params* p, double fnh, double virial_out[9]) {
    if (!c || !p) return SMD_EINVAL;
    for (int m = 0; m < c->nmol; ++m)
        if (c->h_mol_start[m + 1] - c->h_mol_start[m] != 3)
            return fail(c, SMD_ENOTSUP, "SHAKE is implemented for 3-site (water) molecules");
    if (!(p->dt > 0) || p->max_count < 1) return fail(c, SMD_EINVAL, "bad SHAKE parameters");
    SMD_HIP(c, hipMemsetAsync(c->d_terms, 0, sizeof(double) * T_COUNT, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_status, 0, sizeof(unsigned), c->stream));
    c->ind_last_converged = false;   // positions move
    int rc = shake_run(c, p, false, fnh);
    if (rc) return rc;
    double t[T_COUNT];
    if ((rc = read_terms(c, t))) return rc;
    if ((rc = check_status(c))) return rc;
    if (virial_out)
        for (int k = 0; k < 9; ++k) virial_out[k] = t[T_SH_VIR + k] / (p->dt * p->dt);
    return SMD_OK;
}

extern "C" int smd_lf_atoms(smd_ctx* c, const smd_integrator_params* p, double fnh) {
    if (!c || !p) return SMD_EINVAL;
    SMD_HIP(c, hipMemsetAsync(c->d_status, 0, sizeof(unsigned), c->stream));
    c->ind_last_converged = false;   // positions move
    int rc = lf_atoms_run(c, p, fnh, false);
    if (rc) return rc;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return check_status(c);
}

extern "C" int smd_mol_kinetics(smd_ctx* c, double out2[2]) {
    if (!c) return SMD_EINVAL;
    SMD_HIP(c, hipMemsetAsync(c->d_terms, 0, sizeof(double) * T_COUNT, c->stream));
    int rc = kinetics_run(c, false);
    if (rc) return rc;
    double t[T_COUNT];
    if ((rc = read_terms(c, t))) return rc;
    if (out2) { out2[0] = t[T_KE_ATOM]; out2[1] = t[T_KE_MOL]; }
    return SMD_OK;
}

extern "C" int smd_compute_mol_centers(smd_ctx* c) {
    if (!c) return SMD_EINVAL;
    hipLaunchKernelGGL(k_mol_centers, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, c->nmol, c->d_mol_start,
                       c->d_mass, c->d_pos, c->d_com);
    SMD_HIP(c, hipGetLastError());
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return SMD_OK;
}

extern "C" int smd_nh_chain(smd_ctx* c, double ekin, double temperature, double dt, double ndf, double press_diff,
                            double volume, double* chain_state, double* scale_out) {
    // chain_state: vnh[4] gnh[4] qnh[4] vlog glog wbar (15 doubles), updated in place
    if (!c || !chain_state) return SMD_EINVAL;
    double st[21];
    for (int k = 0; k < 15; ++k) st[k] = chain_state[k];
    st[15] = ekin; st[16] = temperature; st[17] = dt; st[18] = ndf; st[19] = press_diff; st[20] = volume;
    double* d = nullptr;
    SMD_HIP(c, hipMalloc(&d, sizeof(double) * 22));
    hipError_t e = hipMemcpy(d, st, sizeof st, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_nh_chain, dim3(1), dim3(64), 0, c->stream, d, d + 21);
    hipLaunchKernelGGL(k_scale, dim3((3 * c->n + 255) / 256), dim3(256), 0, c->stream, 3 * c->n, c->d_vel, d + 21);
    if (e == hipSuccess) e = hipGetLastError();
    double back[22];
    if (e == hipSuccess) e = hipMemcpyAsync(back, d, sizeof back, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d);
    if (e != hipSuccess) return hip_fail(c, e, "NH chain");
    for (int k = 0; k < 15; ++k) chain_state[k] = back[k];
    if (scale_out) *scale_out = back[21];
    return SMD_OK;
}

extern "C" int smd_md_init(smd_ctx* c, const smd_md_params* p) {
    if (!c || !p) return SMD_EINVAL;
    if (!c->box_set) return fail(c, SMD_ESTATE, "box not set");
    if (!p->sigma2 || !p->eps4) return fail(c, SMD_EINVAL, "LJ table required");
    if (p->prec != SMD_PREC_FP64 && p->prec != SMD_PREC_MIXED) return fail(c, SMD_EINVAL, "bad precision");
    if (p->constraint_mode != 0 && p->constraint_mode != 2)
        return fail(c, SMD_ENOTSUP, "constraint mode %d (0 none, 2 SHAKE)", p->constraint_mode);
    if (p->constraint_mode == 2)
        for (int m = 0; m < c->nmol; ++m)
            if (c->h_mol_start[m + 1] - c->h_mol_start[m] != 3)
                return fail(c, SMD_ENOTSUP, "SHAKE in the MD step is implemented for rigid 3-site water");
    if (p->constraint_mode == 0 && p->ewald_on)
        return fail(c, SMD_ENOTSUP, "the MD step with Ewald is implemented for rigid water (constraint mode 2)");
    c->md = *p;
    const int nt2 = c->ntypes * c->ntypes;
    c->md_sigma2.assign(p->sigma2, p->sigma2 + nt2);
    c->md_eps4.assign(p->eps4, p->eps4 + nt2);
    c->md.sigma2 = c->md_sigma2.data();
    c->md.eps4 = c->md_eps4.data();
    smd_vdw_params vp{};
    vp.kind = SMD_VDW_LJ;
    vp.sigma2 = c->md.sigma2;
    vp.eps4 = c->md.eps4;
    int rc = upload_vdw_tables(c, &vp);
    if (rc) return rc;
    if (!c->stream2) {
        SMD_HIP(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        SMD_HIP(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        SMD_HIP(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    if (!c->d_frecip) SMD_HIP(c, hipMalloc(&c->d_frecip, sizeof(double) * 3 * c->n));
    if (p->ewald_on && (rc = ewald_setup(c, &p->ewald))) return rc;
    if ((rc = pairlist_build(c, p->cutoff, p->cut_buff, p->use_cells, false))) return rc;
    // Give in-step rebuilds headroom: the row capacity is fixed once a step
    // is captured.
    {
        int64_t np;
        int32_t mx, uc;
        if ((rc = smd_pairlist_info(c, &np, &mx, &uc))) return rc;
        const int32_t want = ((int32_t)(mx * 1.5) + 64 + 63) / 64 * 64;
        if (want > c->stride) {
            if ((rc = pairlist_alloc(c, want))) return rc;
            if ((rc = pairlist_build(c, p->cutoff, p->cut_buff, p->use_cells, false))) return rc;
        }
    }
    if ((rc = smd_compute_mol_centers(c))) return rc;
    // start(): forces at t, vtMinus = v - dt/2 F/m
    SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_terms, 0, sizeof(double) * T_COUNT, c->stream));
    if (p->ewald_on) {
        if ((rc = pairs_lj_ewald(c, nullptr, nullptr, p->cutoff, p->ewald.alpha, p->ewald.real_cut,
                                 p->ewald.use_constraint, 1, p->prec)))
            return rc;
        if ((rc = ewald_intra(c, 1))) return rc;
        if (p->ewald.surface && (rc = ewald_surface(c))) return rc;
        if ((rc = ewald_recip(c, p->prec, 1))) return rc;
    } else if ((rc = pairs_lj(c, p->cutoff, p->switch_on, p->constraint_mode ? 1 : 0, 1, p->prec, false))) {
        return rc;
    }
    hipLaunchKernelGGL(k_start_vtm, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n, c->d_vel, c->d_force,
                       c->d_mass, c->d_vtm, p->integ.dt);
    SMD_HIP(c, hipMemsetAsync(c->d_scalars, 0, sizeof(double) * 8, c->stream));
    if (!c->d_stress) SMD_HIP(c, hipMalloc(&c->d_stress, sizeof(double) * 12));
    SMD_HIP(c, hipMemsetAsync(c->d_stress, 0, sizeof(double) * 12, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_mom, 0, sizeof(double) * 3 * c->nmol, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_disp, 0, sizeof(double) * 3 * c->n, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_status, 0, sizeof(unsigned), c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_flag, 0, sizeof(int) * 8, c->stream));   // rebuild flag, block tickets
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    c->md_step = 0;
    c->md_ready = true;
    return SMD_OK;
}

extern "C" int smd_md_time_family(smd_ctx* c, const char* family, int reps, double* avg_us) {
    if (!c || !family || !avg_us || reps < 1) return SMD_EINVAL;
    if (!c->md_ready) return fail(c, SMD_ESTATE, "smd_md_init not called");
    const std::string f(family);
    const smd_md_params& p = c->md;
    unsigned pass = 0;
    if (f == "recip_sfac") pass = RECIP_SFAC;
    else if (f == "recip_reduce") pass = RECIP_REDUCE;
    else if (f == "recip_force") pass = RECIP_FORCE;
    else if (f != "pairs") return fail(c, SMD_EINVAL, "family '%s' cannot be timed alone", family);
    // The timed kernels overwrite or add to the forces and energy sums; both
    // are saved here and restored below, so the MD state is unchanged.  Their
    // other outputs (k-space partials) are recomputed from unchanged inputs.
    const size_t fb = sizeof(double) * 3 * c->n, tb = sizeof(double) * T_COUNT;
    double* save = nullptr;
    SMD_HIP(c, hipMalloc(&save, fb + tb));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool prof = c->prof.on;
    c->prof.on = false;
    int rc = SMD_OK;
    hipError_t e = hipMemcpyAsync(save, c->d_force, fb, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(save + 3 * c->n, c->d_terms, tb, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // The pair family is timed as its kernel alone: one untimed call packs
    // the site records, which the timed calls reuse.
    const bool pairs_fam = pass == 0;
    if (e == hipSuccess && pairs_fam)
        rc = pairs_lj_ewald(c, nullptr, nullptr, p.cutoff, p.ewald.alpha, p.ewald.real_cut, p.ewald.use_constraint,
                            1, p.prec, true);
    c->pack_reuse = pairs_fam;
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (int r = 0; e == hipSuccess && rc == SMD_OK && r < reps; ++r)
        rc = pass ? ewald_recip(c, p.prec, 1, pass)
                  : pairs_lj_ewald(c, nullptr, nullptr, p.cutoff, p.ewald.alpha, p.ewald.real_cut,
                                   p.ewald.use_constraint, 1, p.prec, true);
    if (e == hipSuccess && rc == SMD_OK) e = hipEventRecord(e1, c->stream);
    c->pack_reuse = false;
    // restore whatever happened above
    hipError_t er = hipMemcpyAsync(c->d_force, save, fb, hipMemcpyDeviceToDevice, c->stream);
    if (er == hipSuccess) er = hipMemcpyAsync(c->d_terms, save + 3 * c->n, tb, hipMemcpyDeviceToDevice, c->stream);
    if (er == hipSuccess) er = hipStreamSynchronize(c->stream);
    float ms = 0.0f;
    if (e == hipSuccess && rc == SMD_OK) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(save);
    c->prof.on = prof;
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(c, e, "timing run");
    if (er != hipSuccess) return hip_fail(c, er, "restoring forces");
    *avg_us = 1e3 * (double)ms / reps;
    return SMD_OK;
}

extern "C" int smd_md_run(smd_ctx* c, int64_t nsteps, smd_md_report* out) {
    if (!c || nsteps < 0) return SMD_EINVAL;
    if (!c->md_ready) return fail(c, SMD_ESTATE, "smd_md_init not called");
    c->ind_last_converged = false;   // positions move
    const int ls = c->md.long_steps > 1 ? c->md.long_steps : 1;
    const bool graph = c->md.use_graph && ls == 1 && !c->prof.on;
    if (graph && !c->graph_exec) {
        SMD_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = md_step_enqueue(c, true);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
        c->graph = g;
        SMD_HIP(c, hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
    }
    for (int64_t s = 0; s < nsteps; ++s) {
        if (graph) {
            SMD_HIP(c, hipGraphLaunch(c->graph_exec, c->stream));
        } else {
            // (i % longSteps) == 0 with i counted within this call, as run() does
            int rc = md_step_enqueue(c, (s % ls) == 0);
            if (rc) return rc;
        }
        c->md_step++;
    }
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    int rc = check_status(c);
    if (rc) return rc;
    if (out) {
        double t[T_COUNT];
        if ((rc = read_terms(c, t))) return rc;
        double sc[8];
        SMD_HIP(c, hipMemcpy(sc, c->d_scalars, sizeof sc, hipMemcpyDeviceToHost));
        *out = smd_md_report{};
        out->steps = (int64_t)sc[3];
        out->list_updates = (int64_t)sc[2];
        out->e_lj = t[T_LJ_E];
        out->e_real = t[T_EW_E];
        out->e_long = t[T_RC_E];
        out->e_self = c->self_energy + c->charged_energy;
        out->e_intra = -t[T_IN_E];
        out->atom_ke2 = t[T_KE_ATOM];
        out->mol_ke2 = t[T_KE_MOL];
        out->fnh = sc[0];
        for (int k = 0; k < 9; ++k) out->virial[k] = t[T_PR_VIR + k];
        if (c->md.ewald.surface) {
            const double vol = c->box.L[0] * c->box.L[1] * c->box.L[2];
            const double* qr = t + T_SU_QR;
            out->e_surf = 2.0 * PI * (1.0 / vol) / 3.0 * (qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2]);
        }
    }
    return SMD_OK;
}

extern "C" int smd_md_pressure(smd_ctx* c, smd_md_pressure_out* out) {
    if (!c || !out) return SMD_EINVAL;
    if (!c->md_ready) return fail(c, SMD_ESTATE, "smd_md_init not called");
    double t[T_COUNT], st[12];
    int rc = read_terms(c, t);
    if (rc) return rc;
    SMD_HIP(c, hipMemcpy(st, c->d_stress, sizeof st, hipMemcpyDeviceToHost));
    *out = smd_md_pressure_out{};
    for (int k = 0; k < 9; ++k) {
        out->virial[k] = t[T_PR_VIR + k];
        out->mol_virial[k] = t[T_PR_MOLVIR + k];
        out->atom_pressure[k] = t[T_PR_ATOM + k];
        out->mol_pressure[k] = t[T_PR_MOL + k];
    }
    out->mol_pt = t[T_PR_MOLPT];
    out->lj_dudv = t[T_PR_LUS];
    out->real_deriv = t[T_PR_LUS + 1];
    out->long_deriv = t[T_PR_LUS + 2];
    out->mol_deriv = t[T_PR_LUS + 3];
    out->surf_deriv = t[T_PR_LUS + 4];
    out->counter = (int64_t)st[0];
    out->samples = (int64_t)st[1];
    for (int k = 0; k < 5; ++k) {
        out->atom_stress[k] = st[2 + k];
        out->mol_stress[k] = st[7 + k];
    }
    return SMD_OK;
}

extern "C" int smd_set_real_positions(smd_ctx* c, const double* r, int64_t n) {
    if (!c || !r || n != c->n) return SMD_EINVAL;
    SMD_HIP(c, hipMemcpy(c->d_real, r, sizeof(double) * 3 * c->n, hipMemcpyHostToDevice));
    return SMD_OK;
}

extern "C" int smd_get_real_positions(smd_ctx* c, double* r, int64_t n) {
    if (!c || !r || n != c->n) return SMD_EINVAL;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemcpy(r, c->d_real, sizeof(double) * 3 * c->n, hipMemcpyDeviceToHost));
    return SMD_OK;
}

extern "C" int smd_md_get_thermostat(smd_ctx* c, double* fnh, double* fnh_minus) {
    if (!c || !fnh || !fnh_minus) return SMD_EINVAL;
    if (!c->md_ready) return fail(c, SMD_ESTATE, "smd_md_init not called");
    double sc[2];
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemcpy(sc, c->d_scalars, sizeof sc, hipMemcpyDeviceToHost));
    *fnh = sc[0];
    *fnh_minus = sc[1];
    return SMD_OK;
}

extern "C" int smd_md_set_thermostat(smd_ctx* c, double fnh, double fnh_minus) {
    if (!c) return SMD_EINVAL;
    if (!c->md_ready) return fail(c, SMD_ESTATE, "smd_md_init not called");
    const double sc[2] = {fnh, fnh_minus};
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemcpy(c->d_scalars, sc, sizeof sc, hipMemcpyHostToDevice));
    return SMD_OK;
}
