// pair_forces.hip -- short-range pair terms over the half pair list.
//
// One kernel template covers the reference's pair loops:
//   LJForce::compute          src/forces/LJForce.cpp:58-255
//   MieForce::compute         src/forces/MieForce.cpp:62-241
//   BuckinghamForce::compute  src/forces/BuckinghamForce.cpp:79-287
//   BuffvdWForce::compute     src/forces/BuffvdWForce.cpp:93-343
//   EwaldForce::real_term     src/forces/EwaldForce.cpp:296-527
// and the fused LJ + Ewald real-space pass the MD step uses.
//
// Work mapping (gfx950, wave64): G consecutive lanes share one owner atom i
// and stride through its ELL row, so a wave covers 64/G owners.  The owner's
// force stays in registers and is combined across the G lanes with
// cross-lane shuffles at the end; partner forces go out as FP64 atomics
// (global_atomic_add_f64).  Energies, Lustig derivatives and virials are
// reduced per block and added once per block.  The pair math runs in R
// (double, or float for the mixed-precision form); coordinate differences and
// all accumulation stay FP64.
#include "smd_internal.h"

namespace smd {

enum VdwKind { V_NONE = 0, V_LJ = 1, V_MIE = 2, V_BUCK = 3, V_BUFF = 4 };

struct PairArgs {
    int n;
    Box box;
    const double* pos;
    const double* epos;     // virtual-site positions (BUCK, BUFF)
    const double* q;
    const int32_t* type;
    const int32_t* mol;
    const double* com;
    const int32_t* nb;
    const int32_t* nb_cnt;
    const int32_t* nb_own;
    int store_forces;        // 1: write the atom forces (first term of a step), 0: add
    int32_t stride;
    double* force;
    double* fvirt;
    double* terms;
    Red red;
    // vdW
    int ntypes;
    const double* ta;
    const double* tb;
    const double* tc;
    const double* fa;
    const double* fb;
    const int32_t* host;
    double cutoff, cut2, rcut_2;
    int switch_on, use_constraint, coulomb;
    double sw_dist2;
    double c1, c2, c3;            // CHARMM-type switch (Mie, Buckingham)
    double s0, s1, s2, s3, s4, s5; // 5th-order taper (buffered 14-7)
    double n_mie, m_mie, cn_mie;
    double n_buff, m_buff, delta, gamma;
    int buff_147;                  // n_buff = 14 and m_buff = 7: integer powers
    // Ewald
    double alpha, alpha2, alpha2PI, realcut2;
    // mixed buffered 14-7: fixed-point virtual-site records (x, y, z as 32-bit
    // fractions of the box, type), or null
    const int4* vrec;
    float fscale[3];          // L / 2^32 per dimension
};

// CHARMM-type switch of MieForce/BuckinghamForce (src/forces/MieForce.cpp:145-154)
// in FP64 from the FP64 displacement: its terms cancel near the cutoff.
__device__ __forceinline__ void charmm_switch(const PairArgs& a, double dx, double dy, double dz, double& sv,
                                              double& dsv) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double cc = a.cut2 - r2;
    const double c4 = cc * (2 * r2 + a.c2);
    sv = cc * c4 * a.c1;
    dsv = a.c3 * (cc * cc - c4);
}

template <typename R>
__device__ __forceinline__ R rsqrt_(R x) { return R(1) / sqrt(x); }

template <typename R> __device__ __forceinline__ R erfc_(R x);
template <> __device__ __forceinline__ double erfc_<double>(double x) { return erfc(x); }
template <> __device__ __forceinline__ float erfc_<float>(float x) { return erfcf(x); }
template <typename R> __device__ __forceinline__ R exp_(R x);
template <> __device__ __forceinline__ double exp_<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float exp_<float>(float x) { return expf(x); }
// a / b: the IEEE quotient in FP64, a times the hardware reciprocal (1 ulp)
// in FP32
template <typename R> __device__ __forceinline__ R div_(R a, R b);
template <> __device__ __forceinline__ double div_<double>(double a, double b) { return a / b; }
template <> __device__ __forceinline__ float div_<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename R> __device__ __forceinline__ R pow_(R x, R y);
template <> __device__ __forceinline__ double pow_<double>(double x, double y) { return pow(x, y); }
template <> __device__ __forceinline__ float pow_<float>(float x, float y) { return powf(x, y); }

// Accumulator layout inside the kernel (per lane, FP64).
enum {
    A_VE = 0, A_VELEC, A_VDUDV, A_VD2UDV,   // vdW energy, shifted Coulomb, Lustig
    A_VVIR,                                  // 9
    A_EE = A_VVIR + 9, A_ED1, A_ED2,
    A_EVIR,                                  // 9
    A_N = A_EVIR + 9
};

// Per-owner data the pair term reads.
struct Owner {
    double px, py, pz;      // position
    double ex, ey, ez;      // virtual-site position (BUCK, BUFF)
    double cx, cy, cz;      // molecule centre (Lustig terms)
    double qi;
    int ti, mi;
    unsigned ux, uy, uz;    // fixed-point virtual site (the packed mixed buffered 14-7)
};

// One entry of row i.  FULL: a pair of the reference's half list in its
// owner's row -- energy, volume derivatives, virial and force on i.
// Otherwise: the same pair seen from its other atom -- force on i only.
// Buckingham and buffered 14-7 on the virtual sites: one pair given the
// partner's virtual-site position and type (loaded by the caller, so a row
// walk can fetch them ahead).
// The pair at rij = ri - rj on the virtual sites (minimum image, FP64).
template <int VDW, bool LUSTIG, typename R, bool FULL>
__device__ __forceinline__ void vsite_core(const PairArgs& a, const Owner& o, double dx, double dy, double dz, int tj,
                                           double qj, double (&acc)[A_N], double& fix, double& fiy, double& fiz) {
        const R rx = (R)dx, ry = (R)dy, rz = (R)dz;
        const R r = sqrt(rx * rx + ry * ry + rz * rz);
        const R rij2 = r * r;
        if (rij2 <= (R)a.cut2) {
            const int t = o.ti * a.ntypes + tj;
            R tmpE = 0, tmpf = 0, dulj = 0, d2ulj = 0;
            bool live = true;
            if constexpr (VDW == V_BUCK) {
                const R Aa = (R)a.ta[t], Bb = (R)a.tb[t], Cc = (R)a.tc[t];
                const R r_1 = R(1) / r;
                const R rij_2 = r_1 * r_1;
                const R rij6 = rij_2 * rij_2 * rij_2;
                const R ex = exp_<R>(-Cc * r);
                tmpE = Bb * ex - Aa * rij6;
                tmpf = -Bb * Cc * r_1 * ex + R(6) * Aa * rij6 * rij_2;
                if constexpr (FULL && LUSTIG) {
                    dulj = -Bb * Cc * r * ex + R(6) * Aa * rij6;
                    d2ulj = (R(2) * r + Cc * rij2) * Bb * Cc * ex - R(54) * Aa * rij6;
                }
                if (a.switch_on && rij2 > (R)a.sw_dist2) {
                    double sv, dsv;
                    charmm_switch(a, dx, dy, dz, sv, dsv);
                    tmpf = (R)((double)tmpf * sv - (double)tmpE * dsv);
                    tmpE = (R)((double)tmpE * sv);
                }
                if (a.coulomb) {
                    const R qij = (R)(o.qi * qj);
                    const R e = qij * r_1;
                    const R f = e * rij_2;
                    const R r_rc = r / (R)a.cutoff;
                    const R r4 = rij2 * rij2 * (R)a.rcut_2 * (R)a.rcut_2;
                    if constexpr (FULL)
                        acc[A_VELEC] += (double)(e * (R(1) - R(1.666667) * r_rc + R(1.666667) * r4 - r_rc * r4));
                    tmpf += f * (R(1) - R(5) * r4 + R(4) * r_rc * r4);
                }
            } else {
                const R sigma = (R)a.ta[t], eps = (R)a.tb[t];
                if (eps == R(0)) {
                    live = false;   // tmpE = tmpf = 0 (BuffvdWForce.cpp:190-192)
                } else {
                    const R invs = div_<R>(R(1), sigma);
                    const R m = (R)a.m_buff, nn = (R)a.n_buff, del = (R)a.delta, gam = (R)a.gamma;
                    const R rho = r * invs;
                    const R expnm = nn - m;
                    const R par1 = div_<R>(R(1) + del, rho + del);
                    R rho_m, rho_m1, p1n;
                    if (a.buff_147) {
                        // the 14-7 exponents by products: rho^6, rho^7, par1^7
                        const R rho2 = rho * rho, rho4 = rho2 * rho2;
                        rho_m1 = rho2 * rho4;
                        rho_m = rho_m1 * rho;
                        const R q2 = par1 * par1, q4 = q2 * q2;
                        p1n = q4 * q2 * par1;
                    } else {
                        rho_m = pow_<R>(rho, m);
                        rho_m1 = pow_<R>(rho, m - R(1));
                        p1n = pow_<R>(par1, expnm);
                    }
                    const R par2 = div_<R>(R(1) + gam, rho_m + gam);
                    tmpE = eps * p1n * (par2 - R(2));
                    tmpf = div_<R>(-eps * expnm * p1n * (par2 - R(2)) * invs, rho + del) -
                           div_<R>(eps * p1n * par2 * m * rho_m1 * invs, rho_m + gam);
                    if (a.switch_on && rij2 > (R)a.sw_dist2) {
                        // The taper's coefficients scale as 1/(rc - rsw)^5:
                        // evaluated in FP64 from the FP64 distance.
                        const double rd = sqrt(dx * dx + dy * dy + dz * dz);
                        const double r2d = rd * rd, r3 = r2d * rd, r4 = r2d * r2d, r5 = r2d * r3;
                        const double sv = a.s5 * r5 + a.s4 * r4 + a.s3 * r3 + a.s2 * r2d + a.s1 * rd + a.s0;
                        const double dsv = 5.0 * a.s5 * r4 + 4.0 * a.s4 * r3 + 3.0 * a.s3 * r2d + 2.0 * a.s2 * rd + a.s1;
                        tmpf = (R)((double)tmpf * sv + (double)tmpE * dsv);
                        tmpE = (R)((double)tmpE * sv);
                    }
                    dulj = tmpf * r;   // always accumulated (BuffvdWForce.cpp:242-244)
                }
            }
            if (live) {
                double fx, fy, fz;
                if constexpr (VDW == V_BUCK) {
                    fx = (double)(-tmpf * rx); fy = (double)(-tmpf * ry); fz = (double)(-tmpf * rz);
                } else {
                    const R r_1 = div_<R>(R(1), r);
                    fx = (double)(-tmpf * rx * r_1); fy = (double)(-tmpf * ry * r_1); fz = (double)(-tmpf * rz * r_1);
                }
                fix += fx; fiy += fy; fiz += fz;
                if constexpr (FULL) {
                    acc[A_VE] += (double)tmpE;
                    acc[A_VDUDV] += (double)dulj;
                    if constexpr (LUSTIG) acc[A_VD2UDV] += (double)d2ulj;
                    // virial += -fij * rij for the buffered 14-7 (BuffvdWForce.cpp:276-284)
                    const double sg = VDW == V_BUCK ? 1.0 : -1.0;
                    const double vx = sg * fx, vy = sg * fy, vz = sg * fz;
                    double* v = acc + A_VVIR;
                    v[0] += vx * dx; v[1] += vy * dx; v[2] += vz * dx;
                    v[3] += vx * dy; v[4] += vy * dy; v[5] += vz * dy;
                    v[6] += vx * dz; v[7] += vy * dz; v[8] += vz * dz;
                }
            }
        }
}

template <int VDW, bool LUSTIG, typename R, bool FULL>
__device__ __forceinline__ void vsite_term(const PairArgs& a, const Owner& o, double ejx, double ejy, double ejz, int tj,
                                           double qj, double (&acc)[A_N], double& fix, double& fiy, double& fiz) {
    // rij = ri - rj on the virtual sites (BuckinghamForce.cpp:156,
    // BuffvdWForce.cpp:172)
    double dx = o.ex - ejx, dy = o.ey - ejy, dz = o.ez - ejz;
    pbc_near(dx, dy, dz, a.box);
    vsite_core<VDW, LUSTIG, R, FULL>(a, o, dx, dy, dz, tj, qj, acc, fix, fiy, fiz);
}

template <int VDW, bool EWALD, bool LUSTIG, typename R, bool FULL>
__device__ __forceinline__ void pair_term(const PairArgs& a, const Owner& o, int j, double (&acc)[A_N], double& fix,
                                          double& fiy, double& fiz) {
    const int mj = a.mol[j];
    // rijm = massCenter_j - massCenter_i, minimum image (LJForce.cpp:112,116)
    R mx = 0, my = 0, mz = 0;
    if constexpr (FULL && LUSTIG && VDW != V_BUCK && VDW != V_BUFF) {
        double dx = a.com[3 * mj] - o.cx, dy = a.com[3 * mj + 1] - o.cy, dz = a.com[3 * mj + 2] - o.cz;
        pbc_near(dx, dy, dz, a.box);
        mx = (R)dx; my = (R)dy; mz = (R)dz;
    }
    if constexpr (VDW == V_LJ || VDW == V_MIE || EWALD) {
        double dx = a.pos[3 * j] - o.px, dy = a.pos[3 * j + 1] - o.py, dz = a.pos[3 * j + 2] - o.pz;
        pbc_near(dx, dy, dz, a.box);
        const R rx = (R)dx, ry = (R)dy, rz = (R)dz;
        const R r2 = rx * rx + ry * ry + rz * rz;
        R s = R(0);   // force on i = s * rij
        constexpr bool F32 = sizeof(R) == sizeof(float);
        R r, r_1;
        if constexpr (F32) {
            r_1 = __builtin_amdgcn_rsqf(r2);
            r = r2 * r_1;
        } else {
            r = sqrt(r2);
            r_1 = R(1) / r;
        }

        if constexpr (VDW == V_LJ || VDW == V_MIE) {
            if (o.mi != mj) {
                const R rij2 = F32 ? r2 : r * r;   // the reference tests |r|*|r| <= cut^2
                if (rij2 <= (R)a.cut2) {
                    const int t = o.ti * a.ntypes + a.type[j];
                    const R rij_2 = r_1 * r_1;
                    R tmpE, tmpf, dulj = 0, d2ulj = 0;
                    const R rijRij = (LUSTIG && a.use_constraint) ? (rx * mx + ry * my + rz * mz) : rij2;
                    if constexpr (VDW == V_LJ) {
                        const R sigma2 = (R)a.ta[t], eps4 = (R)a.tb[t];
                        const R sr2 = sigma2 * rij_2;
                        const R sr6 = sr2 * sr2 * sr2;
                        const R sr12 = sr6 * sr6;
                        tmpE = eps4 * (sr12 - sr6);
                        tmpf = (sr12 + sr12 - sr6) * R(6) * eps4 * rij_2;
                        if constexpr (FULL && LUSTIG) {
                            dulj = tmpf * rijRij;
                            d2ulj = eps4 * (R(156) * sr12 - R(42) * sr6) * rijRij * rijRij * rij_2 * rij_2 + R(2) * dulj;
                        }
                        if (FULL && a.switch_on) {   // energy shift only (LJForce.cpp:152-168)
                            const R sr2c = sigma2 * (R)a.rcut_2;
                            const R sr6c = sr2c * sr2c * sr2c;
                            tmpE = tmpE - eps4 * (sr6c * sr6c - sr6c);
                        }
                    } else {
                        const R sigma = sqrt((R)a.ta[t]);
                        const R eps = R(0.25) * (R)a.tb[t];
                        const R sg = sigma * r_1;
                        const R nm = (R)a.n_mie, mm = (R)a.m_mie, cn = (R)a.cn_mie;
                        const R srn = pow_<R>(sg, nm), srm = pow_<R>(sg, mm);
                        tmpE = cn * eps * (srn - srm);
                        tmpf = cn * eps * (nm * srn - mm * srm) * rij_2;
                        if constexpr (FULL && LUSTIG) {
                            const R rr = rx * mx + ry * my + rz * mz;   // MieForce.cpp:140 (always rij.rijm)
                            dulj = tmpf * rr;
                            d2ulj = cn * eps * (nm * (nm + 1) * srn - mm * (mm + 1) * srm) * rr * rr * rij_2 * rij_2 + R(2) * dulj;
                        }
                        if (a.switch_on && rij2 > (R)a.sw_dist2) {
                            double sv, dsv;
                            charmm_switch(a, dx, dy, dz, sv, dsv);
                            tmpf = (R)((double)tmpf * sv - (double)tmpE * dsv);
                            tmpE = (R)((double)tmpE * sv);
                        }
                    }
                    if constexpr (FULL) {
                        acc[A_VE] += (double)tmpE;
                        if constexpr (LUSTIG) { acc[A_VDUDV] += (double)dulj; acc[A_VD2UDV] += (double)d2ulj; }
                    }
                    if (a.coulomb) {   // GROMACS shift (LJForce.cpp:188-201)
                        const R qij = (R)(o.qi * a.q[j]);
                        const R e = qij * r_1;
                        const R f = e * rij_2;
                        const R r_rc = r / (R)a.cutoff;
                        const R r4 = rij2 * rij2 * (R)a.rcut_2 * (R)a.rcut_2;
                        if constexpr (FULL)
                            acc[A_VELEC] += (double)(e * (R(1) - R(1.666667) * r_rc + R(1.666667) * r4 - r_rc * r4));
                        tmpf += f * (R(1) - R(5) * r4 + R(4) * r_rc * r4);
                    }
                    s -= tmpf;
                    if constexpr (FULL) {
                        const double fx = (double)(-tmpf * rx), fy = (double)(-tmpf * ry), fz = (double)(-tmpf * rz);
                        double* v = acc + A_VVIR;
                        v[0] += fx * dx; v[1] += fy * dx; v[2] += fz * dx;
                        v[3] += fx * dy; v[4] += fy * dy; v[5] += fz * dy;
                        v[6] += fx * dz; v[7] += fy * dz; v[8] += fz * dz;
                    }
                }
            }
        }
        if constexpr (EWALD) {
            if (r2 < (R)a.realcut2) {
                const R qij = (R)(o.qi * a.q[j]);
                R e, ex, f;
                if constexpr (F32) {
                    ewald_erfc_f32((float)a.alpha, (float)a.alpha2, r2, r, r_1, e, ex);
                    f = (e + (R)a.alpha2PI * ex) * (r_1 * r_1);
                } else {
                    e = erfc_<R>((R)a.alpha * r) / r;
                    ex = exp_<R>(-(R)a.alpha2 * r2);
                    f = (e + (R)a.alpha2PI * ex) / r2;
                }
                const R se = -qij * f;
                if constexpr (FULL) {
                    acc[A_EE] += (double)(qij * e);
                    if constexpr (LUSTIG) {
                        const R rr = a.use_constraint ? (rx * mx + ry * my + rz * mz) : r2;
                        acc[A_ED1] += (double)(-qij * f * rr);
                        acc[A_ED2] += (double)(qij * (R(2) * f + R(2) * (R)a.alpha2PI * (R)a.alpha2 * ex) * rr * rr / r2 +
                                               R(2) * qij * f * rr);
                    }
                    const double fx = (double)(se * rx), fy = (double)(se * ry), fz = (double)(se * rz);
                    double* v = acc + A_EVIR;
                    v[0] += fx * dx; v[1] += fy * dx; v[2] += fz * dx;
                    v[3] += fx * dy; v[4] += fy * dy; v[5] += fz * dy;
                    v[6] += fx * dz; v[7] += fy * dz; v[8] += fz * dz;
                }
                s += se;
            }
        }
        fix += (double)(s * rx); fiy += (double)(s * ry); fiz += (double)(s * rz);
    } else if constexpr (VDW == V_BUCK || VDW == V_BUFF) {
        vsite_term<VDW, LUSTIG, R, FULL>(a, o, a.epos[3 * j], a.epos[3 * j + 1], a.epos[3 * j + 2], a.type[j],
                                         VDW == V_BUCK ? a.q[j] : 0.0, acc, fix, fiy, fiz);
    }
}

// Entries k = first, first + G, ... < count of a row segment (partner at(k))
// through vsite_term, with the next entry's virtual site and type loaded
// before the current one is evaluated.
template <int VDW, bool LUSTIG, typename R, bool FULL, int G, typename At>
__device__ __forceinline__ void vsite_walk(const PairArgs& a, const Owner& o, int first, int count, At at,
                                           double (&acc)[A_N], double& fix, double& fiy, double& fiz) {
    if (first >= count) return;
    int j = at(first);
    int jn = first + G < count ? at(first + G) : j;
    double ex = a.epos[3 * j], ey = a.epos[3 * j + 1], ez = a.epos[3 * j + 2];
    int tj = a.type[j];
    double qj = VDW == V_BUCK ? a.q[j] : 0.0;   // Buckingham's optional Coulomb term
    for (int k = first; k < count; k += G) {
        const int jnn = k + 2 * G < count ? at(k + 2 * G) : jn;
        const double nx = a.epos[3 * jn], ny = a.epos[3 * jn + 1], nz = a.epos[3 * jn + 2];
        const int nt = a.type[jn];
        const double nq = VDW == V_BUCK ? a.q[jn] : 0.0;
        vsite_term<VDW, LUSTIG, R, FULL>(a, o, ex, ey, ez, tj, qj, acc, fix, fiy, fiz);
        jn = jnn;
        ex = nx; ey = ny; ez = nz; tj = nt; qj = nq;
    }
}

// The mixed buffered 14-7's walk over fixed-point virtual-site records: one
// 16-byte record per partner; the 32-bit coordinate differences are exact
// and already the minimum image, then one rounding to FP32.
template <int VDW, bool LUSTIG, bool FULL, int G, typename At>
__device__ __forceinline__ void vsite_walk_packed(const PairArgs& a, const Owner& o, int first, int count, At at,
                                                  double (&acc)[A_N], double& fix, double& fiy, double& fiz) {
    if (first >= count) return;
    int jn = first + G < count ? at(first + G) : at(first);
    int4 r = a.vrec[at(first)];
    for (int k = first; k < count; k += G) {
        const int jnn = k + 2 * G < count ? at(k + 2 * G) : jn;
        const int4 rn = a.vrec[jn];
        const float dx = (float)(int)(o.ux - (unsigned)r.x) * a.fscale[0];
        const float dy = (float)(int)(o.uy - (unsigned)r.y) * a.fscale[1];
        const float dz = (float)(int)(o.uz - (unsigned)r.z) * a.fscale[2];
        vsite_core<VDW, LUSTIG, float, FULL>(a, o, dx, dy, dz, r.w, 0.0, acc, fix, fiy, fiz);
        jn = jnn;
        r = rn;
    }
}

template <int VDW, bool EWALD, bool LUSTIG, typename R, int G>
__global__ __launch_bounds__(256) void k_pairs(PairArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G;
    const int sub = tid % G;
    double acc[A_N];
#pragma unroll
    for (int k = 0; k < A_N; ++k) acc[k] = 0.0;
    double fix = 0.0, fiy = 0.0, fiz = 0.0;

    if (i < a.n) {
        const int32_t* row = a.nb + (int64_t)i * a.stride;
        const int own = min(a.nb_own[i], a.stride);
        const int non = min(a.nb_cnt[i] - a.nb_own[i], a.stride - own);
        Owner o;
        o.px = a.pos[3 * i]; o.py = a.pos[3 * i + 1]; o.pz = a.pos[3 * i + 2];
        o.ex = o.px; o.ey = o.py; o.ez = o.pz;
        if constexpr (VDW == V_BUCK || VDW == V_BUFF) {
            o.ex = a.epos[3 * i]; o.ey = a.epos[3 * i + 1]; o.ez = a.epos[3 * i + 2];
        }
        o.qi = a.q[i];
        o.ti = a.type[i];
        o.mi = a.mol[i];
        o.cx = o.cy = o.cz = 0.0;
        if constexpr (LUSTIG) { o.cx = a.com[3 * o.mi]; o.cy = a.com[3 * o.mi + 1]; o.cz = a.com[3 * o.mi + 2]; }
        const int32_t* back = row + a.stride - 1;
        if constexpr (VDW == V_BUFF && !EWALD && sizeof(R) == sizeof(float)) {
            if (a.vrec) {
                const int4 ri = a.vrec[i];
                o.ux = (unsigned)ri.x; o.uy = (unsigned)ri.y; o.uz = (unsigned)ri.z;
                vsite_walk_packed<VDW, LUSTIG, true, G>(a, o, sub, own, [row](int k) { return row[k]; }, acc, fix,
                                                        fiy, fiz);
                vsite_walk_packed<VDW, LUSTIG, false, G>(a, o, sub, non, [back](int k) { return back[-k]; }, acc,
                                                         fix, fiy, fiz);
            } else {
                vsite_walk<VDW, LUSTIG, R, true, G>(a, o, sub, own, [row](int k) { return row[k]; }, acc, fix, fiy,
                                                    fiz);
                vsite_walk<VDW, LUSTIG, R, false, G>(a, o, sub, non, [back](int k) { return back[-k]; }, acc, fix,
                                                     fiy, fiz);
            }
        } else if constexpr ((VDW == V_BUCK || VDW == V_BUFF) && !EWALD) {
            // the partner's virtual site and type one entry ahead, its index two
            vsite_walk<VDW, LUSTIG, R, true, G>(a, o, sub, own, [row](int k) { return row[k]; }, acc, fix, fiy, fiz);
            vsite_walk<VDW, LUSTIG, R, false, G>(a, o, sub, non, [back](int k) { return back[-k]; }, acc, fix, fiy,
                                                 fiz);
        } else {
            // the reference's row i: everything
            for (int k = sub; k < own; k += G) pair_term<VDW, EWALD, LUSTIG, R, true>(a, o, row[k], acc, fix, fiy, fiz);
            // partners whose reference row holds the pair: force on i only
            for (int k = sub; k < non; k += G)
                pair_term<VDW, EWALD, LUSTIG, R, false>(a, o, back[-k], acc, fix, fiy, fiz);
        }
    }

    // Owner force: combine the G lanes.  Row i holds every partner of i, so
    // lane 0 of the group is the only writer of atom i in this kernel.
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        fix += __shfl_xor(fix, o, G);
        fiy += __shfl_xor(fiy, o, G);
        fiz += __shfl_xor(fiz, o, G);
    }
    if (i < a.n && sub == 0) {
        if constexpr (VDW == V_BUCK || VDW == V_BUFF) {
            // force on the virtual site; k_redistribute moves it to atoms
            a.fvirt[3 * i] = fix; a.fvirt[3 * i + 1] = fiy; a.fvirt[3 * i + 2] = fiz;
        } else {
            if (a.store_forces) {
                a.force[3 * i] = fix; a.force[3 * i + 1] = fiy; a.force[3 * i + 2] = fiz;
            } else {
                a.force[3 * i] += fix; a.force[3 * i + 1] += fiy; a.force[3 * i + 2] += fiz;
            }
        }
    }

    // energies / derivatives / virials: block sums
    if constexpr (VDW != V_NONE) {
        double v[4 + 9];
#pragma unroll
        for (int k = 0; k < 13; ++k) v[k] = acc[A_VE + k];
        block_part<13>(v, a.red, RS_PAIR_LJ);
    }
    if constexpr (EWALD) {
        __syncthreads();
        double v[3 + 9];
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = acc[A_EE + k];
        block_part<12>(v, a.red, RS_PAIR_EW);
    }
}

// ---------------------------------------------------------------------------

template <int VDW, bool EWALD, int G>
static void launch_pairs_g(Ctx* c, const PairArgs& a, int prec, int lustig) {
    const int threads = 256;
    const int64_t total = (int64_t)c->n * G;
    const dim3 grid((unsigned)((total + threads - 1) / threads));
    if (prec == SMD_PREC_MIXED) {
        if (lustig) hipLaunchKernelGGL((k_pairs<VDW, EWALD, true, float, G>), grid, dim3(threads), 0, c->stream, a);
        else hipLaunchKernelGGL((k_pairs<VDW, EWALD, false, float, G>), grid, dim3(threads), 0, c->stream, a);
    } else {
        if (lustig) hipLaunchKernelGGL((k_pairs<VDW, EWALD, true, double, G>), grid, dim3(threads), 0, c->stream, a);
        else hipLaunchKernelGGL((k_pairs<VDW, EWALD, false, double, G>), grid, dim3(threads), 0, c->stream, a);
    }
}

template <int VDW, bool EWALD>
static hipError_t launch_pairs(Ctx* c, const PairArgs& a, int prec, int lustig) {
    ProfScope ps(c, "pairs");
    // measured best on the 12,288-atom water list: 16 lanes for FP32 pair
    // math, 32 for FP64
    const int lanes = c->pair_lanes ? c->pair_lanes : (prec == SMD_PREC_MIXED ? 16 : 32);
    switch (lanes) {
        case 4: launch_pairs_g<VDW, EWALD, 4>(c, a, prec, lustig); break;
        case 8: launch_pairs_g<VDW, EWALD, 8>(c, a, prec, lustig); break;
        case 32: launch_pairs_g<VDW, EWALD, 32>(c, a, prec, lustig); break;
        case 64: launch_pairs_g<VDW, EWALD, 64>(c, a, prec, lustig); break;
        default: launch_pairs_g<VDW, EWALD, 16>(c, a, prec, lustig); break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the energies' and virials' block partials, added in block order
    const int g = lanes == 4 || lanes == 8 || lanes == 32 || lanes == 64 ? lanes : 16;
    const unsigned nb = (unsigned)(((int64_t)c->n * g + 255) / 256);
    if (VDW != V_NONE && red_finish(c, RS_PAIR_LJ, nb, 13, a.terms + T_LJ_E)) return hipErrorInvalidValue;
    if (EWALD && red_finish(c, RS_PAIR_EW, nb, 12, a.terms + T_EW_E)) return hipErrorInvalidValue;
    return hipSuccess;
}

static void base_args(Ctx* c, PairArgs& a) {
    a = PairArgs{};
    a.n = c->n;
    a.box = c->box;
    a.pos = c->d_pos;
    a.epos = c->d_epos;
    a.q = c->d_q;
    a.type = c->d_type;
    a.mol = c->d_mol;
    a.com = c->d_com;
    a.nb = c->d_nb;
    a.nb_cnt = c->d_nb_cnt;
    a.nb_own = c->d_nb_own;
    a.stride = c->stride;
    a.force = c->d_force;
    a.fvirt = c->d_fvirt;
    a.terms = c->d_terms;
    a.red = c->red;
    a.ntypes = c->ntypes;
    a.ta = c->d_vdw_a;
    a.tb = c->d_vdw_b;
    a.tc = c->d_vdw_c;
    a.fa = c->d_fa;
    a.fb = c->d_fb;
    a.host = c->d_host;
}

static void vdw_constants(PairArgs& a, const smd_vdw_params* p) {
    a.cutoff = p->cutoff;
    a.cut2 = p->cutoff * p->cutoff;
    a.rcut_2 = 1.0 / a.cut2;
    a.switch_on = p->switch_on;
    a.use_constraint = p->use_constraint;
    a.coulomb = p->coulomb;
    double sd2 = a.cut2;
    if (p->switch_on) sd2 = p->switch_dist * p->switch_dist;
    a.sw_dist2 = sd2;
    if (p->switch_on && a.cut2 > sd2) {   // CHARMM switch (MieForce ctor)
        double c1 = 1.0 / (a.cut2 - sd2);
        a.c1 = c1 * c1 * c1;
        a.c2 = a.cut2 - 3.0 * sd2;
        a.c3 = 4 * a.c1;
        // 5th-order taper (BuffvdWForce ctor :65-72)
        const double rc = p->cutoff, rc2 = a.cut2, sw = p->switch_dist;
        const double off = rc - sw;
        const double den = off * off * off * off * off;
        a.s0 = rc * rc2 * (rc2 - 5.0 * rc * sw + 10.0 * sd2) / den;
        a.s1 = -30.0 * rc2 * sd2 / den;
        a.s2 = 30.0 * (rc2 * sw + rc * sd2) / den;
        a.s3 = -10.0 * (rc2 + 4.0 * rc * sw + sd2) / den;
        a.s4 = 15.0 * (rc + sw) / den;
        a.s5 = -6.0 / den;
    }
    a.n_mie = p->n_mie;
    a.m_mie = p->m_mie;
    if (p->kind == SMD_VDW_MIE) {   // MieForce ctor (src/forces/MieForce.cpp:32-35)
        const double Cbase = p->n_mie / p->m_mie;
        const double Cexp = p->m_mie / (p->n_mie - p->m_mie);
        a.cn_mie = (p->n_mie / (p->n_mie - p->m_mie)) * pow(Cbase, Cexp);
    }
    a.n_buff = p->n_buff;
    a.m_buff = p->m_buff;
    a.buff_147 = p->n_buff == 14.0 && p->m_buff == 7.0;
    a.delta = p->delta;
    a.gamma = p->gamma;
}

// Uploads the per-kind tables and per-atom virtual-site data.
int upload_vdw_tables(Ctx* c, const smd_vdw_params* p) {
    const int nt = c->ntypes, nt2 = nt * nt;
    std::vector<double> A(nt2, 0.0), B(nt2, 0.0), Cc(nt2, 0.0);
    std::vector<double> fa(c->n, 1.0), fb(c->n, 0.0);
    std::vector<int32_t> host(c->n, 0);
    for (int i = 0; i < c->n; ++i) host[i] = i;
    switch (p->kind) {
    case SMD_VDW_LJ:
    case SMD_VDW_MIE:
        if (!p->sigma2 || !p->eps4) return fail(c, SMD_EINVAL, "LJ/Mie need sigma2 and eps4 tables");
        for (int k = 0; k < nt2; ++k) { A[k] = p->sigma2[k]; B[k] = p->eps4[k]; }
        break;
    case SMD_VDW_BUCK: {
        if (!p->buck_a || !p->buck_b || !p->buck_c) return fail(c, SMD_EINVAL, "Buckingham needs A, B, C");
        for (int k = 0; k < nt2; ++k) { A[k] = p->buck_a[k]; B[k] = p->buck_b[k]; Cc[k] = p->buck_c[k]; }
        // O0, O1 -> 0.91 on the site and 0.09 on the carbon (myAtoms[2]);
        // carbon keeps its force (BuckinghamForce.cpp:217-238).
        for (int m = 0; m < c->nmol; ++m) {
            const int a0 = c->h_mol_start[m];
            if (c->h_mol_start[m + 1] - a0 != 3) return fail(c, SMD_ENOTSUP, "Buckingham expects 3-site CO2");
            for (int s = 0; s < 3; ++s) {
                const int id = a0 + s;
                const int t = c->h_type[id];
                if (t == 0 || t == 1) { fa[id] = 0.91; fb[id] = 0.09; host[id] = a0 + 2; }
                else if (t == 2) { fa[id] = 1.0; fb[id] = 0.0; }
                else { fa[id] = 0.0; fb[id] = 0.0; }
            }
        }
        break;
    }
    case SMD_VDW_BUFF147: {
        if (!p->type_sigma || !p->type_eps) return fail(c, SMD_EINVAL, "buffered 14-7 needs per-type R and eps");
        // BuffvdWForce::set_buff_param (src/forces/BuffvdWForce.cpp:392-433)
        for (int i = 0; i < nt; ++i)
            for (int j = 0; j < nt; ++j) {
                const double ei = p->type_eps[i], si = p->type_sigma[i];
                const double ej = p->type_eps[j], sj = p->type_sigma[j];
                const double si2 = si * si, sj2 = sj * sj;
                const double es = sqrt(ei) + sqrt(ej);
                double R = 0.0000001, E = 0;
                if (!(ei == 0 || ej == 0)) {
                    R = (si2 * si + sj2 * sj) / (si2 + sj2);
                    E = (4.0 * ei * ej) / (es * es);
                }
                A[i * nt + j] = R;
                B[i * nt + j] = E;
            }
        for (int m = 0; m < c->nmol; ++m) {
            const int a0 = c->h_mol_start[m];
            const int o = a0 + 1;
            if (c->h_mol_start[m + 1] - a0 < 2 || c->h_type[o] != 1)
                return fail(c, SMD_EINVAL, "buffered 14-7 expects type 1 (oxygen) as each molecule's second site");
            for (int id = a0; id < c->h_mol_start[m + 1]; ++id)
                if (c->h_type[id] == 0) { fa[id] = p->hfactor; fb[id] = 1.0 - p->hfactor; host[id] = o; }
        }
        break;
    }
    default:
        return fail(c, SMD_EINVAL, "unknown vdW kind %d", p->kind);
    }
    SMD_HIP(c, hipMemcpyAsync(c->d_vdw_a, A.data(), nt2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipMemcpyAsync(c->d_vdw_b, B.data(), nt2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipMemcpyAsync(c->d_vdw_c, Cc.data(), nt2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipMemcpyAsync(c->d_fa, fa.data(), c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipMemcpyAsync(c->d_fb, fb.data(), c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipMemcpyAsync(c->d_host, host.data(), c->n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return SMD_OK;
}

__global__ void k_virtual_sites(int nmol, const int32_t* __restrict__ mol_start, const int32_t* __restrict__ type,
                                const double* __restrict__ pos, double* __restrict__ epos, int kind, double hfactor,
                                double co2_bond) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const int a0 = mol_start[m], a1 = mol_start[m + 1];
    for (int id = a0; id < a1; ++id)
        for (int d = 0; d < 3; ++d) epos[3 * id + d] = pos[3 * id + d];
    if (kind == SMD_VDW_BUFF147) {
        // BuffvdWForce::VirtualPositions (src/forces/BuffvdWForce.cpp:358-385)
        const int o = a0 + 1;
        for (int id = a0; id < a1; ++id)
            if (type[id] == 0)
                for (int d = 0; d < 3; ++d) epos[3 * id + d] = hfactor * pos[3 * id + d] + (1.0 - hfactor) * pos[3 * o + d];
    } else if (kind == SMD_VDW_BUCK) {
        // BuckinghamForce::VirtualPositions (src/forces/BuckinghamForce.cpp:372-398)
        const double ll = co2_bond * 0.5;
        const double cc = 0.91 * ll;
        const double bb = ll - cc;
        const int cid = a0 + 2;
        for (int s = 0; s < 2; ++s) {
            const int id = a0 + s;
            for (int d = 0; d < 3; ++d) epos[3 * id + d] = (bb / ll) * pos[3 * cid + d] + (cc / ll) * pos[3 * id + d];
        }
    }
}

// Moves the force on each virtual site back onto its atom and host atom
// (BuffvdWForce.cpp:262-274, BuckinghamForce.cpp:217-238); one thread per
// molecule, so the molecule's atoms have a single writer.
__global__ __launch_bounds__(256) void k_redistribute(int nmol, const int32_t* __restrict__ mol_start,
                                                      const double* __restrict__ fa, const double* __restrict__ fb,
                                                      const int32_t* __restrict__ host,
                                                      const double* __restrict__ fvirt, double* __restrict__ force) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    for (int a = mol_start[m]; a < mol_start[m + 1]; ++a) {
        const double x = fvirt[3 * a], y = fvirt[3 * a + 1], z = fvirt[3 * a + 2];
        force[3 * a] += fa[a] * x; force[3 * a + 1] += fa[a] * y; force[3 * a + 2] += fa[a] * z;
        if (fb[a] != 0.0) {
            const int h = host[a];
            force[3 * h] += fb[a] * x; force[3 * h + 1] += fb[a] * y; force[3 * h + 2] += fb[a] * z;
        }
    }
}

int vdw_run(Ctx* c, const smd_vdw_params* p, int prec, smd_vdw_out* out, bool read_back) {
    if (!c->list_valid) return fail(c, SMD_ESTATE, "no pair list");
    int rc = upload_vdw_tables(c, p);
    if (rc) return rc;
    PairArgs a;
    base_args(c, a);
    vdw_constants(a, p);
    const int lustig = p->lustig ? 1 : 0;
    SMD_HIP(c, hipMemsetAsync(c->d_terms, 0, sizeof(double) * T_COUNT, c->stream));
    hipError_t e = hipSuccess;
    switch (p->kind) {
    case SMD_VDW_LJ: e = launch_pairs<V_LJ, false>(c, a, prec, lustig); break;
    case SMD_VDW_MIE: e = launch_pairs<V_MIE, false>(c, a, prec, lustig); break;
    case SMD_VDW_BUCK:
    case SMD_VDW_BUFF147:
        hipLaunchKernelGGL(k_virtual_sites, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, c->nmol,
                           c->d_mol_start, c->d_type, c->d_pos, c->d_epos, p->kind, p->hfactor, p->co2_bond);
        e = (p->kind == SMD_VDW_BUCK) ? launch_pairs<V_BUCK, false>(c, a, prec, lustig)
                                      : launch_pairs<V_BUFF, false>(c, a, prec, lustig);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_redistribute, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, c->nmol,
                               c->d_mol_start, c->d_fa, c->d_fb, c->d_host, c->d_fvirt, c->d_force);
            e = hipGetLastError();
        }
        break;
    default:
        return fail(c, SMD_EINVAL, "unknown vdW kind");
    }
    if (e != hipSuccess) return hip_fail(c, e, "pair kernel launch");
    if (!read_back) return SMD_OK;
    double t[T_COUNT];
    rc = read_terms(c, t);
    if (rc) return rc;
    const double vol = c->box.L[0] * c->box.L[1] * c->box.L[2];
    const double vol2 = vol * vol;
    double energy = t[T_LJ_E];
    double vir[9];
    for (int k = 0; k < 9; ++k) vir[k] = t[T_LJ_VIR + k];
    const double elec = t[T_LJ_ELEC];
    double dudv = t[T_LJ_DUDV], d2udv = t[T_LJ_D2UDV];
    *out = smd_vdw_out{};
    if (p->kind == SMD_VDW_BUFF147) {
        // BuffvdWForce::compute tail (:320-336)
        vir[0] += p->v_lrc; vir[4] += p->v_lrc; vir[8] += p->v_lrc;
        dudv = dudv / (3.0 * vol);
        out->energy = energy;
        out->lj_energy = energy + p->e_lrc;
        out->ulus = energy + p->e_lrc;
        out->dudv = dudv + p->du_lrc;
        out->d2udv = p->lustig ? d2udv / (9.0 * vol2) + p->d2u_lrc : 0.0;
    } else {
        // LJForce::compute tail (:232-254); Buckingham has +dudv (:262)
        const double ulus = energy;
        energy += p->e_lrc;
        vir[0] += p->v_lrc; vir[4] += p->v_lrc; vir[8] += p->v_lrc;
        out->energy = energy;
        out->lj_energy = energy + elec;
        if (p->lustig) {
            dudv = (p->kind == SMD_VDW_BUCK) ? dudv / (3 * vol) : -dudv / (3.0 * vol);
            d2udv = d2udv / (9.0 * vol2);
            out->ulus = ulus + p->e_lrc;
            out->dudv = dudv + p->du_lrc;
            out->d2udv = d2udv + p->d2u_lrc;
        }
    }
    for (int k = 0; k < 9; ++k) out->virial[k] = vir[k];
    return SMD_OK;
}

int ewald_real(Ctx* c, int prec, int lustig) {
    PairArgs a;
    base_args(c, a);
    const double al = c->ew.alpha;
    a.alpha = al;
    a.alpha2 = al * al;
    a.alpha2PI = 2.0 * al / SQRT_PI;
    a.realcut2 = c->ew.real_cut * c->ew.real_cut;
    a.use_constraint = c->ew.use_constraint;
    hipError_t e = launch_pairs<V_NONE, true>(c, a, prec, lustig);
    if (e != hipSuccess) return hip_fail(c, e, "Ewald real kernel launch");
    return SMD_OK;
}

// ---------------------------------------------------------------------------
// Mixed-precision LJ + Ewald pairs of the MD step from packed site records.
// Every row reads one 16-byte record and one word per partner instead of
// separate position, charge, type and molecule gathers:
//   rec[j] = (x, y, z as 32-bit fixed-point fractions of the box, q)
//   tm[j]  = type << 24 | mol
// and, for the Lustig terms of owned pairs, off[j] = float(r_j - R_mol(j)).
// The difference of two fixed-point coordinates is exact and, read as a
// signed integer, is already the minimum image; one rounding to FP32
// follows (the quantum is L / 2^32).  The molecule vector of the Lustig
// terms is r_ij + off_i - off_j.

__global__ __launch_bounds__(256) void k_pack_sites(int n, const double* __restrict__ pos, const double* __restrict__ q,
                                                    const int32_t* __restrict__ type, const int32_t* __restrict__ mol,
                                                    const double* __restrict__ com, Box box, int4* __restrict__ rec,
                                                    int32_t* __restrict__ tm, float4* __restrict__ off) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xs[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    unsigned u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double f = xs[k] * box.inv[k];
        f -= floor(f);
        u[k] = (unsigned)((unsigned long long)(f * 4294967296.0) & 0xffffffffull);
    }
    const int m = mol[i];
    rec[i] = make_int4((int)u[0], (int)u[1], (int)u[2], __float_as_int((float)q[i]));
    tm[i] = (type[i] << 24) | (m & 0xffffff);
    off[i] = make_float4((float)(xs[0] - com[3 * m]), (float)(xs[1] - com[3 * m + 1]), (float)(xs[2] - com[3 * m + 2]),
                         0.0f);
}

struct PackedArgs {
    int n, stride, ntypes, store_forces, use_constraint, switch_on;
    const int32_t* nb;
    const int32_t* nb_cnt;
    const int32_t* nb_own;
    const int4* rec;
    const int32_t* tm;
    const float4* off;
    const double* ta;   // sigma^2 [ntypes^2]
    const double* tb;   // 4 eps
    double* force;
    double* terms;
    Red red;
    float fscale[3];    // L / 2^32 per dimension
    float cut2, rcut_2, alpha, alpha2, alpha2PI, realcut2;
};

// Largest type count whose (sigma^2, 4 eps) table the packed kernel stages in LDS.
constexpr int PACKED_MAX_TYPES = 16;

struct PackedOwner {
    unsigned ux, uy, uz;
    float qi, ox, oy, oz;
    int ti, mi;
};

// Per-lane sums of the packed kernel.  A pair's virial s d d^T is symmetric,
// so each part keeps its 6 upper-triangle components (xx xy xz yy yz zz).
enum {
    PK_VE = 0, PK_VDUDV, PK_VD2UDV,            // vdW energy, Lustig
    PK_EE, PK_ED1, PK_ED2,                     // Ewald real energy, Lustig
    PK_VIR,                                    // virial of both terms (6)
    PK_N = PK_VIR + 6
};

// Pair arithmetic in FP32; each pair's energy, Lustig and virial terms are
// added to per-lane FP64 sums, which the wave, block and grid sums keep in
// FP64 (the grid's in a fixed block order).
template <bool LUSTIG, bool FULL>
__device__ __forceinline__ void packed_term(const PackedArgs& a, const float2* __restrict__ ljt, const PackedOwner& o,
                                            int j, int4 r0, int tm, double (&acc)[PK_N], double& fix,
                                            double& fiy, double& fiz) {
    const float dx = (float)(int)((unsigned)r0.x - o.ux) * a.fscale[0];
    const float dy = (float)(int)((unsigned)r0.y - o.uy) * a.fscale[1];
    const float dz = (float)(int)((unsigned)r0.z - o.uz) * a.fscale[2];
    const float qj = __int_as_float(r0.w);
    const int mj = tm & 0xffffff, tj = (int)((unsigned)tm >> 24);
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float r_1 = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * r_1;
    const float rij_2 = r_1 * r_1;
    float mx = 0.f, my = 0.f, mz = 0.f;
    if constexpr (FULL && LUSTIG) {
        const float4 oj = a.off[j];
        mx = dx + (o.ox - oj.x);
        my = dy + (o.oy - oj.y);
        mz = dz + (o.oz - oj.z);
    }
    float dd[6];
    if constexpr (FULL) {
        dd[0] = dx * dx; dd[1] = dx * dy; dd[2] = dx * dz;
        dd[3] = dy * dy; dd[4] = dy * dz; dd[5] = dz * dz;
    }
    float s = 0.f;
    if (o.mi != mj && r2 <= a.cut2) {
        const float2 lj = ljt[o.ti * a.ntypes + tj];
        const float sigma2 = lj.x, eps4 = lj.y;
        const float sr2 = sigma2 * rij_2;
        const float sr6 = sr2 * sr2 * sr2;
        const float sr12 = sr6 * sr6;
        float tmpE = eps4 * (sr12 - sr6);
        const float tmpf = (sr12 + sr12 - sr6) * 6.f * eps4 * rij_2;
        if constexpr (FULL) {
            if constexpr (LUSTIG) {
                const float rijRij = a.use_constraint ? (dx * mx + dy * my + dz * mz) : r2;
                const float dulj = tmpf * rijRij;
                acc[PK_VDUDV] += (double)dulj;
                acc[PK_VD2UDV] += (double)(eps4 * (156.f * sr12 - 42.f * sr6) * rijRij * rijRij * rij_2 * rij_2 + 2.f * dulj);
            }
            if (a.switch_on) {   // energy shift only (LJForce.cpp:152-168)
                const float sr2c = sigma2 * a.rcut_2;
                const float sr6c = sr2c * sr2c * sr2c;
                tmpE = tmpE - eps4 * (sr6c * sr6c - sr6c);
            }
            acc[PK_VE] += (double)tmpE;
        }
        s -= tmpf;
    }
    if (r2 < a.realcut2) {
        const float qij = o.qi * qj;
        float e, ex;
        ewald_erfc_f32(a.alpha, a.alpha2, r2, r, r_1, e, ex);
        const float f = (e + a.alpha2PI * ex) * rij_2;
        const float se = -qij * f;
        if constexpr (FULL) {
            acc[PK_EE] += (double)(qij * e);
            if constexpr (LUSTIG) {
                const float rr = a.use_constraint ? (dx * mx + dy * my + dz * mz) : r2;
                acc[PK_ED1] += (double)(-qij * f * rr);
                acc[PK_ED2] += (double)(qij * (2.f * f + 2.f * a.alpha2PI * a.alpha2 * ex) * rr * rr * rij_2 +
                                        2.f * qij * f * rr);
            }
        }
        s += se;
    }
    if constexpr (FULL) {
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[PK_VIR + k] += (double)(s * dd[k]);
    }
    fix += (double)(s * dx);
    fiy += (double)(s * dy);
    fiz += (double)(s * dz);
}

// Walks entries k = first, first + G, ... < count of one row segment, where
// entry k's partner index is at(k).  The partner index two entries ahead and
// the record one entry ahead are loaded before the current entry is evaluated.
template <bool LUSTIG, bool FULL, int G, typename At>
__device__ __forceinline__ void packed_segment(const PackedArgs& a, const float2* __restrict__ ljt,
                                               const PackedOwner& o, int first, int count, At at,
                                               double (&acc)[PK_N], double& fix, double& fiy, double& fiz) {
    if (first >= count) return;
    int j = at(first);
    int jn = first + G < count ? at(first + G) : j;
    int4 r0 = a.rec[j];
    int t0 = a.tm[j];
    for (int k = first; k < count; k += G) {
        const int jnn = k + 2 * G < count ? at(k + 2 * G) : jn;
        const int4 n0 = a.rec[jn];
        const int nt = a.tm[jn];
        packed_term<LUSTIG, FULL>(a, ljt, o, j, r0, t0, acc, fix, fiy, fiz);
        j = jn; jn = jnn; r0 = n0; t0 = nt;
    }
}

template <bool LUSTIG, int G>
__global__ __launch_bounds__(256) void k_pairs_packed(PackedArgs a) {
    __shared__ float2 ljt[PACKED_MAX_TYPES * PACKED_MAX_TYPES];
    __shared__ double red[PK_N][256 / WAVE];
    for (int t = threadIdx.x; t < a.ntypes * a.ntypes; t += blockDim.x)
        ljt[t] = make_float2((float)a.ta[t], (float)a.tb[t]);
    __syncthreads();
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G;
    const int sub = tid % G;
    double acc[PK_N];
#pragma unroll
    for (int k = 0; k < PK_N; ++k) acc[k] = 0.0;
    double fix = 0.0, fiy = 0.0, fiz = 0.0;
    if (i < a.n) {
        const int32_t* row = a.nb + (int64_t)i * a.stride;
        const int own = min(a.nb_own[i], a.stride);
        const int non = min(a.nb_cnt[i] - a.nb_own[i], a.stride - own);
        const int4 r0 = a.rec[i];
        PackedOwner o;
        o.ux = (unsigned)r0.x; o.uy = (unsigned)r0.y; o.uz = (unsigned)r0.z; o.qi = __int_as_float(r0.w);
        const int tm = a.tm[i];
        o.mi = tm & 0xffffff;
        o.ti = (int)((unsigned)tm >> 24);
        o.ox = o.oy = o.oz = 0.f;
        if constexpr (LUSTIG) {
            const float4 oi = a.off[i];
            o.ox = oi.x; o.oy = oi.y; o.oz = oi.z;
        }
        // owned partners from the front of the row, the others from the back
        packed_segment<LUSTIG, true, G>(a, ljt, o, sub, own, [row](int k) { return row[k]; }, acc, fix, fiy, fiz);
        const int32_t* back = row + a.stride - 1;
        packed_segment<LUSTIG, false, G>(a, ljt, o, sub, non, [back](int k) { return back[-k]; }, acc, fix, fiy,
                                         fiz);
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        fix += __shfl_xor(fix, o, G);
        fiy += __shfl_xor(fiy, o, G);
        fiz += __shfl_xor(fiz, o, G);
    }
    if (i < a.n && sub == 0) {
        if (a.store_forces) {
            a.force[3 * i] = fix; a.force[3 * i + 1] = fiy; a.force[3 * i + 2] = fiz;
        } else {
            a.force[3 * i] += fix; a.force[3 * i + 1] += fiy; a.force[3 * i + 2] += fiz;
        }
    }
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < PK_N; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < PK_N; ++k) red[k][wid] = (double)acc[k];
    }
    __syncthreads();
    // Thread t < 25 forms term T_LJ_E + t: the LJ block (E, ELEC, DUDV, D2UDV,
    // VIR[9]) then the Ewald block (E, D1, D2, VIR[9]); VIR[3a+b] is the
    // symmetric component (a, b).  The fused kernel keeps one virial for both
    // terms (the step uses only their sum): it goes to the LJ block, the
    // Ewald block's virial stays zero.
    constexpr int NT = (T_EW_VIR + 9) - T_LJ_E;
    static_assert(NT == 25, "LJ and Ewald term blocks are contiguous");
    const int t = threadIdx.x;
    if (t < NT) {
        int src;
        if (t < 4) src = t == 0 ? PK_VE : t == 1 ? -1 : t == 2 ? PK_VDUDV : PK_VD2UDV;
        else if (t < 13) {
            const int v = t - 4;
            const int ra = v / 3, cb = v % 3;
            const int lo = min(ra, cb), hi = max(ra, cb);
            src = PK_VIR + (lo == 0 ? hi : lo == 1 ? 2 + hi : 5);
        } else if (t < 16) src = PK_EE + (t - 13);
        else src = -1;
        double s = 0.0;
        if (src >= 0)
            for (int w = 0; w < (int)(blockDim.x / WAVE); ++w) s += red[src][w];
        // the block's sums; red_finish adds the grid's in block order
        if ((int)gridDim.x <= a.red.cap) red_part(a.red, RS_PACKED)[(size_t)t * a.red.cap + blockIdx.x] = s;
    }
}

static int pairs_packed(Ctx* c, const PairArgs& pa, int lustig) {
    const int n = c->n;
    if (!c->d_pack) {
        SMD_HIP(c, hipMalloc(&c->d_pack, sizeof(float4) * 3 * (size_t)n));   // rec [n], off [n], tm [n]
    }
    int4* rec = reinterpret_cast<int4*>(c->d_pack);
    float4* off = reinterpret_cast<float4*>(c->d_pack) + (size_t)n;
    int32_t* tm = reinterpret_cast<int32_t*>(reinterpret_cast<float4*>(c->d_pack) + 2 * (size_t)n);
    PackedArgs a{};
    a.n = n;
    a.stride = pa.stride;
    a.ntypes = pa.ntypes;
    a.store_forces = pa.store_forces;
    a.use_constraint = pa.use_constraint;
    a.switch_on = pa.switch_on;
    a.nb = pa.nb;
    a.nb_cnt = pa.nb_cnt;
    a.nb_own = pa.nb_own;
    a.rec = rec;
    a.tm = tm;
    a.off = off;
    a.ta = pa.ta;
    a.tb = pa.tb;
    a.force = pa.force;
    a.terms = pa.terms;
    a.red = pa.red;
    for (int d = 0; d < 3; ++d) a.fscale[d] = (float)(c->box.L[d] / 4294967296.0);
    a.cut2 = (float)pa.cut2;
    a.rcut_2 = (float)pa.rcut_2;
    a.alpha = (float)pa.alpha;
    a.alpha2 = (float)pa.alpha2;
    a.alpha2PI = (float)pa.alpha2PI;
    a.realcut2 = (float)pa.realcut2;
    ProfScope ps(c, "pairs");
    if (!c->pack_reuse)
        hipLaunchKernelGGL(k_pack_sites, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->d_pos, c->d_q,
                           c->d_type, c->d_mol, c->d_com, c->box, rec, tm, off);
    const int pl = c->pair_lanes;
    const int lanes = pl == 4 || pl == 8 || pl == 32 || pl == 64 ? pl : 16;
    const dim3 grid((unsigned)(((int64_t)n * lanes + 255) / 256));
#define SMD_PACKED(G)                                                                              \
    do {                                                                                           \
        if (lustig) hipLaunchKernelGGL((k_pairs_packed<true, G>), grid, dim3(256), 0, c->stream, a);  \
        else hipLaunchKernelGGL((k_pairs_packed<false, G>), grid, dim3(256), 0, c->stream, a);        \
    } while (0)
    if (lanes == 64) SMD_PACKED(64);
    else if (lanes == 32) SMD_PACKED(32);
    else if (lanes == 8) SMD_PACKED(8);
    else if (lanes == 4) SMD_PACKED(4);
    else SMD_PACKED(16);
#undef SMD_PACKED
    SMD_HIP(c, hipGetLastError());
    return red_finish(c, RS_PACKED, grid.x, (T_EW_VIR + 9) - T_LJ_E, a.terms + T_LJ_E);
}

// Fused LJ + Ewald real space for the MD step (one pass over the list).
int pairs_lj_ewald(Ctx* c, const double* sigma2, const double* eps4, double cutoff, double alpha, double real_cut,
                   int use_constraint, int lustig, int prec, bool store_forces) {
    (void)sigma2; (void)eps4;   // tables uploaded once by the MD setup
    PairArgs a;
    base_args(c, a);
    a.store_forces = store_forces ? 1 : 0;
    a.cutoff = cutoff;
    a.cut2 = cutoff * cutoff;
    a.rcut_2 = 1.0 / a.cut2;
    a.use_constraint = use_constraint;
    a.switch_on = c->md.switch_on;   // the MD step's LJ energy shift
    a.sw_dist2 = a.cut2;
    a.alpha = alpha;
    a.alpha2 = alpha * alpha;
    a.alpha2PI = 2.0 * alpha / SQRT_PI;
    a.realcut2 = real_cut * real_cut;
    if (prec == SMD_PREC_MIXED && c->pair_packed && c->n < (1 << 24) && c->ntypes <= PACKED_MAX_TYPES && !a.coulomb)
        return pairs_packed(c, a, lustig);
    hipError_t e = launch_pairs<V_LJ, true>(c, a, prec, lustig);
    if (e != hipSuccess) return hip_fail(c, e, "LJ+Ewald kernel launch");
    return SMD_OK;
}

// Fixed-point virtual-site records of the mixed buffered 14-7 (the
// fractional coordinates as in k_pack_sites, and the type).
__global__ __launch_bounds__(256) void k_pack_vsites(int n, const double* __restrict__ epos,
                                                     const int32_t* __restrict__ type, Box box, int4* __restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double f = epos[3 * i + k] * box.inv[k];
        f -= floor(f);
        u[k] = (unsigned)((unsigned long long)(f * 4294967296.0) & 0xffffffffull);
    }
    rec[i] = make_int4((int)u[0], (int)u[1], (int)u[2], type[i]);
}

// The mixed buffered 14-7 reads packed virtual-site records once this buffer
// exists (allocated outside any capture).
int vdw_packed_alloc(Ctx* c) {
    if (!c->d_vrec && hipMalloc(&c->d_vrec, sizeof(int4) * (size_t)(c->n > 0 ? c->n : 1)) != hipSuccess)
        return fail(c, SMD_ENOMEM, "virtual-site records of %d sites", c->n);
    return SMD_OK;
}

// vdW of the polarizable MD step: the kernels of vdw_run without the table
// upload, the accumulator clear and the read-back (capturable).  Forces are
// added to d_force; energies, Lustig sums and the virial to the step's terms.
int vdw_enqueue(Ctx* c, const smd_vdw_params* p, int prec) {
    PairArgs a;
    base_args(c, a);
    vdw_constants(a, p);
    const int lustig = p->lustig ? 1 : 0;
    hipError_t e = hipSuccess;
    switch (p->kind) {
    case SMD_VDW_LJ: e = launch_pairs<V_LJ, false>(c, a, prec, lustig); break;
    case SMD_VDW_BUFF147:
        hipLaunchKernelGGL(k_virtual_sites, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, c->nmol,
                           c->d_mol_start, c->d_type, c->d_pos, c->d_epos, p->kind, p->hfactor, p->co2_bond);
        if (prec == SMD_PREC_MIXED && c->d_vrec) {
            hipLaunchKernelGGL(k_pack_vsites, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n, c->d_epos,
                               c->d_type, c->box, c->d_vrec);
            a.vrec = c->d_vrec;
            for (int d = 0; d < 3; ++d) a.fscale[d] = (float)(c->box.L[d] / 4294967296.0);
        }
        e = launch_pairs<V_BUFF, false>(c, a, prec, lustig);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_redistribute, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, c->nmol,
                               c->d_mol_start, c->d_fa, c->d_fb, c->d_host, c->d_fvirt, c->d_force);
            e = hipGetLastError();
        }
        break;
    default:
        return fail(c, SMD_ENOTSUP, "the polarizable MD step takes LJ or buffered 14-7 vdW");
    }
    if (e != hipSuccess) return hip_fail(c, e, "vdW kernel launch");
    return SMD_OK;
}

// LJ alone over the list (no Ewald): the MD step of unconstrained LJ atoms.
// Tables uploaded by the MD setup; the energy is shifted by E(rc) when
// switch_on (LJForce.cpp:152-168).
int pairs_lj(Ctx* c, double cutoff, int switch_on, int use_constraint, int lustig, int prec, bool store_forces) {
    PairArgs a;
    base_args(c, a);
    a.store_forces = store_forces ? 1 : 0;
    a.cutoff = cutoff;
    a.cut2 = cutoff * cutoff;
    a.rcut_2 = 1.0 / a.cut2;
    a.switch_on = switch_on;
    a.sw_dist2 = a.cut2;
    a.use_constraint = use_constraint;
    hipError_t e = launch_pairs<V_LJ, false>(c, a, prec, lustig);
    if (e != hipSuccess) return hip_fail(c, e, "LJ kernel launch");
    return SMD_OK;
}

}  // namespace smd

using namespace smd;

extern "C" int smd_vdw_compute(smd_ctx* c, const smd_vdw_params* p, int prec, double* f_accum, smd_vdw_out* out) {
    if (!c || !p || !out) return SMD_EINVAL;
    if (prec != SMD_PREC_FP64 && prec != SMD_PREC_MIXED) return fail(c, SMD_EINVAL, "bad precision %d", prec);
    if (f_accum) SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    int rc = vdw_run(c, p, prec, out, true);
    if (rc) return rc;
    if (f_accum) {
        std::vector<double> f(3 * (size_t)c->n);
        SMD_HIP(c, hipMemcpyAsync(f.data(), c->d_force, f.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        SMD_HIP(c, hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < f.size(); ++k) f_accum[k] += f[k];
    }
    return SMD_OK;
}

extern "C" int smd_lj_long_range_correct(smd_ctx* c, const double* sigma2, const double* eps4, double cutoff,
                                         double out4[4]) {
    // LJForce::long_range_correct (src/forces/LJForce.cpp:260-370), csix
    // started at 0 (the reference leaves it uninitialised, :268).
    if (!c || !sigma2 || !eps4 || !out4) return SMD_EINVAL;
    if (!c->box_set) return fail(c, SMD_ESTATE, "box not set");
    const int nt = c->ntypes, n = c->n;
    std::vector<int> num(nt, 0);
    for (int i = 0; i < n; ++i) num[c->h_type[i]] += 1;
    const double cutoff3 = cutoff * cutoff * cutoff;
    const double volrc = c->box.L[0] * c->box.L[1] * c->box.L[2];
    const double invvol = 1.0 / volrc;
    double csix = 0.0;
    int npair = 0, nexcl = 0;
    for (int i = 0; i < nt; ++i)
        for (int j = i; j < nt; ++j) {
            const double s2 = sigma2[i * nt + j], e4 = eps4[i * nt + j];
            const double s6 = s2 * s2 * s2;
            const int np = (i != j) ? num[i] * num[j] : (int)(num[i] * (num[i] - 1) * 0.5);
            csix += np * e4 * s6;
            npair += np;
        }
    for (int i = 0; i < n; ++i)
        for (int k = c->h_excl_off[i]; k < c->h_excl_off[i + 1]; ++k) {
            const int j = c->h_excl[k];
            if (j < i) continue;
            const int t = c->h_type[i] * nt + c->h_type[j];
            const double s6 = sigma2[t] * sigma2[t] * sigma2[t];
            csix -= eps4[t] * s6;
            nexcl += 1;
        }
    if (npair - nexcl <= 0) csix = 0;
    else csix /= npair - nexcl;
    out4[0] = -2.0 * PI * n * n * invvol * csix / (3.0 * cutoff3);
    out4[2] = 4.0 * PI * n * n * invvol * invvol * csix / (3.0 * cutoff3);
    out4[3] = -4.0 * PI * n * n * invvol * invvol * invvol * csix / cutoff3;
    out4[1] = -2.0 * out4[0];
    return SMD_OK;
}
