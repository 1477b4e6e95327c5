// ewald.hip -- point-charge Ewald pieces other than the real-space pair loop.
//
//   set_up / ctor             src/forces/EwaldForce.cpp:21-89, :198-257  (host)
//   intraMolCorrect           :709-813   one thread per molecule
//   surfDipoleCorrect         :816-937   two passes over atoms
//   reciprocal_term           :529-698   structure factors, then forces
//
// Reciprocal space.  exp(i k.r) with k = 2 pi (nx/Lx, ny/Ly, nz/Lz) factors
// into X^nx Y^ny Z^nz with X = exp(i 2 pi x / Lx), ...; the k-vectors of the
// reference's list come in (nx, ny) rows of consecutive nz, so along a row each
// term is the previous one times Z.  Rows are cut into groups of at most GK
// consecutive nz.
//   pass 0 (k_recip_phase): exp(i 2 pi r_p / L_p) per atom and axis.
//   pass 1 (k_recip_sfac): a block of SF_WAVES waves owns 64 k-groups (one
//     per lane) and a chunk of atoms.  Per 64-atom tile the block writes each
//     atom's powers X^0..X^nmax, Y^0.., Z^0.. to LDS (negative powers are
//     conjugates); a lane then starts its group at q X^nx Y^ny Z^nz0 and walks
//     the row with one complex product per k.  Per k it sums q sin, q cos and,
//     for the volume derivatives, q cos d, q sin d (d = r - R_mol) and
//     q cos |d|^2, q sin |d|^2.  The waves of a block split the tile's atoms
//     and are added in LDS, so one partial per (chunk, k) reaches memory.
//   pass 1b (k_recip_reduce): partials summed over chunks in a fixed order,
//     then per k: energy, volume derivatives and virial (:596-607); the
//     reference's per-atom Lustig sums of the force loop (:609-658) are sums
//     over atoms of products with the per-k sums above, so they are formed
//     here per k; and (b C, b S) for the force pass.
//   pass 2 (k_recip_force): a block of RF_WAVES waves owns 64 atoms (one per
//     lane); wave w takes the w-th range of k-groups, with the groups and
//     (b C, b S) staged in LDS, walks the rows as pass 1 does, and the waves'
//     partial forces are added in LDS in wave order.
// Every sum is taken in a fixed order: results do not depend on scheduling.
#include <algorithm>
#include <climits>

#include <mutex>

#include "smd_internal.h"

namespace smd {

constexpr int GK = 8;          // k-vectors per group: consecutive nz within one (nx, ny) row
constexpr int SF_TILE = 64;    // atoms per LDS tile in the structure-factor pass
constexpr int SF_WAVES = 8;    // waves per structure-factor block (same groups, different atoms)
constexpr int RF_WAVES = 16;   // waves per force block (same atoms, different k ranges)
constexpr int NSUM = 10;       // per-k sums of the structure-factor pass
constexpr size_t LDS_MAX = 65536;

// d_e1 holds the phases [n][3] (complex) and then, 32-byte aligned, the
// per-atom (q, d) of the structure-factor pass.
static inline size_t e1_am_offset(int n) { return (6 * (size_t)n + 3) & ~(size_t)3; }

// ---- host setup -----------------------------------------------------------

int ewald_setup(Ctx* c, const smd_ewald_params* p) {
    if (!c->box_set) return fail(c, SMD_ESTATE, "box not set");
    if (!(p->alpha > 0) || !(p->real_cut > 0) || !(p->kcut > 0)) return fail(c, SMD_EINVAL, "bad Ewald parameters");
    c->ew = *p;
    const double kCut2 = p->kcut * p->kcut;
    const double alpha2 = p->alpha * p->alpha;
    const double alphaR4 = 1.0 / (4.0 * alpha2);
    int nmax[3];
    for (int d = 0; d < 3; ++d) nmax[d] = (int)(p->kcut * c->box.L[d] / (2 * PI));
    const double vx = 2.0 * PI / c->box.L[0], vy = 2.0 * PI / c->box.L[1], vz = 2.0 * PI / c->box.L[2];
    std::vector<double> kv, k2v, ek;
    std::vector<int32_t> kn;
    for (int nx = 0; nx <= nmax[0]; nx++) {
        const double kx = nx * vx;
        for (int ny = (nx == 0 ? 0 : -nmax[1]); ny <= nmax[1]; ny++) {
            const double ky = ny * vy;
            for (int nz = ((nx == 0 && ny == 0) ? 1 : -nmax[2]); nz <= nmax[2]; nz++) {
                const double kz = nz * vz;
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 < kCut2) {
                    kv.push_back(kx); kv.push_back(ky); kv.push_back(kz);
                    k2v.push_back(k2);
                    ek.push_back(exp(-k2 * alphaR4) / k2);
                    kn.push_back(nx); kn.push_back(ny); kn.push_back(nz);
                }
            }
        }
    }
    const int K = (int)k2v.size();
    if (K == 0) return fail(c, SMD_EINVAL, "no k-vectors below kcut");
    for (int d = 0; d < 3; ++d) c->nmax[d] = nmax[d];
    auto realloc = [&](auto*& ptr, size_t bytes) -> int {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        if (hipMalloc(&ptr, bytes) != hipSuccess) return fail(c, SMD_ENOMEM, "Ewald allocation of %zu bytes", bytes);
        return SMD_OK;
    };
    int rc;
    if ((rc = realloc(c->d_kvec, sizeof(double) * 3 * K))) return rc;
    if ((rc = realloc(c->d_kmod2, sizeof(double) * K))) return rc;
    if ((rc = realloc(c->d_expk, sizeof(double) * K))) return rc;
    if ((rc = realloc(c->d_kn, sizeof(int32_t) * 3 * K))) return rc;
    // k-groups: runs of at most GK consecutive nz inside one (nx, ny) row,
    // stored as (first k, count, nx << 16 | (ny & 0xffff), nz of the first k)
    std::vector<int4> grp;
    for (int k = 0; k < K;) {
        int e = k + 1;
        while (e < K && e - k < GK && kn[3 * e] == kn[3 * k] && kn[3 * e + 1] == kn[3 * k + 1] &&
               kn[3 * e + 2] == kn[3 * (e - 1) + 2] + 1)
            ++e;
        grp.push_back(make_int4(k, e - k, (kn[3 * k] << 16) | (kn[3 * k + 1] & 0xffff), kn[3 * k + 2]));
        k = e;
    }
    const int ng = (int)grp.size();
    // force-pass ranges: RF_WAVES contiguous group ranges of similar k count
    std::vector<int32_t> split(RF_WAVES + 1, ng);
    split[0] = 0;
    {
        int r = 1;
        int64_t acc = 0;
        for (int g = 0; g < ng && r < RF_WAVES; ++g) {
            acc += grp[g].y;
            if (acc >= (int64_t)K * r / RF_WAVES) split[r++] = g + 1;
        }
        for (; r < RF_WAVES; ++r) split[r] = ng;
    }
    // LDS of the two passes (FP64 and FP32 forms)
    const int ne = nmax[0] + nmax[1] + nmax[2] + 3;
    for (int f = 0; f < 2; ++f) {
        const size_t rsz = f ? sizeof(float) : sizeof(double);
        const size_t tables = (size_t)ne * (SF_TILE + 1) * 2 * rsz + 5 * SF_TILE * rsz;
        const size_t red = (size_t)(SF_WAVES - 1) * WAVE * NSUM * sizeof(double);
        c->sfac_lds[f] = std::max(tables, red);
        const size_t kdat = ((size_t)K * 2 * rsz + 15) / 16 * 16 + (size_t)ng * sizeof(int4);
        c->force_lds[f] = std::max(kdat, (size_t)(RF_WAVES - 1) * WAVE * 3 * sizeof(double));
        if (c->sfac_lds[f] > LDS_MAX || c->force_lds[f] > LDS_MAX)
            return fail(c, SMD_ENOTSUP, "reciprocal space too large for the LDS-resident passes (%d k-vectors, nmax %d %d %d)",
                        K, nmax[0], nmax[1], nmax[2]);
    }
    c->n_kgrp = ng;
    if ((rc = realloc(c->d_kgrp, sizeof(int4) * grp.size()))) return rc;
    if ((rc = realloc(c->d_ksplit, sizeof(int32_t) * (RF_WAVES + 1)))) return rc;
    if ((rc = realloc(c->d_kf, sizeof(double) * 2 * (size_t)K))) return rc;
    // structure-factor pass: about one block per CU over (group tile, atom chunk)
    const int gt = (ng + WAVE - 1) / WAVE;
    int chunks = std::max(1, 256 / gt);   // at most one block per CU: a single round
    int per = (c->n + chunks - 1) / chunks;
    if (per < 32) per = 32;
    chunks = (c->n + per - 1) / per;
    c->sfac_chunks = chunks;
    if ((rc = realloc(c->d_sfac_part, sizeof(double) * NSUM * (size_t)K * chunks))) return rc;
    if ((rc = realloc(c->d_e1, sizeof(double) * (e1_am_offset(c->n) + 4 * (size_t)c->n)))) return rc;
    SMD_HIP(c, hipMemcpy(c->d_kgrp, grp.data(), sizeof(int4) * grp.size(), hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_ksplit, split.data(), sizeof(int32_t) * (RF_WAVES + 1), hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_kvec, kv.data(), sizeof(double) * 3 * K, hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_kmod2, k2v.data(), sizeof(double) * K, hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_expk, ek.data(), sizeof(double) * K, hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_kn, kn.data(), sizeof(int32_t) * 3 * K, hipMemcpyHostToDevice));
    c->K = K;
    // self and charged-system energies (ctor :69-79)
    double qSum = 0.0, q2Sum = 0.0;
    for (int i = 0; i < c->n; ++i) { qSum += c->h_q[i]; q2Sum += c->h_q[i] * c->h_q[i]; }
    const double volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
    c->self_energy = -q2Sum * p->alpha / SQRT_PI;
    c->charged_energy = 0.0;
    if (fabs(qSum) > 1.0e-5) c->charged_energy = -PI * qSum * qSum / (2.0 * volume * alpha2);
    c->ewald_ready = true;
    return SMD_OK;
}

// ---- intramolecular correction ---------------------------------------------

struct IntraArgs {
    int nmol;
    Box box;
    const int32_t* mol_start;
    const double* pos;
    const double* q;
    const int32_t* excl_off;
    const int32_t* excl;
    double* force;
    double* terms;
    Red red;
    double alpha, alpha2, alpha2PI, volume;
    int use_constraint;
};

__global__ __launch_bounds__(256) void k_ewald_intra(IntraArgs a) {
    // EwaldForce::intraMolCorrect (src/forces/EwaldForce.cpp:709-813)
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};   // sum qij erf/r, U_i, molderiv, mol2ndDeriv
    if (m < a.nmol) {
        const int s = a.mol_start[m], e = a.mol_start[m + 1];
        const double vol2 = a.volume * a.volume;
        for (int i = s; i < e; ++i) {
            const double qi = a.q[i];
            const double rix = a.pos[3 * i], riy = a.pos[3 * i + 1], riz = a.pos[3 * i + 2];
            for (int j = i + 1; j < e; ++j) {
                bool ex = false;
                for (int k = a.excl_off[i]; k < a.excl_off[i + 1]; ++k) ex = ex || (a.excl[k] == j);
                if (!ex) continue;
                double dx = a.pos[3 * j] - rix, dy = a.pos[3 * j + 1] - riy, dz = a.pos[3 * j + 2] - riz;
                pbc(dx, dy, dz, a.box);
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double r = sqrt(r2);
                const double qij = qi * a.q[j];
                const double er = erf(a.alpha * r) / r;
                const double ex2 = exp(-a.alpha2 * r2);
                v[0] += qij * er;
                v[1] += a.alpha2PI * qij * ex2;
                if (!a.use_constraint) {
                    const double dudr = er - a.alpha2PI * ex2;
                    v[2] += qij * dudr / (3.0 * a.volume);
                    v[3] += (-2.0 * qij * dudr + qij * (-2.0 * dudr + 2.0 * a.alpha2PI * a.alpha2 * ex2 * r2)) / (9.0 * vol2);
                    const double s2 = qij * dudr / r2;
                    // atoms of one molecule belong to this thread only
                    a.force[3 * i] += s2 * dx; a.force[3 * i + 1] += s2 * dy; a.force[3 * i + 2] += s2 * dz;
                    a.force[3 * j] -= s2 * dx; a.force[3 * j + 1] -= s2 * dy; a.force[3 * j + 2] -= s2 * dz;
                }
            }
        }
    }
    block_part<4>(v, a.red, RS_EW_INTRA);
}

int ewald_intra(Ctx* c, int lustig) {
    (void)lustig;
    IntraArgs a;
    a.nmol = c->nmol;
    a.box = c->box;
    a.mol_start = c->d_mol_start;
    a.pos = c->d_pos;
    a.q = c->d_q;
    a.excl_off = c->d_excl_off;
    a.excl = c->d_excl;
    a.force = c->d_force;
    a.terms = c->d_terms;
    a.red = c->red;
    a.alpha = c->ew.alpha;
    a.alpha2 = a.alpha * a.alpha;
    a.alpha2PI = 2.0 * a.alpha / SQRT_PI;
    a.volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
    a.use_constraint = c->ew.use_constraint;
    ProfScope ps(c, "ewald_intra");
    hipLaunchKernelGGL(k_ewald_intra, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, a);
    SMD_HIP(c, hipGetLastError());
    return red_finish(c, RS_EW_INTRA, (c->nmol + 255) / 256, 4, c->d_terms + T_IN_E);
}

// ---- surface dipole ---------------------------------------------------------

// The dipole sums feed every atom's force: block partials added in block
// order by a finisher launch, so the forces do not depend on scheduling.
__global__ __launch_bounds__(256) void k_surf_sums(int n, const double* __restrict__ pos, const double* __restrict__ q,
                                                   const int32_t* __restrict__ mol, const double* __restrict__ com,
                                                   double* part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double qi = q[i];
        const int m = mol[i];
        for (int d = 0; d < 3; ++d) { v[d] = qi * pos[3 * i + d]; v[3 + d] = qi * com[3 * m + d]; }
    }
    block_store<6>(v, part, gridDim.x);
}

__global__ __launch_bounds__(256) void k_surf_force(int n, const double* __restrict__ pos, const double* __restrict__ q,
                                                    double* __restrict__ force, double* terms, double coeff,
                                                    Red red) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double qx = coeff * terms[T_SU_QR], qy = coeff * terms[T_SU_QR + 1], qz = coeff * terms[T_SU_QR + 2];
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double qi = q[i];
        const double tx = -qi * qx, ty = -qi * qy, tz = -qi * qz;
        const double rx = pos[3 * i], ry = pos[3 * i + 1], rz = pos[3 * i + 2];
        const double s = 2.0 * qi;
        force[3 * i] -= s * qx; force[3 * i + 1] -= s * qy; force[3 * i + 2] -= s * qz;
        v[0] = tx * rx; v[1] = ty * rx; v[2] = tz * rx;
        v[3] = tx * ry; v[4] = ty * ry; v[5] = tz * ry;
        v[6] = tx * rz; v[7] = ty * rz; v[8] = tz * rz;
    }
    block_part<9>(v, red, RS_EW_SURF);
}

int ewald_surface(Ctx* c) {
    const double volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
    const double coeff = 2.0 * PI * (1.0 / volume) / 3.0;
    const dim3 g((c->n + 255) / 256);
    hipLaunchKernelGGL(k_surf_sums, g, dim3(256), 0, c->stream, c->n, c->d_pos, c->d_q, c->d_mol, c->d_com,
                       c->part_su);
    int rc = rows_finish(c, c->stream, c->part_su, (int)g.x, (int)g.x, 6, c->d_terms + T_SU_QR);
    if (rc) return rc;
    hipLaunchKernelGGL(k_surf_force, g, dim3(256), 0, c->stream, c->n, c->d_pos, c->d_q, c->d_force, c->d_terms, coeff,
                       c->red);
    SMD_HIP(c, hipGetLastError());
    return red_finish(c, RS_EW_SURF, g.x, 9, c->d_terms + T_SU_VIR);
}

// ---- reciprocal space ---------------------------------------------------------

struct RecipArgs {
    int n, K, ng, chunks, per;
    int nmax[3];
    double vk[3];              // 2 pi / L per axis
    const double* pos;
    const double* q;
    const int32_t* mol;
    const double* com;
    const double* kvec;
    const double* kmod2;
    const double* expk;
    const int4* grp;           // [ng] (first k, count, nx << 16 | ny, nz0)
    const int32_t* split;      // [RF_WAVES+1] group ranges of the force pass
    double* part;              // [chunk][K][NSUM]
    void* kf;                  // [K] (b C, b S) as R pairs
    double2* e1;               // [n][3] exp(i 2 pi r_p / L_p) per atom (pass 0)
    double4* am;               // [n] (q, d) per atom, d = r - R_mol (pass 0)
    double* force;
    int store;                 // 1: the force pass stores (separate array), 0: adds
    double* terms;
    Red red;
    double alphaR4, pi4V, pi8V;
};

template <typename R>
struct cplx { R re, im; };

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// b^n for |n| <= 2^15 by binary powering; negative n through the conjugate
// (|b| = 1).
__device__ __forceinline__ cplx<double> cpow(cplx<double> b, int n) {
    if (n < 0) { b.im = -b.im; n = -n; }
    cplx<double> r{1.0, 0.0};
    for (int it = 0; it < 16 && n; ++it) {
        if (n & 1) r = cmul(r, b);
        b = cmul(b, b);
        n >>= 1;
    }
    return r;
}

__device__ __forceinline__ int grp_nx(int v) { return v >> 16; }
__device__ __forceinline__ int grp_ny(int v) { return (int)(short)(v & 0xffff); }

// Pass 0: exp(i 2 pi r_p / L_p) per atom and axis, used by both passes, and
// each atom's charge and offset from its molecule's centre, read by pass 1.
__global__ __launch_bounds__(256) void k_recip_phase(int n, const double* __restrict__ pos, double vx, double vy,
                                                     double vz, double2* __restrict__ e1, const double* __restrict__ q,
                                                     const int32_t* __restrict__ mol, const double* __restrict__ com,
                                                     double4* __restrict__ am) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        const int m = mol[k];
        am[k] = make_double4(q[k], pos[3 * k] - com[3 * m], pos[3 * k + 1] - com[3 * m + 1],
                             pos[3 * k + 2] - com[3 * m + 2]);
    }
    if (k >= 3 * n) return;
    const int p = k % 3;
    const double v = p == 0 ? vx : (p == 1 ? vy : vz);
    double s, c;
    sincos(v * pos[k], &s, &c);
    e1[k] = make_double2(c, s);
}

// Pass 1.  Sums per k (layout of part): 0 q sin, 1 q cos,
// 2-4 q cos d, 5-7 q sin d, 8 q cos |d|^2, 9 q sin |d|^2 (reciprocal_term
// :568-594 and the moments its force loop needs).
template <typename R, bool LUSTIG>
__global__ __launch_bounds__(SF_WAVES * WAVE) void k_recip_sfac(RecipArgs a) {
    extern __shared__ double lds[];
    constexpr int TS = SF_TILE + 1;   // padded table stride: lanes reading different powers spread over banks
    cplx<R>* tab = reinterpret_cast<cplx<R>*>(lds);
    const int ey0 = a.nmax[0] + 1, ez0 = ey0 + a.nmax[1] + 1, ne = ez0 + a.nmax[2] + 1;
    R* misc = reinterpret_cast<R*>(tab + ne * TS);   // [5][SF_TILE]: q, dx, dy, dz, |d|^2
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    const int g = blockIdx.x * WAVE + lane;
    const int ch = blockIdx.y;
    const int i0 = ch * a.per, i1 = min(a.n, i0 + a.per);
    int4 gg = make_int4(0, 0, 0, 0);
    if (g < a.ng) gg = a.grp[g];
    const int gnx = grp_nx(gg.z), gny = grp_ny(gg.z);
    const int ix = gnx, iy = ey0 + abs(gny), iz = ez0 + abs(gg.w);
    const R sy = gny < 0 ? R(-1) : R(1), sz = gg.w < 0 ? R(-1) : R(1);
    // Sums of this lane's k-group over the wave's atoms of the chunk (a few
    // dozen), in the pass's precision; they enter FP64 at the cross-wave sum,
    // and chunks are added in FP64 by k_recip_reduce.
    R acc[GK][NSUM];
#pragma unroll
    for (int t = 0; t < GK; ++t)
#pragma unroll
        for (int u = 0; u < NSUM; ++u) acc[t][u] = R(0);

    // Tile fill: thread fu + SF_TILE fp writes atom fu's powers along axis fp
    // (fp < 3) or its moments (fp == 3).  Its inputs for the next tile are
    // loaded before the current tile is evaluated.
    static_assert(SF_WAVES * WAVE >= 4 * SF_TILE, "one fill task per thread");
    const int fu = threadIdx.x % SF_TILE, fp = threadIdx.x / SF_TILE;
    const bool filler = fp < 4;
    double2 pe = make_double2(1.0, 0.0);
    double4 pm = make_double4(0.0, 0.0, 0.0, 0.0);
    auto fetch = [&](int t0) {
        const int ia = t0 + fu;
        if (filler && ia < i1) {
            if (fp < 3) pe = a.e1[3 * ia + fp];
            else pm = a.am[ia];
        }
    };
    fetch(i0);
    for (int t0 = i0; t0 < i1; t0 += SF_TILE) {
        const int na = min(SF_TILE, i1 - t0);
        __syncthreads();
        if (filler && fu < na) {
            const int u = fu;
            if (fp < 3) {
                // powers 0..nmax of exp(i 2 pi r_p / L_p), by recurrence in FP64
                const int base = fp == 0 ? 0 : (fp == 1 ? ey0 : ez0);
                const double c1 = pe.x, s1 = pe.y;
                double er = 1.0, ei = 0.0;
                tab[base * TS + u] = {R(1), R(0)};
                for (int k = 1; k <= a.nmax[fp]; ++k) {
                    const double nr = er * c1 - ei * s1, ni = er * s1 + ei * c1;
                    er = nr; ei = ni;
                    tab[(base + k) * TS + u] = {(R)er, (R)ei};
                }
            } else {
                misc[u] = (R)pm.x;
                misc[SF_TILE + u] = (R)pm.y;
                misc[2 * SF_TILE + u] = (R)pm.z;
                misc[3 * SF_TILE + u] = (R)pm.w;
                misc[4 * SF_TILE + u] = (R)(pm.y * pm.y + pm.z * pm.z + pm.w * pm.w);
            }
        }
        __syncthreads();
        fetch(t0 + SF_TILE);
        if (g < a.ng) {
            for (int u = w; u < na; u += SF_WAVES) {
                const R q = misc[u];
                cplx<R> X = tab[ix * TS + u], Y = tab[iy * TS + u], Z = tab[iz * TS + u];
                Y.im *= sy;
                Z.im *= sz;
                cplx<R> e = cmul(cmul(X, Y), Z);
                e.re *= q;
                e.im *= q;
                const cplx<R> z1 = tab[(ez0 + 1) * TS + u];
                const R dx = misc[SF_TILE + u], dy = misc[2 * SF_TILE + u], dz = misc[3 * SF_TILE + u];
                const R d2 = misc[4 * SF_TILE + u];
                // all GK sums are formed (uniform control flow); those past
                // the group's count are never stored
#pragma unroll
                for (int t = 0; t < GK; ++t) {
                    acc[t][0] += e.im;
                    acc[t][1] += e.re;
                    if constexpr (LUSTIG) {
                        acc[t][2] += e.re * dx; acc[t][3] += e.re * dy; acc[t][4] += e.re * dz;
                        acc[t][5] += e.im * dx; acc[t][6] += e.im * dy; acc[t][7] += e.im * dz;
                        acc[t][8] += e.re * d2; acc[t][9] += e.im * d2;
                    }
                    e = cmul(e, z1);
                }
            }
        }
    }
    // add the waves' sums (wave order) and store one partial per (chunk, k)
    double* red = lds;   // [(SF_WAVES-1) * WAVE][NSUM]
    constexpr int NU = LUSTIG ? NSUM : 2;
#pragma unroll
    for (int t = 0; t < GK; ++t) {
        __syncthreads();
        if (w > 0)
#pragma unroll
            for (int u = 0; u < NU; ++u) red[((w - 1) * WAVE + lane) * NSUM + u] = (double)acc[t][u];
        __syncthreads();
        if (w == 0 && g < a.ng && t < gg.y) {
            double* p = a.part + ((size_t)ch * a.K + gg.x + t) * NSUM;
#pragma unroll
            for (int u = 0; u < NSUM; ++u) {
                double s = (double)acc[t][u];
                if (u < NU)
                    for (int v = 0; v < SF_WAVES - 1; ++v) s += red[(v * WAVE + lane) * NSUM + u];
                p[u] = s;
            }
        }
    }
}

// Pass 1b: per block, RK_PER_BLOCK k-vectors.  Their chunk partials are
// summed in chunk order (four chunk quarters added in quarter order), then
// per k: energy, volume derivatives and virial (:596-607), the force loop's
// Lustig sums (:609-658) in per-k form, and (b C, b S) for the force pass.
constexpr int RK_PER_BLOCK = WAVE / NSUM;   // 6 k-vectors: 60 of the 64 lanes

template <typename R, bool LUSTIG>
__global__ __launch_bounds__(256) void k_recip_reduce(RecipArgs a) {
    __shared__ double red[3][WAVE];
    __shared__ double sums[WAVE];
    const int lane = threadIdx.x % WAVE, qt = threadIdx.x / WAVE;
    const int nv = a.K * NSUM;
    const int k0 = blockIdx.x * RK_PER_BLOCK;
    const int v = k0 * NSUM + lane;
    const bool live = lane < RK_PER_BLOCK * NSUM && v < nv;
    const int c0 = (a.chunks * qt) / 4, c1 = (a.chunks * (qt + 1)) / 4;
    double s = 0.0;
    if (live) {
#pragma unroll 4
        for (int ch = c0; ch < c1; ++ch) s += a.part[(size_t)ch * nv + v];
    }
    if (qt > 0) red[qt - 1][lane] = s;
    __syncthreads();
    if (qt == 0) sums[lane] = live ? ((s + red[0][lane]) + red[1][lane]) + red[2][lane] : 0.0;
    __syncthreads();

    double o[3 + 9 + 4] = {0};   // energy, longDeriv, long2ndDeriv, virial[9], Ff, Ffk2, d2UV1, d2UrR
    const int k = k0 + (int)threadIdx.x;
    if (threadIdx.x < RK_PER_BLOCK && k < a.K) {
        const double* sk = sums + threadIdx.x * NSUM;
        const double kx = a.kvec[3 * k], ky = a.kvec[3 * k + 1], kz = a.kvec[3 * k + 2];
        const double k2 = a.kmod2[k];
        const double sinSum = sk[0], cosSum = sk[1];
        const double te = a.pi4V * a.expk[k] * (sinSum * sinSum + cosSum * cosSum);
        const double coeff = 2.0 * (a.alphaR4 + 1.0 / k2);
        const double kk = (te * kx) * kx + (te * ky) * ky + (te * kz) * kz;
        o[0] = te;
        o[1] = -te + kk * 2.0 * a.alphaR4;
        const double t7 = ((7.0 * te * kx) * kx + (7.0 * te * ky) * ky + (7.0 * te * kz) * kz) * 2.0 * a.alphaR4;
        const double dd = kk * 2.0 * a.alphaR4;
        const double t2 = ((dd * kx) * kx + (dd * ky) * ky + (dd * kz) * kz) * 2.0 * a.alphaR4;
        o[2] = 4.0 * te - t7 + t2;
        o[3 + 0] = te - te * coeff * kx * kx;
        o[3 + 4] = te - te * coeff * ky * ky;
        o[3 + 8] = te - te * coeff * kz * kz;
        o[3 + 1] = -te * coeff * kx * ky;
        o[3 + 2] = -te * coeff * kx * kz;
        o[3 + 5] = -te * coeff * ky * kz;
        o[3 + 3] = o[3 + 1];
        o[3 + 6] = o[3 + 2];
        o[3 + 7] = o[3 + 5];
        const double b = a.pi8V * a.expk[k];
        if constexpr (LUSTIG) {
            // sum_i val_i (k.d_i) with val_i = b (q_i sin_i C - q_i cos_i S)
            const double kSD = kx * sk[5] + ky * sk[6] + kz * sk[7];
            const double kCD = kx * sk[2] + ky * sk[3] + kz * sk[4];
            const double ff = b * (cosSum * kSD - sinSum * kCD);
            o[12] = ff;
            o[13] = ff * (4.0 * a.alphaR4 * k2);
            const double bkk = (b * kx) * kx + (b * ky) * ky + (b * kz) * kz;
            o[14] = bkk * (cosSum * sk[8] + sinSum * sk[9]);
            o[15] = bkk * ((sk[2] * sk[2] + sk[3] * sk[3] + sk[4] * sk[4]) + (sk[5] * sk[5] + sk[6] * sk[6] + sk[7] * sk[7]));
        }
        R* kf = reinterpret_cast<R*>(a.kf);
        kf[2 * k] = (R)(b * cosSum);
        kf[2 * k + 1] = (R)(b * sinSum);
    }
    __syncthreads();
    block_part<16>(o, a.red, RS_RECIP);
}

// Bytes of the (b C, b S) table in the force pass's LDS, rounded up so the
// k-groups that follow it are 16-byte aligned.
template <typename R>
__host__ __device__ constexpr size_t kf_bytes(int K) { return ((size_t)K * 2 * sizeof(R) + 15) / 16 * 16; }

// Pass 2: force on atom i, sum_k (b C q sin - b S q cos) k (:609-658).
template <typename R>
__global__ __launch_bounds__(RF_WAVES * WAVE) void k_recip_force(RecipArgs a) {
    extern __shared__ double lds[];
    R* kf = reinterpret_cast<R*>(lds);                                   // [K][2]
    int4* grp = reinterpret_cast<int4*>(reinterpret_cast<char*>(lds) + kf_bytes<R>(a.K));   // [ng]
    {
        const R* src = reinterpret_cast<const R*>(a.kf);
        for (int idx = threadIdx.x; idx < 2 * a.K; idx += blockDim.x) kf[idx] = src[idx];
        for (int idx = threadIdx.x; idx < a.ng; idx += blockDim.x) grp[idx] = a.grp[idx];
    }
    __syncthreads();
    const int lane = threadIdx.x % WAVE;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int i = blockIdx.x * WAVE + lane;
    const int ic = min(i, a.n - 1);   // idle lanes repeat the last atom and store nothing
    double F[3] = {0.0, 0.0, 0.0};
    {
        const double q = a.q[ic];
        const double vx = a.vk[0], vy = a.vk[1], vz = a.vk[2];
        const double2 ex = a.e1[3 * ic], ey = a.e1[3 * ic + 1], ez = a.e1[3 * ic + 2];
        const cplx<double> X1{ex.x, ex.y}, Y1{ey.x, ey.y}, Z1{ez.x, ez.y};
        const cplx<double> Z1c{Z1.re, -Z1.im};
        const cplx<R> z1{(R)Z1.re, (R)Z1.im};
        cplx<double> XY{1.0, 0.0}, Zp{1.0, 0.0};
        cplx<R> e{R(0), R(0)};
        int cnx = INT_MIN, cny = 0, zp = 0, next_nz = INT_MIN;
        const int gs = a.split[w], ge = a.split[w + 1];
        for (int gi = gs; gi < ge; ++gi) {
            const int4 gg = grp[gi];
            const int nx = grp_nx(gg.z), ny = grp_ny(gg.z), nz0 = gg.w;
            if (!(nx == cnx && ny == cny && nz0 == next_nz)) {
                // a new row: X^nx Y^ny Z^nz0 q
                if (nx != cnx) {
                    XY = cmul(cpow(X1, nx), cpow(Y1, ny));
                    Zp = cpow(Z1, nz0);
                    cnx = nx; cny = ny; zp = nz0;
                } else {
                    for (; cny < ny; ++cny) XY = cmul(XY, Y1);
                    for (; zp < nz0; ++zp) Zp = cmul(Zp, Z1);
                    for (; zp > nz0; --zp) Zp = cmul(Zp, Z1c);
                }
                const cplx<double> e0 = cmul(XY, Zp);
                e = {(R)(q * e0.re), (R)(q * e0.im)};
            }
            next_nz = nz0 + GK;   // a following group continues this row only after a full group
            R sv = R(0), szv = R(0);
            R nzf = (R)nz0;
            const R* kd = kf + 2 * gg.x;
#pragma unroll
            for (int t = 0; t < GK; ++t) {
                if (t < gg.y) {
                    const R val = e.im * kd[2 * t] - e.re * kd[2 * t + 1];
                    sv += val;
                    szv += val * nzf;
                }
                nzf += R(1);
                e = cmul(e, z1);
            }
            if (gg.y < GK) next_nz = INT_MIN;
            F[0] += (double)sv * (nx * vx);
            F[1] += (double)sv * (ny * vy);
            F[2] += (double)szv * vz;
        }
    }
    // add the waves' forces in wave order
    __syncthreads();
    double* red = lds;   // [(RF_WAVES-1) * WAVE][3]
    if (w > 0) {
        red[((w - 1) * WAVE + lane) * 3] = F[0];
        red[((w - 1) * WAVE + lane) * 3 + 1] = F[1];
        red[((w - 1) * WAVE + lane) * 3 + 2] = F[2];
    }
    __syncthreads();
    if (w == 0 && i < a.n) {
        for (int v = 0; v < RF_WAVES - 1; ++v) {
            F[0] += red[(v * WAVE + lane) * 3];
            F[1] += red[(v * WAVE + lane) * 3 + 1];
            F[2] += red[(v * WAVE + lane) * 3 + 2];
        }
        if (a.store) {
            a.force[3 * i] = F[0];
            a.force[3 * i + 1] = F[1];
            a.force[3 * i + 2] = F[2];
        } else {
            a.force[3 * i] += F[0];
            a.force[3 * i + 1] += F[1];
            a.force[3 * i + 2] += F[2];
        }
    }
}

// Scratch of the reciprocal-space passes: per-k coefficients, (chunk, k)
// partials, per-atom phases and moments.
struct RecipScratch { double* kf; double* part; double* e1; };

static int ewald_recip_with(Ctx* c, int prec, int lustig, unsigned passes, const RecipScratch& sc);

int ewald_recip(Ctx* c, int prec, int lustig, unsigned passes) {
    return ewald_recip_with(c, prec, lustig, passes, RecipScratch{c->d_kf, c->d_sfac_part, c->d_e1});
}

// ---- the allocation-free form's scratch -------------------------------------
// One static device array per GPU holds the reciprocal-space scratch of
// smd_ewald_compute_long_static.  In host code `scratch` is only a host-side
// stand-in: its device address, one per GPU, comes from hipGetSymbolAddress
// with that GPU current (static_scratch_init, called by smd_ctx_create) and is
// the only form ever passed to a kernel or a copy.
constexpr size_t STATIC_SCRATCH_DOUBLES = SMD_STATIC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[STATIC_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[MAX_DEVICES] = {};

int static_scratch_init(Ctx* c) {
    if (c->device < 0 || c->device >= MAX_DEVICES) return fail(c, SMD_EINVAL, "device %d out of range", c->device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return hip_fail(c, e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[c->device] = static_cast<double*>(ptr);
    }
    c->static_scratch = g_scratch_ptr[c->device];
    return SMD_OK;
}

static int ewald_recip_with(Ctx* c, int prec, int lustig, unsigned passes, const RecipScratch& sc) {
    if (!c->ewald_ready) return fail(c, SMD_ESTATE, "Ewald not set up");
    RecipArgs a;
    a.n = c->n;
    a.K = c->K;
    a.ng = c->n_kgrp;
    a.chunks = c->sfac_chunks;
    a.per = (c->n + a.chunks - 1) / a.chunks;
    for (int d = 0; d < 3; ++d) {
        a.nmax[d] = c->nmax[d];
        a.vk[d] = 2.0 * PI / c->box.L[d];
    }
    a.pos = c->d_pos;
    a.q = c->d_q;
    a.mol = c->d_mol;
    a.com = c->d_com;
    a.kvec = c->d_kvec;
    a.kmod2 = c->d_kmod2;
    a.expk = c->d_expk;
    a.grp = c->d_kgrp;
    a.split = c->d_ksplit;
    a.part = sc.part;
    a.kf = sc.kf;
    a.e1 = reinterpret_cast<double2*>(sc.e1);
    a.am = reinterpret_cast<double4*>(sc.e1 + e1_am_offset(c->n));
    a.store = c->recip_store ? 1 : 0;
    a.force = c->recip_store ? c->d_frecip : c->d_force;
    a.terms = c->d_terms;
    a.red = c->red;
    a.alphaR4 = 1.0 / (4.0 * (c->ew.alpha * c->ew.alpha));
    const double volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
    a.pi4V = 4.0 * PI * (1.0 / volume);
    a.pi8V = 2.0 * a.pi4V;
    const bool mixed = prec == SMD_PREC_MIXED;
    const int f = mixed ? 1 : 0;
    if (passes & RECIP_SFAC) {
        ProfScope ps(c, "recip_sfac");
        hipLaunchKernelGGL(k_recip_phase, dim3((3 * c->n + 255) / 256), dim3(256), 0, c->stream, c->n, c->d_pos,
                           a.vk[0], a.vk[1], a.vk[2], a.e1, c->d_q, c->d_mol, c->d_com, a.am);
        const dim3 g1((a.ng + WAVE - 1) / WAVE, a.chunks);
        const dim3 b1(SF_WAVES * WAVE);
        // mixed: FP32 phases, products and per-wave sums over the chunk's
        // atoms; FP64 from the cross-wave sum on
        const size_t lds = c->sfac_lds[f];
        if (mixed) {
            if (lustig) hipLaunchKernelGGL((k_recip_sfac<float, true>), g1, b1, lds, c->stream, a);
            else hipLaunchKernelGGL((k_recip_sfac<float, false>), g1, b1, lds, c->stream, a);
        } else {
            if (lustig) hipLaunchKernelGGL((k_recip_sfac<double, true>), g1, b1, lds, c->stream, a);
            else hipLaunchKernelGGL((k_recip_sfac<double, false>), g1, b1, lds, c->stream, a);
        }
    }
    if (passes & RECIP_REDUCE) {
        ProfScope ps(c, "recip_reduce");
        const dim3 gk((a.K + RK_PER_BLOCK - 1) / RK_PER_BLOCK);
        if (mixed) {
            if (lustig) hipLaunchKernelGGL((k_recip_reduce<float, true>), gk, dim3(256), 0, c->stream, a);
            else hipLaunchKernelGGL((k_recip_reduce<float, false>), gk, dim3(256), 0, c->stream, a);
        } else {
            if (lustig) hipLaunchKernelGGL((k_recip_reduce<double, true>), gk, dim3(256), 0, c->stream, a);
            else hipLaunchKernelGGL((k_recip_reduce<double, false>), gk, dim3(256), 0, c->stream, a);
        }
        SMD_HIP(c, hipGetLastError());
        const int rc = red_finish(c, RS_RECIP, gk.x, 16, a.terms + T_RC_E);
        if (rc) return rc;
    }
    if (passes & RECIP_FORCE) {
        ProfScope ps(c, "recip_force");
        const dim3 g2((c->n + WAVE - 1) / WAVE);
        const size_t lds = c->force_lds[f];
        if (mixed) hipLaunchKernelGGL((k_recip_force<float>), g2, dim3(RF_WAVES * WAVE), lds, c->stream, a);
        else hipLaunchKernelGGL((k_recip_force<double>), g2, dim3(RF_WAVES * WAVE), lds, c->stream, a);
    }
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

}  // namespace smd

using namespace smd;

extern "C" int smd_ewald_setup(smd_ctx* c, const smd_ewald_params* p, int32_t* n_kvec) {
    if (!c || !p) return SMD_EINVAL;
    int rc = ewald_setup(c, p);
    if (rc) return rc;
    if (n_kvec) *n_kvec = c->K;
    return SMD_OK;
}

static int fetch_forces_into(smd_ctx* c, double* f_accum) {
    std::vector<double> f(3 * (size_t)c->n);
    SMD_HIP(c, hipMemcpyAsync(f.data(), c->d_force, f.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < f.size(); ++k) f_accum[k] += f[k];
    return SMD_OK;
}

extern "C" int smd_ewald_compute(smd_ctx* c, int prec, double* f_accum, smd_ewald_out* out) {
    // EwaldForce::compute (src/forces/EwaldForce.cpp:128-176)
    if (!c || !out) return SMD_EINVAL;
    if (!c->ewald_ready) return fail(c, SMD_ESTATE, "Ewald not set up");
    if (!c->list_valid) return fail(c, SMD_ESTATE, "no pair list");
    if (f_accum) SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_terms, 0, sizeof(double) * T_COUNT, c->stream));
    int rc = ewald_real(c, prec, 1);
    if (!rc) rc = ewald_intra(c, 1);
    if (!rc && c->ew.surface) rc = ewald_surface(c);
    if (rc) return rc;
    double t[T_COUNT];
    if ((rc = read_terms(c, t))) return rc;
    *out = smd_ewald_out{};
    out->real_energy = t[T_EW_E];
    out->real_deriv = t[T_EW_D1];
    out->real_deriv2 = t[T_EW_D2];
    out->self_energy = c->self_energy;
    out->charged_energy = c->charged_energy;
    out->intra_energy = -t[T_IN_E];
    out->intra_vir_correct = -t[T_IN_U];
    out->mol_deriv = t[T_IN_D1];
    out->mol_deriv2 = t[T_IN_D2];
    for (int k = 0; k < 9; ++k) out->virial[k] = t[T_EW_VIR + k];
    if (c->ew.surface) {
        // surfDipoleCorrect energy and derivatives (:854-864)
        const double volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
        const double vr = 1.0 / volume;
        const double coeff = 2.0 * PI * vr / 3.0;
        const double* qr = t + T_SU_QR;
        const double* qR = t + T_SU_QRR;
        const double rr = qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2];
        const double rR = qr[0] * qR[0] + qr[1] * qR[1] + qr[2] * qR[2];
        const double RR = qR[0] * qR[0] + qR[1] * qR[1] + qR[2] * qR[2];
        out->surf_energy = coeff * rr;
        if (c->ew.use_constraint) {
            out->surf_deriv = -coeff * vr * rr + 2.0 * coeff * vr * rR / 3.0;
            out->surf_deriv2 = 2.0 * coeff * vr * vr * rr - 16.0 * coeff * vr * vr * rR / 9.0 + 2.0 * coeff * vr * vr * RR / 9.0;
        } else {
            out->surf_deriv = -coeff * vr * rr / 3.0;
            out->surf_deriv2 = 4.0 * coeff * vr * vr * rr / 9.0;
        }
        for (int k = 0; k < 9; ++k) out->virial[k] += t[T_SU_VIR + k];
    }
    if (f_accum) return fetch_forces_into(c, f_accum);
    return SMD_OK;
}

static int recip_long_out(smd_ctx* c, smd_ewald_long_out* out);

extern "C" int smd_ewald_compute_long(smd_ctx* c, int prec, double* f_accum, smd_ewald_long_out* out) {
    // EwaldForce::compute_long (src/forces/EwaldForce.cpp:178-196)
    if (!c || !out) return SMD_EINVAL;
    if (!c->ewald_ready) return fail(c, SMD_ESTATE, "Ewald not set up");
    if (f_accum) SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_terms, 0, sizeof(double) * T_COUNT, c->stream));
    int rc = ewald_recip(c, prec, 1);
    if (rc) return rc;
    if ((rc = recip_long_out(c, out))) return rc;
    if (f_accum) return fetch_forces_into(c, f_accum);
    return SMD_OK;
}

// Clears the term accumulators with a kernel rather than a memset node.
__global__ __launch_bounds__(256) void k_clear_terms(double* t, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) t[i] = 0.0;
}

// The same passes with their scratch in the static array: no device
// allocation.  Every atom's phases and moments, the per-k coefficients and
// the (chunk, k) partials are written before they are read in each call, so
// nothing in the array needs clearing; the term accumulators (context state)
// are cleared by a fill kernel on the context stream.
extern "C" int smd_ewald_compute_long_static(smd_ctx* c, int prec, smd_ewald_long_out* out) {
    if (!c || !out) return SMD_EINVAL;
    if (!c->ewald_ready) return fail(c, SMD_ESTATE, "Ewald not set up");
    if (prec != SMD_PREC_FP64 && prec != SMD_PREC_MIXED) return fail(c, SMD_EINVAL, "bad precision %d", prec);
    if (!c->static_scratch) return fail(c, SMD_ESTATE, "static scratch not resolved for device %d", c->device);
    auto up = [](size_t v) { return (v + 31) / 32 * 32; };   // 256-byte segments
    const size_t n_e1 = up(e1_am_offset(c->n) + 4 * (size_t)c->n);
    const size_t n_kf = up(2 * (size_t)c->K);
    const size_t n_part = up((size_t)NSUM * c->K * c->sfac_chunks);
    const size_t need = n_e1 + n_kf + n_part;
    if (need > STATIC_SCRATCH_DOUBLES)
        return fail(c, SMD_ECAPACITY,
                    "static scratch holds %zu bytes; %d atoms and %d k-vectors need %zu (use smd_ewald_compute_long)",
                    (size_t)SMD_STATIC_SCRATCH_BYTES, c->n, c->K, need * sizeof(double));
    const RecipScratch sc{c->static_scratch + n_e1, c->static_scratch + n_e1 + n_kf, c->static_scratch};
    hipLaunchKernelGGL(k_clear_terms, dim3(1), dim3(256), 0, c->stream, c->d_terms, (int)T_COUNT);
    SMD_HIP(c, hipGetLastError());
    int rc = ewald_recip_with(c, prec, 1, RECIP_ALL, sc);
    if (rc) return rc;
    return recip_long_out(c, out);
}

// compute_long's outputs from the term accumulators (reads them back).
static int recip_long_out(smd_ctx* c, smd_ewald_long_out* out) {
    int rc;
    double t[T_COUNT];
    if ((rc = read_terms(c, t))) return rc;
    out->energy = t[T_RC_E];
    out->deriv = t[T_RC_D1];
    out->deriv2 = t[T_RC_D2];
    if (c->ew.use_constraint) {   // :671-676
        out->deriv += t[T_RC_FF];
        out->deriv2 += -6.0 * t[T_RC_FF] + t[T_RC_FFK2] - t[T_RC_D2UV1] + t[T_RC_D2URR];
    }
    for (int k = 0; k < 9; ++k) out->virial[k] = t[T_RC_VIR + k];
    return SMD_OK;
}
