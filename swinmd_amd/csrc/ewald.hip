// ewald.hip -- point-charge Ewald pieces other than the real-space pair loop.
//
//   set_up / ctor             src/forces/EwaldForce.cpp:21-89, :198-257  (host)
//   intraMolCorrect           :709-813   one thread per molecule
//   surfDipoleCorrect         :816-937   two passes over atoms
//   reciprocal_term           :529-698   structure factors, then forces
//
// Reciprocal space.  exp(i k.r) with k = (nx,ny,nz)*2pi/L factors into
// X[nx] Y[ny] Z[nz] with X[n] = exp(i n 2pi x/L); each atom's tables (n up
// to nmax per axis, negative n by conjugation) are built once per step, so
// every (atom, k) term costs two complex products instead of a sin and a cos.
//   pass 1 (k_recip_sfac): one lane per k-vector, atoms in chunks along the
//     grid's second dimension; eight FP64 sums per k per chunk
//     (sum q sin, sum q cos, and their moments q*(r - R_mol)).
//   pass 1b (k_recip_reduce): chunk partials summed in chunk order (the
//     result does not depend on scheduling), per-k energy, virial and volume
//     derivatives.
//   pass 2 (k_recip_force): G lanes per atom split the k-vectors; each lane
//     accumulates its share of the force and the Lustig sums, then the lanes
//     are combined with shuffles.
#include <algorithm>

#include "smd_internal.h"

namespace smd {

constexpr int GK = 8;          // k-vectors per group: consecutive nz within one (nx, ny) row
constexpr int SF_TILE = 64;    // atoms per LDS tile in the structure-factor pass
constexpr int RF_KSPLIT = 8;   // k-group ranges per atom in the force pass
constexpr int KD = 14;         // doubles staged per k-vector in the force pass

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
    // k-groups: runs of at most GK consecutive nz inside one (nx, ny) row
    std::vector<int32_t> grp;   // (first k, count) pairs
    for (int k = 0; k < K; ++k) {
        const bool cont = k > 0 && kn[3 * k] == kn[3 * (k - 1)] && kn[3 * k + 1] == kn[3 * (k - 1) + 1] &&
                          kn[3 * k + 2] == kn[3 * (k - 1) + 2] + 1 && grp.back() < GK;
        if (cont) grp.back() += 1;
        else { grp.push_back(k); grp.push_back(1); }
    }
    const int ng = (int)grp.size() / 2;
    // force-pass ranges: RF_KSPLIT contiguous group ranges of similar k count
    std::vector<int32_t> split(RF_KSPLIT + 1, ng);
    split[0] = 0;
    {
        int r = 1, acc = 0;
        for (int g = 0; g < ng && r < RF_KSPLIT; ++g) {
            acc += grp[2 * g + 1];
            if (acc >= (int64_t)K * r / RF_KSPLIT) split[r++] = g + 1;
        }
        for (; r < RF_KSPLIT; ++r) split[r] = ng;
    }
    int kmax_range = 1;
    for (int r = 0; r < RF_KSPLIT; ++r) {
        if (split[r + 1] <= split[r]) continue;
        const int kb = grp[2 * split[r]], ke = grp[2 * (split[r + 1] - 1)] + grp[2 * (split[r + 1] - 1) + 1];
        kmax_range = std::max(kmax_range, ke - kb);
    }
    c->n_kgrp = ng;
    c->kmax_range = kmax_range;
    if ((rc = realloc(c->d_kgrp, sizeof(int32_t) * grp.size()))) return rc;
    if ((rc = realloc(c->d_ksplit, sizeof(int32_t) * (RF_KSPLIT + 1)))) return rc;
    if ((rc = realloc(c->d_fpart, sizeof(double) * 3 * (size_t)RF_KSPLIT * c->n))) return rc;
    // structure-factor pass: about 2048 wavefronts over (k-group, atom chunk)
    const int gt = (ng + 63) / 64;
    int chunks = (2048 + gt - 1) / gt;
    int per = (c->n + chunks - 1) / chunks;
    if (per < 16) per = 16;
    chunks = (c->n + per - 1) / per;
    c->sfac_chunks = chunks;
    if ((rc = realloc(c->d_sfac_part, sizeof(double) * 8 * (size_t)K * chunks))) return rc;
    if ((rc = realloc(c->d_sfac, sizeof(double) * 16 * (size_t)K))) return rc;
    SMD_HIP(c, hipMemcpy(c->d_kgrp, grp.data(), sizeof(int32_t) * grp.size(), hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_ksplit, split.data(), sizeof(int32_t) * (RF_KSPLIT + 1), hipMemcpyHostToDevice));
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
    block_add<4>(v, a.terms + T_IN_E);
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
    a.alpha = c->ew.alpha;
    a.alpha2 = a.alpha * a.alpha;
    a.alpha2PI = 2.0 * a.alpha / SQRT_PI;
    a.volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
    a.use_constraint = c->ew.use_constraint;
    ProfScope ps(c, "ewald_intra");
    hipLaunchKernelGGL(k_ewald_intra, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, a);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// ---- surface dipole ---------------------------------------------------------

__global__ __launch_bounds__(256) void k_surf_sums(int n, const double* __restrict__ pos, const double* __restrict__ q,
                                                   const int32_t* __restrict__ mol, const double* __restrict__ com,
                                                   double* terms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double qi = q[i];
        const int m = mol[i];
        for (int d = 0; d < 3; ++d) { v[d] = qi * pos[3 * i + d]; v[3 + d] = qi * com[3 * m + d]; }
    }
    block_add<6>(v, terms + T_SU_QR);
}

__global__ __launch_bounds__(256) void k_surf_force(int n, const double* __restrict__ pos, const double* __restrict__ q,
                                                    double* __restrict__ force, double* terms, double coeff) {
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
    block_add<9>(v, terms + T_SU_VIR);
}

int ewald_surface(Ctx* c) {
    const double volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
    const double coeff = 2.0 * PI * (1.0 / volume) / 3.0;
    const dim3 g((c->n + 255) / 256);
    hipLaunchKernelGGL(k_surf_sums, g, dim3(256), 0, c->stream, c->n, c->d_pos, c->d_q, c->d_mol, c->d_com, c->d_terms);
    hipLaunchKernelGGL(k_surf_force, g, dim3(256), 0, c->stream, c->n, c->d_pos, c->d_q, c->d_force, c->d_terms, coeff);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// ---- reciprocal space ---------------------------------------------------------


struct RecipArgs {
    int n, K, ng, chunks, per;
    double vz;                 // 2 pi / Lz
    const double* pos;
    const double* q;
    const int32_t* mol;
    const double* com;
    const double* kvec;
    const double* kmod2;
    const double* expk;
    const int2* grp;           // [ng] (first k, count)
    const int32_t* split;      // [RF_KSPLIT+1] group ranges of the force pass
    double* part;              // [chunk][K][8]
    double* sfac;              // [K][16]
    double* fpart;             // [RF_KSPLIT][n][3]
    double* force;
    double* terms;
    double alphaR4, pi4V, pi8V;
};

template <typename R>
struct cplx { R re, im; };

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// exp(i ph) in R; the FP32 form reduces the phase to [-pi, pi] in FP64 first.
template <typename R>
__device__ __forceinline__ cplx<R> expi(double ph);
template <>
__device__ __forceinline__ cplx<double> expi<double>(double ph) {
    double s, c;
    sincos(ph, &s, &c);
    return {c, s};
}
template <>
__device__ __forceinline__ cplx<float> expi<float>(double ph) {
    const double red = ph - 6.283185307179586 * rint(ph * 0.15915494309189535);
    float s, c;
    sincosf((float)red, &s, &c);
    return {c, s};
}

// Pass 1: lane = k-group, atoms of one chunk streamed through LDS.  Sums per
// k: sum q sin(k.r), sum q cos(k.r) and, for the volume derivatives, their
// moments with r - R_mol (reciprocal_term :568-594).
template <typename R, bool LUSTIG>
__global__ __launch_bounds__(64) void k_recip_sfac(RecipArgs a) {
    __shared__ double tile[9][SF_TILE];   // x y z q dx dy dz cos(vz z) sin(vz z)
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int ch = blockIdx.y;
    const int i0 = ch * a.per, i1 = min(a.n, i0 + a.per);
    int k0 = 0, cnt = 0;
    double kx = 0, ky = 0, kz = 0;
    if (g < a.ng) {
        const int2 gg = a.grp[g];
        k0 = gg.x; cnt = gg.y;
        kx = a.kvec[3 * k0]; ky = a.kvec[3 * k0 + 1]; kz = a.kvec[3 * k0 + 2];
    }
    double acc[GK][8];
#pragma unroll
    for (int t = 0; t < GK; ++t)
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[t][u] = 0.0;
    for (int t0 = i0; t0 < i1; t0 += SF_TILE) {
        const int ia = t0 + threadIdx.x;
        __syncthreads();
        if (ia < i1) {
            const int m = a.mol[ia];
            const double x = a.pos[3 * ia], y = a.pos[3 * ia + 1], z = a.pos[3 * ia + 2];
            tile[0][threadIdx.x] = x; tile[1][threadIdx.x] = y; tile[2][threadIdx.x] = z;
            tile[3][threadIdx.x] = a.q[ia];
            tile[4][threadIdx.x] = x - a.com[3 * m];
            tile[5][threadIdx.x] = y - a.com[3 * m + 1];
            tile[6][threadIdx.x] = z - a.com[3 * m + 2];
            double s, c;
            sincos(a.vz * z, &s, &c);
            tile[7][threadIdx.x] = c; tile[8][threadIdx.x] = s;
        }
        __syncthreads();
        const int na = min(SF_TILE, i1 - t0);
        if (g < a.ng) {
            for (int u = 0; u < na; ++u) {
                const double x = tile[0][u], y = tile[1][u], z = tile[2][u];
                const R q = (R)tile[3][u];
                cplx<R> e = expi<R>(x * kx + y * ky + z * kz);
                const cplx<R> z1{(R)tile[7][u], (R)tile[8][u]};
                const double dx = tile[4][u], dy = tile[5][u], dz = tile[6][u];
#pragma unroll
                for (int t = 0; t < GK; ++t) {
                    if (t < cnt) {
                        const double sq = (double)(q * e.im), cq = (double)(q * e.re);
                        acc[t][0] += sq;
                        acc[t][1] += cq;
                        if constexpr (LUSTIG) {
                            acc[t][2] += cq * dx; acc[t][3] += cq * dy; acc[t][4] += cq * dz;
                            acc[t][5] += sq * dx; acc[t][6] += sq * dy; acc[t][7] += sq * dz;
                        }
                    }
                    e = cmul(e, z1);
                }
            }
        }
    }
    if (g < a.ng) {
#pragma unroll
        for (int t = 0; t < GK; ++t)
            if (t < cnt) {
                double* p = a.part + ((size_t)ch * a.K + k0 + t) * 8;
#pragma unroll
                for (int u = 0; u < 8; ++u) p[u] = acc[t][u];
            }
    }
}

// Pass 1b: chunk partials summed in chunk order; per-k energy, volume
// derivatives and virial (:596-607).  sfac row: 0-2 k, 3 b, 4 b|k|^2,
// 5 4 aR4 |k|^2, 6 C, 7 S, 8-10 cosrR, 11-13 sinrR.
__global__ __launch_bounds__(256) void k_recip_reduce(RecipArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v[3 + 9] = {0};   // energy, longderiv, long2ndDeriv, virial[9]
    if (k < a.K) {
        double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int ch = 0; ch < a.chunks; ++ch) {
            const double* p = a.part + ((size_t)ch * a.K + k) * 8;
#pragma unroll
            for (int u = 0; u < 8; ++u) s[u] += p[u];
        }
        const double kx = a.kvec[3 * k], ky = a.kvec[3 * k + 1], kz = a.kvec[3 * k + 2];
        const double k2 = a.kmod2[k];
        const double sinSum = s[0], cosSum = s[1];
        const double te = a.pi4V * a.expk[k] * (sinSum * sinSum + cosSum * cosSum);
        const double coeff = 2.0 * (a.alphaR4 + 1.0 / k2);
        const double kk = (te * kx) * kx + (te * ky) * ky + (te * kz) * kz;
        v[0] = te;
        v[1] = -te + kk * 2.0 * a.alphaR4;
        const double t7 = ((7.0 * te * kx) * kx + (7.0 * te * ky) * ky + (7.0 * te * kz) * kz) * 2.0 * a.alphaR4;
        const double dd = kk * 2.0 * a.alphaR4;
        const double t2 = ((dd * kx) * kx + (dd * ky) * ky + (dd * kz) * kz) * 2.0 * a.alphaR4;
        v[2] = 4.0 * te - t7 + t2;
        v[3 + 0] = te - te * coeff * kx * kx;
        v[3 + 4] = te - te * coeff * ky * ky;
        v[3 + 8] = te - te * coeff * kz * kz;
        v[3 + 1] = -te * coeff * kx * ky;
        v[3 + 2] = -te * coeff * kx * kz;
        v[3 + 5] = -te * coeff * ky * kz;
        v[3 + 3] = v[3 + 1];
        v[3 + 6] = v[3 + 2];
        v[3 + 7] = v[3 + 5];
        const double b = a.pi8V * a.expk[k];
        double* r = a.sfac + (size_t)k * 16;
        r[0] = kx; r[1] = ky; r[2] = kz; r[3] = b;
        r[4] = (b * kx) * kx + (b * ky) * ky + (b * kz) * kz;
        r[5] = 4.0 * a.alphaR4 * k2;
        r[6] = cosSum; r[7] = sinSum;
        r[8] = s[2]; r[9] = s[3]; r[10] = s[4];
        r[11] = s[5]; r[12] = s[6]; r[13] = s[7];
    }
    block_add<12>(v, a.terms + T_RC_E);
}

// Pass 2: thread per atom over one range of k-groups whose per-k data sit in
// LDS (every lane reads the same k, so the reads broadcast).  Force on atom i
// from k: (8 pi/V) expK q_i (sin_i S_c - cos_i S_s) k (:609-658).
template <typename R, bool LUSTIG>
__global__ __launch_bounds__(256) void k_recip_force(RecipArgs a) {
    extern __shared__ double kd[];
    const int ks = blockIdx.y;
    const int gs = a.split[ks], ge = a.split[ks + 1];
    const int kbeg = gs < ge ? a.grp[gs].x : 0;
    const int kend = gs < ge ? a.grp[ge - 1].x + a.grp[ge - 1].y : 0;
    for (int idx = threadIdx.x; idx < (kend - kbeg) * KD; idx += blockDim.x) {
        const int kk = idx / KD, u = idx % KD;
        kd[idx] = a.sfac[(size_t)(kbeg + kk) * 16 + u];
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double F[3] = {0, 0, 0};
    double l[4] = {0, 0, 0, 0};   // Ff, Ffk2, d2UV1, d2UrR
    if (i < a.n) {
        const int m = a.mol[i];
        const double x = a.pos[3 * i], y = a.pos[3 * i + 1], z = a.pos[3 * i + 2];
        const R q = (R)a.q[i];
        const double dxd = x - a.com[3 * m], dyd = y - a.com[3 * m + 1], dzd = z - a.com[3 * m + 2];
        const R dx = (R)dxd, dy = (R)dyd, dz = (R)dzd;
        const R d2 = dx * dx + dy * dy + dz * dz;
        const cplx<R> z1 = expi<R>(a.vz * z);
        for (int g = gs; g < ge; ++g) {
            const int2 gg = a.grp[g];
            const double* r0 = kd + (gg.x - kbeg) * KD;
            cplx<R> e = expi<R>(x * r0[0] + y * r0[1] + z * r0[2]);
#pragma unroll
            for (int t = 0; t < GK; ++t) {
                if (t < gg.y) {
                    const double* r = r0 + t * KD;
                    const R s = q * e.im, c = q * e.re;
                    const R C = (R)r[6], S = (R)r[7];
                    const R val = (R)r[3] * (s * C - c * S);
                    const R kx = (R)r[0], ky = (R)r[1], kz = (R)r[2];
                    F[0] += (double)(val * kx); F[1] += (double)(val * ky); F[2] += (double)(val * kz);
                    if constexpr (LUSTIG) {
                        const R kdd = kx * dx + ky * dy + kz * dz;
                        l[0] += (double)(val * kdd);
                        l[1] += (double)(val * (R)r[5] * kdd);
                        const R bkk = (R)r[4];
                        l[2] += (double)(bkk * (c * C + s * S) * d2);
                        l[3] += (double)(bkk * ((c * (R)r[8] + s * (R)r[11]) * dx + (c * (R)r[9] + s * (R)r[12]) * dy +
                                                (c * (R)r[10] + s * (R)r[13]) * dz));
                    }
                }
                e = cmul(e, z1);
            }
        }
        double* fp = a.fpart + ((size_t)ks * a.n + i) * 3;
        fp[0] = F[0]; fp[1] = F[1]; fp[2] = F[2];
    }
    if constexpr (LUSTIG) block_add<4>(l, a.terms + T_RC_FF);
}

// Adds the RF_KSPLIT partial forces in a fixed order.
__global__ __launch_bounds__(256) void k_recip_fsum(int n, const double* __restrict__ fpart, double* __restrict__ force) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f[3] = {0, 0, 0};
    for (int ks = 0; ks < RF_KSPLIT; ++ks)
        for (int d = 0; d < 3; ++d) f[d] += fpart[((size_t)ks * n + i) * 3 + d];
    for (int d = 0; d < 3; ++d) force[3 * i + d] += f[d];
}

int ewald_recip(Ctx* c, int prec, int lustig) {
    if (!c->ewald_ready) return fail(c, SMD_ESTATE, "Ewald not set up");
    RecipArgs a;
    a.n = c->n;
    a.K = c->K;
    a.ng = c->n_kgrp;
    a.chunks = c->sfac_chunks;
    a.per = (c->n + a.chunks - 1) / a.chunks;
    a.vz = 2.0 * PI / c->box.L[2];
    a.pos = c->d_pos;
    a.q = c->d_q;
    a.mol = c->d_mol;
    a.com = c->d_com;
    a.kvec = c->d_kvec;
    a.kmod2 = c->d_kmod2;
    a.expk = c->d_expk;
    a.grp = reinterpret_cast<const int2*>(c->d_kgrp);
    a.split = c->d_ksplit;
    a.part = c->d_sfac_part;
    a.sfac = c->d_sfac;
    a.fpart = c->d_fpart;
    a.force = c->d_force;
    a.terms = c->d_terms;
    a.alphaR4 = 1.0 / (4.0 * (c->ew.alpha * c->ew.alpha));
    const double volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
    a.pi4V = 4.0 * PI * (1.0 / volume);
    a.pi8V = 2.0 * a.pi4V;
    const bool mixed = prec == SMD_PREC_MIXED;
    {
        ProfScope ps(c, "recip_sfac");
        const dim3 g1((a.ng + 63) / 64, a.chunks);
        if (mixed) {
            if (lustig) hipLaunchKernelGGL((k_recip_sfac<float, true>), g1, dim3(64), 0, c->stream, a);
            else hipLaunchKernelGGL((k_recip_sfac<float, false>), g1, dim3(64), 0, c->stream, a);
        } else {
            if (lustig) hipLaunchKernelGGL((k_recip_sfac<double, true>), g1, dim3(64), 0, c->stream, a);
            else hipLaunchKernelGGL((k_recip_sfac<double, false>), g1, dim3(64), 0, c->stream, a);
        }
        hipLaunchKernelGGL(k_recip_reduce, dim3((a.K + 255) / 256), dim3(256), 0, c->stream, a);
    }
    {
        ProfScope ps(c, "recip_force");
        const dim3 g2((c->n + 255) / 256, RF_KSPLIT);
        const size_t lds = sizeof(double) * KD * (size_t)c->kmax_range;
        if (mixed) {
            if (lustig) hipLaunchKernelGGL((k_recip_force<float, true>), g2, dim3(256), lds, c->stream, a);
            else hipLaunchKernelGGL((k_recip_force<float, false>), g2, dim3(256), lds, c->stream, a);
        } else {
            if (lustig) hipLaunchKernelGGL((k_recip_force<double, true>), g2, dim3(256), lds, c->stream, a);
            else hipLaunchKernelGGL((k_recip_force<double, false>), g2, dim3(256), lds, c->stream, a);
        }
        hipLaunchKernelGGL(k_recip_fsum, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n, c->d_fpart, c->d_force);
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

extern "C" int smd_ewald_compute_long(smd_ctx* c, int prec, double* f_accum, smd_ewald_long_out* out) {
    // EwaldForce::compute_long (src/forces/EwaldForce.cpp:178-196)
    if (!c || !out) return SMD_EINVAL;
    if (!c->ewald_ready) return fail(c, SMD_ESTATE, "Ewald not set up");
    if (f_accum) SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_terms, 0, sizeof(double) * T_COUNT, c->stream));
    int rc = ewald_recip(c, prec, 1);
    if (rc) return rc;
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
    if (f_accum) return fetch_forces_into(c, f_accum);
    return SMD_OK;
}
