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
#include "smd_internal.h"

namespace smd {

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
    c->tab_stride = (nmax[0] + 1) + (nmax[1] + 1) + (nmax[2] + 1);
    if ((rc = realloc(c->d_tab, sizeof(double) * 2 * (size_t)c->tab_stride * c->n))) return rc;
    // Chunking of pass 1: enough blocks to fill the device.
    const int kt = (K + WAVE - 1) / WAVE;
    int chunks = (2048 + kt - 1) / kt;
    int per = (c->n + chunks - 1) / chunks;
    if (per < 32) per = 32;
    chunks = (c->n + per - 1) / per;
    c->sfac_chunks = chunks;
    if ((rc = realloc(c->d_sfac_part, sizeof(double) * 8 * (size_t)K * chunks))) return rc;
    if ((rc = realloc(c->d_sfac, sizeof(double) * 16 * (size_t)K))) return rc;
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
    int n, K, chunks, per, tab_stride;
    int nmax[3];
    double v[3];       // 2 pi / L
    const double* pos;
    const double* q;
    const int32_t* mol;
    const double* com;
    const double* kvec;
    const double* kmod2;
    const double* expk;
    const int32_t* kn;
    double* tab;
    double* part;      // [chunk][K][8]
    double* sfac;      // [K][16]
    double* force;
    double* terms;
    double alphaR4, pi4V, pi8V;
    int use_constraint;
};

template <typename R>
struct cplx { R re, im; };

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Per-atom tables X[0..nx], Y[0..ny], Z[0..nz] with X[n] = exp(i n 2pi x/L).
__global__ __launch_bounds__(256) void k_recip_tables(RecipArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double* t = a.tab + (size_t)i * 2 * a.tab_stride;
    int o = 0;
    for (int d = 0; d < 3; ++d) {
        const double th = a.v[d] * a.pos[3 * i + d];
        for (int m = 0; m <= a.nmax[d]; ++m, ++o) {
            double s, cc;
            sincos((double)m * th, &s, &cc);
            t[2 * o] = cc;
            t[2 * o + 1] = s;
        }
    }
}

__device__ __forceinline__ cplx<double> phase(const double* t, int nx, int ny, int nz, int oy, int oz) {
    cplx<double> X{t[2 * nx], t[2 * nx + 1]};
    const int ay = ny < 0 ? -ny : ny, az = nz < 0 ? -nz : nz;
    cplx<double> Y{t[2 * (oy + ay)], t[2 * (oy + ay) + 1]};
    cplx<double> Z{t[2 * (oz + az)], t[2 * (oz + az) + 1]};
    if (ny < 0) Y.im = -Y.im;
    if (nz < 0) Z.im = -Z.im;
    return cmul(cmul(X, Y), Z);
}

__global__ __launch_bounds__(64) void k_recip_sfac(RecipArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int ch = blockIdx.y;
    if (k >= a.K) return;
    const int nx = a.kn[3 * k], ny = a.kn[3 * k + 1], nz = a.kn[3 * k + 2];
    const int oy = a.nmax[0] + 1, oz = oy + a.nmax[1] + 1;
    double S = 0, Cs = 0, cx = 0, cy = 0, cz = 0, sx = 0, sy = 0, sz = 0;
    const int i0 = ch * a.per, i1 = min(a.n, i0 + a.per);
    for (int i = i0; i < i1; ++i) {
        const double* t = a.tab + (size_t)i * 2 * a.tab_stride;
        const cplx<double> e = phase(t, nx, ny, nz, oy, oz);
        const double qi = a.q[i];
        const double sq = qi * e.im, cq = qi * e.re;   // sinAqi, cosAqi (:586-587)
        const int m = a.mol[i];
        const double dx = a.pos[3 * i] - a.com[3 * m], dy = a.pos[3 * i + 1] - a.com[3 * m + 1],
                     dz = a.pos[3 * i + 2] - a.com[3 * m + 2];
        S += sq; Cs += cq;
        cx += cq * dx; cy += cq * dy; cz += cq * dz;
        sx += sq * dx; sy += sq * dy; sz += sq * dz;
    }
    double* p = a.part + ((size_t)ch * a.K + k) * 8;
    p[0] = S; p[1] = Cs; p[2] = cx; p[3] = cy; p[4] = cz; p[5] = sx; p[6] = sy; p[7] = sz;
}

// sfac row layout: 0 kx 1 ky 2 kz 3 b 4 b|k|^2 5 4 aR4 |k|^2 6 C 7 S 8-10 cosrR 11-13 sinrR
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
        // reciprocal_term (:596-607)
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

template <bool LUSTIG, int G>
__global__ __launch_bounds__(256) void k_recip_force(RecipArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double fx = 0, fy = 0, fz = 0;
    double l[4] = {0, 0, 0, 0};   // Ff, Ffk2, d2UV1, d2UrR
    if (i < a.n) {
        const double* t = a.tab + (size_t)i * 2 * a.tab_stride;
        const int oy = a.nmax[0] + 1, oz = oy + a.nmax[1] + 1;
        const double qi = a.q[i];
        const int m = a.mol[i];
        const double dx = a.pos[3 * i] - a.com[3 * m], dy = a.pos[3 * i + 1] - a.com[3 * m + 1],
                     dz = a.pos[3 * i + 2] - a.com[3 * m + 2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        for (int k = sub; k < a.K; k += G) {
            const int nx = a.kn[3 * k], ny = a.kn[3 * k + 1], nz = a.kn[3 * k + 2];
            const cplx<double> e = phase(t, nx, ny, nz, oy, oz);
            const double s = qi * e.im, c = qi * e.re;
            const double* r = a.sfac + (size_t)k * 16;
            const double b = r[3], C = r[6], S = r[7];
            const double val = b * (s * C - c * S);   // :613
            fx += val * r[0]; fy += val * r[1]; fz += val * r[2];
            if constexpr (LUSTIG) {
                const double kd = r[0] * dx + r[1] * dy + r[2] * dz;
                l[0] += val * kd;
                l[1] += val * r[5] * kd;
                const double bkk = r[4];
                l[2] += bkk * (c * C + s * S) * d2;
                l[3] += bkk * ((c * r[8] + s * r[11]) * dx + (c * r[9] + s * r[12]) * dy + (c * r[10] + s * r[13]) * dz);
            }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        fx += __shfl_xor(fx, o, G);
        fy += __shfl_xor(fy, o, G);
        fz += __shfl_xor(fz, o, G);
    }
    if (i < a.n && sub == 0) atomic_add3(a.force + 3 * i, fx, fy, fz);
    if constexpr (LUSTIG) block_add<4>(l, a.terms + T_RC_FF);
}

int ewald_recip(Ctx* c, int prec, int lustig) {
    (void)prec;
    if (!c->ewald_ready) return fail(c, SMD_ESTATE, "Ewald not set up");
    RecipArgs a;
    a.n = c->n;
    a.K = c->K;
    a.chunks = c->sfac_chunks;
    a.per = (c->n + a.chunks - 1) / a.chunks;
    a.tab_stride = c->tab_stride;
    for (int d = 0; d < 3; ++d) { a.nmax[d] = c->nmax[d]; a.v[d] = 2.0 * PI / c->box.L[d]; }
    a.pos = c->d_pos;
    a.q = c->d_q;
    a.mol = c->d_mol;
    a.com = c->d_com;
    a.kvec = c->d_kvec;
    a.kmod2 = c->d_kmod2;
    a.expk = c->d_expk;
    a.kn = c->d_kn;
    a.tab = c->d_tab;
    a.part = c->d_sfac_part;
    a.sfac = c->d_sfac;
    a.force = c->d_force;
    a.terms = c->d_terms;
    a.alphaR4 = 1.0 / (4.0 * (c->ew.alpha * c->ew.alpha));
    const double volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
    a.pi4V = 4.0 * PI * (1.0 / volume);
    a.pi8V = 2.0 * a.pi4V;
    a.use_constraint = c->ew.use_constraint;
    {
        ProfScope ps(c, "recip_tables");
        hipLaunchKernelGGL(k_recip_tables, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, a);
    }
    {
        ProfScope ps(c, "recip_sfac");
        hipLaunchKernelGGL(k_recip_sfac, dim3((a.K + 63) / 64, a.chunks), dim3(64), 0, c->stream, a);
        hipLaunchKernelGGL(k_recip_reduce, dim3((a.K + 255) / 256), dim3(256), 0, c->stream, a);
    }
    {
        ProfScope ps(c, "recip_force");
        constexpr int G = 16;
        const dim3 grid((unsigned)(((int64_t)c->n * G + 255) / 256));
        if (lustig) hipLaunchKernelGGL((k_recip_force<true, G>), grid, dim3(256), 0, c->stream, a);
        else hipLaunchKernelGGL((k_recip_force<false, G>), grid, dim3(256), 0, c->stream, a);
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
