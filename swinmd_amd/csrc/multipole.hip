// multipole.hip -- permanent-multipole Ewald (charges, dipoles and traced
// quadrupoles with torques) on gfx950.
//
//   Multipole::compute          src/forces/Multipole.cpp:163-216
//   Multipole::compute_long     :218-240
//   self_correct                :338-395
//   real_term                   :398-949
//   reciprocal_term             :951-1279
//   intraMolCorrect             :1290-1750
//   surfDipoleCorrect           :1753-1945
//   torques2forces              :1946-2136
//
// Sites are water atoms H1, O, H2 with lab-frame multipoles from k_frames.
//  * Real space walks the full neighbour rows of the pair list (both
//    segments), so each row gathers the force and the torque on its own atom
//    with no atomics.  The pair terms are written once, for the site that owns
//    the row: the torque the reference puts on the partner of a pair is the
//    same expression with the roles swapped, and the force is its negative.
//    The owned segment adds the energy, virial and volume derivative.
//  * Each part's torques (real space, intramolecular, surface, k-space) are
//    kept apart and turned into forces per molecule, so each part reports its
//    own dU/dV as the reference does.
//  * k-space: a structure-factor pass forms 20 sums per k -- the six the
//    reference forms and the moments from which its per-k force-pass sums
//    (Ff, d2UV1, d2UrR) follow in closed form -- then a per-k finish, then a
//    per-atom pass over all k with two coefficients per k:
//      P = Sc - Ms,  Q = Ss + Mc  give  val = s (A P + Mu Q) + c (Mu P - A Q)
//    and the torques (c Q - s P) k x mu and -2 (c P + s Q) (kk^T x Qi).
// All arithmetic is FP64.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "smd_internal.h"

namespace smd {

constexpr double COULOMBS = 138.93548501;      // COUlOMBS, src/NEMD_defs.h:104
constexpr double SQRT_COULOMB_MP = 11.78708976;
constexpr int MP_LANES = 16;                   // lanes per atom row in real space
constexpr int MP_NSUM = 20;                    // structure-factor sums per k
constexpr int MP_SF_THREADS = 256;             // k per block of the structure-factor pass
constexpr int MP_RF_THREADS = 256;             // atoms per block of the force pass
constexpr int MP_KTILE = 128;                  // k staged per LDS tile in the force pass

// Accumulator slots (FP64, zeroed per call).
enum MpSlot {
    M_SELF = 0, M_REAL_E, M_REAL_D1, M_REAL_VIR,                 // VIR: 9
    M_INTRA_E = M_REAL_VIR + 9, M_MOL_D1,
    // surface sums: q r, q R_mol, q r + mu (multipoles) or mu (induction),
    // the self sum, then the induced dipoles' sum and their squares
    M_SU_QR, M_SU_QRR = M_SU_QR + 3, M_SU_QM = M_SU_QRR + 3, M_SU_SELF = M_SU_QM + 3, M_SU_PI,
    M_SU_PP = M_SU_PI + 3, M_SU_VIR,                              // VIR: 9
    M_DUDV = M_SU_VIR + 9,                                        // 5 parts: real, intra, surface, long, self
    M_TRQ = M_DUDV + 5,                                           // 5: |torque| sums per part
    M_LONG_E = M_TRQ + 5, M_LONG_D1, M_LONG_D2, M_LONG_VIR,      // VIR: 6 (xx yy zz xy xz yz)
    M_COUNT = M_LONG_VIR + 6
};

struct MpArgs {
    int n, nmol, stride, K, chunks, per;
    Box box;
    // the mixed solve's coupling table (the partners inside cutoffInd, the
    // intramolecular ones included), when Induction's real space walks it
    const int32_t* tcnt;
    const int32_t* tj;
    int tstride;
    const int* kstep;           // the k set's nz-step flags (see KSet)
    double kz_step;             // 2 pi / Lz
    const int32_t* mol_start;
    const int32_t* mol;
    const double* pos;
    const double* com;
    const double* q;            // scaled charges
    const double* dip;          // lab frame [3n]
    const double* quad;         // lab frame, traced / 3 [6n]: xx xy xz yy yz zz
    const int32_t* nb;
    const int32_t* nb_cnt;
    const int32_t* nb_own;
    const int32_t* excl_off;
    const int32_t* excl;
    double* force;
    double* trq;                // [5][3n]: real, intra, surface, k-space, self
    double* terms;              // [M_COUNT]
    Red red;
    const double* kvec;         // [3K]
    const double* kmod2;        // [K]
    const double* expk;         // [K]
    double* part;               // [chunks][K][sums]
    double* kf;                 // [K][4]: force-pass coefficients
    double* fpart;              // [ksplit][n][6]: force and torque of each k-range
    int ksplit, kper;           // k-ranges of the force pass and their length
    double alpha, alpha2, alpha2PI, realcut2, volume, alphaR4;
    int use_constraint, torques;
    // induction terms
    const double* pind;         // induced dipoles [3n]
    const double* polar;        // polarizabilities [n]
    double a_damp;
    int damp_amoeba, do_min;
    // fixed-order partials of the per-site sums (Ctx::part_su) and their ticket
    double* spart;
    // mixed form: FP32 site records [3n] and induced dipoles [n]
    const float4* site32;
    const float4* pind32;
};

// Vector, symmetric-matrix and site types in the pair math's precision R
// (double for the plain form, float for the mixed form's pair math).
template <typename R> struct V3T { R x, y, z; };
using V3 = V3T<double>;
template <typename R> __device__ __forceinline__ V3T<R> vec3(R x, R y, R z) { return {x, y, z}; }
__device__ __forceinline__ V3 v3(double x, double y, double z) { return {x, y, z}; }
template <typename R> __device__ __forceinline__ V3T<R> operator+(V3T<R> a, V3T<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> __device__ __forceinline__ V3T<R> operator-(V3T<R> a, V3T<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> __device__ __forceinline__ V3T<R> operator-(V3T<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> __device__ __forceinline__ V3T<R> operator*(R f, V3T<R> a) { return {f * a.x, f * a.y, f * a.z}; }
template <typename R> __device__ __forceinline__ R dot(V3T<R> a, V3T<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> __device__ __forceinline__ V3T<R> cross(V3T<R> a, V3T<R> b) {
    return {a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ V3 ldv(const double* p) { return {p[0], p[1], p[2]}; }
// FP64 accumulation of an R-precision vector
template <typename R> __device__ __forceinline__ V3 to64(V3T<R> a) { return {(double)a.x, (double)a.y, (double)a.z}; }

// Symmetric matrix as xx xy xz yy yz zz.
template <typename R> struct S6T { R xx, xy, xz, yy, yz, zz; };
using S6 = S6T<double>;
__device__ __forceinline__ S6 lds6(const double* p) { return {p[0], p[1], p[2], p[3], p[4], p[5]}; }
template <typename R> __device__ __forceinline__ V3T<R> smul(const S6T<R>& Q, V3T<R> r) {
    return {Q.xx * r.x + Q.xy * r.y + Q.xz * r.z, Q.xy * r.x + Q.yy * r.y + Q.yz * r.z,
            Q.xz * r.x + Q.yz * r.y + Q.zz * r.z};
}
template <typename R> __device__ __forceinline__ R trace(const S6T<R>& Q) { return Q.xx + Q.yy + Q.zz; }

// cross(A, B, ones) of the reference (src/utils/Vector3.h:220-225) for
// A = a b^T and B symmetric: sum over columns c of the cross product of
// column c of A's rows, i.e. cross(a, B b) with B b the matrix-vector product.
template <typename R>
__device__ __forceinline__ V3T<R> mcross_outer(V3T<R> a, V3T<R> b, const S6T<R>& B) { return cross(a, smul(B, b)); }
// cross(A, B, ones) for A, B symmetric: x = sum_c A[1][c] B[2][c] - A[2][c] B[1][c], ...
template <typename R>
__device__ __forceinline__ V3T<R> mcross_sym(const S6T<R>& A, const S6T<R>& B) {
    return {A.xy * B.xz - A.xz * B.xy + A.yy * B.yz - A.yz * B.yy + A.yz * B.zz - A.zz * B.yz,
            A.xz * B.xx - A.xx * B.xz + A.yz * B.xy - A.xy * B.yz + A.zz * B.xz - A.xz * B.zz,
            A.xx * B.xy - A.xy * B.xx + A.xy * B.yy - A.yy * B.xy + A.xz * B.yz - A.yz * B.xz};
}

template <typename R> struct SiteT { R q; V3T<R> d; S6T<R> Q; };
using Site = SiteT<double>;

__device__ __forceinline__ Site load_site(const MpArgs& a, int i) {
    Site s;
    s.q = a.q[i] / SQRT_COULOMB_MP;
    s.d = ldv(a.dip + 3 * i);
    s.Q = lds6(a.quad + 6 * i);
    return s;
}

template <typename R> __device__ __forceinline__ SiteT<R> site_as(const Site& s) {
    return {(R)s.q, {(R)s.d.x, (R)s.d.y, (R)s.d.z}, {(R)s.Q.xx, (R)s.Q.xy, (R)s.Q.xz, (R)s.Q.yy, (R)s.Q.yz, (R)s.Q.zz}};
}
template <typename R> __device__ __forceinline__ SiteT<R> load_site_as(const MpArgs& a, int i) {
    return site_as<R>(load_site(a, i));
}

// Multipole pair of "self" site I and partner J at r = r_I - r_J: the G_l
// coefficients, the force on I in units of the B_l (or D_l) radial functions,
// and the torque on I (real_term :548-617 and :730-877).  The torque the
// reference assigns to the partner of a pair (Gt_j) is this expression with
// the roles of the two sites swapped.
template <typename R> struct PairOutT { R G[5]; V3T<R> gG[5]; V3T<R> Gt[5]; };
using PairOut = PairOutT<double>;

template <bool GRAD, bool TORQUE, typename R>
__device__ __forceinline__ void mp_pair(const SiteT<R>& I, const SiteT<R>& J, V3T<R> r, PairOutT<R>& o) {
    const R qi = I.q, qj = J.q;
    const V3T<R> mi = I.d, mj = J.d;
    const R QiI = trace(I.Q), QjI = trace(J.Q);
    const V3T<R> Qir = smul(I.Q, r), Qjr = smul(J.Q, r);
    const R QiR = dot(r, Qir), QjR = dot(r, Qjr);
    const R QiQj = I.Q.xx * J.Q.xx + I.Q.yy * J.Q.yy + I.Q.zz * J.Q.zz +
                   R(2) * (I.Q.xy * J.Q.xy + I.Q.xz * J.Q.xz + I.Q.yz * J.Q.yz);
    const R mir = dot(mi, r), mjr = dot(mj, r);
    const R QimjXr = dot(mj, Qir), QjmiXr = dot(mi, Qjr);
    const R QirQj = dot(Qir, Qjr);   // r Qi Qj r
    o.G[0] = qi * qj;
    o.G[1] = qj * mir - qi * mjr - dot(mi, mj) + qj * QiI + qi * QjI;
    o.G[2] = qj * QiR + qi * QjR - mir * mjr - mjr * QiI + mir * QjI - R(2) * QimjXr + R(2) * QjmiXr + QiI * QjI +
             R(2) * QiQj;
    o.G[3] = mir * QjR - mjr * QiR + QiI * QjR + QjI * QiR + R(4) * QirQj;
    o.G[4] = QiR * QjR;
    if constexpr (GRAD) {
        const V3T<R> Qimj = smul(I.Q, mj), Qjmi = smul(J.Q, mi);
        const V3T<R> QiQjr = smul(I.Q, Qjr), QjQir = smul(J.Q, Qir);
        o.gG[1] = qj * mi - qi * mj;
        o.gG[2] = (R(2) * qj) * Qir - mjr * mi - mir * mj - R(2) * Qimj - QiI * mj + (R(2) * qi) * Qjr + R(2) * Qjmi +
                  QjI * mi;
        o.gG[3] = (-QiR) * mj - (R(2) * mjr) * Qir + QjR * mi + (R(2) * mir) * Qjr + (R(2) * QiI) * Qjr +
                  (R(2) * QjI) * Qir + R(4) * QiQjr + R(4) * QjQir;
        o.gG[4] = (R(2) * QjR) * Qir + (R(2) * QiR) * Qjr;
    }
    if constexpr (TORQUE) {
        const V3T<R> rXmi = cross(r, mi);
        const V3T<R> mjXmi = cross(mj, mi);
        const V3T<R> QjrXmi = cross(Qjr, mi);
        const V3T<R> RXQi = mcross_outer(r, r, I.Q);          // cross(r r^T, Qi)
        const V3T<R> mjrXQi = mcross_outer(mj, r, I.Q);       // cross(mj r^T, Qi)
        const V3T<R> rmjXQi = mcross_outer(r, mj, I.Q);       // cross(r mj^T, Qi)
        const V3T<R> QjXQi = mcross_sym(J.Q, I.Q);
        const V3T<R> rQjrXQi = mcross_outer(r, Qjr, I.Q);     // cross(r (Qj r)^T, Qi)
        const V3T<R> QjrrXQi = mcross_outer(Qjr, r, I.Q);     // cross((Qj r) r^T, Qi)
        o.Gt[1] = qj * rXmi - mjXmi;
        o.Gt[2] = (-mjr) * rXmi + R(2) * QjrXmi + QjI * rXmi + (R(2) * qj) * RXQi - R(2) * mjrXQi - R(2) * rmjXQi +
                  R(4) * QjXQi;
        o.Gt[3] = QjR * rXmi - (R(2) * mjr) * RXQi + (R(2) * QjI) * RXQi + R(4) * rQjrXQi + R(4) * QjrrXQi;
        o.Gt[4] = (R(2) * QjR) * RXQi;
    }
}

// erfc-screened radial functions B_0..B_5 (:629-634).  FP32: exp2 and the
// fitted erfc that shares exp(-alpha^2 r^2), a reciprocal of r^2.
template <typename R>
__device__ __forceinline__ void radial_b(const MpArgs& a, R r2, R (&B)[6]) {
    if constexpr (sizeof(R) == sizeof(double)) {
        const double r = sqrt(r2);
        const double e = exp(-a.alpha2 * r2);
        const double inv = 1.0 / r2;
        B[0] = erfc(a.alpha * r) / r;
        B[1] = (-B[0] - a.alpha2PI * e) * inv;
        B[2] = (-3.0 * B[1] + 2.0 * a.alpha2PI * a.alpha2 * e) * inv;
        B[3] = (-5.0 * B[2] - 4.0 * a.alpha2PI * a.alpha2 * a.alpha2 * e) * inv;
        B[4] = (-7.0 * B[3] + 8.0 * a.alpha2PI * a.alpha2 * a.alpha2 * a.alpha2 * e) * inv;
        B[5] = (-9.0 * B[4] - 16.0 * a.alpha2PI * a.alpha2 * a.alpha2 * a.alpha2 * a.alpha2 * e) * inv;
    } else {
        const float r_1 = __builtin_amdgcn_rsqf(r2), r = r2 * r_1;
        const float al = (float)a.alpha, al2 = (float)a.alpha2;
        float e, b0;
        ewald_erfc_f32(al, al2, r2, r, r_1, b0, e);
        const float inv = r_1 * r_1;
        const float c1 = (float)a.alpha2PI, c2 = (float)(2.0 * a.alpha2PI * a.alpha2),
                    c3 = (float)(4.0 * a.alpha2PI * a.alpha2 * a.alpha2),
                    c4 = (float)(8.0 * a.alpha2PI * a.alpha2 * a.alpha2 * a.alpha2),
                    c5 = (float)(16.0 * a.alpha2PI * a.alpha2 * a.alpha2 * a.alpha2 * a.alpha2);
        B[0] = b0;
        B[1] = (-B[0] - c1 * e) * inv;
        B[2] = (-3.0f * B[1] + c2 * e) * inv;
        B[3] = (-5.0f * B[2] - c3 * e) * inv;
        B[4] = (-7.0f * B[3] + c4 * e) * inv;
        B[5] = (-9.0f * B[4] - c5 * e) * inv;
    }
}

// sin and cos of a k-space phase k.r: FP64, or (mixed form) the phase
// formed and reduced to [-pi, pi] in FP64 and the hardware FP32 sin/cos.
template <typename R> __device__ __forceinline__ void phase_sincos(double ph, R& sn, R& cs) {
    if constexpr (sizeof(R) == sizeof(double)) {
        sincos(ph, &sn, &cs);
    } else {
        constexpr double TWO_PI = 6.283185307179586476925, INV_TWO_PI = 0.159154943091895335769;
        const float t = (float)(ph - TWO_PI * rint(ph * INV_TWO_PI));
        sn = __sinf(t);
        cs = __cosf(t);
    }
}

// ---- real space ---------------------------------------------------------------

// Site j in the pair math's precision: FP64 from the lab-frame arrays, FP32
// from the packed records (q, mu | Qxx Qxy Qxz Qyy | Qyz Qzz alpha -).
template <typename R> __device__ __forceinline__ SiteT<R> site_r(const MpArgs& a, int j) {
    if constexpr (sizeof(R) == sizeof(double)) {
        return load_site(a, j);
    } else {
        const float4 s0 = a.site32[3 * j], s1 = a.site32[3 * j + 1], s2 = a.site32[3 * j + 2];
        return {s0.x, {s0.y, s0.z, s0.w}, {s1.x, s1.y, s1.z, s1.w, s2.x, s2.y}};
    }
}

template <bool FULL, typename R>
__device__ __forceinline__ void real_entry(const MpArgs& a, const SiteT<R>& I, V3 ri, int j, V3& f, V3& t,
                                           double (&acc)[11], double w = 1.0) {
    double dx = ri.x - a.pos[3 * j], dy = ri.y - a.pos[3 * j + 1], dz = ri.z - a.pos[3 * j + 2];
    pbc(dx, dy, dz, a.box);
    const double r2d = dx * dx + dy * dy + dz * dz;
    if (!(r2d < a.realcut2)) return;
    const SiteT<R> J = site_r<R>(a, j);
    const V3T<R> r = vec3<R>((R)dx, (R)dy, (R)dz);
    PairOutT<R> o;
    R B[6];
    radial_b<R>(a, (R)r2d, B);
    if (a.torques) mp_pair<true, true>(I, J, r, o);
    else mp_pair<true, false>(I, J, r, o);
    const R radial = o.G[0] * B[1] + o.G[1] * B[2] + o.G[2] * B[3] + o.G[3] * B[4] + o.G[4] * B[5];
    const V3T<R> fij = -(B[1] * o.gG[1] + B[2] * o.gG[2] + B[3] * o.gG[3] + B[4] * o.gG[4]) - radial * r;
    f = f + to64(fij);
    if (a.torques) t = t + to64(B[1] * o.Gt[1] + B[2] * o.Gt[2] + B[3] * o.Gt[3] + B[4] * o.Gt[4]);
    if constexpr (FULL) {
        acc[0] += w * (double)(o.G[0] * B[0] + o.G[1] * B[1] + o.G[2] * B[2] + o.G[3] * B[3] + o.G[4] * B[4]);
        const V3 fd = w * to64(fij), rd = to64(r);
        acc[1] += -dot(fd, rd);
        acc[2] += fd.x * rd.x; acc[3] += fd.y * rd.x; acc[4] += fd.z * rd.x;
        acc[5] += fd.x * rd.y; acc[6] += fd.y * rd.y; acc[7] += fd.z * rd.y;
        acc[8] += fd.x * rd.z; acc[9] += fd.y * rd.z; acc[10] += fd.z * rd.z;
    }
}

template <int G, typename R>
__global__ __launch_bounds__(256) void k_mp_real(MpArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double acc[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) acc[k] = 0.0;
    V3 f = v3(0, 0, 0), t = v3(0, 0, 0);
    if (i < a.n) {
        const int32_t* row = a.nb + (int64_t)i * a.stride;
        const int own = min(a.nb_own[i], a.stride);
        const int non = min(a.nb_cnt[i] - a.nb_own[i], a.stride - own);
        const SiteT<R> I = site_r<R>(a, i);
        const V3 ri = ldv(a.pos + 3 * i);
        for (int k = sub; k < own; k += G) real_entry<true, R>(a, I, ri, row[k], f, t, acc);
        const int32_t* back = row + a.stride - 1;
        for (int k = sub; k < non; k += G) real_entry<false, R>(a, I, ri, back[-k], f, t, acc);
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        f.x += __shfl_xor(f.x, o, G); f.y += __shfl_xor(f.y, o, G); f.z += __shfl_xor(f.z, o, G);
        t.x += __shfl_xor(t.x, o, G); t.y += __shfl_xor(t.y, o, G); t.z += __shfl_xor(t.z, o, G);
    }
    if (i < a.n && sub == 0) {
        a.force[3 * i] += COULOMBS * f.x;
        a.force[3 * i + 1] += COULOMBS * f.y;
        a.force[3 * i + 2] += COULOMBS * f.z;
        double* tr = a.trq;   // real-space part
        tr[3 * i] = COULOMBS * t.x;
        tr[3 * i + 1] = COULOMBS * t.y;
        tr[3 * i + 2] = COULOMBS * t.z;
    }
    acc[0] *= COULOMBS;
    acc[1] *= COULOMBS;
    block_part<11>(acc, a.red, RS_MP_REAL);
}

// The mixed real space over the solve's coupling table instead of the list
// (the table's rows hold exactly the partners inside cutoffInd, equal to the
// real-space cutoff when this kernel is used): the intramolecular entries
// are skipped (k_mp_intra has them), and every pair, seen from both rows,
// adds energy, virial and dU/dV at half weight.
template <int G>
__global__ __launch_bounds__(256) void k_mp_real_tab(MpArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double acc[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) acc[k] = 0.0;
    V3 f = v3(0, 0, 0), t = v3(0, 0, 0);
    if (i < a.n) {
        const int cnt = a.tcnt[i];
        const int32_t* row = a.tj + (int64_t)i * a.tstride;
        const SiteT<float> I = site_r<float>(a, i);
        const V3 ri = ldv(a.pos + 3 * i);
        const int mi = i / 3;
        for (int k = sub; k < cnt; k += G) {
            const int j = row[k];
            if (j / 3 != mi) real_entry<true, float>(a, I, ri, j, f, t, acc, 0.5);
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        f.x += __shfl_xor(f.x, o, G); f.y += __shfl_xor(f.y, o, G); f.z += __shfl_xor(f.z, o, G);
        t.x += __shfl_xor(t.x, o, G); t.y += __shfl_xor(t.y, o, G); t.z += __shfl_xor(t.z, o, G);
    }
    if (i < a.n && sub == 0) {
        a.force[3 * i] += COULOMBS * f.x;
        a.force[3 * i + 1] += COULOMBS * f.y;
        a.force[3 * i + 2] += COULOMBS * f.z;
        double* tr = a.trq;   // real-space part
        tr[3 * i] = COULOMBS * t.x;
        tr[3 * i + 1] = COULOMBS * t.y;
        tr[3 * i + 2] = COULOMBS * t.z;
    }
    acc[0] *= COULOMBS;
    acc[1] *= COULOMBS;
    block_part<11>(acc, a.red, RS_MP_REAL);
}

template <int G, typename R> __global__ __launch_bounds__(256) void k_ind_real(MpArgs a);
template <int G> __global__ __launch_bounds__(256) void k_ind_real_tab(MpArgs a);
__global__ __launch_bounds__(256) void k_pack_sites32(MpArgs a, float4* __restrict__ site32,
                                                      float4* __restrict__ pind32, int with_pind);

// Real-space kernels in the context's precision for these classes (FP64, or
// FP32 pair math from packed records with FP64 accumulation); the mixed form
// first packs the records (and, for induction, the induced dipoles).
static int launch_real(Ctx* c, const MpArgs& a, bool induction) {
    const dim3 gr((unsigned)(((int64_t)c->n * MP_LANES + 255) / 256)), b256(256);
    if (c->mp_prec == SMD_PREC_MIXED) {
        hipLaunchKernelGGL(k_pack_sites32, dim3((c->n + 255) / 256), b256, 0, c->stream, a, c->d_site32, c->d_pind32,
                           induction ? 1 : 0);
        if (induction && a.tj) hipLaunchKernelGGL((k_ind_real_tab<MP_LANES>), gr, b256, 0, c->stream, a);
        else if (induction) hipLaunchKernelGGL((k_ind_real<MP_LANES, float>), gr, b256, 0, c->stream, a);
        else if (a.tj) hipLaunchKernelGGL((k_mp_real_tab<MP_LANES>), gr, b256, 0, c->stream, a);
        else hipLaunchKernelGGL((k_mp_real<MP_LANES, float>), gr, b256, 0, c->stream, a);
    } else {
        if (induction) hipLaunchKernelGGL((k_ind_real<MP_LANES, double>), gr, b256, 0, c->stream, a);
        else hipLaunchKernelGGL((k_mp_real<MP_LANES, double>), gr, b256, 0, c->stream, a);
    }
    return red_finish(c, induction ? RS_IND_REAL : RS_MP_REAL, gr.x, 11, a.terms + M_REAL_E);
}

__global__ __launch_bounds__(256) void k_mp_torques(MpArgs a, unsigned parts, int site);

// Torques to forces, then the sums of its dU/dV and torque terms (sites
// `site` and `site + 1`: each launch of a step has its own pair).
static int launch_torques(Ctx* c, const MpArgs& a, dim3 g, unsigned parts, int site) {
    hipLaunchKernelGGL(k_mp_torques, g, dim3(256), 0, c->stream, a, parts, site);
    int rc = red_finish(c, site, g.x, 5, a.terms + M_DUDV);
    if (!rc) rc = red_finish(c, site + 1, g.x, 5, a.terms + M_TRQ);
    return rc;
}

// The site-sums kernels' totals: row 0 the self energy, the rest the surface
// sums (10 for the permanent multipoles, 14 with the induced dipoles).
static int site_sums_finish(Ctx* c, const MpArgs& a, unsigned nb, bool induced) {
    return rows_finish(c, c->stream, a.spart, (int)nb, (int)nb, induced ? 15 : 11, a.terms + M_SELF, 1,
                       a.terms + M_SU_QR);
}

// FP32 records of each site's lab-frame multipoles and polarizability for the
// mixed form's real-space partner gathers (3 x 16 bytes per site), and of the
// induced dipoles (16 bytes per site).
__global__ __launch_bounds__(256) void k_pack_sites32(MpArgs a, float4* __restrict__ site32,
                                                      float4* __restrict__ pind32, int with_pind) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double* d = a.dip + 3 * i;
    const double* Q = a.quad + 6 * i;
    site32[3 * i] = make_float4((float)(a.q[i] / SQRT_COULOMB_MP), (float)d[0], (float)d[1], (float)d[2]);
    site32[3 * i + 1] = make_float4((float)Q[0], (float)Q[1], (float)Q[2], (float)Q[3]);
    site32[3 * i + 2] = make_float4((float)Q[4], (float)Q[5], (float)a.polar[i], 0.0f);
    if (with_pind)
        pind32[i] = make_float4((float)a.pind[3 * i], (float)a.pind[3 * i + 1], (float)a.pind[3 * i + 2], 0.0f);
}

// ---- self, intramolecular, surface --------------------------------------------

// Per site: the self term (:356-392) and the surface sums (:1785-1830).
__global__ __launch_bounds__(256) void k_mp_site_sums(MpArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < a.n) {
        const Site I = load_site(a, i);
        const double QiI = trace(I.Q);
        const double QiQi = I.Q.xx * I.Q.xx + I.Q.yy * I.Q.yy + I.Q.zz * I.Q.zz +
                            2.0 * (I.Q.xy * I.Q.xy + I.Q.xz * I.Q.xz + I.Q.yz * I.Q.yz);
        const double selfterm = I.q * I.q * a.alpha2PI +
                                (dot(I.d, I.d) - 2.0 * I.q * QiI) * 2.0 * a.alpha2PI * a.alpha2 / 3.0 +
                                (QiI * QiI + 2.0 * QiQi) * 4.0 * a.alpha2PI * a.alpha2 * a.alpha2 / 5.0;
        v[0] = -0.5 * selfterm * COULOMBS;
        const V3 r = ldv(a.pos + 3 * i), rim = ldv(a.com + 3 * a.mol[i]);
        v[1] = I.q * r.x; v[2] = I.q * r.y; v[3] = I.q * r.z;              // q r
        v[4] = I.q * rim.x; v[5] = I.q * rim.y; v[6] = I.q * rim.z;        // q R_mol
        v[7] = I.q * r.x + I.d.x; v[8] = I.q * r.y + I.d.y; v[9] = I.q * r.z + I.d.z;   // q r + mu
        v[10] = 2.0 * I.q * QiI - dot(I.d, I.d);
    }
    // the surface sums feed the forces: block partials added in block order
    // by site_sums_finish before the next kernel
    block_store<11>(v, a.spart, gridDim.x);
}

// One thread per molecule: excluded pairs inside it with erf-screened D_l
// (intraMolCorrect :1350-1600).  Forces and torques only without constraints.
__global__ __launch_bounds__(256) void k_mp_intra(MpArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (m < a.nmol) {
        const int a0 = a.mol_start[m], ns = a.mol_start[m + 1] - a0;
        V3 tq[3] = {v3(0, 0, 0), v3(0, 0, 0), v3(0, 0, 0)};
        for (int x = 0; x < ns; ++x) {
            const int idi = a0 + x;
            const Site I = load_site(a, idi);
            for (int y = x + 1; y < ns; ++y) {
                const int idj = a0 + y;
                bool ex = false;
                for (int k = a.excl_off[idi]; k < a.excl_off[idi + 1]; ++k) ex = ex || a.excl[k] == idj;
                if (!ex) continue;
                double dx = a.pos[3 * idi] - a.pos[3 * idj], dy = a.pos[3 * idi + 1] - a.pos[3 * idj + 1],
                       dz = a.pos[3 * idi + 2] - a.pos[3 * idj + 2];
                pbc(dx, dy, dz, a.box);
                const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
                const V3 rv = v3(dx, dy, dz);
                const Site J = load_site(a, idj);
                const double e = exp(-a.alpha2 * r2), inv = 1.0 / r2;
                double D[6];
                D[0] = erf(a.alpha * r) / r;
                D[1] = (-D[0] + a.alpha2PI * e) * inv;
                D[2] = (-3.0 * D[1] - 2.0 * a.alpha2PI * a.alpha2 * e) * inv;
                D[3] = (-5.0 * D[2] + 4.0 * a.alpha2PI * a.alpha2 * a.alpha2 * e) * inv;
                D[4] = (-7.0 * D[3] - 8.0 * a.alpha2PI * a.alpha2 * a.alpha2 * a.alpha2 * e) * inv;
                D[5] = (-9.0 * D[4] + 16.0 * a.alpha2PI * a.alpha2 * a.alpha2 * a.alpha2 * a.alpha2 * e) * inv;
                PairOut o;
                if (a.use_constraint) {
                    mp_pair<false, false>(I, J, rv, o);
                } else {
                    if (a.torques) mp_pair<true, true>(I, J, rv, o);
                    else mp_pair<true, false>(I, J, rv, o);
                }
                v[0] += o.G[0] * D[0] + o.G[1] * D[1] + o.G[2] * D[2] + o.G[3] * D[3] + o.G[4] * D[4];
                if (!a.use_constraint) {
                    const double radial = o.G[0] * D[1] + o.G[1] * D[2] + o.G[2] * D[3] + o.G[3] * D[4] + o.G[4] * D[5];
                    const V3 fij = D[1] * o.gG[1] + D[2] * o.gG[2] + D[3] * o.gG[3] + D[4] * o.gG[4] + radial * rv;
                    v[1] += -dot(fij, rv) * COULOMBS / (3.0 * a.volume);
                    a.force[3 * idi] += COULOMBS * fij.x;
                    a.force[3 * idi + 1] += COULOMBS * fij.y;
                    a.force[3 * idi + 2] += COULOMBS * fij.z;
                    a.force[3 * idj] -= COULOMBS * fij.x;
                    a.force[3 * idj + 1] -= COULOMBS * fij.y;
                    a.force[3 * idj + 2] -= COULOMBS * fij.z;
                    if (a.torques) {
                        tq[x] = tq[x] - COULOMBS * (D[1] * o.Gt[1] + D[2] * o.Gt[2] + D[3] * o.Gt[3] + D[4] * o.Gt[4]);
                        PairOut oj;
                        mp_pair<false, true>(J, I, -rv, oj);
                        tq[y] = tq[y] -
                                COULOMBS * (D[1] * oj.Gt[1] + D[2] * oj.Gt[2] + D[3] * oj.Gt[3] + D[4] * oj.Gt[4]);
                    }
                }
            }
        }
        double* tr = a.trq + 3 * (size_t)a.n;   // intramolecular part
        for (int x = 0; x < ns; ++x) {
            tr[3 * (a0 + x)] = tq[x].x;
            tr[3 * (a0 + x) + 1] = tq[x].y;
            tr[3 * (a0 + x) + 2] = tq[x].z;
        }
    }
    v[0] *= -COULOMBS;
    block_part<2>(v, a.red, RS_MP_INTRA);
}

// Surface force, virial and torque per site (:1846-1900), from the sums.
__global__ __launch_bounds__(256) void k_mp_surface(MpArgs a, double coeff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < a.n) {
        const V3 qr = ldv(a.terms + M_SU_QR), qm = ldv(a.terms + M_SU_QM);
        const Site I = load_site(a, i);
        const V3 ri = ldv(a.pos + 3 * i);
        const V3 fi = (-2.0 * coeff * I.q) * qm;
        const V3 temp = (-a.q[i]) * qr;
        a.force[3 * i] += COULOMBS * fi.x;
        a.force[3 * i + 1] += COULOMBS * fi.y;
        a.force[3 * i + 2] += COULOMBS * fi.z;
        v[0] = temp.x * ri.x; v[1] = temp.y * ri.x; v[2] = temp.z * ri.x;
        v[3] = temp.x * ri.y; v[4] = temp.y * ri.y; v[5] = temp.z * ri.y;
        v[6] = temp.x * ri.z; v[7] = temp.y * ri.z; v[8] = temp.z * ri.z;
        const V3 tq = (2.0 * coeff * COULOMBS) * cross(qm, I.d);
        double* tr = a.trq + 6 * (size_t)a.n;   // surface part
        tr[3 * i] = a.torques ? tq.x : 0.0;
        tr[3 * i + 1] = a.torques ? tq.y : 0.0;
        tr[3 * i + 2] = a.torques ? tq.z : 0.0;
    }
    block_part<9>(v, a.red, RS_MP_SURF);
}

// ---- torques to forces ----------------------------------------------------------

// Frame of a hydrogen site: axis from O, second vector x (y when the axis is
// within ~30 deg of x).  dU/du of the torque (:1988-2022).
__device__ __forceinline__ V3 h_site(V3 rh, V3 ro, V3 tau, double& dudv) {
    V3 u = rh - ro;
    const double ul = sqrt(dot(u, u));
    V3 v = fabs(u.x / ul) > 0.8666 ? v3(0.0, 1.0, 0.0) : v3(1.0, 0.0, 0.0);
    V3 w = cross(u, v);
    const double wl = sqrt(dot(w, w));
    u = (1.0 / ul) * u;
    w = (1.0 / wl) * w;
    const V3 vxu = cross(v, u), wxu = cross(w, u);
    const double vxuL2 = dot(vxu, vxu), wxuL2 = dot(wxu, wxu);
    const V3 dUdu = (dot(-v, tau) / (vxuL2 * ul)) * vxu + (dot(-w, tau) / (wxuL2 * ul)) * wxu;
    dudv += dot(dUdu, u) * ul;
    return dUdu;
}

// One thread per water; the torque-array parts set in `parts` (bit p for
// part p), each part's dU/dV summed into its own slot.
__global__ __launch_bounds__(256) void k_mp_torques(MpArgs a, unsigned parts, int site) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double dv[5] = {0, 0, 0, 0, 0}, tn[5] = {0, 0, 0, 0, 0};
    if (m < a.nmol) {
        const int h1 = a.mol_start[m], o = h1 + 1, h2 = h1 + 2;
        const V3 r1 = ldv(a.pos + 3 * h1), ro = ldv(a.pos + 3 * o), r2 = ldv(a.pos + 3 * h2);
        V3 F1 = v3(0, 0, 0), FO = v3(0, 0, 0), F2 = v3(0, 0, 0);
        for (int p = 0; p < 5; ++p) {
            if (!(parts >> p & 1u)) continue;
            const double* tr = a.trq + 3 * (size_t)a.n * p;
            const V3 t1 = ldv(tr + 3 * h1), to = ldv(tr + 3 * o), t2 = ldv(tr + 3 * h2);
            tn[p] = sqrt(dot(t1, t1)) + sqrt(dot(to, to)) + sqrt(dot(t2, t2));
            // H1
            V3 d = h_site(r1, ro, t1, dv[p]);
            F1 = F1 - d;
            FO = FO + d;
            // O: bisector-type frame from the two O-H vectors (:2024-2072)
            {
                V3 u = ro - r1, v = ro - r2;
                V3 w = cross(u, v);
                const double ul = sqrt(dot(u, u)), vl = sqrt(dot(v, v)), wl = sqrt(dot(w, w));
                u = (1.0 / ul) * u;
                v = (1.0 / vl) * v;
                w = (1.0 / wl) * w;
                const V3 vxu = cross(v, u), wxu = cross(w, u), wxv = cross(w, v);
                const double vxuL2 = dot(vxu, vxu), wxuL2 = dot(wxu, wxu), wxvL2 = dot(wxv, wxv);
                const V3 dUdu = (dot(-v, to) / (vxuL2 * ul)) * vxu + (0.5 * dot(-w, to) / (wxuL2 * ul)) * wxu;
                const V3 dUdv = (dot(-u, to) / (vxuL2 * vl)) * (-vxu) + (0.5 * dot(-w, to) / (wxvL2 * vl)) * wxv;
                FO = FO - dUdu - dUdv;
                F1 = F1 + dUdu;
                F2 = F2 + dUdv;
                dv[p] += dot(dUdu, u) * ul + dot(dUdv, v) * vl;
            }
            // H2
            d = h_site(r2, ro, t2, dv[p]);
            F2 = F2 - d;
            FO = FO + d;
        }
        a.force[3 * h1] += F1.x; a.force[3 * h1 + 1] += F1.y; a.force[3 * h1 + 2] += F1.z;
        a.force[3 * o] += FO.x; a.force[3 * o + 1] += FO.y; a.force[3 * o + 2] += FO.z;
        a.force[3 * h2] += F2.x; a.force[3 * h2 + 1] += F2.y; a.force[3 * h2 + 2] += F2.z;
    }
    block_part<5>(dv, a.red, site);
    __syncthreads();
    block_part<5>(tn, a.red, site + 1);
}

// ---- k-space ---------------------------------------------------------------------

// Pass 1: thread per k over one chunk of sites; 20 partial sums per (chunk, k):
//   0 Sc = sum (q - kQk) cos      1 Ss = sum (q - kQk) sin
//   2 Mc = sum (mu.k) cos         3 Ms = sum (mu.k) sin
//   4 Vc = sum 2 kQk cos          5 Vs = sum 2 kQk sin
//   6..8 sum (q - kQk) cos d     9..11 sum (q - kQk) sin d      (d = r - R_mol)
//  12..14 sum (mu.k) cos d      15..17 sum (mu.k) sin d
//  18 sum (q - kQk) cos |d|^2   19 sum (q - kQk) sin |d|^2
template <typename R>
__global__ __launch_bounds__(MP_SF_THREADS) void k_mp_sfac(MpArgs a) {
    // per atom of the tile: r in FP64 (the phase); q, mu, Q, d = r - R_mol in R
    __shared__ double shr[MP_SF_THREADS][3];
    __shared__ R shm[MP_SF_THREADS][13];
    const int k = blockIdx.x * MP_SF_THREADS + threadIdx.x;
    const int chunk = blockIdx.y;
    const int i0 = chunk * a.per, i1 = min(a.n, i0 + a.per);
    V3 kv = v3(0, 0, 0);
    if (k < a.K) kv = ldv(a.kvec + 3 * k);
    const R kx = (R)kv.x, ky = (R)kv.y, kz = (R)kv.z;
    double S[MP_NSUM];
#pragma unroll
    for (int x = 0; x < MP_NSUM; ++x) S[x] = 0.0;
    for (int t0 = i0; t0 < i1; t0 += MP_SF_THREADS) {
        const int i = t0 + threadIdx.x;
        __syncthreads();
        if (i < i1) {
            shr[threadIdx.x][0] = a.pos[3 * i]; shr[threadIdx.x][1] = a.pos[3 * i + 1]; shr[threadIdx.x][2] = a.pos[3 * i + 2];
            R* p = shm[threadIdx.x];
            p[0] = (R)(a.q[i] / SQRT_COULOMB_MP);
            p[1] = (R)a.dip[3 * i]; p[2] = (R)a.dip[3 * i + 1]; p[3] = (R)a.dip[3 * i + 2];
            for (int c = 0; c < 6; ++c) p[4 + c] = (R)a.quad[6 * i + c];
            const int m = a.mol[i];
            p[10] = (R)(a.pos[3 * i] - a.com[3 * m]);
            p[11] = (R)(a.pos[3 * i + 1] - a.com[3 * m + 1]);
            p[12] = (R)(a.pos[3 * i + 2] - a.com[3 * m + 2]);
        }
        __syncthreads();
        const int cnt = min(MP_SF_THREADS, i1 - t0);
        if (k < a.K) {
            // FP64: the sums accumulate directly; FP32: the tile's sums, added
            // to the FP64 totals after the tile
            auto tile = [&](auto& s) {
            for (int x = 0; x < cnt; ++x) {
                const R* p = shm[x];
                R sn, cs;
                phase_sincos<R>(shr[x][0] * kv.x + shr[x][1] * kv.y + shr[x][2] * kv.z, sn, cs);
                const R QK = p[4] * kx * kx + p[7] * ky * ky + p[9] * kz * kz +
                             R(2) * (p[5] * kx * ky + p[6] * kx * kz + p[8] * ky * kz);
                const R A = p[0] - QK;
                const R Mu = p[1] * kx + p[2] * ky + p[3] * kz;
                const R ac = A * cs, as = A * sn, mc = Mu * cs, ms = Mu * sn;
                const R dx = p[10], dy = p[11], dz = p[12];
                const R d2 = dx * dx + dy * dy + dz * dz;
                s[0] += ac; s[1] += as; s[2] += mc; s[3] += ms;
                s[4] += R(2) * QK * cs; s[5] += R(2) * QK * sn;
                s[6] += ac * dx; s[7] += ac * dy; s[8] += ac * dz;
                s[9] += as * dx; s[10] += as * dy; s[11] += as * dz;
                s[12] += mc * dx; s[13] += mc * dy; s[14] += mc * dz;
                s[15] += ms * dx; s[16] += ms * dy; s[17] += ms * dz;
                s[18] += ac * d2; s[19] += as * d2;
            }
            };
            if constexpr (sizeof(R) == sizeof(double)) {
                tile(S);
            } else {
                R s[MP_NSUM];
#pragma unroll
                for (int x = 0; x < MP_NSUM; ++x) s[x] = R(0);
                tile(s);
#pragma unroll
                for (int x = 0; x < MP_NSUM; ++x) S[x] += (double)s[x];
            }
        }
    }
    if (k < a.K) {
        double* out = a.part + ((size_t)chunk * a.K + k) * MP_NSUM;
#pragma unroll
        for (int x = 0; x < MP_NSUM; ++x) out[x] = S[x];
    }
}

// Pass 2: per k, the chunk sums in fixed order, then energy, virial, volume
// derivatives (:1098-1138 and the force-pass sums :1180-1188 from the
// moments) and the force-pass coefficients (b P, b Q).
// Chunk sums of one k by one wave: lane l adds chunks l, l + 64, ... in
// order, then a butterfly over the wave (every lane ends with the same total).
template <int NS>
__device__ __forceinline__ void wave_chunk_sum(const MpArgs& a, int k, double (&S)[NS]) {
    const int lane = threadIdx.x % WAVE;
#pragma unroll
    for (int x = 0; x < NS; ++x) S[x] = 0.0;
    for (int c = lane; c < a.chunks; c += WAVE) {
        const double* p = a.part + ((size_t)c * a.K + k) * NS;
#pragma unroll
        for (int x = 0; x < NS; ++x) S[x] += p[x];
    }
#pragma unroll
    for (int x = 0; x < NS; ++x) S[x] = wave_sum(S[x]);
}

__device__ __forceinline__ void mp_finish(const MpArgs& a, int k, const double (&S)[MP_NSUM], double (&v)[9]) {
    {
        const V3 kv = ldv(a.kvec + 3 * k);
        const double pi4V = 4.0 * PI / a.volume, pi8V = 2.0 * pi4V;
        const double ek = a.expk[k];
        const double Sc = S[0], Ss = S[1], Mc = S[2], Ms = S[3], Vc = S[4], Vs = S[5];
        const double UqQ = pi4V * ek * (Sc * Sc + Ss * Ss);
        const double UmqQ = pi4V * ek * (Ss * Mc - Sc * Ms);
        const double Umm = pi4V * ek * (Mc * Mc + Ms * Ms);
        const double t = (UqQ + 2.0 * UmqQ + Umm) * COULOMBS;
        const double dUdVqQ = pi8V * ek * (Vc * Sc + Vs * Ss);
        const double dUdVqQm = pi8V * ek * (Vs * Mc - Vc * Ms);
        const double d3 = (dUdVqQ - 2.0 * UmqQ + dUdVqQm - 2.0 * Umm) * COULOMBS;
        const double k2 = dot(kv, kv);
        const double x = k2 * 2.0 * a.alphaR4;
        double d1 = -t + t * x + d3;
        double d2 = 4.0 * t - 7.0 * t * x + t * x * x;
        const double coeff = 2.0 * (a.alphaR4 + 1.0 / a.kmod2[k]);
        const double b = pi8V * ek;
        const double P = Sc - Ms, Q = Ss + Mc;
        if (a.use_constraint) {
            const V3 cosrR = v3(S[6], S[7], S[8]), sinrR = v3(S[9], S[10], S[11]);
            const V3 mcR = v3(S[12], S[13], S[14]), msR = v3(S[15], S[16], S[17]);
            const double Ff = COULOMBS * b * (P * dot(kv, sinrR + mcR) + Q * dot(kv, msR - cosrR));
            const double Ffk2 = 2.0 * k2 * 2.0 * a.alphaR4 * Ff;
            const double d2UV1 = b * k2 * (Sc * S[18] + Ss * S[19]);
            const double d2UrR = b * k2 * (dot(cosrR, cosrR) + dot(sinrR, sinrR));
            d1 += Ff;
            d2 += -6.0 * Ff + Ffk2 - d2UV1 + d2UrR;
        }
        v[0] = t; v[1] = d1; v[2] = d2;
        v[3] = t - t * coeff * kv.x * kv.x;
        v[4] = t - t * coeff * kv.y * kv.y;
        v[5] = t - t * coeff * kv.z * kv.z;
        v[6] = -t * coeff * kv.x * kv.y;
        v[7] = -t * coeff * kv.x * kv.z;
        v[8] = -t * coeff * kv.y * kv.z;
        a.kf[4 * k] = b * P;
        a.kf[4 * k + 1] = b * Q;
    }
}

// One wave per k (four per block).
__global__ __launch_bounds__(256) void k_mp_kfinish(MpArgs a) {
    const int k = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (k < a.K) {
        double S[MP_NSUM];
        wave_chunk_sum<MP_NSUM>(a, k, S);
        if (threadIdx.x % WAVE == 0) mp_finish(a, k, S, v);
    }
    block_part<9>(v, a.red, RS_MP_LONG);
}

// Pass 3: force val k and the torque of the dipole and the quadrupole
// (:1149-1210).  Block (64 sites, one k-range); each lane sums its site over
// the range and writes a partial; k_kforce_sum adds the ranges in order.
template <typename R>
__global__ __launch_bounds__(WAVE) void k_mp_kforce(MpArgs a) {
    __shared__ double sk[MP_KTILE][5];
    const int i = blockIdx.x * WAVE + threadIdx.x;
    const int kb = blockIdx.y * a.kper, ke = min(a.K, kb + a.kper);
    Site I64{};
    V3 ri = v3(0, 0, 0);
    if (i < a.n) {
        I64 = load_site(a, i);
        ri = ldv(a.pos + 3 * i);
    }
    const SiteT<R> I = site_as<R>(I64);
    V3 F = v3(0, 0, 0), T = v3(0, 0, 0);
    for (int k0 = kb; k0 < ke; k0 += MP_KTILE) {
        const int cnt = min(MP_KTILE, ke - k0);
        __syncthreads();
        for (int x = threadIdx.x; x < cnt; x += WAVE) {
            sk[x][0] = a.kvec[3 * (k0 + x)];
            sk[x][1] = a.kvec[3 * (k0 + x) + 1];
            sk[x][2] = a.kvec[3 * (k0 + x) + 2];
            sk[x][3] = a.kf[4 * (k0 + x)];
            sk[x][4] = a.kf[4 * (k0 + x) + 1];
        }
        __syncthreads();
        if (i < a.n) {
            // the tile's force and torque in R, added to FP64 after the tile
            V3T<R> f = vec3<R>(R(0), R(0), R(0)), t = f;
            for (int x = 0; x < cnt; ++x) {
                R sn, cs;
                phase_sincos<R>(ri.x * sk[x][0] + ri.y * sk[x][1] + ri.z * sk[x][2], sn, cs);
                const V3T<R> kv = vec3<R>((R)sk[x][0], (R)sk[x][1], (R)sk[x][2]);
                const R bP = (R)sk[x][3], bQ = (R)sk[x][4];
                const V3T<R> Qk = smul(I.Q, kv);
                const R A = I.q - dot(kv, Qk);
                const R Mu = dot(I.d, kv);
                const R val = sn * (A * bP + Mu * bQ) + cs * (Mu * bP - A * bQ);
                f = f + val * kv;
                if (a.torques) {
                    // (kk^T x Qi) = k x (Qi k); tau = (c bQ - s bP) k x mu - 2 (c bP + s bQ) k x Qk
                    t = t + (cs * bQ - sn * bP) * cross(kv, I.d) - (R(2) * (cs * bP + sn * bQ)) * cross(kv, Qk);
                }
            }
            F = F + to64(f);
            T = T + to64(t);
        }
    }
    if (i < a.n) {
        double* p = a.fpart + ((size_t)blockIdx.y * a.n + i) * 6;
        p[0] = F.x; p[1] = F.y; p[2] = F.z;
        p[3] = T.x; p[4] = T.y; p[5] = T.z;
    }
}

// The k-range partials of each site in order: force added, k-space torque
// stored (torque-array part 3).
__global__ __launch_bounds__(256) void k_kforce_sum(MpArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double F[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < a.ksplit; ++s) {
        const double* p = a.fpart + ((size_t)s * a.n + i) * 6;
#pragma unroll
        for (int x = 0; x < 6; ++x) F[x] += p[x];
    }
    a.force[3 * i] += COULOMBS * F[0];
    a.force[3 * i + 1] += COULOMBS * F[1];
    a.force[3 * i + 2] += COULOMBS * F[2];
    double* tr = a.trq + 9 * (size_t)a.n;
    tr[3 * i] = COULOMBS * F[3];
    tr[3 * i + 1] = COULOMBS * F[4];
    tr[3 * i + 2] = COULOMBS * F[5];
}

// ---- host side ---------------------------------------------------------------------

int frames_run(Ctx* c);

static_assert(M_COUNT <= POLAR_TERMS, "polar accumulator too small");

__global__ __launch_bounds__(256) void k_zero_terms(double* t) {
    for (int k = threadIdx.x; k < M_COUNT; k += blockDim.x) t[k] = 0.0;
}

static MpArgs mp_args(Ctx* c, const Ctx::KSet& ks, double alpha, double real_cut, int use_constraint,
                      int torques) {
    MpArgs a{};
    a.n = c->n;
    a.nmol = c->nmol;
    a.stride = c->stride;
    a.K = ks.K;
    a.chunks = ks.chunks;
    a.per = ks.K > 0 ? (c->n + ks.chunks - 1) / ks.chunks : 0;
    a.box = c->box;
    a.mol_start = c->d_mol_start;
    a.mol = c->d_mol;
    a.pos = c->d_pos;
    a.com = c->d_com;
    a.q = c->d_q;
    a.dip = c->d_dipg;
    a.quad = c->d_quadg;
    a.nb = c->d_nb;
    a.nb_cnt = c->d_nb_cnt;
    a.nb_own = c->d_nb_own;
    a.excl_off = c->d_excl_off;
    a.excl = c->d_excl;
    a.force = c->d_force;
    a.trq = c->d_mp_trq;
    a.terms = c->d_mp_terms;
    a.red = c->red;
    a.kvec = ks.kvec;
    a.kmod2 = ks.kmod2;
    a.expk = ks.expk;
    a.part = ks.part;
    a.kf = ks.kf;
    a.fpart = ks.fpart;
    a.kstep = ks.kstep;
    a.kz_step = 2.0 * PI / c->box.L[2];
    a.ksplit = ks.ksplit;
    a.kper = ks.ksplit > 0 ? (ks.K + ks.ksplit - 1) / ks.ksplit : 0;
    a.alpha = alpha;
    a.alpha2 = alpha * alpha;
    a.alpha2PI = 2.0 * alpha / SQRT_PI;
    a.realcut2 = real_cut * real_cut;
    a.volume = c->box.L[0] * c->box.L[1] * c->box.L[2];
    a.alphaR4 = 1.0 / (4.0 * a.alpha2);
    a.use_constraint = use_constraint;
    a.torques = torques;
    a.pind = c->d_mu;
    a.polar = c->d_polar;
    a.spart = c->part_su;
    a.site32 = c->d_site32;
    a.pind32 = c->d_pind32;
    return a;
}

static MpArgs mp_args(Ctx* c) {
    return mp_args(c, c->mp_ks, c->mp.alpha, c->mp.real_cut, c->mp.use_constraint, c->mp.torques);
}

static int mp_check(Ctx* c) {
    if (!c->mp_ready) return fail(c, SMD_ESTATE, "smd_multipole_setup not called");
    if (!c->multipoles_set) return fail(c, SMD_ESTATE, "multipoles not set");
    return SMD_OK;
}

static int mp_read_terms(Ctx* c, double* t) {
    SMD_HIP(c, hipMemcpyAsync(t, c->d_mp_terms, sizeof(double) * M_COUNT, hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return SMD_OK;
}

int multipole_short(Ctx* c, smd_mp_out* out) {
    int rc = mp_check(c);
    if (rc) return rc;
    if (!c->list_valid) return fail(c, SMD_ESTATE, "no pair list");
    if (c->mp.real_cut > c->list_cutoff + c->list_buff)
        return fail(c, SMD_ENOTSUP, "real-space cutoff %g exceeds the pair-list radius %g", c->mp.real_cut,
                    c->list_cutoff + c->list_buff);
    MpArgs a = mp_args(c);
    const dim3 b256(256);
    const dim3 ga((c->n + 255) / 256), gm((c->nmol + 255) / 256);
    const dim3 gr((unsigned)(((int64_t)c->n * MP_LANES + 255) / 256));
    ProfScope ps(c, "mp_short");
    SMD_HIP(c, hipMemsetAsync(c->d_mp_terms, 0, sizeof(double) * M_COUNT, c->stream));
    if ((rc = frames_run(c))) return rc;
    hipLaunchKernelGGL(k_mp_site_sums, ga, b256, 0, c->stream, a);
    if ((rc = site_sums_finish(c, a, ga.x, false))) return rc;
    if ((rc = launch_real(c, a, false))) return rc;
    hipLaunchKernelGGL(k_mp_intra, gm, b256, 0, c->stream, a);
    if ((rc = red_finish(c, RS_MP_INTRA, gm.x, 2, a.terms + M_INTRA_E))) return rc;
    const double volume = a.volume, volumer = 1.0 / volume;
    const double coeff = 2.0 * PI * volumer / 3.0;
    if (c->mp.surface) {
        hipLaunchKernelGGL(k_mp_surface, ga, b256, 0, c->stream, a, coeff);
        if ((rc = red_finish(c, RS_MP_SURF, ga.x, 9, a.terms + M_SU_VIR))) return rc;
    }
    if (c->mp.torques && (rc = launch_torques(c, a, gm, c->mp.surface ? 7u : 3u, RS_MP_DUDV))) return rc;
    SMD_HIP(c, hipGetLastError());
    double t[M_COUNT];
    if ((rc = mp_read_terms(c, t))) return rc;
    *out = smd_mp_out{};
    double qsum = 0.0;
    for (int i = 0; i < c->n; ++i) qsum += c->h_q[i];
    if (fabs(qsum) > 1.0e-5) out->charged_energy = -PI * qsum * qsum / (2.0 * volume * a.alpha2);
    out->self_energy = t[M_SELF];
    out->real_energy = t[M_REAL_E];
    out->real_deriv = t[M_REAL_D1] + t[M_DUDV];
    out->intra_energy = t[M_INTRA_E];
    out->mol_deriv = t[M_MOL_D1] + t[M_DUDV + 1] / (3.0 * volume);
    for (int k = 0; k < 9; ++k) out->virial[k] = t[M_REAL_VIR + k];
    out->tot_torque = t[M_TRQ] + t[M_TRQ + 1] + t[M_TRQ + 2];
    if (c->mp.surface) {
        const double* qr = t + M_SU_QR;
        const double* qR = t + M_SU_QRR;
        const double* qm = t + M_SU_QM;
        const double selfterm = qm[0] * qm[0] + qm[1] * qm[1] + qm[2] * qm[2] + t[M_SU_SELF];
        out->surf_energy = coeff * selfterm * COULOMBS;
        const double qmqr = qm[0] * qr[0] + qm[1] * qr[1] + qm[2] * qr[2];
        const double qrqr = qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2];
        const double qrqR = qr[0] * qR[0] + qr[1] * qR[1] + qr[2] * qR[2];
        const double qRqR = qR[0] * qR[0] + qR[1] * qR[1] + qR[2] * qR[2];
        out->surf_deriv = -out->surf_energy * volumer + 2.0 * coeff * volumer * qmqr * COULOMBS / 3.0;
        if (c->mp.use_constraint)
            out->surf_deriv2 = 2.0 * coeff * volumer * volumer * qrqr - 16.0 * coeff * volumer * volumer * qrqR / 9.0 +
                               2.0 * coeff * volumer * volumer * qRqR / 9.0;
        else
            out->surf_deriv2 = 4.0 * coeff * volumer * volumer * qrqr / 9.0;
        out->surf_deriv += t[M_DUDV + 2] * volumer / 3.0;
        for (int k = 0; k < 9; ++k) out->virial[k] += t[M_SU_VIR + k];
    }
    return SMD_OK;
}

int multipole_long(Ctx* c, smd_mp_long_out* out) {
    int rc = mp_check(c);
    if (rc) return rc;
    MpArgs a = mp_args(c);
    const dim3 b256(256);
    ProfScope ps(c, "mp_long");
    SMD_HIP(c, hipMemsetAsync(c->d_mp_terms, 0, sizeof(double) * M_COUNT, c->stream));
    if ((rc = frames_run(c))) return rc;
    const dim3 g1((a.K + MP_SF_THREADS - 1) / MP_SF_THREADS, a.chunks);
    if (c->mp_prec == SMD_PREC_MIXED) hipLaunchKernelGGL((k_mp_sfac<float>), g1, dim3(MP_SF_THREADS), 0, c->stream, a);
        else hipLaunchKernelGGL((k_mp_sfac<double>), g1, dim3(MP_SF_THREADS), 0, c->stream, a);
    hipLaunchKernelGGL(k_mp_kfinish, dim3((a.K + 3) / 4), b256, 0, c->stream, a);
    if ((rc = red_finish(c, RS_MP_LONG, (a.K + 3) / 4, 9, a.terms + M_LONG_E))) return rc;
    if (c->mp_prec == SMD_PREC_MIXED) hipLaunchKernelGGL((k_mp_kforce<float>), dim3((c->n + WAVE - 1) / WAVE, a.ksplit), dim3(WAVE), 0, c->stream, a);
        else hipLaunchKernelGGL((k_mp_kforce<double>), dim3((c->n + WAVE - 1) / WAVE, a.ksplit), dim3(WAVE), 0, c->stream, a);
    hipLaunchKernelGGL(k_kforce_sum, dim3((c->n + 255) / 256), b256, 0, c->stream, a);
    if (c->mp.torques && (rc = launch_torques(c, a, dim3((c->nmol + 255) / 256), 8u, RS_MP_DUDV_L))) return rc;
    SMD_HIP(c, hipGetLastError());
    double t[M_COUNT];
    if ((rc = mp_read_terms(c, t))) return rc;
    *out = smd_mp_long_out{};
    out->energy = t[M_LONG_E];
    out->deriv = t[M_LONG_D1] + t[M_DUDV + 3];
    out->deriv2 = t[M_LONG_D2];
    const double* lv = t + M_LONG_VIR;   // xx yy zz xy xz yz
    out->virial[0] = lv[0]; out->virial[4] = lv[1]; out->virial[8] = lv[2];
    out->virial[1] = out->virial[3] = lv[3];
    out->virial[2] = out->virial[6] = lv[4];
    out->virial[5] = out->virial[7] = lv[5];
    out->tot_torque = t[M_TRQ + 3];
    return SMD_OK;
}

// Half-space k set, k^2 < kCut^2, and expK (Multipole::set_up :242-302;
// Induction::set_up builds the same set, src/forces/Induction.cpp:592-652),
// with structure-factor scratch for `nsum` sums per k.
static int kset_build(Ctx* c, Ctx::KSet& ks, double alpha, double kcut, int nsum) {
    if (!c->box_set) return fail(c, SMD_ESTATE, "box not set");
    for (int m = 0; m < c->nmol; ++m)
        if (c->h_mol_start[m + 1] - c->h_mol_start[m] != 3)
            return fail(c, SMD_ENOTSUP, "multipole and induction Ewald are implemented for 3-site water (H1, O, H2)");
    std::vector<double> kv, k2v, ekv;
    std::vector<int> kst;
    int pnx = 0, pny = 0, pnz = 0;
    const double kCut2 = kcut * kcut, alphaR4 = 1.0 / (4.0 * (alpha * alpha));
    const int nxMax = (int)(kcut * c->box.L[0] / (2 * PI));
    const int nyMax = (int)(kcut * c->box.L[1] / (2 * PI));
    const int nzMax = (int)(kcut * c->box.L[2] / (2 * PI));
    const double vx = 2.0 * PI / c->box.L[0], vy = 2.0 * PI / c->box.L[1], vz = 2.0 * PI / c->box.L[2];
    for (int nx = 0; nx <= nxMax; nx++)
        for (int ny = (nx == 0 ? 0 : -nyMax); ny <= nyMax; ny++)
            for (int nz = ((nx == 0 && ny == 0) ? 1 : -nzMax); nz <= nzMax; nz++) {
                const double kx = nx * vx, ky = ny * vy, kz = nz * vz;
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 < kCut2) {
                    kst.push_back(!kst.empty() && nx == pnx && ny == pny && nz == pnz + 1 ? 1 : 0);
                    pnx = nx; pny = ny; pnz = nz;
                    kv.push_back(kx); kv.push_back(ky); kv.push_back(kz);
                    k2v.push_back(k2);
                    ekv.push_back(exp(-k2 * alphaR4) / k2);
                }
            }
    const int K = (int)k2v.size();
    if (K == 0) return fail(c, SMD_EINVAL, "no k-vectors below kcut");
    // chunks of sites so that the structure-factor grid covers the chip
    const int kblocks = (K + MP_SF_THREADS - 1) / MP_SF_THREADS;
    int chunks = std::max(1, std::min((512 + kblocks - 1) / kblocks, (c->n + 31) / 32));
    auto R = [&](double*& ptr, size_t bytes) -> int {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        if (hipMalloc(&ptr, bytes) != hipSuccess) return fail(c, SMD_ENOMEM, "Ewald k-space allocation of %zu bytes", bytes);
        return SMD_OK;
    };
    int rc;
    // k-ranges of the force pass: enough (site tile, k-range) blocks to fill the chip
    const int tiles = (c->n + WAVE - 1) / WAVE;
    const int ksplit = std::max(1, std::min(std::min(64, K), (2048 + tiles - 1) / tiles));
    if ((rc = R(ks.kvec, sizeof(double) * 3 * K)) || (rc = R(ks.kmod2, sizeof(double) * K)) ||
        (rc = R(ks.expk, sizeof(double) * K)) || (rc = R(ks.kf, sizeof(double) * 4 * K)) ||
        (rc = R(ks.part, sizeof(double) * nsum * (size_t)K * chunks)) ||
        (rc = R(ks.fpart, sizeof(double) * 6 * (size_t)ksplit * c->n)))
        return rc;
    if (!c->d_mp_trq && (rc = R(c->d_mp_trq, sizeof(double) * 15 * (size_t)c->n))) return rc;
    if (!c->d_mp_terms && (rc = R(c->d_mp_terms, sizeof(double) * M_COUNT))) return rc;
    SMD_HIP(c, hipMemset(c->d_mp_trq, 0, sizeof(double) * 15 * (size_t)c->n));
    SMD_HIP(c, hipMemcpy(ks.kvec, kv.data(), sizeof(double) * 3 * K, hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(ks.kmod2, k2v.data(), sizeof(double) * K, hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(ks.expk, ekv.data(), sizeof(double) * K, hipMemcpyHostToDevice));
    if (ks.kstep) (void)hipFree(ks.kstep);
    ks.kstep = nullptr;
    if (hipMalloc(&ks.kstep, sizeof(int) * K) != hipSuccess) return fail(c, SMD_ENOMEM, "k-step flags of %d k", K);
    SMD_HIP(c, hipMemcpy(ks.kstep, kst.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    ks.K = K;
    ks.chunks = chunks;
    ks.ksplit = ksplit;
    ks.alpha = alpha;
    ks.kcut = kcut;
    return SMD_OK;
}

static int mp_setup(Ctx* c, const smd_mp_params* p, double kcut, int32_t* n_kvec) {
    if (!(p->alpha > 0) || !(p->real_cut > 0) || !(kcut > 0)) return fail(c, SMD_EINVAL, "bad multipole parameters");
    int rc = kset_build(c, c->mp_ks, p->alpha, kcut, MP_NSUM);
    if (rc) return rc;
    c->mp = *p;
    c->mp_ready = true;
    if (n_kvec) *n_kvec = c->mp_ks.K;
    return SMD_OK;
}


// =====================================================================================
// Induced dipoles against the permanent multipoles (Induction::compute after
// the dipole solve, and compute_long; src/forces/Induction.cpp).
// =====================================================================================

constexpr int IN_NSUM = 28;   // structure-factor sums per k of the induction pass

// Site I's permanent multipoles with partner J's induced dipole pj at
// r = r_I - r_J (the "i" side of real_term :940-1028 and the torque vectors
// :1276-1281).  The reference's "j" side is this with the roles swapped.
template <typename R> struct IndSideT { R G1, G2, G3; V3T<R> g1, g2, g3; V3T<R> Gt1, Gt2, Gt3; };
using IndSide = IndSideT<double>;

template <bool GRAD, bool TORQUE, typename R>
__device__ __forceinline__ void ind_side(const SiteT<R>& I, V3T<R> pj, V3T<R> r, IndSideT<R>& t) {
    const R QiI = trace(I.Q);
    const V3T<R> Qir = smul(I.Q, r);
    const R QiR = dot(r, Qir);
    const R pjr = dot(pj, r), mir = dot(I.d, r);
    t.G1 = -I.q * pjr - dot(I.d, pj);
    t.G2 = -mir * pjr - R(2) * dot(pj, Qir) - pjr * QiI;
    t.G3 = -pjr * QiR;
    if constexpr (GRAD) {
        t.g1 = (-I.q) * pj;
        t.g2 = (-pjr) * I.d - mir * pj - R(2) * smul(I.Q, pj) - QiI * pj;
        t.g3 = (-QiR) * pj - (R(2) * pjr) * Qir;
    }
    if constexpr (TORQUE) {
        // cross(Q, a b^T, ones) = cross(Q b, a)
        t.Gt1 = cross(I.d, pj);
        t.Gt2 = pjr * cross(I.d, r) + R(2) * cross(Qir, pj) + R(2) * cross(smul(I.Q, pj), r);
        t.Gt3 = (R(2) * pjr) * cross(Qir, r);
    }
}

// Damped radial functions of one pair (:1010-1066).
template <typename R> struct IndRadialT { R B1d, B2d, B3d, B4d, gB1, gB2, gB3, B2du, gB1u, gB2u; };
using IndRadial = IndRadialT<double>;

template <typename R>
__device__ __forceinline__ void ind_radial(const MpArgs& a, R r2, R pi_, R pj_, R scale, IndRadialT<R>& o) {
    R B[6];
    radial_b<R>(a, r2, B);
    R r, u3, expau3, inv_r2;
    const R ad = (R)a.a_damp;
    if constexpr (sizeof(R) == sizeof(double)) {
        r = sqrt(r2);
        u3 = r * r * r / sqrt(pi_ * pj_);
        expau3 = exp(-a.a_damp * u3);
        inv_r2 = R(1);   // unused: the FP64 form divides as the reference does
    } else {
        const float r_1 = __builtin_amdgcn_rsqf(r2);
        r = r2 * r_1;
        u3 = r * r * r * __builtin_amdgcn_rsqf(pi_ * pj_);
        expau3 = __builtin_amdgcn_exp2f(-ad * u3 * LOG2E_F);
        inv_r2 = r_1 * r_1;
    }
    const R au3 = ad * u3, a2u6 = au3 * au3, a3u9 = a2u6 * au3;
    const R l3 = R(1) - expau3;
    const R l5 = R(1) - (R(1) + au3) * expau3;
    const R l7 = R(1) - (R(1) + au3 + R(3) * a2u6 / R(5)) * expau3;
    R gl3, gl5, gl7;
    if constexpr (sizeof(R) == sizeof(double)) {
        gl3 = 3.0 * au3 * expau3 / r2;
        gl5 = 3.0 * a2u6 * expau3 / r2;
        gl7 = (-3.0 * a2u6 + 9.0 * a3u9) * expau3 / (5.0 * r2);
    } else {
        gl3 = 3.0f * au3 * expau3 * inv_r2;
        gl5 = 3.0f * a2u6 * expau3 * inv_r2;
        gl7 = (-3.0f * a2u6 + 9.0f * a3u9) * expau3 * (0.2f * inv_r2);
    }
    if (a.damp_amoeba) {
        R invr;
        if constexpr (sizeof(R) == sizeof(double)) invr = 1.0 / r;
        else invr = __builtin_amdgcn_rsqf(r2);
        const R invr2 = invr * invr, invr3 = invr2 * invr;
        const R invr5 = R(3) * invr3 * invr2, invr7 = R(5) * invr5 * invr2, invr9 = R(7) * invr7 * invr2;
        o.B1d = B[1] + (R(1) - scale * l3) * invr3;
        o.B2d = B[2] - (R(1) - scale * l5) * invr5;
        o.B3d = B[3] + (R(1) - scale * l7) * invr7;
        o.B4d = R(0);
        o.gB1 = B[2] - scale * gl3 * invr3 - (R(1) - scale * l3) * invr5;
        o.gB2 = B[3] + scale * gl5 * invr5 + (R(1) - scale * l5) * invr7;
        o.gB3 = B[4] - scale * gl7 * invr7 - (R(1) - scale * l7) * invr9;
        o.B2du = B[2] - (R(1) - l5) * invr5;
        o.gB1u = B[2] - gl3 * invr3 - (R(1) - l3) * invr5;
        o.gB2u = B[3] + gl5 * invr5 + (R(1) - l5) * invr7;
    } else {
        o.B1d = B[1] * l3;
        o.B2d = B[2] * l3 + B[1] * gl3;
        if constexpr (sizeof(R) == sizeof(double)) {
            o.B3d = B[3] * l3 + B[2] * 2.0 * gl3 + B[1] * (gl3 - 3.0 * gl5) / r2;
            o.B4d = B[4] * l3 + B[3] * 3.0 * gl3 + B[2] * (3.0 * gl3 - 9.0 * gl5) / r2 +
                    B[1] * (-gl3 - 6.0 * gl5 + 15.0 * gl7) / (r2 * r2);
        } else {
            o.B3d = B[3] * l3 + B[2] * 2.0f * gl3 + B[1] * (gl3 - 3.0f * gl5) * inv_r2;
            o.B4d = B[4] * l3 + B[3] * 3.0f * gl3 + B[2] * (3.0f * gl3 - 9.0f * gl5) * inv_r2 +
                    B[1] * (-gl3 - 6.0f * gl5 + 15.0f * gl7) * (inv_r2 * inv_r2);
        }
        o.gB1 = o.gB2 = o.gB3 = o.B2du = o.gB1u = o.gB2u = R(0);
    }
}

// Site j's induced dipole and polarizability in the pair math's precision.
template <typename R> __device__ __forceinline__ V3T<R> pind_r(const MpArgs& a, int j) {
    if constexpr (sizeof(R) == sizeof(double)) {
        return ldv(a.pind + 3 * j);
    } else {
        const float4 p = a.pind32[j];
        return {p.x, p.y, p.z};
    }
}
template <typename R> __device__ __forceinline__ R polar_r(const MpArgs& a, int j) {
    if constexpr (sizeof(R) == sizeof(double)) return a.polar[j];
    else return a.site32[3 * j + 2].z;
}

// One pair seen from row atom i (partner j, r = r_i - r_j): force on i, torque
// on i, and the pair's energy / virial / dU/dV contributions at half weight
// (every pair is visited from both rows).
template <typename R>
__device__ __forceinline__ void ind_entry(const MpArgs& a, const SiteT<R>& I, V3T<R> pi, R poli, V3 ri, int j,
                                          R scale, V3& f, V3& t, double (&acc)[11]) {
    double dx = ri.x - a.pos[3 * j], dy = ri.y - a.pos[3 * j + 1], dz = ri.z - a.pos[3 * j + 2];
    pbc(dx, dy, dz, a.box);
    const double r2d = dx * dx + dy * dy + dz * dz;
    if (!(r2d < a.realcut2)) return;
    const SiteT<R> J = site_r<R>(a, j);
    const V3T<R> pj = pind_r<R>(a, j);
    const V3T<R> r = vec3<R>((R)dx, (R)dy, (R)dz), rji = -r;
    IndSideT<R> si, sj;
    if (a.torques) ind_side<true, true>(I, pj, r, si);
    else ind_side<true, false>(I, pj, r, si);
    ind_side<true, false>(J, pi, rji, sj);
    IndRadialT<R> Rd;
    ind_radial<R>(a, (R)r2d, poli, polar_r<R>(a, j), scale, Rd);
    const R Gsum = (si.G1 + sj.G1) * Rd.B1d + (si.G2 + sj.G2) * Rd.B2d + (si.G3 + sj.G3) * Rd.B3d;
    const double uen = a.damp_amoeba ? 0.5 * (double)Gsum * COULOMBS : 0.5 * (double)Gsum * (double)scale * COULOMBS;
    const R G1ip = -dot(pi, pj);
    const R G2ip = -dot(pi, r) * dot(pj, r);
    const V3T<R> g2ip = (-dot(pj, r)) * pi;
    const V3T<R> g2jp = (-dot(pi, rji)) * pj;
    V3T<R> fi, fj;
    if (a.damp_amoeba) {
        V3T<R> gi = Rd.B1d * si.g1 + Rd.B2d * si.g2 + Rd.B3d * si.g3;
        V3T<R> gj = Rd.B1d * sj.g1 + Rd.B2d * sj.g2 + Rd.B3d * sj.g3;
        R ri_ = si.G1 * Rd.gB1 + si.G2 * Rd.gB2 + si.G3 * Rd.gB3;
        const R rj_ = sj.G1 * Rd.gB1 + sj.G2 * Rd.gB2 + sj.G3 * Rd.gB3;
        if (a.do_min) {
            gi = gi + Rd.B2du * g2ip;
            gj = gj + Rd.B2du * g2jp;
            ri_ += G1ip * Rd.gB1u + G2ip * Rd.gB2u;
        }
        fi = -gi - ri_ * r;
        fj = -gj - rj_ * rji;
    } else {
        V3T<R> gi = scale * (Rd.B1d * si.g1 + Rd.B2d * si.g2 + Rd.B3d * si.g3);
        V3T<R> gj = scale * (Rd.B1d * sj.g1 + Rd.B2d * sj.g2 + Rd.B3d * sj.g3);
        R ri_ = scale * (si.G1 * Rd.B2d + si.G2 * Rd.B3d + si.G3 * Rd.B4d);
        const R rj_ = scale * (sj.G1 * Rd.B2d + sj.G2 * Rd.B3d + sj.G3 * Rd.B4d);
        if (a.do_min) {
            gi = gi + Rd.B2d * g2ip;
            gj = gj + Rd.B2d * g2jp;
            ri_ += G1ip * Rd.B2d + G2ip * Rd.B3d;
        }
        fi = -gi - ri_ * r;
        fj = -gj - rj_ * rji;
    }
    const V3 fij = to64(fi - fj);
    f = f + fij;
    if (a.torques) {
        const V3T<R> ti = Rd.B1d * si.Gt1 + Rd.B2d * si.Gt2 + Rd.B3d * si.Gt3;
        t = t + to64((a.damp_amoeba ? R(1) : scale) * ti);
    }
    const V3 rd = to64(r);
    acc[0] += 0.5 * uen;
    acc[1] += -0.5 * dot(fij, rd);
    acc[2] += 0.5 * fij.x * rd.x; acc[3] += 0.5 * fij.y * rd.x; acc[4] += 0.5 * fij.z * rd.x;
    acc[5] += 0.5 * fij.x * rd.y; acc[6] += 0.5 * fij.y * rd.y; acc[7] += 0.5 * fij.z * rd.y;
    acc[8] += 0.5 * fij.x * rd.z; acc[9] += 0.5 * fij.y * rd.z; acc[10] += 0.5 * fij.z * rd.z;
}

// Real space over every pair within cutoffEw: a row's pair-list partners
// (both segments, not excluded) and its excluded partners from the exclusion
// table (scale 0).
template <int G, typename R>
__global__ __launch_bounds__(256) void k_ind_real(MpArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double acc[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) acc[k] = 0.0;
    V3 f = v3(0, 0, 0), t = v3(0, 0, 0);
    if (i < a.n) {
        const int32_t* row = a.nb + (int64_t)i * a.stride;
        const int own = min(a.nb_own[i], a.stride);
        const int non = min(a.nb_cnt[i] - a.nb_own[i], a.stride - own);
        const SiteT<R> I = site_r<R>(a, i);
        const V3T<R> pi = pind_r<R>(a, i);
        const R poli = polar_r<R>(a, i);
        const V3 ri = ldv(a.pos + 3 * i);
        for (int k = sub; k < own; k += G) ind_entry<R>(a, I, pi, poli, ri, row[k], R(1), f, t, acc);
        const int32_t* back = row + a.stride - 1;
        for (int k = sub; k < non; k += G) ind_entry<R>(a, I, pi, poli, ri, back[-k], R(1), f, t, acc);
        const int e0 = a.excl_off[i], e1 = a.excl_off[i + 1];
        for (int k = e0 + sub; k < e1; k += G) ind_entry<R>(a, I, pi, poli, ri, a.excl[k], R(0), f, t, acc);
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        f.x += __shfl_xor(f.x, o, G); f.y += __shfl_xor(f.y, o, G); f.z += __shfl_xor(f.z, o, G);
        t.x += __shfl_xor(t.x, o, G); t.y += __shfl_xor(t.y, o, G); t.z += __shfl_xor(t.z, o, G);
    }
    if (i < a.n && sub == 0) {
        a.force[3 * i] += COULOMBS * f.x;
        a.force[3 * i + 1] += COULOMBS * f.y;
        a.force[3 * i + 2] += COULOMBS * f.z;
        a.trq[3 * i] = COULOMBS * t.x;
        a.trq[3 * i + 1] = COULOMBS * t.y;
        a.trq[3 * i + 2] = COULOMBS * t.z;
    }
    acc[1] *= COULOMBS;
    block_part<11>(acc, a.red, RS_IND_REAL);
}

// The mixed real space over the solve's coupling table instead of the list:
// the table's rows hold exactly the partners inside cutoffInd (equal to the
// real-space cutoff when this kernel is used), so no lane walks the list
// buffer's partners beyond it; the intramolecular partners (the list's
// excluded ones) are the same-molecule entries, at scale 0.  Every entry
// keeps the list kernel's cutoff test and half-weight sums.
template <int G>
__global__ __launch_bounds__(256) void k_ind_real_tab(MpArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double acc[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) acc[k] = 0.0;
    V3 f = v3(0, 0, 0), t = v3(0, 0, 0);
    if (i < a.n) {
        const int cnt = a.tcnt[i];
        const int32_t* row = a.tj + (int64_t)i * a.tstride;
        const SiteT<float> I = site_r<float>(a, i);
        const V3T<float> pi = pind_r<float>(a, i);
        const float poli = polar_r<float>(a, i);
        const V3 ri = ldv(a.pos + 3 * i);
        const int mi = i / 3;
        for (int k = sub; k < cnt; k += G) {
            const int j = row[k];
            ind_entry<float>(a, I, pi, poli, ri, j, j / 3 == mi ? 0.0f : 1.0f, f, t, acc);
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        f.x += __shfl_xor(f.x, o, G); f.y += __shfl_xor(f.y, o, G); f.z += __shfl_xor(f.z, o, G);
        t.x += __shfl_xor(t.x, o, G); t.y += __shfl_xor(t.y, o, G); t.z += __shfl_xor(t.z, o, G);
    }
    if (i < a.n && sub == 0) {
        a.force[3 * i] += COULOMBS * f.x;
        a.force[3 * i + 1] += COULOMBS * f.y;
        a.force[3 * i + 2] += COULOMBS * f.z;
        a.trq[3 * i] = COULOMBS * t.x;
        a.trq[3 * i + 1] = COULOMBS * t.y;
        a.trq[3 * i + 2] = COULOMBS * t.z;
    }
    acc[1] *= COULOMBS;
    block_part<11>(acc, a.red, RS_IND_REAL);
}

// Per site: self energy and torque (:688-760) and the surface sums
// (:2222-2250).
__global__ __launch_bounds__(256) void k_ind_site_sums(MpArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) v[k] = 0.0;
    if (i < a.n) {
        const Site I = load_site(a, i);
        const V3 p = ldv(a.pind + 3 * i);
        v[0] = -(dot(I.d, p) * a.alpha2PI * a.alpha2 / 3.0) * COULOMBS;
        const V3 tq = (COULOMBS * (2.0 / 3.0) * a.alpha2PI * a.alpha2) * cross(I.d, p);
        double* tr = a.trq + 12 * (size_t)a.n;   // self part
        tr[3 * i] = a.torques ? tq.x : 0.0;
        tr[3 * i + 1] = a.torques ? tq.y : 0.0;
        tr[3 * i + 2] = a.torques ? tq.z : 0.0;
        const V3 r = ldv(a.pos + 3 * i), rim = ldv(a.com + 3 * a.mol[i]);
        v[1] = I.q * r.x; v[2] = I.q * r.y; v[3] = I.q * r.z;
        v[4] = I.q * rim.x; v[5] = I.q * rim.y; v[6] = I.q * rim.z;
        v[7] = I.d.x; v[8] = I.d.y; v[9] = I.d.z;
        v[10] = dot(I.d, p);
        v[11] = p.x; v[12] = p.y; v[13] = p.z;
        v[14] = dot(p, p);
    }
    block_store<15>(v, a.spart, gridDim.x);
}

// One thread per molecule: excluded pairs inside it (intraMolCorrect
// :1850-1985).  Forces and torques only without constraints.
__global__ __launch_bounds__(256) void k_ind_intra(MpArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (m < a.nmol) {
        const int a0 = a.mol_start[m], ns = a.mol_start[m + 1] - a0;
        V3 tq[3] = {v3(0, 0, 0), v3(0, 0, 0), v3(0, 0, 0)};
        for (int x = 0; x < ns; ++x) {
            const int idi = a0 + x;
            const Site I = load_site(a, idi);
            const V3 pi = ldv(a.pind + 3 * idi);
            for (int y = x + 1; y < ns; ++y) {
                const int idj = a0 + y;
                bool ex = false;
                for (int k = a.excl_off[idi]; k < a.excl_off[idi + 1]; ++k) ex = ex || a.excl[k] == idj;
                if (!ex) continue;
                double dx = a.pos[3 * idi] - a.pos[3 * idj], dy = a.pos[3 * idi + 1] - a.pos[3 * idj + 1],
                       dz = a.pos[3 * idi + 2] - a.pos[3 * idj + 2];
                pbc(dx, dy, dz, a.box);
                const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
                const V3 rv = v3(dx, dy, dz);
                const Site J = load_site(a, idj);
                const V3 pj = ldv(a.pind + 3 * idj);
                const double e = exp(-a.alpha2 * r2), inv = 1.0 / r2;
                double D[5];
                D[0] = erf(a.alpha * r) / r;
                D[1] = (-D[0] + a.alpha2PI * e) * inv;
                D[2] = (-3.0 * D[1] - 2.0 * a.alpha2PI * a.alpha2 * e) * inv;
                D[3] = (-5.0 * D[2] + 4.0 * a.alpha2PI * a.alpha2 * a.alpha2 * e) * inv;
                D[4] = (-7.0 * D[3] - 8.0 * a.alpha2PI * a.alpha2 * a.alpha2 * a.alpha2 * e) * inv;
                IndSide si, sj;
                if (!a.use_constraint && a.torques) {
                    ind_side<true, true>(I, pj, rv, si);
                    ind_side<true, true>(J, pi, -rv, sj);
                } else {
                    ind_side<true, false>(I, pj, rv, si);
                    ind_side<true, false>(J, pi, -rv, sj);
                }
                v[0] += (si.G1 + sj.G1) * D[1] + (si.G2 + sj.G2) * D[2] + (si.G3 + sj.G3) * D[3];
                if (!a.use_constraint) {
                    const V3 fi = D[1] * si.g1 + D[2] * si.g2 + D[3] * si.g3 + (si.G1 * D[2] + si.G2 * D[3] + si.G3 * D[4]) * rv;
                    const V3 fj =
                        D[1] * sj.g1 + D[2] * sj.g2 + D[3] * sj.g3 + (sj.G1 * D[2] + sj.G2 * D[3] + sj.G3 * D[4]) * (-rv);
                    const V3 fij = fi - fj;
                    v[1] += -dot(fij, rv) * COULOMBS / (3.0 * a.volume);
                    a.force[3 * idi] += COULOMBS * fij.x;
                    a.force[3 * idi + 1] += COULOMBS * fij.y;
                    a.force[3 * idi + 2] += COULOMBS * fij.z;
                    a.force[3 * idj] -= COULOMBS * fij.x;
                    a.force[3 * idj + 1] -= COULOMBS * fij.y;
                    a.force[3 * idj + 2] -= COULOMBS * fij.z;
                    if (a.torques) {
                        tq[x] = tq[x] - COULOMBS * (D[1] * si.Gt1 + D[2] * si.Gt2 + D[3] * si.Gt3);
                        tq[y] = tq[y] - COULOMBS * (D[1] * sj.Gt1 + D[2] * sj.Gt2 + D[3] * sj.Gt3);
                    }
                }
            }
        }
        double* tr = a.trq + 3 * (size_t)a.n;   // intramolecular part
        for (int x = 0; x < ns; ++x) {
            tr[3 * (a0 + x)] = tq[x].x;
            tr[3 * (a0 + x) + 1] = tq[x].y;
            tr[3 * (a0 + x) + 2] = tq[x].z;
        }
    }
    v[0] *= -0.5 * COULOMBS;
    block_part<2>(v, a.red, RS_IND_INTRA);
}

// Surface force, virial and torque per site (:2290-2345).
__global__ __launch_bounds__(256) void k_ind_surface(MpArgs a, double coeff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < a.n) {
        const V3 qr = ldv(a.terms + M_SU_QR), ps = ldv(a.terms + M_SU_PI);
        const Site I = load_site(a, i);
        const V3 p = ldv(a.pind + 3 * i);
        const V3 ri = ldv(a.pos + 3 * i);
        const V3 fi = (-2.0 * coeff * I.q) * ps;
        const V3 temp = (-a.q[i]) * qr;
        a.force[3 * i] += COULOMBS * fi.x;
        a.force[3 * i + 1] += COULOMBS * fi.y;
        a.force[3 * i + 2] += COULOMBS * fi.z;
        v[0] = temp.x * ri.x; v[1] = temp.y * ri.x; v[2] = temp.z * ri.x;
        v[3] = temp.x * ri.y; v[4] = temp.y * ri.y; v[5] = temp.z * ri.y;
        v[6] = temp.x * ri.z; v[7] = temp.y * ri.z; v[8] = temp.z * ri.z;
        const V3 tq = (COULOMBS * 2.0 * coeff) * (cross(ps, I.d) - cross(p, I.d));
        double* tr = a.trq + 6 * (size_t)a.n;   // surface part
        tr[3 * i] = a.torques ? tq.x : 0.0;
        tr[3 * i + 1] = a.torques ? tq.y : 0.0;
        tr[3 * i + 2] = a.torques ? tq.z : 0.0;
    }
    block_part<9>(v, a.red, RS_IND_SURF);
}

// k-space pass 1: thread per k over a chunk of sites; 28 sums per (chunk, k):
//   0 Sc  1 Ss   (q - kQk) cos/sin        2 Mc  3 Ms   (mu.k) cos/sin
//   4 Pc  5 Ps   (p.k) cos/sin            6 Q2c 7 Q2s  2 kQk cos/sin
//   8-10 (q - kQk) cos d   11-13 (q - kQk) sin d   14-16 (mu.k) cos d
//  17-19 (mu.k) sin d     20-22 (p.k) cos d       23-25 (p.k) sin d
//  26 (q - kQk) cos |d|^2  27 (q - kQk) sin |d|^2
template <typename R>
__global__ __launch_bounds__(MP_SF_THREADS) void k_ind_sfac(MpArgs a) {
    // per atom of the tile: r in FP64 (the phase); q, mu, Q, p, d in R
    __shared__ double shr[MP_SF_THREADS][3];
    __shared__ R shm[MP_SF_THREADS][16];
    const int k = blockIdx.x * MP_SF_THREADS + threadIdx.x;
    const int chunk = blockIdx.y;
    const int i0 = chunk * a.per, i1 = min(a.n, i0 + a.per);
    V3 kv = v3(0, 0, 0);
    if (k < a.K) kv = ldv(a.kvec + 3 * k);
    const R kx = (R)kv.x, ky = (R)kv.y, kz = (R)kv.z;
    double S[IN_NSUM];
#pragma unroll
    for (int x = 0; x < IN_NSUM; ++x) S[x] = 0.0;
    for (int t0 = i0; t0 < i1; t0 += MP_SF_THREADS) {
        const int i = t0 + threadIdx.x;
        __syncthreads();
        if (i < i1) {
            shr[threadIdx.x][0] = a.pos[3 * i]; shr[threadIdx.x][1] = a.pos[3 * i + 1]; shr[threadIdx.x][2] = a.pos[3 * i + 2];
            R* p = shm[threadIdx.x];
            p[0] = (R)(a.q[i] / SQRT_COULOMB_MP);
            p[1] = (R)a.dip[3 * i]; p[2] = (R)a.dip[3 * i + 1]; p[3] = (R)a.dip[3 * i + 2];
            for (int c = 0; c < 6; ++c) p[4 + c] = (R)a.quad[6 * i + c];
            p[10] = (R)a.pind[3 * i]; p[11] = (R)a.pind[3 * i + 1]; p[12] = (R)a.pind[3 * i + 2];
            const int m = a.mol[i];
            p[13] = (R)(a.pos[3 * i] - a.com[3 * m]);
            p[14] = (R)(a.pos[3 * i + 1] - a.com[3 * m + 1]);
            p[15] = (R)(a.pos[3 * i + 2] - a.com[3 * m + 2]);
        }
        __syncthreads();
        const int cnt = min(MP_SF_THREADS, i1 - t0);
        if (k < a.K) {
            // FP64: the sums accumulate directly; FP32: the tile's sums, added
            // to the FP64 totals after the tile
            auto tile = [&](auto& s) {
            for (int x = 0; x < cnt; ++x) {
                const R* p = shm[x];
                R sn, cs;
                phase_sincos<R>(shr[x][0] * kv.x + shr[x][1] * kv.y + shr[x][2] * kv.z, sn, cs);
                const R QK = p[4] * kx * kx + p[7] * ky * ky + p[9] * kz * kz +
                             R(2) * (p[5] * kx * ky + p[6] * kx * kz + p[8] * ky * kz);
                const R A = p[0] - QK;
                const R Mu = p[1] * kx + p[2] * ky + p[3] * kz;
                const R Pk = p[10] * kx + p[11] * ky + p[12] * kz;
                const R ac = A * cs, as = A * sn, mc = Mu * cs, ms = Mu * sn, pc = Pk * cs, ps = Pk * sn;
                const R dx = p[13], dy = p[14], dz = p[15];
                s[0] += ac; s[1] += as; s[2] += mc; s[3] += ms; s[4] += pc; s[5] += ps;
                s[6] += R(2) * QK * cs; s[7] += R(2) * QK * sn;
                s[8] += ac * dx; s[9] += ac * dy; s[10] += ac * dz;
                s[11] += as * dx; s[12] += as * dy; s[13] += as * dz;
                s[14] += mc * dx; s[15] += mc * dy; s[16] += mc * dz;
                s[17] += ms * dx; s[18] += ms * dy; s[19] += ms * dz;
                s[20] += pc * dx; s[21] += pc * dy; s[22] += pc * dz;
                s[23] += ps * dx; s[24] += ps * dy; s[25] += ps * dz;
                const R d2 = dx * dx + dy * dy + dz * dz;
                s[26] += ac * d2; s[27] += as * d2;
            }
            };
            if constexpr (sizeof(R) == sizeof(double)) {
                tile(S);
            } else {
                R s[IN_NSUM];
#pragma unroll
                for (int x = 0; x < IN_NSUM; ++x) s[x] = R(0);
                tile(s);
#pragma unroll
                for (int x = 0; x < IN_NSUM; ++x) S[x] += (double)s[x];
            }
        }
    }
    if (k < a.K) {
        double* out = a.part + ((size_t)chunk * a.K + k) * IN_NSUM;
#pragma unroll
        for (int x = 0; x < IN_NSUM; ++x) out[x] = S[x];
    }
}

// k-space pass 2 (:1564-1600, the force-pass sums :1640-1650 from moments).
// Force-pass coefficients per k: X = b (Sc - Ms'), Y = b (Ss + Mc'), b Pc,
// b Ps with Mc' = Mc (+ Pc), Ms' = Ms (+ Ps) when dipoles are minimised.
__device__ __forceinline__ void ind_finish(const MpArgs& a, int k, const double (&S)[IN_NSUM], double (&v)[9]) {
    {
        const V3 kv = ldv(a.kvec + 3 * k);
        const double pi4V = 4.0 * PI / a.volume, pi8V = 2.0 * pi4V;
        const double ek = a.expk[k];
        const double Sc = S[0], Ss = S[1], Mc = S[2], Ms = S[3], Pc = S[4], Ps = S[5], Q2c = S[6], Q2s = S[7];
        const double UmqQ = pi4V * ek * (Ss * Pc - Sc * Ps);
        const double Ump = pi4V * ek * (Pc * Mc + Ps * Ms);
        const double t = (UmqQ + Ump) * COULOMBS;
        const double dUdVpQ = pi4V * ek * (Q2s * Pc - Q2c * Ps);
        const double Upp = a.do_min ? pi4V * ek * (Pc * Pc + Ps * Ps) : 0.0;
        const double Sk = 2.0 * UmqQ + 2.0 * Ump + Upp;
        const double k2 = dot(kv, kv);
        const double x = k2 * 2.0 * a.alphaR4;
        double d1 = (-Sk + Sk * x + (-2.0 * UmqQ + 2.0 * dUdVpQ - 4.0 * Ump - 2.0 * Upp)) * COULOMBS;
        double d2 = 4.0 * t - 7.0 * t * x + t * x * x;
        const double coeff = 2.0 * (a.alphaR4 + 1.0 / a.kmod2[k]);
        const double b = pi8V * ek;
        const double Mcp = a.do_min ? Mc + Pc : Mc, Msp = a.do_min ? Ms + Ps : Ms;
        const double X = b * (Sc - Msp), Y = b * (Ss + Mcp), bPc = b * Pc, bPs = b * Ps;
        if (a.use_constraint) {
            const V3 cosrR = v3(S[8], S[9], S[10]), sinrR = v3(S[11], S[12], S[13]);
            const V3 mcR = v3(S[14], S[15], S[16]), msR = v3(S[17], S[18], S[19]);
            const V3 pcR = v3(S[20], S[21], S[22]), psR = v3(S[23], S[24], S[25]);
            const double Ff = COULOMBS * dot(kv, X * pcR + Y * psR - bPc * cosrR - bPs * mcR - bPs * sinrR + bPc * msR);
            const double Ffk2 = 2.0 * k2 * 2.0 * a.alphaR4 * Ff;
            const double d2UV1 = b * k2 * (Sc * S[26] + Ss * S[27]);
            const double d2UrR = b * k2 * (dot(cosrR, cosrR) + dot(sinrR, sinrR));
            d1 += Ff;
            d2 += -6.0 * Ff + Ffk2 - d2UV1 + d2UrR;
        }
        v[0] = t; v[1] = d1; v[2] = d2;
        v[3] = t - t * coeff * kv.x * kv.x;
        v[4] = t - t * coeff * kv.y * kv.y;
        v[5] = t - t * coeff * kv.z * kv.z;
        v[6] = -t * coeff * kv.x * kv.y;
        v[7] = -t * coeff * kv.x * kv.z;
        v[8] = -t * coeff * kv.y * kv.z;
        a.kf[4 * k] = X;
        a.kf[4 * k + 1] = Y;
        a.kf[4 * k + 2] = bPc;
        a.kf[4 * k + 3] = bPs;
    }
}

__global__ __launch_bounds__(256) void k_ind_kfinish(MpArgs a) {
    const int k = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (k < a.K) {
        double S[IN_NSUM];
        wave_chunk_sum<IN_NSUM>(a, k, S);
        if (threadIdx.x % WAVE == 0) ind_finish(a, k, S, v);
    }
    block_part<9>(v, a.red, RS_IND_LONG);
}

// k-space pass 3 (:1603-1690), split over k-ranges like k_mp_kforce.
template <typename R>
__global__ __launch_bounds__(WAVE) void k_ind_kforce(MpArgs a) {
    __shared__ double sk[MP_KTILE][7];
    const int i = blockIdx.x * WAVE + threadIdx.x;
    const int kb = blockIdx.y * a.kper, ke = min(a.K, kb + a.kper);
    Site I64{};
    V3 ri = v3(0, 0, 0), p64 = v3(0, 0, 0);
    if (i < a.n) {
        I64 = load_site(a, i);
        ri = ldv(a.pos + 3 * i);
        p64 = ldv(a.pind + 3 * i);
    }
    const SiteT<R> I = site_as<R>(I64);
    const V3T<R> p = vec3<R>((R)p64.x, (R)p64.y, (R)p64.z);
    V3 F = v3(0, 0, 0), T = v3(0, 0, 0);
    for (int k0 = kb; k0 < ke; k0 += MP_KTILE) {
        const int cnt = min(MP_KTILE, ke - k0);
        __syncthreads();
        for (int x = threadIdx.x; x < cnt; x += WAVE) {
            sk[x][0] = a.kvec[3 * (k0 + x)];
            sk[x][1] = a.kvec[3 * (k0 + x) + 1];
            sk[x][2] = a.kvec[3 * (k0 + x) + 2];
            sk[x][3] = a.kf[4 * (k0 + x)];
            sk[x][4] = a.kf[4 * (k0 + x) + 1];
            sk[x][5] = a.kf[4 * (k0 + x) + 2];
            sk[x][6] = a.kf[4 * (k0 + x) + 3];
        }
        __syncthreads();
        if (i < a.n) {
            V3T<R> f = vec3<R>(R(0), R(0), R(0)), t = f;
            for (int x = 0; x < cnt; ++x) {
                R sn, cs;
                phase_sincos<R>(ri.x * sk[x][0] + ri.y * sk[x][1] + ri.z * sk[x][2], sn, cs);
                const V3T<R> kv = vec3<R>((R)sk[x][0], (R)sk[x][1], (R)sk[x][2]);
                const R X = (R)sk[x][3], Y = (R)sk[x][4], bPc = (R)sk[x][5], bPs = (R)sk[x][6];
                const V3T<R> Qk = smul(I.Q, kv);
                const R A = I.q - dot(kv, Qk), Mu = dot(I.d, kv), Pk = dot(p, kv);
                const R val = cs * (Pk * X - A * bPc - Mu * bPs) + sn * (Pk * Y - A * bPs + Mu * bPc);
                f = f + val * kv;
                if (a.torques)
                    t = t + (cs * bPc + sn * bPs) * cross(kv, I.d) + (R(2) * (cs * bPs - sn * bPc)) * cross(kv, Qk);
            }
            F = F + to64(f);
            T = T + to64(t);
        }
    }
    if (i < a.n) {
        double* q = a.fpart + ((size_t)blockIdx.y * a.n + i) * 6;
        q[0] = F.x; q[1] = F.y; q[2] = F.z;
        q[3] = T.x; q[4] = T.y; q[5] = T.z;
    }
}

// ---- the polarizable step's joint k-space pass ----------------------------------
//
// Multipole and Induction on the same k set (same alpha, kcut and box): the
// induction pass's 28 sums per k hold the 20 of the multipole pass (rows 0-3,
// 2 kQk as 6-7, the constraint moments as 8-19 and 26-27), so one
// structure-factor pass serves both classes' energies, virials and volume
// derivatives, and one force pass applies both classes' coefficients:
//   val   = c (P.k X - A U + Mu W) + s (P.k Y + A W + Mu U)
//   tau   = (c U - s W) k x mu - 2 (c W + s U) k x Qk
// with U = b Q + b Pc and W = b P - b Ps (b P, b Q the multipole pass's
// coefficients; X, Y, b Pc, b Ps the induction pass's).

// The induced dipoles' own sums per (chunk, k), for the joint pass when the
// permanent multipoles' 20 came from k_mp_sfac (before the solve):
//   0 Pc  1 Ps  (p.k) cos/sin     2-4 (p.k) cos d     5-7 (p.k) sin d
constexpr int INP_NSUM = 8;

template <typename R>
__global__ __launch_bounds__(MP_SF_THREADS) void k_indp_sfac(MpArgs a) {
    __shared__ double shr[MP_SF_THREADS][3];
    __shared__ R shm[MP_SF_THREADS][6];
    const int k = blockIdx.x * MP_SF_THREADS + threadIdx.x;
    const int chunk = blockIdx.y;
    const int i0 = chunk * a.per, i1 = min(a.n, i0 + a.per);
    V3 kv = v3(0, 0, 0);
    if (k < a.K) kv = ldv(a.kvec + 3 * k);
    const R kx = (R)kv.x, ky = (R)kv.y, kz = (R)kv.z;
    double S[INP_NSUM];
#pragma unroll
    for (int x = 0; x < INP_NSUM; ++x) S[x] = 0.0;
    for (int t0 = i0; t0 < i1; t0 += MP_SF_THREADS) {
        const int i = t0 + threadIdx.x;
        __syncthreads();
        if (i < i1) {
            shr[threadIdx.x][0] = a.pos[3 * i]; shr[threadIdx.x][1] = a.pos[3 * i + 1]; shr[threadIdx.x][2] = a.pos[3 * i + 2];
            R* p = shm[threadIdx.x];
            p[0] = (R)a.pind[3 * i]; p[1] = (R)a.pind[3 * i + 1]; p[2] = (R)a.pind[3 * i + 2];
            const int m = a.mol[i];
            p[3] = (R)(a.pos[3 * i] - a.com[3 * m]);
            p[4] = (R)(a.pos[3 * i + 1] - a.com[3 * m + 1]);
            p[5] = (R)(a.pos[3 * i + 2] - a.com[3 * m + 2]);
        }
        __syncthreads();
        const int cnt = min(MP_SF_THREADS, i1 - t0);
        if (k < a.K) {
            auto tile = [&](auto& s) {
            for (int x = 0; x < cnt; ++x) {
                const R* p = shm[x];
                R sn, cs;
                phase_sincos<R>(shr[x][0] * kv.x + shr[x][1] * kv.y + shr[x][2] * kv.z, sn, cs);
                const R Pk = p[0] * kx + p[1] * ky + p[2] * kz;
                const R pc = Pk * cs, ps = Pk * sn;
                s[0] += pc; s[1] += ps;
                s[2] += pc * p[3]; s[3] += pc * p[4]; s[4] += pc * p[5];
                s[5] += ps * p[3]; s[6] += ps * p[4]; s[7] += ps * p[5];
            }
            };
            if constexpr (sizeof(R) == sizeof(double)) {
                tile(S);
            } else {
                R s[INP_NSUM];
#pragma unroll
                for (int x = 0; x < INP_NSUM; ++x) s[x] = R(0);
                tile(s);
#pragma unroll
                for (int x = 0; x < INP_NSUM; ++x) S[x] += (double)s[x];
            }
        }
    }
    if (k < a.K) {
        double* out = a.part + ((size_t)chunk * a.K + k) * INP_NSUM;
#pragma unroll
        for (int x = 0; x < INP_NSUM; ++x) out[x] = S[x];
    }
}

// One wave per k: both classes' per-k terms from one set of sums (SPLIT: the
// permanent multipoles' 20 from the multipole pass's partials and the induced
// dipoles' 8 from k_indp_sfac's; otherwise all 28 from k_ind_sfac's); the
// joint force coefficients (X, Y, U, W) replace the induction pass's.
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_joint_kfinish(MpArgs mp, MpArgs in) {
    const int k = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    double vm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, vi[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (k < in.K) {
        double S[IN_NSUM], M[MP_NSUM];
        if constexpr (SPLIT) {
            double P[INP_NSUM];
            wave_chunk_sum<MP_NSUM>(mp, k, M);
            wave_chunk_sum<INP_NSUM>(in, k, P);
            S[0] = M[0]; S[1] = M[1]; S[2] = M[2]; S[3] = M[3]; S[4] = P[0]; S[5] = P[1]; S[6] = M[4]; S[7] = M[5];
#pragma unroll
            for (int x = 0; x < 12; ++x) S[8 + x] = M[6 + x];
#pragma unroll
            for (int x = 0; x < 6; ++x) S[20 + x] = P[2 + x];
            S[26] = M[18]; S[27] = M[19];
        } else {
            wave_chunk_sum<IN_NSUM>(in, k, S);
            M[0] = S[0]; M[1] = S[1]; M[2] = S[2]; M[3] = S[3]; M[4] = S[6]; M[5] = S[7];
#pragma unroll
            for (int x = 0; x < 12; ++x) M[6 + x] = S[8 + x];
            M[18] = S[26]; M[19] = S[27];
        }
        if (threadIdx.x % WAVE == 0) {
            mp_finish(mp, k, M, vm);
            ind_finish(in, k, S, vi);
            const double bP = mp.kf[4 * k], bQ = mp.kf[4 * k + 1];
            const double bPc = in.kf[4 * k + 2], bPs = in.kf[4 * k + 3];
            in.kf[4 * k + 2] = bQ + bPc;
            in.kf[4 * k + 3] = bP - bPs;
        }
    }
    block_part<9>(vm, mp.red, RS_MP_LONG);
    __syncthreads();
    block_part<9>(vi, in.red, RS_IND_LONG);
}

// The joint force pass: the induction pass's layout (site per lane, k-range
// per block row) with both classes' force and torque per (site, k).
// The phase e^{i k.r} of a k that continues the previous one's (nx, ny)
// row with nz + 1 is the previous phase times e^{i 2 pi z / Lz} (a rotation,
// uniform across the wave since every lane takes the same k); every eighth
// k of a tile, and every k that starts a row, is evaluated directly.
template <typename R>
__global__ __launch_bounds__(WAVE) void k_joint_kforce(MpArgs a) {
    __shared__ double sk[MP_KTILE][7];
    __shared__ int sst[MP_KTILE];
    const int i = blockIdx.x * WAVE + threadIdx.x;
    const int kb = blockIdx.y * a.kper, ke = min(a.K, kb + a.kper);
    Site I64{};
    V3 ri = v3(0, 0, 0), p64 = v3(0, 0, 0);
    if (i < a.n) {
        I64 = load_site(a, i);
        ri = ldv(a.pos + 3 * i);
        p64 = ldv(a.pind + 3 * i);
    }
    const SiteT<R> I = site_as<R>(I64);
    const V3T<R> p = vec3<R>((R)p64.x, (R)p64.y, (R)p64.z);
    R sz, cz;
    phase_sincos<R>(ri.z * a.kz_step, sz, cz);
    V3 F = v3(0, 0, 0), T = v3(0, 0, 0);
    for (int k0 = kb; k0 < ke; k0 += MP_KTILE) {
        const int cnt = min(MP_KTILE, ke - k0);
        __syncthreads();
        for (int x = threadIdx.x; x < cnt; x += WAVE) {
            sk[x][0] = a.kvec[3 * (k0 + x)];
            sk[x][1] = a.kvec[3 * (k0 + x) + 1];
            sk[x][2] = a.kvec[3 * (k0 + x) + 2];
            sk[x][3] = a.kf[4 * (k0 + x)];
            sk[x][4] = a.kf[4 * (k0 + x) + 1];
            sk[x][5] = a.kf[4 * (k0 + x) + 2];
            sk[x][6] = a.kf[4 * (k0 + x) + 3];
            sst[x] = a.kstep[k0 + x];
        }
        __syncthreads();
        if (i < a.n) {
            V3T<R> f = vec3<R>(R(0), R(0), R(0)), t = f;
            R sn = R(0), cs = R(1);
            for (int x = 0; x < cnt; ++x) {
                if ((x & 7) != 0 && sst[x]) {
                    const R c2 = cs * cz - sn * sz;
                    sn = sn * cz + cs * sz;
                    cs = c2;
                } else {
                    phase_sincos<R>(ri.x * sk[x][0] + ri.y * sk[x][1] + ri.z * sk[x][2], sn, cs);
                }
                const V3T<R> kv = vec3<R>((R)sk[x][0], (R)sk[x][1], (R)sk[x][2]);
                const R X = (R)sk[x][3], Y = (R)sk[x][4], U = (R)sk[x][5], W = (R)sk[x][6];
                const V3T<R> Qk = smul(I.Q, kv);
                const R A = I.q - dot(kv, Qk), Mu = dot(I.d, kv), Pk = dot(p, kv);
                const R val = cs * (Pk * X - A * U + Mu * W) + sn * (Pk * Y + A * W + Mu * U);
                f = f + val * kv;
                if (a.torques) t = t + (cs * U - sn * W) * cross(kv, I.d) - (R(2) * (cs * W + sn * U)) * cross(kv, Qk);
            }
            F = F + to64(f);
            T = T + to64(t);
        }
    }
    if (i < a.n) {
        double* q = a.fpart + ((size_t)blockIdx.y * a.n + i) * 6;
        q[0] = F.x; q[1] = F.y; q[2] = F.z;
        q[3] = T.x; q[4] = T.y; q[5] = T.z;
    }
}

static MpArgs ind_args(Ctx* c) {
    const smd_indf_params& p = c->indf;
    MpArgs a = mp_args(c, c->ind_ks, p.alpha, p.real_cut, p.use_constraint, p.torques);
    a.a_damp = p.a_damp;
    a.damp_amoeba = p.damp_amoeba;
    a.do_min = p.do_min;
    return a;
}

static int ind_check(Ctx* c) {
    if (!c->indf_ready) return fail(c, SMD_ESTATE, "smd_induction_setup not called");
    if (!c->multipoles_set) return fail(c, SMD_ESTATE, "multipoles not set");
    return SMD_OK;
}

int induction_short(Ctx* c, smd_indf_out* out) {
    int rc = ind_check(c);
    if (rc) return rc;
    if (!c->list_valid) return fail(c, SMD_ESTATE, "no pair list");
    if (c->indf.real_cut > c->list_cutoff + c->list_buff)
        return fail(c, SMD_ENOTSUP, "real-space cutoff %g exceeds the pair-list radius %g", c->indf.real_cut,
                    c->list_cutoff + c->list_buff);
    MpArgs a = ind_args(c);
    const dim3 b256(256);
    const dim3 ga((c->n + 255) / 256), gm((c->nmol + 255) / 256);
    const dim3 gr((unsigned)(((int64_t)c->n * MP_LANES + 255) / 256));
    ProfScope ps(c, "ind_short");
    SMD_HIP(c, hipMemsetAsync(c->d_mp_terms, 0, sizeof(double) * M_COUNT, c->stream));
    if ((rc = frames_run(c))) return rc;
    hipLaunchKernelGGL(k_ind_site_sums, ga, b256, 0, c->stream, a);
    if ((rc = site_sums_finish(c, a, ga.x, true))) return rc;
    if ((rc = launch_real(c, a, true))) return rc;
    hipLaunchKernelGGL(k_ind_intra, gm, b256, 0, c->stream, a);
    if ((rc = red_finish(c, RS_IND_INTRA, gm.x, 2, a.terms + M_INTRA_E))) return rc;
    const double volume = a.volume, volumer = 1.0 / volume;
    const double coeff = 2.0 * PI * volumer / 3.0;
    if (c->indf.surface) {
        hipLaunchKernelGGL(k_ind_surface, ga, b256, 0, c->stream, a, coeff);
        if ((rc = red_finish(c, RS_IND_SURF, ga.x, 9, a.terms + M_SU_VIR))) return rc;
    }
    // torques of the self term, real space, intramolecular, surface
    if (c->indf.torques && (rc = launch_torques(c, a, gm, 16u | 3u | (c->indf.surface ? 4u : 0u), RS_IND_DUDV))) return rc;
    SMD_HIP(c, hipGetLastError());
    double t[M_COUNT];
    if ((rc = mp_read_terms(c, t))) return rc;
    *out = smd_indf_out{};
    out->self_energy = t[M_SELF];
    out->self_deriv = t[M_DUDV + 4];
    out->real_energy = t[M_REAL_E];
    out->real_deriv = t[M_REAL_D1] + t[M_DUDV];
    out->intra_energy = t[M_INTRA_E];
    out->mol_deriv = t[M_MOL_D1] + t[M_DUDV + 1] / (3.0 * volume);
    for (int k = 0; k < 9; ++k) out->virial[k] = t[M_REAL_VIR + k];
    out->tot_torque = t[M_TRQ] + t[M_TRQ + 1] + t[M_TRQ + 2] + t[M_TRQ + 4];
    if (c->indf.surface) {
        const double* qr = t + M_SU_QR;
        const double* qR = t + M_SU_QRR;
        const double* mi = t + M_SU_QM;
        const double* ps = t + M_SU_PI;
        auto d3 = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
        const double selfSum = t[M_SU_SELF], selfpi = t[M_SU_PP];
        out->surf_energy = coeff * (d3(qr, ps) + d3(mi, ps) - selfSum) * COULOMBS;
        const double dUdV1 = c->indf.do_min
                                 ? -coeff * volumer * (2.0 * d3(qr, ps) + 2.0 * d3(mi, ps) + d3(ps, ps) - 2.0 * selfSum - selfpi)
                                 : -coeff * volumer * (2.0 * d3(qr, ps) + 2.0 * d3(mi, ps) - 2.0 * selfSum);
        const double dUdV2 = 2.0 * coeff * volumer * d3(qr, ps) / 3.0;
        out->surf_deriv = COULOMBS * (dUdV1 + dUdV2) + t[M_DUDV + 2] * volumer / 3.0;
        if (c->indf.use_constraint)
            out->surf_deriv2 = 2.0 * coeff * volumer * volumer * d3(qr, qr) - 16.0 * coeff * volumer * volumer * d3(qr, qR) / 9.0 +
                               2.0 * coeff * volumer * volumer * d3(qR, qR) / 9.0;
        else
            out->surf_deriv2 = 4.0 * coeff * volumer * volumer * d3(qr, qr) / 9.0;
        for (int k = 0; k < 9; ++k) out->virial[k] += t[M_SU_VIR + k];
    }
    return SMD_OK;
}

int induction_long(Ctx* c, smd_mp_long_out* out) {
    int rc = ind_check(c);
    if (rc) return rc;
    MpArgs a = ind_args(c);
    const dim3 b256(256);
    ProfScope ps(c, "ind_long");
    SMD_HIP(c, hipMemsetAsync(c->d_mp_terms, 0, sizeof(double) * M_COUNT, c->stream));
    if ((rc = frames_run(c))) return rc;
    const dim3 g1((a.K + MP_SF_THREADS - 1) / MP_SF_THREADS, a.chunks);
    if (c->mp_prec == SMD_PREC_MIXED) hipLaunchKernelGGL((k_ind_sfac<float>), g1, dim3(MP_SF_THREADS), 0, c->stream, a);
        else hipLaunchKernelGGL((k_ind_sfac<double>), g1, dim3(MP_SF_THREADS), 0, c->stream, a);
    hipLaunchKernelGGL(k_ind_kfinish, dim3((a.K + 3) / 4), b256, 0, c->stream, a);
    if ((rc = red_finish(c, RS_IND_LONG, (a.K + 3) / 4, 9, a.terms + M_LONG_E))) return rc;
    if (c->mp_prec == SMD_PREC_MIXED) hipLaunchKernelGGL((k_ind_kforce<float>), dim3((c->n + WAVE - 1) / WAVE, a.ksplit), dim3(WAVE), 0, c->stream, a);
        else hipLaunchKernelGGL((k_ind_kforce<double>), dim3((c->n + WAVE - 1) / WAVE, a.ksplit), dim3(WAVE), 0, c->stream, a);
    hipLaunchKernelGGL(k_kforce_sum, dim3((c->n + 255) / 256), b256, 0, c->stream, a);
    if (c->indf.torques && (rc = launch_torques(c, a, dim3((c->nmol + 255) / 256), 8u, RS_IND_DUDV_L))) return rc;
    SMD_HIP(c, hipGetLastError());
    double t[M_COUNT];
    if ((rc = mp_read_terms(c, t))) return rc;
    *out = smd_mp_long_out{};
    out->energy = t[M_LONG_E];
    out->deriv = t[M_LONG_D1] + t[M_DUDV + 3];
    out->deriv2 = t[M_LONG_D2];
    const double* lv = t + M_LONG_VIR;
    out->virial[0] = lv[0]; out->virial[4] = lv[1]; out->virial[8] = lv[2];
    out->virial[1] = out->virial[3] = lv[3];
    out->virial[2] = out->virial[6] = lv[4];
    out->virial[5] = out->virial[7] = lv[5];
    out->tot_torque = t[M_TRQ + 3];
    return SMD_OK;
}

// The four classes' kernels without read-back for the polarizable MD step:
// terms go to the given accumulator (cleared here by a fill kernel, so the
// sequence can be captured), forces and torque-derived forces into d_force.
// The lab-frame multipoles must be current (frames_run).
// The k-space passes of the step up to the per-site partial forces
// (structure factors, per-k sums and coefficients, the force pass over
// k-ranges) for the permanent multipoles (mp) or the induced dipoles.  They
// write only their own partial buffers and, through the reduction arena,
// the long-range terms.
static int mp_long_sums(Ctx* c, const MpArgs& a, bool mp) {
    const char* fam = mp ? "mp_long" : "ind_long";
    ProfScope ps(c, fam);
    const dim3 g1((a.K + MP_SF_THREADS - 1) / MP_SF_THREADS, a.chunks), b256(256);
    const dim3 gk((c->n + WAVE - 1) / WAVE, a.ksplit);
    const bool mixed = c->mp_prec == SMD_PREC_MIXED;
    if (mp) {
        if (mixed) hipLaunchKernelGGL((k_mp_sfac<float>), g1, dim3(MP_SF_THREADS), 0, c->stream, a);
        else hipLaunchKernelGGL((k_mp_sfac<double>), g1, dim3(MP_SF_THREADS), 0, c->stream, a);
        hipLaunchKernelGGL(k_mp_kfinish, dim3((a.K + 3) / 4), b256, 0, c->stream, a);
        const int rc = red_finish(c, RS_MP_LONG, (a.K + 3) / 4, 9, a.terms + M_LONG_E);
        if (rc) return rc;
        if (mixed) hipLaunchKernelGGL((k_mp_kforce<float>), gk, dim3(WAVE), 0, c->stream, a);
        else hipLaunchKernelGGL((k_mp_kforce<double>), gk, dim3(WAVE), 0, c->stream, a);
    } else {
        if (mixed) hipLaunchKernelGGL((k_ind_sfac<float>), g1, dim3(MP_SF_THREADS), 0, c->stream, a);
        else hipLaunchKernelGGL((k_ind_sfac<double>), g1, dim3(MP_SF_THREADS), 0, c->stream, a);
        hipLaunchKernelGGL(k_ind_kfinish, dim3((a.K + 3) / 4), b256, 0, c->stream, a);
        const int rc = red_finish(c, RS_IND_LONG, (a.K + 3) / 4, 9, a.terms + M_LONG_E);
        if (rc) return rc;
        if (mixed) hipLaunchKernelGGL((k_ind_kforce<float>), gk, dim3(WAVE), 0, c->stream, a);
        else hipLaunchKernelGGL((k_ind_kforce<double>), gk, dim3(WAVE), 0, c->stream, a);
    }
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// Fork of the step's k-space passes onto the second stream (when the
// context has one and reciprocal overlap is on): the second stream waits for
// everything enqueued so far and becomes the context's stream until
// polar_fork_end; polar_join makes the step's stream wait for it.
bool polar_fork_begin(Ctx* c) {
    if (!c->recip_overlap || !c->stream2 || c->red_defer == false) return false;
    if (hipEventRecord(c->ev_fork, c->stream) != hipSuccess) return false;
    if (hipStreamWaitEvent(c->stream2, c->ev_fork, 0) != hipSuccess) return false;
    c->polar_main = c->stream;
    c->stream = c->stream2;
    return true;
}

void polar_fork_end(Ctx* c) {
    (void)hipEventRecord(c->ev_join, c->stream2);
    c->stream = c->polar_main;
}

void polar_join(Ctx* c) { (void)hipStreamWaitEvent(c->stream, c->ev_join, 0); }

// The step's short-range kernels of each class (self, site sums, real
// space, intramolecular, surface, their torques), into a.terms and d_force.
static int mp_short_enqueue(Ctx* c, const MpArgs& a) {
    int rc;
    const dim3 b256(256), ga((c->n + 255) / 256), gm((c->nmol + 255) / 256);
    ProfScope ps(c, "mp_short");
    hipLaunchKernelGGL(k_mp_site_sums, ga, b256, 0, c->stream, a);
    if ((rc = site_sums_finish(c, a, ga.x, false))) return rc;
    if ((rc = launch_real(c, a, false))) return rc;
    hipLaunchKernelGGL(k_mp_intra, gm, b256, 0, c->stream, a);
    if ((rc = red_finish(c, RS_MP_INTRA, gm.x, 2, a.terms + M_INTRA_E))) return rc;
    if (c->mp.surface) {
        hipLaunchKernelGGL(k_mp_surface, ga, b256, 0, c->stream, a, 2.0 * PI * (1.0 / a.volume) / 3.0);
        if ((rc = red_finish(c, RS_MP_SURF, ga.x, 9, a.terms + M_SU_VIR))) return rc;
    }
    if (c->mp.torques && (rc = launch_torques(c, a, gm, c->mp.surface ? 7u : 3u, RS_MP_DUDV))) return rc;
    return SMD_OK;
}

static int ind_short_enqueue(Ctx* c, const MpArgs& a) {
    int rc;
    const dim3 b256(256), ga((c->n + 255) / 256), gm((c->nmol + 255) / 256);
    ProfScope ps(c, "ind_short");
    hipLaunchKernelGGL(k_ind_site_sums, ga, b256, 0, c->stream, a);
    if ((rc = site_sums_finish(c, a, ga.x, true))) return rc;
    if ((rc = launch_real(c, a, true))) return rc;
    hipLaunchKernelGGL(k_ind_intra, gm, b256, 0, c->stream, a);
    if ((rc = red_finish(c, RS_IND_INTRA, gm.x, 2, a.terms + M_INTRA_E))) return rc;
    if (c->indf.surface) {
        hipLaunchKernelGGL(k_ind_surface, ga, b256, 0, c->stream, a, 2.0 * PI * (1.0 / a.volume) / 3.0);
        if ((rc = red_finish(c, RS_IND_SURF, ga.x, 9, a.terms + M_SU_VIR))) return rc;
    }
    if (c->indf.torques && (rc = launch_torques(c, a, gm, 16u | 3u | (c->indf.surface ? 4u : 0u), RS_IND_DUDV)))
        return rc;
    return SMD_OK;
}

int multipole_enqueue(Ctx* c, double* terms) {
    int rc;
    MpArgs a = mp_args(c);
    a.terms = terms;
    const dim3 b256(256);
    const dim3 ga((c->n + 255) / 256), gm((c->nmol + 255) / 256);
    const dim3 gr((unsigned)(((int64_t)c->n * MP_LANES + 255) / 256));
    hipLaunchKernelGGL(k_zero_terms, dim3(1), b256, 0, c->stream, terms);
    // the k-space sums and per-site partial forces on the second stream,
    // beside the real-space kernels (they touch none of the same buffers)
    const bool fork = polar_fork_begin(c);
    int rc_long = fork ? mp_long_sums(c, a, true) : SMD_OK;
    if (fork) polar_fork_end(c);
    if (rc_long) return rc_long;
    if ((rc = mp_short_enqueue(c, a))) return rc;
    if (fork) polar_join(c);
    else if ((rc = mp_long_sums(c, a, true))) return rc;
    {
        ProfScope ps(c, "mp_long");
        hipLaunchKernelGGL(k_kforce_sum, ga, b256, 0, c->stream, a);
        if (c->mp.torques && (rc = launch_torques(c, a, gm, 8u, RS_MP_DUDV_L))) return rc;
    }
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

int induction_enqueue(Ctx* c, double* terms) {
    int rc;
    MpArgs a = ind_args(c);
    a.terms = terms;
    const dim3 b256(256);
    const dim3 ga((c->n + 255) / 256), gm((c->nmol + 255) / 256);
    const dim3 gr((unsigned)(((int64_t)c->n * MP_LANES + 255) / 256));
    hipLaunchKernelGGL(k_zero_terms, dim3(1), b256, 0, c->stream, terms);
    const bool fork = polar_fork_begin(c);
    int rc_long = fork ? mp_long_sums(c, a, false) : SMD_OK;
    if (fork) polar_fork_end(c);
    if (rc_long) return rc_long;
    if ((rc = ind_short_enqueue(c, a))) return rc;
    if (fork) polar_join(c);
    else if ((rc = mp_long_sums(c, a, false))) return rc;
    {
        ProfScope ps(c, "ind_long");
        hipLaunchKernelGGL(k_kforce_sum, ga, b256, 0, c->stream, a);
        if (c->indf.torques && (rc = launch_torques(c, a, gm, 8u, RS_IND_DUDV_L))) return rc;
    }
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// The joint pass applies when both classes share the k set and the torque
// switch (and the tuning switch allows it).
bool polar_joint_long(Ctx* c) {
    const Ctx::KSet &m = c->mp_ks, &i = c->ind_ks;
    return c->polar_joint && c->mp_ready && c->indf_ready && m.K == i.K && m.alpha == i.alpha && m.kcut == i.kcut &&
           m.chunks == i.chunks && m.ksplit == i.ksplit && (c->mp.torques != 0) == (c->indf.torques != 0);
}

// Multipole's short-range kernels of the polarizable step with its terms
// cleared first, and (with_sums) the permanent multipoles' structure-factor
// sums: none needs the dipoles, so the step can run them beside the solve.
// Points a's real space at the solve's coupling table when the caller allows it.
static void use_table(Ctx* c, MpArgs& a, bool allowed) {
    if (!allowed || c->mp_prec != SMD_PREC_MIXED) return;
    a.tcnt = c->d_it_cnt;
    a.tj = c->d_it_j;
    a.tstride = c->it_stride;
}

int polar_mp_short_enqueue(Ctx* c, double* mp_t, bool with_sums) {
    MpArgs am = mp_args(c);
    am.terms = mp_t;
    use_table(c, am, c->mp_real_table);
    hipLaunchKernelGGL(k_zero_terms, dim3(1), dim3(256), 0, c->stream, mp_t);
    int rc = mp_short_enqueue(c, am);
    if (rc || !with_sums) return rc;
    // the permanent multipoles' structure-factor sums (the joint pass's other
    // 8 per k follow the solve)
    ProfScope ps(c, "mp_long");
    const dim3 g1((am.K + MP_SF_THREADS - 1) / MP_SF_THREADS, am.chunks);
    if (c->mp_prec == SMD_PREC_MIXED) hipLaunchKernelGGL((k_mp_sfac<float>), g1, dim3(MP_SF_THREADS), 0, c->stream, am);
    else hipLaunchKernelGGL((k_mp_sfac<double>), g1, dim3(MP_SF_THREADS), 0, c->stream, am);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// Both classes after the solve: the joint k-space pass (structure factors with
// the converged dipoles, both classes' per-k terms, the joint force pass) on
// the second stream beside the short-range kernels (Induction's, and
// Multipole's when with_mp_short; otherwise polar_mp_short_enqueue ran them
// and cleared Multipole's terms), then the k-space forces and torques once.
// mp_sums_done: polar_mp_short_enqueue also summed the permanent multipoles'
// structure factors, so only the induced dipoles' 8 sums per k remain.
// The k-space torque's volume-derivative share is booked to Induction's
// terms (the step reports the sum).
int polar_joint_enqueue(Ctx* c, double* mp_t, double* ind_t, bool with_mp_short, bool mp_sums_done) {
    int rc;
    MpArgs am = mp_args(c), ai = ind_args(c);
    am.terms = mp_t;
    ai.terms = ind_t;
    use_table(c, am, c->mp_real_table);
    use_table(c, ai, c->ind_real_table);
    const dim3 b256(256), ga((c->n + 255) / 256), gm((c->nmol + 255) / 256);
    if (with_mp_short) hipLaunchKernelGGL(k_zero_terms, dim3(1), b256, 0, c->stream, mp_t);
    hipLaunchKernelGGL(k_zero_terms, dim3(1), b256, 0, c->stream, ind_t);
    auto long_sums = [&]() -> int {
        ProfScope ps(c, "ind_long");
        const dim3 g1((ai.K + MP_SF_THREADS - 1) / MP_SF_THREADS, ai.chunks);
        const dim3 gk((c->n + WAVE - 1) / WAVE, ai.ksplit), gf((ai.K + 3) / 4);
        const bool mixed = c->mp_prec == SMD_PREC_MIXED;
        if (!mp_sums_done) {
            if (mixed) hipLaunchKernelGGL((k_ind_sfac<float>), g1, dim3(MP_SF_THREADS), 0, c->stream, ai);
            else hipLaunchKernelGGL((k_ind_sfac<double>), g1, dim3(MP_SF_THREADS), 0, c->stream, ai);
            hipLaunchKernelGGL(k_joint_kfinish<false>, gf, b256, 0, c->stream, am, ai);
        } else {
            if (mixed) hipLaunchKernelGGL((k_indp_sfac<float>), g1, dim3(MP_SF_THREADS), 0, c->stream, ai);
            else hipLaunchKernelGGL((k_indp_sfac<double>), g1, dim3(MP_SF_THREADS), 0, c->stream, ai);
            hipLaunchKernelGGL(k_joint_kfinish<true>, gf, b256, 0, c->stream, am, ai);
        }
        int r = red_finish(c, RS_MP_LONG, gf.x, 9, mp_t + M_LONG_E);
        if (!r) r = red_finish(c, RS_IND_LONG, gf.x, 9, ind_t + M_LONG_E);
        if (r) return r;
        if (mixed) hipLaunchKernelGGL((k_joint_kforce<float>), gk, dim3(WAVE), 0, c->stream, ai);
        else hipLaunchKernelGGL((k_joint_kforce<double>), gk, dim3(WAVE), 0, c->stream, ai);
        SMD_HIP(c, hipGetLastError());
        return SMD_OK;
    };
    const bool fork = polar_fork_begin(c);
    const int rc_long = fork ? long_sums() : SMD_OK;
    if (fork) polar_fork_end(c);
    if (rc_long) return rc_long;
    if (with_mp_short && (rc = mp_short_enqueue(c, am))) return rc;
    if ((rc = ind_short_enqueue(c, ai))) return rc;
    if (fork) polar_join(c);
    else if ((rc = long_sums())) return rc;
    {
        ProfScope ps(c, "ind_long");
        hipLaunchKernelGGL(k_kforce_sum, ga, b256, 0, c->stream, ai);
        if (c->indf.torques && (rc = launch_torques(c, ai, gm, 8u, RS_IND_DUDV_L))) return rc;
    }
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// Checks the polarizable step's classes once, outside any capture.
int polar_classes_check(Ctx* c) {
    int rc = mp_check(c);
    if (rc) return rc;
    if ((rc = ind_check(c))) return rc;
    const double rl = c->list_cutoff + c->list_buff;
    if (c->mp.real_cut > rl || c->indf.real_cut > rl)
        return fail(c, SMD_ENOTSUP, "real-space cutoff exceeds the pair-list radius %g", rl);
    return SMD_OK;
}

// Energies of the multipole and induction accumulators of the MD step.
int polar_terms_energies(Ctx* c, const double* mp_t, const double* ind_t, double* e) {
    e[0] = mp_t[M_REAL_E];
    e[1] = mp_t[M_SELF];
    e[2] = mp_t[M_INTRA_E];
    e[3] = mp_t[M_LONG_E];
    e[4] = ind_t[M_REAL_E];
    e[5] = ind_t[M_SELF];
    e[6] = ind_t[M_INTRA_E];
    e[7] = ind_t[M_LONG_E];
    return SMD_OK;
}

static int ind_setup(Ctx* c, const smd_indf_params* p, double kcut, int32_t* n_kvec) {
    if (!(p->alpha > 0) || !(p->real_cut > 0) || !(kcut > 0) || !(p->a_damp >= 0))
        return fail(c, SMD_EINVAL, "bad induction parameters");
    int rc = kset_build(c, c->ind_ks, p->alpha, kcut, IN_NSUM);
    if (rc) return rc;
    c->indf = *p;
    c->indf_ready = true;
    if (n_kvec) *n_kvec = c->ind_ks.K;
    return SMD_OK;
}

}  // namespace smd

using namespace smd;

extern "C" int smd_multipole_setup(smd_ctx* c, const smd_mp_params* p, double kcut, int32_t* n_kvec) {
    if (!c || !p) return SMD_EINVAL;
    return mp_setup(c, p, kcut, n_kvec);
}

static int mp_forces_out(smd_ctx* c, double* f_accum) {
    std::vector<double> f(3 * (size_t)c->n);
    SMD_HIP(c, hipMemcpyAsync(f.data(), c->d_force, f.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < f.size(); ++k) f_accum[k] += f[k];
    return SMD_OK;
}

extern "C" int smd_multipole_compute(smd_ctx* c, int prec, double* f_accum, smd_mp_out* out) {
    if (!c || !out) return SMD_EINVAL;
    if (prec != SMD_PREC_FP64 && prec != SMD_PREC_MIXED) return fail(c, SMD_EINVAL, "bad precision %d", prec);
    c->mp_prec = prec;
    if (f_accum) SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    int rc = multipole_short(c, out);
    if (rc) return rc;
    return f_accum ? mp_forces_out(c, f_accum) : SMD_OK;
}

extern "C" int smd_multipole_compute_long(smd_ctx* c, int prec, double* f_accum, smd_mp_long_out* out) {
    if (!c || !out) return SMD_EINVAL;
    if (prec != SMD_PREC_FP64 && prec != SMD_PREC_MIXED) return fail(c, SMD_EINVAL, "bad precision %d", prec);
    c->mp_prec = prec;
    if (f_accum) SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    int rc = multipole_long(c, out);
    if (rc) return rc;
    return f_accum ? mp_forces_out(c, f_accum) : SMD_OK;
}

extern "C" int smd_induction_setup(smd_ctx* c, const smd_indf_params* p, double kcut, int32_t* n_kvec) {
    if (!c || !p) return SMD_EINVAL;
    return ind_setup(c, p, kcut, n_kvec);
}

extern "C" int smd_induction_compute(smd_ctx* c, int prec, double* f_accum, smd_indf_out* out) {
    if (!c || !out) return SMD_EINVAL;
    if (prec != SMD_PREC_FP64 && prec != SMD_PREC_MIXED) return fail(c, SMD_EINVAL, "bad precision %d", prec);
    c->mp_prec = prec;
    if (f_accum) SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    int rc = induction_short(c, out);
    if (rc) return rc;
    return f_accum ? mp_forces_out(c, f_accum) : SMD_OK;
}

extern "C" int smd_induction_compute_long(smd_ctx* c, int prec, double* f_accum, smd_mp_long_out* out) {
    if (!c || !out) return SMD_EINVAL;
    if (prec != SMD_PREC_FP64 && prec != SMD_PREC_MIXED) return fail(c, SMD_EINVAL, "bad precision %d", prec);
    c->mp_prec = prec;
    if (f_accum) SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    int rc = induction_long(c, out);
    if (rc) return rc;
    return f_accum ? mp_forces_out(c, f_accum) : SMD_OK;
}

// reps launches of one real-space kernel of the polarizable step on the
// current state, between two events on the context stream.  The kernels
// write their forces and sums to scratch (the vdW virtual-site force buffer
// and a spare accumulator), so the step's state is unchanged; the torque
// scratch is rewritten by every step anyway.
extern "C" int smd_polar_time_kernel(smd_ctx* c, const char* name, int reps, double* avg_us) {
    if (!c || !name || reps < 1 || !avg_us) return SMD_EINVAL;
    if (!c->pmd_ready) return fail(c, SMD_ESTATE, "smd_polar_md_init not called");
    const std::string f(name);
    const bool ind = f == "ind_real";
    if (!ind && f != "mp_real") return fail(c, SMD_EINVAL, "kernel '%s' is not mp_real or ind_real", name);
    MpArgs a = ind ? ind_args(c) : mp_args(c);
    a.force = c->d_fvirt;
    a.terms = c->d_polar_terms + 2 * POLAR_TERMS;
    // the mixed induction real space the step runs: over the solve's table
    // when its cutoff is the real-space one
    const bool tab = c->mp_prec == SMD_PREC_MIXED && c->ind_table_built && c->pmd.induce.do_min &&
                     c->pmd.induce.prec == SMD_PREC_MIXED &&
                     c->pmd.induce.cutoff_ind == (ind ? c->indf.real_cut : c->mp.real_cut);
    if (tab) {
        a.tcnt = c->d_it_cnt;
        a.tj = c->d_it_j;
        a.tstride = c->it_stride;
    }
    const dim3 gr((unsigned)(((int64_t)c->n * MP_LANES + 255) / 256)), b256(256);
    const bool mixed = c->mp_prec == SMD_PREC_MIXED;
    if (mixed)
        hipLaunchKernelGGL(k_pack_sites32, dim3((c->n + 255) / 256), b256, 0, c->stream, a, c->d_site32, c->d_pind32, 1);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SMD_HIP(c, hipEventCreate(&e0));
    SMD_HIP(c, hipEventCreate(&e1));
    SMD_HIP(c, hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) {
        if (ind) {
            if (tab) hipLaunchKernelGGL((k_ind_real_tab<MP_LANES>), gr, b256, 0, c->stream, a);
            else if (mixed) hipLaunchKernelGGL((k_ind_real<MP_LANES, float>), gr, b256, 0, c->stream, a);
            else hipLaunchKernelGGL((k_ind_real<MP_LANES, double>), gr, b256, 0, c->stream, a);
        } else {
            if (tab) hipLaunchKernelGGL((k_mp_real_tab<MP_LANES>), gr, b256, 0, c->stream, a);
            else if (mixed) hipLaunchKernelGGL((k_mp_real<MP_LANES, float>), gr, b256, 0, c->stream, a);
            else hipLaunchKernelGGL((k_mp_real<MP_LANES, double>), gr, b256, 0, c->stream, a);
        }
    }
    SMD_HIP(c, hipEventRecord(e1, c->stream));
    SMD_HIP(c, hipEventSynchronize(e1));
    float ms = 0.f;
    SMD_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SMD_HIP(c, hipGetLastError());
    *avg_us = 1e3 * (double)ms / reps;
    return SMD_OK;
}
