// induction.hip -- induced dipoles of polarizable water on gfx950.
//
//   Local2Global::FrameofReference  src/utils/Local2Global.cpp:90-225
//   Induction::InducedDipoles       src/forces/Induction.cpp:181-455
//   Induction::ConjugateGradient    :2401-2467
//
// The reference assembles A = diag(1/alpha) - T as a dense (3N)^2 matrix every
// step and runs CG on it.  Here A is never stored: each product A v recomputes
// the Thole-damped dipole tensor of every pair inside cutoffInd from the
// positions.
//  * Non-excluded pairs come from the full neighbour rows of the pair list.
//    Both of a row's segments are used, so a row holds every partner of its
//    atom.
//  * Excluded (intramolecular) pairs come from the exclusion table.  A couples
//    them too; E0 does not.
//  * The permanent field E0 uses the same rows.
// CG follows the reference's recurrence:
//  * Every reduction is a per-block partial followed by a fixed-order sum, so
//    results do not depend on scheduling.
//  * The scalar of iteration k is read from one partial buffer while the next
//    one is written to the other.
//  * The host launches iterations in batches and stops when the device flag
//    says the residual test passed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "smd_internal.h"

namespace smd {

constexpr int IND_LANES = 16;      // lanes per atom row in the field and T.mu kernels
constexpr int CG_THREADS = 256;    // threads per block of the vector kernels
constexpr int BLK_MOLS = 64;       // molecules per block of the block preconditioner's vector kernels
constexpr int BLK_THREADS = 9 * BLK_MOLS;   // a thread per dipole component (nine waves)
constexpr double SQRT_COULOMB = 11.78708976;   // SQRTCOULOMBCONSTANT, src/NEMD_defs.h:103
constexpr double POLMIN = 0.00000001;          // Induction.cpp:195

struct IndArgs {
    int n, nmol, stride;
    Box box;
    const int32_t* mol_start;
    const double* pos;
    const double* q;            // scaled charges
    const int32_t* nb;
    const int32_t* nb_cnt;
    const int32_t* nb_own;
    const int32_t* excl_off;
    const int32_t* excl;
    const double* polar;
    const double* dipl;         // local-frame dipoles [3n]
    const double* quadl;        // local-frame quadrupoles [6n] xx xy xz yy yz zz
    double* dip;                // lab frame [3n]
    double* quad;               // lab frame, traced / 3 [6n]
    double* e0;                 // [3n]
    double cut2, a;
    int damp;
    // coupling table of the solve: per site the partners inside cutoffInd
    // (list partners, then excluded ones) with 3 lambda5 / r^5 and lambda3 / r^3
    int tstride;
    int32_t* tcnt;              // [n]
    int32_t* tj;                // [n][tstride]
    double2* tab;               // [n][tstride] (a, b): one 16-byte load per entry (FP64 solve)
    // mixed solve: the factors are recomputed in FP32 from hi/lo site records
    // rec32[2i] = (x_hi, y_hi, z_hi, 1/sqrt(alpha)), rec32[2i+1] = (x_lo, y_lo, z_lo, q / sqrt(C))
    int mixed;
    float4* rec32;
    // mixed A v: positions as 32-bit fixed-point fractions of the box (the
    // difference of two wraps to the minimum image) and 1/sqrt(alpha)
    int4* fix;
    float fscale[3];            // L / 2^32 per dimension
    float4* site32;             // mixed solve: (q / sqrt(C), mu), (Q xx xy xz yy), (Q yz zz, alpha, 0) per site
    float Lh[3], Ll[3], invL[3];
};

// ---- frames -------------------------------------------------------------------

struct d3 { double x, y, z; };
__device__ __forceinline__ d3 ld3(const double* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ d3 sub3(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 add3(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 div3(d3 a, double f) { return {a.x / f, a.y / f, a.z / f}; }
__device__ __forceinline__ d3 scl3(double f, d3 a) { return {f * a.x, f * a.y, f * a.z}; }
__device__ __forceinline__ double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double len3(d3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// One thread per water (sites H1, O, H2).
__global__ __launch_bounds__(256) void k_frames(IndArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.nmol) return;
    const int a0 = a.mol_start[m];
    const d3 h1 = ld3(a.pos + 3 * a0), o = ld3(a.pos + 3 * (a0 + 1)), h2 = ld3(a.pos + 3 * (a0 + 2));
    const d3 ex{1.0, 0.0, 0.0}, ey{0.0, 1.0, 0.0};
    d3 p1[3], p2[3];
    // hydrogens: axis to O, second vector x (y when the axis is within 30 deg of x)
    p1[0] = sub3(o, h1);
    p2[0] = fabs(p1[0].x / len3(p1[0])) > 0.8660 ? ey : ex;
    // oxygen: sum of the unit vectors to the hydrogens, second vector to H2
    p1[1] = div3(sub3(h1, o), len3(sub3(h1, o)));
    p2[1] = div3(sub3(h2, o), len3(sub3(h2, o)));
    p1[1] = add3(p1[1], p2[1]);
    p1[2] = sub3(o, h2);
    p2[2] = fabs(p1[2].x / len3(p1[2])) > 0.8660 ? ey : ex;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const d3 w = div3(p1[s], len3(p1[s]));
        const d3 ss = sub3(p2[s], scl3(dot3(p2[s], w), w));
        const d3 u = div3(ss, len3(ss));
        const d3 v{w.y * u.z - u.y * w.z, u.x * w.z - w.x * u.z, w.x * u.y - u.x * w.y};
        const int id = a0 + s;
        const double dx = a.dipl[3 * id], dy = a.dipl[3 * id + 1], dz = a.dipl[3 * id + 2];
        a.dip[3 * id] = u.x * dx + v.x * dy + w.x * dz;
        a.dip[3 * id + 1] = u.y * dx + v.y * dy + w.y * dz;
        a.dip[3 * id + 2] = u.z * dx + v.z * dy + w.z * dz;
        const double* ql = a.quadl + 6 * id;
        const double qxx = ql[0], qxy = ql[1], qxz = ql[2], qyy = ql[3], qyz = ql[4], qzz = ql[5];
        const double R[3][3] = {{u.x, v.x, w.x}, {u.y, v.y, w.y}, {u.z, v.z, w.z}};
        const int ab[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double* ra = R[ab[k][0]];
            const double* rb = R[ab[k][1]];
            const double qk = ra[0] * (rb[0] * qxx + rb[1] * qxy + rb[2] * qxz) +
                              ra[1] * (rb[0] * qxy + rb[1] * qyy + rb[2] * qyz) +
                              ra[2] * (rb[0] * qxz + rb[1] * qyz + rb[2] * qzz);
            a.quad[6 * id + k] = qk / 3;   // traced form (Local2Global.cpp:210-215)
        }
    }
}

// ---- pair geometry --------------------------------------------------------------

// Thole factors lambda3, lambda5, lambda7 and r^-3, r^-5, r^-7 (:296-313).
struct Thole { double l3, l5, l7, r3, r5, r7; };

__device__ __forceinline__ bool thole(const IndArgs& a, double rij2, double r, double pi, double pj, Thole& t) {
    if (!(rij2 <= a.cut2)) return false;
    const double rij_1 = 1.0 / r;
    t.r3 = rij_1 * rij_1 * rij_1;
    t.r5 = t.r3 * rij_1 * rij_1;
    t.r7 = t.r5 * rij_1 * rij_1;
    const double u3 = rij2 * r / (sqrt(pi * pj));
    const double e = exp(-a.a * u3);
    t.l3 = 1.0 - e;
    t.l5 = 1.0 - (1.0 + a.a * u3) * e;
    t.l7 = 1.0 - (1.0 + a.a * u3 + 3.0 * a.a * a.a * u3 * u3 / 5.0) * e;
    return true;
}

// Entry k of row i's partners: the front segment, then the back segment.
__device__ __forceinline__ int row_partner(const IndArgs& a, int i, int own, int k) {
    const int32_t* row = a.nb + (int64_t)i * a.stride;
    return k < own ? row[k] : row[a.stride - 1 - (k - own)];
}

// ---- permanent field E0 --------------------------------------------------------

// Field at i of site j's charge, dipole and quadrupole with r = r_i - r_j: the
// reference's b_i of a pair (:423-425; b_j is the same with the roles
// swapped, :430-432).
__device__ __forceinline__ void field_from(const IndArgs& a, int j, const double r[3], const Thole& t, double f[3]) {
    const double qj = a.q[j] / SQRT_COULOMB;
    const double* dj = a.dip + 3 * j;
    const double* Q = a.quad + 6 * j;
    const double Qxx = Q[0], Qxy = Q[1], Qxz = Q[2], Qyy = Q[3], Qyz = Q[4], Qzz = Q[5];
    double RM[3], QR[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) RM[x] = r[x] * r[0] * dj[0] + r[x] * r[1] * dj[1] + r[x] * r[2] * dj[2];
    QR[0] = Qxx * r[0] + Qxy * r[1] + Qxz * r[2];
    QR[1] = Qxy * r[0] + Qyy * r[1] + Qyz * r[2];
    QR[2] = Qxz * r[0] + Qyz * r[1] + Qzz * r[2];
    const double QI = Qxx + Qyy + Qzz;
    const double QRR = Qxx * r[0] * r[0] + Qxy * r[0] * r[1] + Qxz * r[0] * r[2] + Qxy * r[1] * r[0] + Qyy * r[1] * r[1] +
                       Qyz * r[1] * r[2] + Qxz * r[2] * r[0] + Qyz * r[2] * r[1] + Qzz * r[2] * r[2];
#pragma unroll
    for (int x = 0; x < 3; ++x)
        f[x] += qj * t.l3 * t.r3 * r[x] + 3.0 * t.l5 * t.r5 * RM[x] - t.l3 * t.r3 * dj[x] +
                15.0 * t.l7 * t.r7 * QRR * r[x] - 6.0 * t.l5 * t.r5 * QR[x] - 3.0 * t.l5 * t.r5 * QI * r[x];
}

template <int G>
__global__ __launch_bounds__(256) void k_field(IndArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double f[3] = {0.0, 0.0, 0.0};
    if (i < a.n) {
        const int own = min(a.nb_own[i], a.stride);
        const int cnt = min(a.nb_cnt[i], a.stride);
        const double xi = a.pos[3 * i], yi = a.pos[3 * i + 1], zi = a.pos[3 * i + 2];
        const double pi = a.polar[i] > POLMIN ? a.polar[i] : POLMIN;
        for (int k = sub; k < cnt; k += G) {
            const int j = row_partner(a, i, own, k);
            double dx = xi - a.pos[3 * j], dy = yi - a.pos[3 * j + 1], dz = zi - a.pos[3 * j + 2];
            pbc_near(dx, dy, dz, a.box);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const double pj = a.polar[j] > POLMIN ? a.polar[j] : POLMIN;
            Thole t;
            if (!thole(a, r * r, r, pi, pj, t)) continue;
            const double rr[3] = {dx, dy, dz};
            field_from(a, j, rr, t, f);
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int x = 0; x < 3; ++x) f[x] += __shfl_xor(f[x], o, G);
    if (i < a.n && sub == 0)
        for (int x = 0; x < 3; ++x) a.e0[3 * i + x] = f[x];
}

// ---- A v without A ------------------------------------------------------------

// (A v)_i = v_i / alpha_i - sum_j T_ij v_j over every pair inside cutoffInd
// (Induction.cpp:248-256, :309-361).  With a dot output, each block also
// stores its partial of dot . (A v).
__device__ __forceinline__ void tmu_pair(const IndArgs& a, double dx, double dy, double dz, double pi, int j,
                                         const double* v, double acc[3]) {
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double pj = a.polar[j] > POLMIN ? a.polar[j] : POLMIN;
    Thole t;
    if (!thole(a, r * r, r, pi, pj, t)) return;
    const double rr[3] = {dx, dy, dz};
    const double vj[3] = {v[3 * j], v[3 * j + 1], v[3 * j + 2]};
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        double s = 0.0;
#pragma unroll
        for (int y = 0; y < 3; ++y) {
            double T = a.damp ? t.l5 * 3.0 * t.r5 * rr[x] * rr[y] : 3.0 * t.r5 * rr[x] * rr[y];
            if (x == y) T = a.damp ? T - t.l3 * t.r3 : T - t.r3;
            s += T * vj[y];
        }
        acc[x] += s;
    }
}

// A v recomputing every pair's Thole factors (outside CG, where no coupling
// table has been built).
template <int G>
__global__ __launch_bounds__(256) void k_apply_a(IndArgs a, const double* __restrict__ v, double* __restrict__ out) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double acc[3] = {0.0, 0.0, 0.0};
    double pi = 1.0;
    if (i < a.n) {
        const int own = min(a.nb_own[i], a.stride);
        const int cnt = min(a.nb_cnt[i], a.stride);
        const double xi = a.pos[3 * i], yi = a.pos[3 * i + 1], zi = a.pos[3 * i + 2];
        pi = a.polar[i] > POLMIN ? a.polar[i] : POLMIN;
        for (int k = sub; k < cnt; k += G) {
            const int j = row_partner(a, i, own, k);
            double dx = xi - a.pos[3 * j], dy = yi - a.pos[3 * j + 1], dz = zi - a.pos[3 * j + 2];
            pbc_near(dx, dy, dz, a.box);
            tmu_pair(a, dx, dy, dz, pi, j, v, acc);
        }
        // excluded partners are coupled in A as well (:263-361)
        for (int k = a.excl_off[i] + sub; k < a.excl_off[i + 1]; k += G) {
            const int j = a.excl[k];
            double dx = xi - a.pos[3 * j], dy = yi - a.pos[3 * j + 1], dz = zi - a.pos[3 * j + 2];
            pbc_near(dx, dy, dz, a.box);
            tmu_pair(a, dx, dy, dz, pi, j, v, acc);
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int x = 0; x < 3; ++x) acc[x] += __shfl_xor(acc[x], o, G);
    if (i < a.n && sub == 0)
#pragma unroll
        for (int x = 0; x < 3; ++x) out[3 * i + x] = v[3 * i + x] / pi - acc[x];
}

// Sum of n block partials in a fixed order (lane-strided sums, then a
// shuffle tree), computed by the block's first wave and shared through LDS:
// every block gets the same value, whatever the scheduling.
__device__ __forceinline__ double wave_total(const double* p, int n) {
    // eight loads in flight per lane; the additions keep the order k, k + WAVE, ...
    double s = 0.0;
    int k = threadIdx.x % WAVE;
    for (; k + 7 * WAVE < n; k += 8 * WAVE) {
        double q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = p[k + u * WAVE];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += q[u];
    }
    for (; k < n; k += WAVE) s += p[k];
    return wave_sum(s);
}

__device__ __forceinline__ double block_total(const double* p, int n) {
    __shared__ double tot;
    if (threadIdx.x < WAVE) {
        const double s = wave_total(p, n);
        if (threadIdx.x == 0) tot = s;
    }
    __syncthreads();
    const double v = tot;
    __syncthreads();   // tot may be rewritten by the next call
    return v;
}

// Up to four totals at once, wave w summing buffer w in block_total's order
// (the same values, with one load latency chain instead of NB).
template <int NB>
__device__ __forceinline__ void block_totals(const double* p0, int n0, const double* p1, int n1, const double* p2,
                                             int n2, double (&out)[NB]) {
    static_assert(NB >= 2 && NB <= 3, "two or three totals");
    __shared__ double tot[NB];
    const int w = threadIdx.x / WAVE;
    if (w < NB) {
        const double* p = w == 0 ? p0 : (w == 1 ? p1 : p2);
        const int n = w == 0 ? n0 : (w == 1 ? n1 : n2);
        const double s = wave_total(p, n);
        if (threadIdx.x % WAVE == 0) tot[w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) out[b] = tot[b];
    __syncthreads();
}

__device__ __forceinline__ void block_partial(double v, double* part) {
    __shared__ double red[256 / WAVE];
    v = wave_sum(v);
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / WAVE); ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// Loop-top test of a CG iteration (:2431-2434), the same in every block:
// stop when sqrt(scale * sum of p_dv) < tol (the reference's CG: |v| with
// scale 1; the single-reduction PCG: the RMS of alpha r, also stored in
// res[0]).
struct CgTop {
    int* state;          // [0] done, [2] stop reason; null: plain A v
    const double* p_dv;  // partials of the tested sum (null: the done flag only)
    int nparts;
    double tol;
    double scale = 1.0;
    double* res = nullptr;
    int reason = 1;
    int* hdone = nullptr;   // mapped host memory told of the stop (see top_stop), or null
};

// Coupling table of a solve.  A is fixed while CG runs, so the Thole factors
// of every pair are computed once: each group of G lanes compacts its row's
// partners inside cutoffInd, in list order, with a ballot.
template <int G>
__global__ __launch_bounds__(256) void k_ind_table(IndArgs a) {
    static_assert(WAVE % G == 0, "a lane group must not straddle waves");
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    const int gbase = (threadIdx.x % WAVE) - sub;
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    int count = 0;
    if (i < a.n) {
        const int own = min(a.nb_own[i], a.stride);
        const int cnt = min(a.nb_cnt[i], a.stride);
        const int e0 = a.excl_off[i], e1 = a.excl_off[i + 1];
        const double xi = a.pos[3 * i], yi = a.pos[3 * i + 1], zi = a.pos[3 * i + 2];
        const double pi = a.polar[i] > POLMIN ? a.polar[i] : POLMIN;
        const int total = cnt + (e1 - e0);
        for (int kb = 0; kb < total; kb += G) {
            const int k = kb + sub;
            bool ok = false;
            int j = 0;
            double ca = 0.0, cb = 0.0;
            if (k < total) {
                j = k < cnt ? row_partner(a, i, own, k) : a.excl[e0 + (k - cnt)];
                double dx = xi - a.pos[3 * j], dy = yi - a.pos[3 * j + 1], dz = zi - a.pos[3 * j + 2];
                pbc_near(dx, dy, dz, a.box);
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                const double pj = a.polar[j] > POLMIN ? a.polar[j] : POLMIN;
                Thole t;
                if (thole(a, r * r, r, pi, pj, t)) {
                    ok = true;
                    ca = a.damp ? t.l5 * 3.0 * t.r5 : 3.0 * t.r5;
                    cb = a.damp ? t.l3 * t.r3 : t.r3;
                }
            }
            const uint64_t m = (__ballot(ok) >> gbase) & gmask;
            if (ok) {
                const int at = count + __popcll(m & ((1ull << sub) - 1ull));
                const int64_t e = (int64_t)i * a.tstride + at;
                a.tj[e] = j;
                a.tab[e] = make_double2(ca, cb);
            }
            count += __popcll(m);
        }
        if (sub == 0) a.tcnt[i] = count;
    }
}

// The mixed solve's permanent field and coupling table in one pass over the
// rows: which pairs count is decided in FP64 exactly as the FP64 kernels
// decide it (the truncated field jumps by q/rc^2 at cutoffInd, so one pair
// decided differently would move a dipole by a few percent), the pair math
// is FP32 from FP32 multipole records (site32), four entries per lane in
// flight, FP64 field sums.
// The table holds the partners only: the mixed A v recomputes their Thole
// factors.  The field of a pair is field_from's expression gathered by
// powers of r:
//   f = (q c3 + c5 (3 r.d - 3 tr Q) + 15 c7 r.Q.r) r - c3 d - 6 c5 Q r,
// cn = lambda_n / r^n.  With `shells`, the learned guess's charge shells
// (alpha-less sums q_j r / r^3 of the intermolecular partners inside 0.3,
// 0.4, 0.5, 0.6 nm split by the sign of q_j, 24 floats per site) come out of
// the same pass.
template <int G>
__global__ __launch_bounds__(256) void k_field_table32(IndArgs a, float* __restrict__ shells) {
    static_assert(WAVE % G == 0, "a lane group must not straddle waves");
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    const int gbase = (threadIdx.x % WAVE) - sub;
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    double f[3] = {0.0, 0.0, 0.0};
    float sh[8][3];
#pragma unroll
    for (int q = 0; q < 8; ++q) sh[q][0] = sh[q][1] = sh[q][2] = 0.0f;
    int count = 0;
    if (i < a.n) {
        const int own = min(a.nb_own[i], a.stride);
        const int cnt = min(a.nb_cnt[i], a.stride);
        const int e0 = a.excl_off[i], e1 = a.excl_off[i + 1];
        const double xi = a.pos[3 * i], yi = a.pos[3 * i + 1], zi = a.pos[3 * i + 2];
        const float rs_i = a.rec32[2 * i].w;
        const float damp = (float)a.a;
        const int total = cnt + (e1 - e0);
        for (int kb = 0; kb < total; kb += 4 * G) {
            int jj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = kb + u * G + sub;
                jj[u] = k < total ? (k < cnt ? row_partner(a, i, own, k) : a.excl[e0 + (k - cnt)]) : i;
            }
            double px[4], py[4], pz[4];
            float rsj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                px[u] = a.pos[3 * jj[u]];
                py[u] = a.pos[3 * jj[u] + 1];
                pz[u] = a.pos[3 * jj[u] + 2];
                rsj[u] = a.rec32[2 * jj[u]].w;
            }
            bool ok[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = kb + u * G + sub;
                const int j = jj[u];
                double dx = xi - px[u], dy = yi - py[u], dz = zi - pz[u];
                pbc_near(dx, dy, dz, a.box);
                const double rd64 = sqrt(dx * dx + dy * dy + dz * dz);
                ok[u] = k < total && rd64 * rd64 <= a.cut2;   // the FP64 kernels' test, bit for bit
                if (ok[u] && k < cnt) {
                    // the permanent field (E0 leaves the excluded partners out)
                    const float rx = (float)dx, ry = (float)dy, rz = (float)dz, r2 = (float)(rd64 * rd64);
                    const float4 s0 = a.site32[3 * j], s1 = a.site32[3 * j + 1], s2 = a.site32[3 * j + 2];
                    const float r_1 = __builtin_amdgcn_rsqf(r2);
                    const float x = damp * (r2 * r2 * r_1) * (rs_i * rsj[u]);
                    const float e = __expf(-x);
                    const float r3 = r_1 * r_1 * r_1, r5 = r3 * r_1 * r_1;
                    const float l3 = 1.0f - e, l5 = 1.0f - (1.0f + x) * e;
                    const float l7 = 1.0f - (1.0f + x + 0.6f * x * x) * e;
                    const float c7 = l7 * r5 * r_1 * r_1;
                    const float c3f = l3 * r3, c5f = l5 * r5;
                    const float qj = s0.x, djx = s0.y, djy = s0.z, djz = s0.w;
                    const float Qxx = s1.x, Qxy = s1.y, Qxz = s1.z, Qyy = s1.w, Qyz = s2.x, Qzz = s2.y;
                    const float QRx = Qxx * rx + Qxy * ry + Qxz * rz;
                    const float QRy = Qxy * rx + Qyy * ry + Qyz * rz;
                    const float QRz = Qxz * rx + Qyz * ry + Qzz * rz;
                    const float QRR = rx * QRx + ry * QRy + rz * QRz;
                    const float rd = rx * djx + ry * djy + rz * djz;
                    const float sc = qj * c3f + c5f * (3.0f * rd - 3.0f * (Qxx + Qyy + Qzz)) + 15.0f * c7 * QRR;
                    f[0] += (double)(sc * rx - c3f * djx - 6.0f * c5f * QRx);
                    f[1] += (double)(sc * ry - c3f * djy - 6.0f * c5f * QRy);
                    f[2] += (double)(sc * rz - c3f * djz - 6.0f * c5f * QRz);
                    if (shells && r2 < 0.36f) {
                        const int b = 2 * (r2 < 0.09f ? 0 : r2 < 0.16f ? 1 : r2 < 0.25f ? 2 : 3) + (qj > 0.0f ? 0 : 1);
                        const float w = qj * r3;
#pragma unroll
                        for (int q = 0; q < 8; ++q)
                            if (q == b) {
                                sh[q][0] += w * rx;
                                sh[q][1] += w * ry;
                                sh[q][2] += w * rz;
                            }
                    }
                }
            }
            // ordered compaction: entry k = kb + u G + sub, u outermost
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint64_t m = (__ballot(ok[u]) >> gbase) & gmask;
                if (ok[u]) {
                    const int at = count + __popcll(m & ((1ull << sub) - 1ull));
                    a.tj[(int64_t)i * a.tstride + at] = jj[u];
                }
                count += __popcll(m);
            }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int x = 0; x < 3; ++x) f[x] += __shfl_xor(f[x], o, G);
    if (shells) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int x = 0; x < 3; ++x) sh[q][x] += __shfl_xor(sh[q][x], o, G);
    }
    if (i < a.n && sub == 0) {
        a.tcnt[i] = count;
        for (int x = 0; x < 3; ++x) a.e0[3 * i + x] = f[x];
        if (shells)
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int x = 0; x < 3; ++x) shells[24 * (size_t)i + 3 * q + x] = sh[q][x];
    }
}

__device__ __forceinline__ bool top_stop(const CgTop& top) {
    if (!top.state) return false;
    if (top.state[0]) return true;
    if (!top.p_dv) return false;
    const double res = sqrt(block_total(top.p_dv, top.nparts) * top.scale);
    if (top.res && blockIdx.x == 0 && threadIdx.x == 0) top.res[0] = res;
    if (res < top.tol) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            // block 0 alone writes state[0], so it is the block that sees the stop
            top.state[0] = 1;
            top.state[2] = top.reason;
            if (top.hdone) {
                // the host waits on hdone[0] instead of copying the state back
                top.hdone[1] = top.state[1];
                top.hdone[2] = top.reason;
                reinterpret_cast<double*>(top.hdone + 4)[0] = res;
                __threadfence_system();
                top.hdone[0] = 1;
                __threadfence_system();
            }
        }
        return true;
    }
    return false;
}

// A v from the coupling table: (A v)_i = v_i / alpha_i - sum_j (a_ij (r.v_j) r - b_ij v_j).
template <int G>
__global__ __launch_bounds__(256) void k_apply_table(IndArgs a, const double* __restrict__ v,
                                                     double* __restrict__ out, const double* __restrict__ dotv,
                                                     double* part, CgTop top) {
    if (top_stop(top)) return;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double ax = 0.0, ay = 0.0, az = 0.0;
    double pi = 1.0;
    if (i < a.n) {
        const int cnt = a.tcnt[i];
        const double xi = a.pos[3 * i], yi = a.pos[3 * i + 1], zi = a.pos[3 * i + 2];
        pi = a.polar[i] > POLMIN ? a.polar[i] : POLMIN;
        const int32_t* tj = a.tj + (int64_t)i * a.tstride;
        const double2* tab = a.tab + (int64_t)i * a.tstride;
        auto term = [&](int j, double ca, double cb, double px, double py, double pz, double vx, double vy,
                        double vz) {
            double dx = xi - px, dy = yi - py, dz = zi - pz;
            pbc_near(dx, dy, dz, a.box);
            const double rv = ca * (dx * vx + dy * vy + dz * vz);
            ax += rv * dx - cb * vx;
            ay += rv * dy - cb * vy;
            az += rv * dz - cb * vz;
        };
        // four partners per lane in flight: their table entries, then the
        // positions and v of all four, then the arithmetic in partner order
        int k = sub;
        for (; k + 3 * G < cnt; k += 4 * G) {
            int jj[4];
            double ca[4], cb[4], px[4], py[4], pz[4], vx[4], vy[4], vz[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                jj[u] = tj[k + u * G];
                const double2 ab = tab[k + u * G];
                ca[u] = ab.x;
                cb[u] = ab.y;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = jj[u];
                px[u] = a.pos[3 * j]; py[u] = a.pos[3 * j + 1]; pz[u] = a.pos[3 * j + 2];
                vx[u] = v[3 * j]; vy[u] = v[3 * j + 1]; vz[u] = v[3 * j + 2];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) term(jj[u], ca[u], cb[u], px[u], py[u], pz[u], vx[u], vy[u], vz[u]);
        }
        for (; k < cnt; k += G) {
            const int j = tj[k];
            const double2 ab = tab[k];
            term(j, ab.x, ab.y, a.pos[3 * j], a.pos[3 * j + 1], a.pos[3 * j + 2], v[3 * j], v[3 * j + 1],
                 v[3 * j + 2]);
        }
    }
    double acc[3] = {ax, ay, az};
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int x = 0; x < 3; ++x) acc[x] += __shfl_xor(acc[x], o, G);
    double d = 0.0;
    if (i < a.n && sub == 0) {
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double z = v[3 * i + x] / pi - acc[x];
            out[3 * i + x] = z;
            if (dotv) d += dotv[3 * i + x] * z;
        }
    }
    if (part) block_partial(d, part);   // v.Av of the CG step
}

// Hi/lo FP32 site records of the mixed solve: x_hi = float(x), x_lo =
// float(x - x_hi), 1/sqrt(alpha) (alpha clamped at POLMIN) in the first
// record's w and q / sqrt(Coulomb constant) in the second's (the learned
// guess's descriptors read it).
__global__ __launch_bounds__(256) void k_pack_ind32(IndArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double x = a.pos[3 * i], y = a.pos[3 * i + 1], z = a.pos[3 * i + 2];
    const float xh = (float)x, yh = (float)y, zh = (float)z;
    const double p = a.polar[i] > POLMIN ? a.polar[i] : POLMIN;
    a.rec32[2 * i] = make_float4(xh, yh, zh, (float)(1.0 / sqrt(p)));
    a.rec32[2 * i + 1] = make_float4((float)(x - (double)xh), (float)(y - (double)yh), (float)(z - (double)zh),
                                     (float)(a.q[i] / SQRT_COULOMB));
    {
        // fractions of the box in [0, 1) as unsigned 32-bit fixed point
        unsigned u[3];
        const double xs[3] = {x, y, z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double f = xs[k] * a.box.inv[k];
            f -= floor(f);
            u[k] = (unsigned)((unsigned long long)(f * 4294967296.0) & 0xffffffffull);
        }
        a.fix[i] = make_int4((int)u[0], (int)u[1], (int)u[2], __float_as_int((float)(1.0 / sqrt(p))));
    }
    const double* d = a.dip + 3 * i;
    const double* Q = a.quad + 6 * i;
    a.site32[3 * i] = make_float4((float)(a.q[i] / SQRT_COULOMB), (float)d[0], (float)d[1], (float)d[2]);
    a.site32[3 * i + 1] = make_float4((float)Q[0], (float)Q[1], (float)Q[2], (float)Q[3]);
    a.site32[3 * i + 2] = make_float4((float)Q[4], (float)Q[5], (float)a.polar[i], 0.0f);
}

// A v of the mixed solve over the coupling table's partners: each pair's
// Thole factors recomputed in FP32 from the hi/lo records (the displacement
// is (x_hi_j - x_hi_i - n L_hi) + (x_lo_j - x_lo_i - n L_lo), whose first
// difference is exact once the image is taken), an FP32 copy of v (float4
// per site), FP64 accumulation; v_i / alpha_i and the output in FP64.  The
// table stores 4 bytes per entry instead of 12.
template <int G>
__global__ __launch_bounds__(256) void k_apply_pos32(IndArgs a, const float4* __restrict__ v32,
                                                     const double* __restrict__ v, double* __restrict__ out,
                                                     double* part, CgTop top) {
    if (top_stop(top)) return;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double ax = 0.0, ay = 0.0, az = 0.0;
    double pi = 1.0;
    if (i < a.n) {
        const int cnt = a.tcnt[i];
        pi = a.polar[i] > POLMIN ? a.polar[i] : POLMIN;
        const int4 fi = a.fix[i];
        const float rs_i = __int_as_float(fi.w);
        const float damp = (float)a.a;
        const int32_t* tj = a.tj + (int64_t)i * a.tstride;
        auto term = [&](int4 fj, float4 vj) {
            // 32-bit differences wrap to the minimum image; exact, then one rounding
            const float dx = (float)(int)((unsigned)fj.x - (unsigned)fi.x) * a.fscale[0];
            const float dy = (float)(int)((unsigned)fj.y - (unsigned)fi.y) * a.fscale[1];
            const float dz = (float)(int)((unsigned)fj.z - (unsigned)fi.z) * a.fscale[2];
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float r_1 = __builtin_amdgcn_rsqf(r2);
            const float r3 = r_1 * r_1 * r_1, r5 = r3 * r_1 * r_1;
            float ca = 3.0f * r5, cb = r3;
            if (a.damp) {
                const float x = damp * (r2 * r2 * r_1) * (rs_i * __int_as_float(fj.w));
                const float e = __expf(-x);
                cb = (1.0f - e) * r3;
                ca = 3.0f * (1.0f - (1.0f + x) * e) * r5;
            }
            const float rv = ca * (dx * vj.x + dy * vj.y + dz * vj.z);
            ax += (double)(rv * dx - cb * vj.x);
            ay += (double)(rv * dy - cb * vj.y);
            az += (double)(rv * dz - cb * vj.z);
        };
        int k = sub;
        for (; k + 3 * G < cnt; k += 4 * G) {
            int jj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) jj[u] = tj[k + u * G];
            int4 fj[4];
            float4 vj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                fj[u] = a.fix[jj[u]];
                vj[u] = v32[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) term(fj[u], vj[u]);
        }
        for (; k < cnt; k += G) {
            const int j = tj[k];
            term(a.fix[j], v32[j]);
        }
    }
    double acc[3] = {ax, ay, az};
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int x = 0; x < 3; ++x) acc[x] += __shfl_xor(acc[x], o, G);
    double d = 0.0;
    if (i < a.n && sub == 0) {
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double z = v[3 * i + x] / pi - acc[x];
            out[3 * i + x] = z;
            d += v[3 * i + x] * z;
        }
    }
    if (part) block_partial(d, part);   // p.Ap of the CG step
}

// k_apply_table with `lanes` lanes per row (8, 16 or 32).
static void launch_apply(Ctx* c, int lanes, const IndArgs& a, const double* v, double* out, const double* dotv,
                         double* part, CgTop top) {
    const dim3 grid((unsigned)(((int64_t)c->n * lanes + 255) / 256)), b256(256);
    if (lanes == 8) hipLaunchKernelGGL((k_apply_table<8>), grid, b256, 0, c->stream, a, v, out, dotv, part, top);
    else if (lanes == 32) hipLaunchKernelGGL((k_apply_table<32>), grid, b256, 0, c->stream, a, v, out, dotv, part, top);
    else hipLaunchKernelGGL((k_apply_table<16>), grid, b256, 0, c->stream, a, v, out, dotv, part, top);
}

static void launch_apply32(Ctx* c, int lanes, const IndArgs& a, const float4* v32, const double* v, double* out,
                           double* part, CgTop top) {
    const dim3 grid((unsigned)(((int64_t)c->n * lanes + 255) / 256)), b256(256);
    if (lanes == 8) hipLaunchKernelGGL((k_apply_pos32<8>), grid, b256, 0, c->stream, a, v32, v, out, part, top);
    else if (lanes == 32) hipLaunchKernelGGL((k_apply_pos32<32>), grid, b256, 0, c->stream, a, v32, v, out, part, top);
    else hipLaunchKernelGGL((k_apply_pos32<16>), grid, b256, 0, c->stream, a, v32, v, out, part, top);
}

// ---- CG -------------------------------------------------------------------------

// Buffers of block partials, each [max_parts]: VZ (v.Av), DV (v.v), and
// the ping-pong pair RR[0], RR[1] holding r.r of the current and next
// iteration (iteration it reads RR[it & 1]).  The iteration count is kept
// on the device, so a batch of iterations is the same sequence of launches
// whatever iteration it starts at (an even batch keeps the ping-pong phase).
struct CgArgs {
    int dim;          // 3n
    int nparts_mv;    // blocks of the A v kernel
    int nparts_vec;   // blocks of the vector kernels
    double tol;
    double* x;
    double* r;
    double* v;        // CG direction v (PCG: p)
    double* z;        // A v
    const double* e0;
    const double* polar;
    double* p_vz;
    double* p_dv;     // CG: v.v; PCG: |alpha r|^2
    double* p_rr0;
    double* p_rr1;
    int* state;       // [0] done, [1] iterations, [2] stop reason
    double* scal;     // [0] last residual norm
    float4* v32;      // mixed solve: FP32 copy of the direction (else null)
    double inv_npol;  // 1 / (polarizable sites), for the RMS dipole residual
    int nmol;         // block preconditioner: 3-site molecules, sites 3m..3m+2
    const double* blk;  // [nmol][45]: inverse of each molecule's 9x9 block of A, lower triangle (else null)
    double* u;        // block preconditioner: M^-1 r (A applies to it; z holds A u)
    double* s;        // block preconditioner: the recurrence for A p
    int* hdone;       // block preconditioner: mapped host flag of the stop (or null)
};

// x = guess (1 as the reference, or the caller's), then after z = A x:
// r = E0 - z, v = r; partials of r.r (c) and v.v (dotV).
__global__ __launch_bounds__(CG_THREADS) void k_cg_start(CgArgs g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double rr = 0.0;
    if (k < g.dim) {
        const double r = g.e0[k] - g.z[k];
        g.r[k] = r;
        g.v[k] = r;
        rr = r * r;
    }
    block_partial(rr, g.p_rr0);
    if (blockIdx.x == 0 && threadIdx.x == 0) { g.state[0] = 0; g.state[1] = 0; g.state[2] = 0; }
    // dotV equals c at the start: the same partials
    if (threadIdx.x == 0) g.p_dv[blockIdx.x] = g.p_rr0[blockIdx.x];
}

// t = c / v.Av; x += t v; r -= t z; partials of d = r.r (:2446-2454).
template <int PHASE>
__global__ __launch_bounds__(CG_THREADS) void k_cg_step(CgArgs g) {
    if (g.state[0]) return;
    const double* rr_c = PHASE ? g.p_rr1 : g.p_rr0;
    double* rr_d = PHASE ? g.p_rr0 : g.p_rr1;
    double tot[2];   // v.Av, c
    block_totals<2>(g.p_vz, g.nparts_mv, rr_c, g.nparts_vec, nullptr, 0, tot);
    const double c = tot[1];
    const double t = c / tot[0];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (k < g.dim) {
        g.x[k] += t * g.v[k];
        const double r = g.r[k] - t * g.z[k];
        g.r[k] = r;
        d = r * r;
    }
    block_partial(d, rr_d);
    if (blockIdx.x == 0 && threadIdx.x == 0) g.state[1] += 1;
}

// Stop when |r| < tol (:2456-2457); else v = r + (d/c) v and partials of v.v.
template <int PHASE>
__global__ __launch_bounds__(CG_THREADS) void k_cg_dir(CgArgs g) {
    if (g.state[0]) return;
    const double* rr_c = PHASE ? g.p_rr1 : g.p_rr0;
    const double* rr_d = PHASE ? g.p_rr0 : g.p_rr1;
    double tot[2];   // d, c
    block_totals<2>(rr_d, g.nparts_vec, rr_c, g.nparts_vec, nullptr, 0, tot);
    const double d = tot[0], c = tot[1];
    if (blockIdx.x == 0 && threadIdx.x == 0) g.scal[0] = sqrt(d);
    if (sqrt(d) < g.tol) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { g.state[0] = 1; g.state[2] = 2; }
        return;
    }
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double dv = 0.0;
    if (k < g.dim) {
        const double v = g.r[k] + (d / c) * g.v[k];
        g.v[k] = v;
        dv = v * v;
    }
    block_partial(dv, g.p_dv);
}

// ---- preconditioned CG ------------------------------------------------------------
// Jacobi preconditioner M^-1 = diag(alpha): the preconditioned residual
// alpha_i r_i is the dipole step a Jacobi iteration would take, and its RMS
// over the polarizable sites (e nm) is the stopping test.  Same recurrence
// structure as the reference's CG: fixed-order partial sums, iteration count
// on the device, the scalars of iteration k read from one partial buffer
// while the next is written to the other.

__device__ __forceinline__ double site_alpha(const double* polar, int k) {
    const double p = polar[k / 3];
    return p > POLMIN ? p : POLMIN;
}

__device__ __forceinline__ void store_dir32(const CgArgs& g, int k, double v) {
    // the mixed solve's FP32 copy of the direction, one float4 per site
    if (g.v32) reinterpret_cast<float*>(g.v32)[4 * (k / 3) + k % 3] = (float)v;
}

// After z = A x: r = E0 - z, p = alpha r; partials of r.alpha r and |alpha r|^2.
__global__ __launch_bounds__(CG_THREADS) void k_pcg_start(CgArgs g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0.0, zz = 0.0;
    if (k < g.dim) {
        const double r = g.e0[k] - g.z[k];
        const double zr = site_alpha(g.polar, k) * r;
        g.r[k] = r;
        g.v[k] = zr;
        store_dir32(g, k, zr);
        rz = r * zr;
        zz = zr * zr;
    }
    block_partial(rz, g.p_rr0);
    __syncthreads();
    block_partial(zz, g.p_dv);
    if (blockIdx.x == 0 && threadIdx.x == 0) { g.state[0] = 0; g.state[1] = 0; g.state[2] = 0; }
}

// t = rz / p.Ap; x += t p; r -= t Ap; partials of the new r.alpha r and |alpha r|^2.
template <int PHASE>
__global__ __launch_bounds__(CG_THREADS) void k_pcg_step(CgArgs g) {
    if (g.state[0]) return;
    const double* rz_c = PHASE ? g.p_rr1 : g.p_rr0;
    double* rz_d = PHASE ? g.p_rr0 : g.p_rr1;
    double tot[2];   // p.Ap, r.alpha r
    block_totals<2>(g.p_vz, g.nparts_mv, rz_c, g.nparts_vec, nullptr, 0, tot);
    const double t = tot[1] / tot[0];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double rzn = 0.0, zz = 0.0;
    if (k < g.dim) {
        g.x[k] += t * g.v[k];
        const double r = g.r[k] - t * g.z[k];
        g.r[k] = r;
        const double zr = site_alpha(g.polar, k) * r;
        rzn = r * zr;
        zz = zr * zr;
    }
    block_partial(rzn, rz_d);
    __syncthreads();
    block_partial(zz, g.p_dv);
    if (blockIdx.x == 0 && threadIdx.x == 0) g.state[1] += 1;
}

// Stop when the RMS dipole residual is below tol; else p = alpha r + beta p.
template <int PHASE>
__global__ __launch_bounds__(CG_THREADS) void k_pcg_dir(CgArgs g) {
    if (g.state[0]) return;
    const double* rz_c = PHASE ? g.p_rr1 : g.p_rr0;
    const double* rz_d = PHASE ? g.p_rr0 : g.p_rr1;
    double tot[3];   // |alpha r|^2, new r.alpha r, old r.alpha r
    block_totals<3>(g.p_dv, g.nparts_vec, rz_d, g.nparts_vec, rz_c, g.nparts_vec, tot);
    const double res = sqrt(tot[0] * g.inv_npol);
    if (blockIdx.x == 0 && threadIdx.x == 0) g.scal[0] = res;
    if (res < g.tol) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { g.state[0] = 1; g.state[2] = 2; }
        return;
    }
    const double beta = tot[1] / tot[2];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < g.dim) {
        const double v = site_alpha(g.polar, k) * g.r[k] + beta * g.v[k];
        g.v[k] = v;
        store_dir32(g, k, v);
    }
}

// ---- block-Jacobi preconditioner ---------------------------------------------------
// M^-1 = the inverse of each molecule's 9x9 diagonal block of A (its three
// sites' 1/alpha and the intramolecular Thole tensors, which A couples),
// formed once per solve from the block's Cholesky factor (M^-1 = L^-T L^-1)
// and held as its lower triangle (45 doubles, row-major).  The vector
// kernels run a thread per dipole component, BLK_MOLS molecules per block:
// each component's z = M^-1 r is one row of the block times the molecule's
// r, staged in LDS.  The stopping test stays the RMS of alpha r.

__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }
__device__ __forceinline__ int stri(int i, int j) { return i >= j ? tri(i, j) : tri(j, i); }

// Assembles, factors and inverts each molecule's block: thread per molecule.
__global__ __launch_bounds__(256) void k_block_pc(IndArgs a, double* __restrict__ blk, int nmol) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    double B[45];
#pragma unroll
    for (int k = 0; k < 45; ++k) B[k] = 0.0;
    double pl[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const double p = a.polar[3 * m + s];
        pl[s] = p > POLMIN ? p : POLMIN;
#pragma unroll
        for (int x = 0; x < 3; ++x) B[tri(3 * s + x, 3 * s + x)] = 1.0 / pl[s];
    }
    // A_st = -T_st for the pairs s > t of the molecule (inside cutoffInd, as A has them)
#pragma unroll
    for (int s = 1; s < 3; ++s)
#pragma unroll
        for (int t = 0; t < s; ++t) {
            const int i = 3 * m + s, j = 3 * m + t;
            double dx = a.pos[3 * i] - a.pos[3 * j], dy = a.pos[3 * i + 1] - a.pos[3 * j + 1];
            double dz = a.pos[3 * i + 2] - a.pos[3 * j + 2];
            pbc_near(dx, dy, dz, a.box);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            Thole th;
            if (!thole(a, r * r, r, pl[s], pl[t], th)) continue;
            const double ca = a.damp ? th.l5 * 3.0 * th.r5 : 3.0 * th.r5;
            const double cb = a.damp ? th.l3 * th.r3 : th.r3;
            const double d[3] = {dx, dy, dz};
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int y = 0; y < 3; ++y) B[tri(3 * s + x, 3 * t + y)] = -(ca * d[x] * d[y] - (x == y ? cb : 0.0));
        }
    // Cholesky B = L L^T in place (B is a principal block of the positive-definite A)
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        double dj = B[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) dj -= B[tri(j, k)] * B[tri(j, k)];
        dj = sqrt(dj);
        B[tri(j, j)] = dj;
        const double inv = 1.0 / dj;
#pragma unroll
        for (int i = j + 1; i < 9; ++i) {
            double v = B[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) v -= B[tri(i, k)] * B[tri(j, k)];
            B[tri(i, j)] = v * inv;
        }
    }
    // Li = L^-1 (lower triangular), then M^-1 = Li^T Li
    double Li[45];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        Li[tri(j, j)] = 1.0 / B[tri(j, j)];
#pragma unroll
        for (int i = j + 1; i < 9; ++i) {
            double v = 0.0;
#pragma unroll
            for (int k = j; k < i; ++k) v -= B[tri(i, k)] * Li[tri(k, j)];
            Li[tri(i, j)] = v / B[tri(i, i)];
        }
    }
    double* out = blk + 45 * (size_t)m;
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double v = 0.0;
#pragma unroll
            for (int k = i; k < 9; ++k) v += Li[tri(k, i)] * Li[tri(k, j)];
            out[tri(i, j)] = v;
        }
}

// Block totals of the vector kernels (BLK_THREADS threads): wave sums, then
// the waves in order by one thread.
__device__ __forceinline__ void blk_partial(double v, double* part) {
    __shared__ double red[BLK_THREADS / WAVE];
    v = wave_sum(v);
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < BLK_THREADS / WAVE; ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// This thread's component: molecule m (invalid past nmol), component c, the
// molecule's slot in the block.
struct BlkSlot { int m, c, ml, k; bool ok; };
__device__ __forceinline__ BlkSlot blk_slot(const CgArgs& g) {
    BlkSlot b;
    b.ml = threadIdx.x / 9;
    b.c = threadIdx.x % 9;
    b.m = blockIdx.x * BLK_MOLS + b.ml;
    b.k = 9 * b.m + b.c;
    b.ok = b.m < g.nmol;
    return b;
}

// z_c = (M^-1 r)_c for this thread's component: rs holds the block's r.
__device__ __forceinline__ double blk_apply(const CgArgs& g, const BlkSlot& b, const double (*rs)[9]) {
    const double* Mi = g.blk + 45 * (size_t)b.m;
    double z = 0.0;
#pragma unroll
    for (int j = 0; j < 9; ++j) z += Mi[stri(b.c, j)] * rs[b.ml][j];
    return z;
}

// Two or three totals by the whole block, the same in every block: thread t
// adds entries t, t + BLK_THREADS, ... (all loads issued before the adds),
// then the wave sums, then the waves in order.
template <int NB>
__device__ __forceinline__ void blk_totals(const double* p0, int n0, const double* p1, int n1, const double* p2,
                                           int n2, double (&out)[NB]) {
    static_assert(NB >= 2 && NB <= 3, "two or three totals");
    constexpr int PER = 4;   // entries per thread and buffer in one pass of loads
    __shared__ double red[NB][BLK_THREADS / WAVE];
    const double* ps[3] = {p0, p1, p2};
    const int ns[3] = {n0, n1, NB == 3 ? n2 : 0};
    double s[NB];
#pragma unroll
    for (int x = 0; x < NB; ++x) s[x] = 0.0;
    const int nmax = max(n0, max(n1, NB == 3 ? n2 : 0));
    for (int base = 0; base < nmax; base += PER * BLK_THREADS) {
        double v[NB][PER];
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int k = base + u * BLK_THREADS + (int)threadIdx.x;
                v[x][u] = k < ns[x] ? ps[x][k] : 0.0;
            }
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int u = 0; u < PER; ++u) s[x] += v[x][u];
    }
#pragma unroll
    for (int x = 0; x < NB; ++x) {
        const double w = wave_sum(s[x]);
        if (threadIdx.x % WAVE == 0) red[x][threadIdx.x / WAVE] = w;
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < NB; ++x) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < BLK_THREADS / WAVE; ++w) t += red[x][w];
        out[x] = t;
    }
}

// The block-preconditioned CG in the single-reduction form of Chronopoulos
// and Gear (one A product and one vector kernel per iteration; in exact
// arithmetic the iterates of the usual PCG):
//   u = M^-1 r, w = A u, gamma = r.u, delta = w.u,
//   beta = gamma / gamma_prev, alpha = gamma / (delta - beta gamma / alpha_prev),
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s.
// The A product of u runs the stopping test first (the RMS of alpha r from
// the vector kernel's partials).  gamma_prev and alpha_prev pass between
// iterations in scal[2..5] (written by one thread, read from the other half
// of the ping-pong); the first iteration starts from gamma_prev = inf,
// p = s = 0.

// After z = A x: r = E0 - A x, u = M^-1 r, p = s = 0; partials of r.u and
// |alpha r|^2.
__global__ __launch_bounds__(BLK_THREADS) void k_pcg_start_blk(CgArgs g) {
    __shared__ double rs[BLK_MOLS][9];
    const BlkSlot b = blk_slot(g);
    double ru = 0.0, zz = 0.0, r = 0.0;
    if (b.ok) {
        r = g.e0[b.k] - g.z[b.k];
        g.r[b.k] = r;
        g.v[b.k] = 0.0;
        g.s[b.k] = 0.0;
        const double ar = site_alpha(g.polar, b.k) * r;
        zz = ar * ar;
        rs[b.ml][b.c] = r;
    }
    __syncthreads();
    if (b.ok) {
        const double u = blk_apply(g, b, rs);
        g.u[b.k] = u;
        store_dir32(g, b.k, u);
        ru = r * u;
    }
    blk_partial(ru, g.p_rr0);
    __syncthreads();
    blk_partial(zz, g.p_dv);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g.state[0] = 0; g.state[1] = 0; g.state[2] = 0;
        g.scal[2] = __longlong_as_double(0x7ff0000000000000ll);   // gamma_prev = +inf: beta = 0
        g.scal[3] = 1.0;
    }
}

// One iteration after w = A u (in z) and its partials of w.u.
template <int PHASE>
__global__ __launch_bounds__(BLK_THREADS) void k_pcg_step_blk(CgArgs g) {
    if (g.state[0]) return;
    __shared__ double rs[BLK_MOLS][9];
    const double* ru_c = PHASE ? g.p_rr1 : g.p_rr0;
    double* ru_d = PHASE ? g.p_rr0 : g.p_rr1;
    double tot[2];   // gamma = r.u, delta = w.u
    blk_totals<2>(ru_c, g.nparts_vec, g.p_vz, g.nparts_mv, nullptr, 0, tot);
    const double gam = tot[0], del = tot[1];
    const double beta = gam / g.scal[2 + 2 * PHASE];
    const double alpha = gam / (del - beta * gam / g.scal[3 + 2 * PHASE]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g.scal[2 + 2 * (1 - PHASE)] = gam;
        g.scal[3 + 2 * (1 - PHASE)] = alpha;
        g.state[1] += 1;
    }
    const BlkSlot b = blk_slot(g);
    double run = 0.0, zz = 0.0, r = 0.0;
    if (b.ok) {
        const double pk = g.u[b.k] + beta * g.v[b.k];
        const double sk = g.z[b.k] + beta * g.s[b.k];
        g.v[b.k] = pk;
        g.s[b.k] = sk;
        g.x[b.k] += alpha * pk;
        r = g.r[b.k] - alpha * sk;
        g.r[b.k] = r;
        const double ar = site_alpha(g.polar, b.k) * r;
        zz = ar * ar;
        rs[b.ml][b.c] = r;
    }
    __syncthreads();
    if (b.ok) {
        const double u = blk_apply(g, b, rs);
        g.u[b.k] = u;
        store_dir32(g, b.k, u);
        run = r * u;
    }
    blk_partial(run, ru_d);
    __syncthreads();
    blk_partial(zz, g.p_dv);
}

// FP32 copy of x (float4 per site) for the mixed solve's first A x.
__global__ __launch_bounds__(256) void k_to_f32(int dim, const double* __restrict__ x, float* __restrict__ x32) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < dim) x32[4 * (k / 3) + k % 3] = (float)x[k];
}

// Without minimisation: Pi = alpha E0 (:446-453).
__global__ __launch_bounds__(256) void k_direct(int dim, const double* __restrict__ polar, const double* __restrict__ e0,
                                                double* __restrict__ x, float* __restrict__ x32) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dim) return;
    const double v = polar[k / 3] * e0[k];
    x[k] = v;
    if (x32) x32[4 * (k / 3) + k % 3] = (float)v;   // the mixed solve's FP32 copy (float4 per site)
}

// x += alpha (E0 - E0_prev); E0_prev = E0 (the solve's guess 4).
__global__ __launch_bounds__(256) void k_extrapolate(int dim, const double* __restrict__ polar,
                                                     const double* __restrict__ e0, double* __restrict__ e0_prev,
                                                     double* __restrict__ x) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dim) return;
    const double e = e0[k];
    x[k] += polar[k / 3] * (e - e0_prev[k]);
    e0_prev[k] = e;
}

__global__ __launch_bounds__(256) void k_fill(int dim, double val, double* __restrict__ x) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < dim) x[k] = val;
}

static IndArgs ind_args(Ctx* c, const smd_induce_params* p) {
    IndArgs a{};
    a.n = c->n;
    a.nmol = c->nmol;
    a.stride = c->stride;
    a.box = c->box;
    a.mol_start = c->d_mol_start;
    a.pos = c->d_pos;
    a.q = c->d_q;
    a.nb = c->d_nb;
    a.nb_cnt = c->d_nb_cnt;
    a.nb_own = c->d_nb_own;
    a.excl_off = c->d_excl_off;
    a.excl = c->d_excl;
    a.polar = c->d_polar;
    a.dipl = c->d_dipl;
    a.quadl = c->d_quadl;
    a.dip = c->d_dipg;
    a.quad = c->d_quadg;
    a.e0 = c->d_e0;
    a.cut2 = p->cutoff_ind * p->cutoff_ind;
    a.a = p->a_damp;
    a.damp = p->damp_tij;
    a.tstride = c->it_stride;
    a.tcnt = c->d_it_cnt;
    a.tj = c->d_it_j;
    a.tab = reinterpret_cast<double2*>(c->d_it_a);
    a.mixed = p->prec == SMD_PREC_MIXED ? 1 : 0;
    a.rec32 = c->d_ind_rec32;
    a.fix = c->d_ind_fix;
    for (int d = 0; d < 3; ++d) a.fscale[d] = (float)(c->box.L[d] / 4294967296.0);
    a.site32 = c->d_site32_ind;   // its own copy: Multipole's kernels may pack d_site32 alongside
    for (int d = 0; d < 3; ++d) {
        const double L = c->box.L[d];
        a.Lh[d] = (float)L;
        a.Ll[d] = (float)(L - (double)a.Lh[d]);
        a.invL[d] = (float)(1.0 / L);
    }
    return a;
}

// Room for every list partner and every excluded partner of each site.
static int table_alloc(Ctx* c) {
    int max_excl = 0;
    for (int i = 0; i < c->n; ++i) max_excl = std::max(max_excl, c->h_excl_off[i + 1] - c->h_excl_off[i]);
    const int stride = c->stride + max_excl;
    const int64_t need = (int64_t)stride * c->n;
    if (need > c->it_cap) {
        SMD_HIP(c, hipStreamSynchronize(c->stream));
        void* ps[] = {c->d_it_cnt, c->d_it_j, c->d_it_a};
        for (void* q : ps)
            if (q) (void)hipFree(q);
        c->d_it_cnt = nullptr; c->d_it_j = nullptr; c->d_it_a = nullptr;
        c->it_cap = 0;
        SMD_HIP(c, hipMalloc(&c->d_it_cnt, sizeof(int32_t) * c->n));
        SMD_HIP(c, hipMalloc(&c->d_it_j, sizeof(int32_t) * need));
        SMD_HIP(c, hipMalloc(&c->d_it_a, sizeof(double2) * need));
        c->it_cap = need;
    }
    c->it_stride = stride;
    return SMD_OK;
}

// Lab-frame multipoles of every site from the local ones (k_frames).
int frames_run(Ctx* c) {
    smd_induce_params p{};
    p.cutoff_ind = 1.0;
    IndArgs a = ind_args(c, &p);
    hipLaunchKernelGGL(k_frames, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, a);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// x = alpha E0 and z = A x for the permanent field of the last solve.
int induce_direct_apply(Ctx* c, const smd_induce_params* p, double* x, double* z) {
    IndArgs a = ind_args(c, p);
    const int dim = 3 * c->n;
    const dim3 grid_rows((unsigned)(((int64_t)c->n * IND_LANES + 255) / 256));
    hipLaunchKernelGGL(k_direct, dim3((dim + 255) / 256), dim3(256), 0, c->stream, dim, c->d_polar, c->d_e0, x,
                       (float*)nullptr);
    hipLaunchKernelGGL((k_apply_a<IND_LANES>), grid_rows, dim3(256), 0, c->stream, a, x, z);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

static int induce_check(Ctx* c, const smd_induce_params* p) {
    if (!c->multipoles_set) return fail(c, SMD_ESTATE, "multipoles not set");
    if (!c->list_valid) return fail(c, SMD_ESTATE, "no pair list");
    if (!(p->cutoff_ind > 0)) return fail(c, SMD_EINVAL, "cutoff_ind must be positive");
    if (p->cutoff_ind > c->list_cutoff)
        return fail(c, SMD_ENOTSUP, "cutoff_ind %g exceeds the pair-list cutoff %g", p->cutoff_ind, c->list_cutoff);
    if (p->prec != SMD_PREC_FP64 && p->prec != SMD_PREC_MIXED) return fail(c, SMD_EINVAL, "bad precision %d", p->prec);
    if (p->guess < 0 || p->guess > 4) return fail(c, SMD_EINVAL, "guess %d is not 0..4", p->guess);
    if (p->precond < 0 || p->precond > 2) return fail(c, SMD_EINVAL, "precond %d is not 0, 1 or 2", p->precond);
    if (p->tol_rms != (p->precond ? 1 : 0))
        return fail(c, SMD_ENOTSUP, "the RMS dipole tolerance goes with the preconditioned solve, |r| with the reference's");
    if (p->prec == SMD_PREC_MIXED && (!p->precond || !p->do_min))
        return fail(c, SMD_ENOTSUP, "the mixed-precision solve is the preconditioned CG");
    for (int m = 0; m < c->nmol; ++m)
        if (c->h_mol_start[m + 1] - c->h_mol_start[m] != 3 || c->h_mol_start[m] != 3 * m)
            return fail(c, SMD_ENOTSUP, "induced dipoles are implemented for 3-site water");
    return SMD_OK;
}

static CgArgs cg_args(Ctx* c, const smd_induce_params* p) {
    const int n = c->n, dim = 3 * n;
    CgArgs g{};
    g.dim = dim;
    g.tol = p->tol > 0 ? p->tol : 1.0e-14;   // ConjugateGradient's tolerance (:2404)
    g.x = c->d_mu;
    g.r = c->d_cg + 0 * (size_t)dim;
    g.v = c->d_cg + 1 * (size_t)dim;
    g.z = c->d_cg + 2 * (size_t)dim;
    g.u = c->d_cg + 3 * (size_t)dim;
    g.s = c->d_cg + 4 * (size_t)dim;
    g.hdone = c->d_cg_done;
    g.e0 = c->d_e0;
    g.polar = c->d_polar;
    g.nparts_mv = (int)(((int64_t)n * c->cg_lanes + 255) / 256);   // blocks of the A v kernel
    g.nmol = c->nmol;
    g.blk = p->precond == 2 ? c->d_pc_blk : nullptr;
    g.nparts_vec = p->precond == 2 ? (c->nmol + BLK_MOLS - 1) / BLK_MOLS : (dim + CG_THREADS - 1) / CG_THREADS;
    g.p_vz = c->d_cg_part;
    g.p_dv = c->d_cg_part + c->cg_parts;
    g.p_rr0 = c->d_cg_part + 2 * c->cg_parts;
    g.p_rr1 = c->d_cg_part + 3 * c->cg_parts;
    g.state = c->d_cg_state;
    g.scal = c->d_cg_scal;
    g.v32 = p->prec == SMD_PREC_MIXED ? c->d_cg32 : nullptr;
    g.inv_npol = 1.0 / (double)(c->npol > 0 ? c->npol : 1);
    return g;
}

// The block-preconditioned solve's stopping test at the top of A u.
static CgTop blk_top(const CgArgs& g) {
    CgTop t{g.state, g.p_dv, g.nparts_vec, g.tol};
    t.scale = g.inv_npol;
    t.res = g.scal;
    t.reason = 2;
    t.hdone = g.hdone;
    return t;
}

// A v of the solve in its precision; v32 is v's FP32 copy (mixed).
static void solve_apply(Ctx* c, const smd_induce_params* p, const IndArgs& a, const double* v, const float4* v32,
                        double* out, double* part, CgTop top) {
    if (p->prec == SMD_PREC_MIXED) launch_apply32(c, c->cg_lanes, a, v32, v, out, part, top);
    else launch_apply(c, c->cg_lanes, a, v, out, part ? v : nullptr, part, top);
}

// Enqueues everything of a solve up to the first CG iteration: frames (when
// `frames`), the permanent field, the coupling table, the start, A x and the
// residual.  Without minimisation it also sets Pi = alpha E0, and the solve
// is complete.  Launches only (capturable once the table is sized).
// The solve's field pass: E0 and, when minimising, the coupling table (and
// the learned guess's shells).
int induce_field(Ctx* c, const smd_induce_params* p, bool frames) {
    int rc = induce_check(c, p);
    if (rc) return rc;
    c->ind_last_converged = false;
    c->ind_shells_valid = false;
    if (p->do_min && (rc = table_alloc(c))) return rc;
    if (p->do_min) c->ind_table_built = true;
    c->ind_last = *p;   // the solve in progress (the learned guess reads its parameters)
    IndArgs a = ind_args(c, p);
    const int n = c->n, dim = 3 * n;
    const dim3 b256(256);
    const dim3 grid_rows((unsigned)(((int64_t)n * IND_LANES + 255) / 256));
    const dim3 grid_vec((dim + CG_THREADS - 1) / CG_THREADS);
    {
        ProfScope ps(c, "ind_field");
        if (frames) hipLaunchKernelGGL(k_frames, dim3((c->nmol + 255) / 256), b256, 0, c->stream, a);
        if (p->do_min && a.mixed) {
            hipLaunchKernelGGL(k_pack_ind32, dim3((n + 255) / 256), b256, 0, c->stream, a);
            // the learned guess's charge shells come out of the same pass
            float* shells = p->guess == 3 ? c->d_shell32 : nullptr;
            hipLaunchKernelGGL((k_field_table32<IND_LANES>), grid_rows, b256, 0, c->stream, a, shells);
            c->ind_shells_valid = shells != nullptr;
        } else {
            hipLaunchKernelGGL((k_field<IND_LANES>), grid_rows, b256, 0, c->stream, a);
            if (p->do_min) hipLaunchKernelGGL((k_ind_table<IND_LANES>), grid_rows, b256, 0, c->stream, a);
        }
    }
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// The solve's start after induce_field: the guess, A x, the residual and the
// preconditioner (without minimisation, Pi = alpha E0).
int induce_start(Ctx* c, const smd_induce_params* p) {
    int rc;
    IndArgs a = ind_args(c, p);
    const int n = c->n, dim = 3 * n;
    const dim3 b256(256);
    const dim3 grid_vec((dim + CG_THREADS - 1) / CG_THREADS);
    (void)n;
    if (!p->do_min) {
        hipLaunchKernelGGL(k_direct, grid_vec, b256, 0, c->stream, dim, c->d_polar, c->d_e0, c->d_mu, (float*)nullptr);
        SMD_HIP(c, hipGetLastError());
        return SMD_OK;
    }
    CgArgs g = cg_args(c, p);
    const bool mixed = p->prec == SMD_PREC_MIXED;
    const CgTop plain{nullptr, nullptr, 0, 0.0};
    ProfScope ps(c, "ind_cg");
    // the FP32 copy of x for the mixed A x, written by the start's last kernel
    float* x32 = mixed ? reinterpret_cast<float*>(g.v32) : nullptr;
    bool copied = false;
    if (p->guess == 0) hipLaunchKernelGGL(k_fill, grid_vec, b256, 0, c->stream, dim, 1.0, g.x);
    else if (p->guess == 2) {
        hipLaunchKernelGGL(k_direct, grid_vec, b256, 0, c->stream, dim, c->d_polar, c->d_e0, g.x, x32);
        copied = true;
    } else if (p->guess == 3) {
        // learned: x = alpha E0, z = A x feed the descriptors, then x = the network's guess
        hipLaunchKernelGGL(k_direct, grid_vec, b256, 0, c->stream, dim, c->d_polar, c->d_e0, g.x, x32);
        solve_apply(c, p, a, g.x, g.v32, g.z, nullptr, plain);
        if ((rc = mlp_guess(c, g.x, g.z, x32))) return rc;
        copied = true;
    }
    else if (p->guess == 4) {
        // the previous solution moved by the change of the direct term:
        // x = mu_prev + alpha (E0 - E0_prev), and E0 kept for the next solve
        hipLaunchKernelGGL(k_extrapolate, grid_vec, b256, 0, c->stream, dim, c->d_polar, c->d_e0, c->d_e0_prev, g.x);
    }
    // guess 1: keep the previous solution in d_mu
    if (mixed && !copied) hipLaunchKernelGGL(k_to_f32, grid_vec, b256, 0, c->stream, dim, g.x, x32);
    solve_apply(c, p, a, g.x, g.v32, g.z, nullptr, plain);
    if (p->precond == 2) {
        hipLaunchKernelGGL(k_block_pc, dim3((c->nmol + 255) / 256), b256, 0, c->stream, a, c->d_pc_blk, c->nmol);
        hipLaunchKernelGGL(k_pcg_start_blk, dim3(g.nparts_vec), dim3(BLK_THREADS), 0, c->stream, g);
        solve_apply(c, p, a, g.u, g.v32, g.z, g.p_vz, blk_top(g));   // w = A u, with the stopping test
    } else if (p->precond) hipLaunchKernelGGL(k_pcg_start, grid_vec, dim3(CG_THREADS), 0, c->stream, g);
    else hipLaunchKernelGGL(k_cg_start, grid_vec, dim3(CG_THREADS), 0, c->stream, g);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

int induce_begin(Ctx* c, const smd_induce_params* p, bool frames) {
    int rc = induce_field(c, p, frames);
    if (!rc) rc = induce_start(c, p);
    return rc;
}

// count CG iterations (from an even one): A v, the step, the direction (the
// block preconditioner's single-reduction form: the step, then A u).
static void cg_iterations(Ctx* c, const smd_induce_params* p, const IndArgs& a, const CgArgs& g, int count) {
    const dim3 grid_vec(g.nparts_vec), bt(CG_THREADS);
    for (int k = 0; k < count; ++k) {
        if (p->precond == 2) {
            const dim3 bw(BLK_THREADS);
            if (k & 1) hipLaunchKernelGGL(k_pcg_step_blk<1>, grid_vec, bw, 0, c->stream, g);
            else hipLaunchKernelGGL(k_pcg_step_blk<0>, grid_vec, bw, 0, c->stream, g);
            solve_apply(c, p, a, g.u, g.v32, g.z, g.p_vz, blk_top(g));
        } else if (p->precond) {
            solve_apply(c, p, a, g.v, g.v32, g.z, g.p_vz, CgTop{g.state, nullptr, g.nparts_vec, g.tol});
            if (k & 1) {
                hipLaunchKernelGGL(k_pcg_step<1>, grid_vec, bt, 0, c->stream, g);
                hipLaunchKernelGGL(k_pcg_dir<1>, grid_vec, bt, 0, c->stream, g);
            } else {
                hipLaunchKernelGGL(k_pcg_step<0>, grid_vec, bt, 0, c->stream, g);
                hipLaunchKernelGGL(k_pcg_dir<0>, grid_vec, bt, 0, c->stream, g);
            }
        } else {
            launch_apply(c, c->cg_lanes, a, g.v, g.z, g.v, g.p_vz, CgTop{g.state, g.p_dv, g.nparts_vec, g.tol});
            if (k & 1) {
                hipLaunchKernelGGL(k_cg_step<1>, grid_vec, bt, 0, c->stream, g);
                hipLaunchKernelGGL(k_cg_dir<1>, grid_vec, bt, 0, c->stream, g);
            } else {
                hipLaunchKernelGGL(k_cg_step<0>, grid_vec, bt, 0, c->stream, g);
                hipLaunchKernelGGL(k_cg_dir<0>, grid_vec, bt, 0, c->stream, g);
            }
        }
    }
}

// Enqueues `batches` batches of cg_batch iterations (iterations after the
// stop test passed return at once).  A batch is replayed from a captured
// graph, captured again when any kernel argument changes.
int induce_batches(Ctx* c, const smd_induce_params* p, int batches) {
    IndArgs a = ind_args(c, p);
    CgArgs g = cg_args(c, p);
    if (!c->cg_graph) {
        for (int b = 0; b < batches; ++b) cg_iterations(c, p, a, g, c->cg_batch);
        SMD_HIP(c, hipGetLastError());
        return SMD_OK;
    }
    std::string key;
    auto put = [&key](const void* q, size_t bytes) { key.append(static_cast<const char*>(q), bytes); };
    put(&c->cg_lanes, sizeof c->cg_lanes);
    put(&c->cg_batch, sizeof c->cg_batch);
    put(&p->precond, sizeof p->precond);
    put(&p->prec, sizeof p->prec);
    put(&a.n, sizeof a.n);
    put(&a.box, sizeof a.box);
    put(&a.pos, sizeof a.pos);
    put(&a.polar, sizeof a.polar);
    put(&a.tstride, sizeof a.tstride);
    put(&a.tcnt, sizeof a.tcnt);
    put(&a.tj, sizeof a.tj);
    put(&a.tab, sizeof a.tab);
    put(&a.mixed, sizeof a.mixed);
    put(&a.rec32, sizeof a.rec32);
    put(&a.Lh, sizeof a.Lh);
    put(&a.Ll, sizeof a.Ll);
    put(&a.invL, sizeof a.invL);
    put(&g, sizeof g);
    if (c->cg_exec && c->cg_key != key) {
        (void)hipGraphExecDestroy(c->cg_exec);
        c->cg_exec = nullptr;
    }
    if (!c->cg_exec) {
        hipGraph_t graph = nullptr;
        SMD_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        cg_iterations(c, p, a, g, c->cg_batch);
        SMD_HIP(c, hipStreamEndCapture(c->stream, &graph));
        const hipError_t e = hipGraphInstantiate(&c->cg_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            c->cg_exec = nullptr;
            return hip_fail(c, e, "hipGraphInstantiate (CG batch)");
        }
        c->cg_key = key;
    }
    for (int b = 0; b < batches; ++b) SMD_HIP(c, hipGraphLaunch(c->cg_exec, c->stream));
    return SMD_OK;
}

// Waits for the stream and reads the solve's state: done, iterations, stop
// reason, last residual.
int induce_poll(Ctx* c, int* done, int* iters, int* reason, double* res) {
    int st[3];
    double r = 0.0;
    SMD_HIP(c, hipMemcpyAsync(st, c->d_cg_state, sizeof st, hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipMemcpyAsync(&r, c->d_cg_scal, sizeof r, hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    if (done) *done = st[0];
    if (iters) *iters = st[1];
    if (reason) *reason = st[2];
    if (res) *res = r;
    return SMD_OK;
}

int induce_run(Ctx* c, const smd_induce_params* p, smd_induce_report* rep) {
    int rc = induce_begin(c, p, true);
    if (rc) return rc;
    int iters = 0, reason = 0;
    double res = 0.0;
    if (p->do_min) {
        const int max_it = p->max_iter > 0 ? p->max_iter : 3 * c->n;   // the reference's loop bound
        int done = 0;
        const int batch = c->cg_batch;
        for (int it0 = 0; it0 < max_it && !done; it0 += batch) {
            const int it1 = std::min(max_it, it0 + batch);
            if (it1 - it0 == batch) {
                if ((rc = induce_batches(c, p, 1))) return rc;
            } else {
                cg_iterations(c, p, ind_args(c, p), cg_args(c, p), it1 - it0);
            }
            if ((rc = induce_poll(c, &done, &iters, &reason, &res))) return rc;
        }
    }
    SMD_HIP(c, hipGetLastError());
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    c->ind_last_converged = p->do_min && reason != 0;
    if (rep) {
        rep->iterations = iters;
        rep->residual = res;
        rep->converged = p->do_min ? (reason != 0) : 1;
    }
    return SMD_OK;
}

}  // namespace smd

using namespace smd;

// reps launches of the CG's A v (k_apply_table on the last solve's coupling
// table and dipoles) between two events on the context stream.
extern "C" int smd_induce_time_apply(smd_ctx* c, int reps, double* avg_us) {
    if (!c || reps <= 0 || !avg_us) return SMD_EINVAL;
    // timing needs a coupling table of this context, not one of the current
    // positions: the last solve's is used as it is
    if (!c->ind_table_built) return fail(c, SMD_ESTATE, "no minimised smd_induce has built a coupling table");
    IndArgs a = ind_args(c, &c->ind_last);
    double* z = c->d_cg + 2 * (size_t)(3 * c->n);   // CG scratch; the solution in d_mu is untouched
    const bool mixed = c->ind_last.prec == SMD_PREC_MIXED;
    const int dim = 3 * c->n;
    if (mixed)
        hipLaunchKernelGGL(k_to_f32, dim3((dim + 255) / 256), dim3(256), 0, c->stream, dim, c->d_mu,
                           reinterpret_cast<float*>(c->d_cg32));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SMD_HIP(c, hipEventCreate(&e0));
    SMD_HIP(c, hipEventCreate(&e1));
    SMD_HIP(c, hipEventRecord(e0, c->stream));
    for (int k = 0; k < reps; ++k) {
        if (mixed) launch_apply32(c, c->cg_lanes, a, c->d_cg32, c->d_mu, z, nullptr, CgTop{nullptr, nullptr, 0, 0.0});
        else launch_apply(c, c->cg_lanes, a, c->d_mu, z, nullptr, nullptr, CgTop{nullptr, nullptr, 0, 0.0});
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

extern "C" int smd_set_multipoles(smd_ctx* c, const double* dip_local, const double* quad_local, const double* polar,
                                  int64_t n_atoms) {
    if (!c || !dip_local || !quad_local || !polar) return SMD_EINVAL;
    if (n_atoms != c->n) return fail(c, SMD_EINVAL, "count %lld != %d", (long long)n_atoms, c->n);
    const int n = c->n;
    if (!c->d_dipl) {
        const int parts = std::max((int)(((int64_t)n * 32 + 255) / 256), (3 * n + CG_THREADS - 1) / CG_THREADS);   // up to 32 lanes per row
        auto A = [&](auto*& p, size_t bytes) -> int {
            if (hipMalloc(&p, bytes) != hipSuccess) return fail(c, SMD_ENOMEM, "induction allocation of %zu bytes", bytes);
            return SMD_OK;
        };
        int rc;
        if ((rc = A(c->d_dipl, sizeof(double) * 3 * n)) || (rc = A(c->d_quadl, sizeof(double) * 6 * n)) ||
            (rc = A(c->d_polar, sizeof(double) * n)) || (rc = A(c->d_dipg, sizeof(double) * 3 * n)) ||
            (rc = A(c->d_quadg, sizeof(double) * 6 * n)) || (rc = A(c->d_e0, sizeof(double) * 3 * n)) ||
            (rc = A(c->d_mu, sizeof(double) * 3 * n)) || (rc = A(c->d_e0_prev, sizeof(double) * 3 * n)) ||
            (rc = A(c->d_cg, sizeof(double) * 15 * n)) ||
            (rc = A(c->d_cg32, sizeof(float4) * n)) || (rc = A(c->d_ind_rec32, sizeof(float4) * 2 * n)) ||
            (rc = A(c->d_ind_fix, sizeof(int4) * n)) ||
            (rc = A(c->d_pc_blk, sizeof(double) * 45 * (size_t)(c->nmol > 0 ? c->nmol : 1))) ||
            (rc = A(c->d_shell32, sizeof(float) * 24 * n)) ||
            (rc = A(c->d_site32, sizeof(float4) * 3 * n)) || (rc = A(c->d_site32_ind, sizeof(float4) * 3 * n)) ||
            (rc = A(c->d_pind32, sizeof(float4) * n)) ||
            (rc = A(c->d_cg_part, sizeof(double) * 4 * parts)) || (rc = A(c->d_cg_state, sizeof(int) * 4)) ||
            (rc = A(c->d_cg_scal, sizeof(double) * 8)))
            return rc;
        c->cg_parts = parts;
        SMD_HIP(c, hipMemset(c->d_mu, 0, sizeof(double) * 3 * n));
        SMD_HIP(c, hipMemset(c->d_e0_prev, 0, sizeof(double) * 3 * n));
    }
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemcpy(c->d_dipl, dip_local, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_quadl, quad_local, sizeof(double) * 6 * n, hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_polar, polar, sizeof(double) * n, hipMemcpyHostToDevice));
    c->npol = 0;
    for (int i = 0; i < n; ++i) c->npol += polar[i] > POLMIN ? 1 : 0;
    c->multipoles_set = true;
    return SMD_OK;
}

extern "C" int smd_induce(smd_ctx* c, const smd_induce_params* p, smd_induce_report* rep) {
    if (!c || !p) return SMD_EINVAL;
    return induce_run(c, p, rep);
}

extern "C" int smd_get_induction(smd_ctx* c, double* dip, double* quad, double* e0, double* mu, int64_t n_atoms) {
    if (!c) return SMD_EINVAL;
    if (!c->multipoles_set) return fail(c, SMD_ESTATE, "multipoles not set");
    if (n_atoms != c->n) return fail(c, SMD_EINVAL, "count %lld != %d", (long long)n_atoms, c->n);
    const size_t n = (size_t)c->n;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    if (dip) SMD_HIP(c, hipMemcpy(dip, c->d_dipg, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    if (quad) SMD_HIP(c, hipMemcpy(quad, c->d_quadg, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
    if (e0) SMD_HIP(c, hipMemcpy(e0, c->d_e0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    if (mu) SMD_HIP(c, hipMemcpy(mu, c->d_mu, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    return SMD_OK;
}
