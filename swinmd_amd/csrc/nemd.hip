// nemd.hip -- batched sheared replicas of rigid 3-site water on gfx950.
//
// Independent replicas (each its own cubic box, shear rate, Lees-Edwards tilt
// and Nose-Hoover state) are laid end to end in one set of device arrays and
// stepped by one launch of each kernel, so that 64 small replicas fill the
// chip the way one large system would.  Every kernel maps its blocks to
// replicas through a host-built table: a block never straddles two replicas,
// so per-replica sums are block partials added in a fixed order, and a
// replica's trajectory does not depend on block scheduling or on the other
// replicas.
//
// The step of one replica is the reference's leap-frog step
// (LFIntegrator::run, src/integrators/LFIntegrator.cpp:278-941: drift
// removal :354-379, forces :382-441, SHAKE with NH friction :561-586 and
// LFIntegrator::shake :1678-1828, kinetics :650-686, friction :688-709,
// rewrap :897-910) for rigid water with LJ on O and point-charge Ewald
// (LJForce.cpp:122-176, EwaldForce.cpp:296-527 real, :529-698 reciprocal,
// :709-813 intramolecular, ctor :69-79 self), with molecular SLLOD and
// Lees-Edwards images added (the reference has no shear flow):
//   * the peculiar update takes -dt gamma P_y / M along x, P the molecule's
//     peculiar momentum of the previous step, M its mass;
//   * the move adds dt gamma Y along x, Y the molecule's centre y (the same
//     for its three sites, so SHAKE's bond vectors are unchanged);
//   * the image across +y is offset by the tilt along x; the tilt advances by
//     gamma L dt after the move and stays in [-L/2, L/2);
//   * the k set is the reciprocal lattice of the sheared cell with |k| < kcut;
//   * the pair list (brute force per replica) is rebuilt when
//     2 max|d| + |gamma| rlist t > cut_buff, d the non-affine displacement
//     since the build and t the time since it.
// Zero shear rate gives the reference's step.
//
// Precision: FP64, or MIXED (FP32 pair and reciprocal-space arithmetic with
// FP64 displacements and phases, FP64 accumulation of every sum).  SHAKE,
// the thermostat and the integration are FP64 in both.

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <string>
#include <vector>

#include "smd_internal.h"

namespace smd {
namespace nemd {

constexpr int NPAIR = 8;     // pair partials: e_lj, e_real, virial xx xy xz yy yz zz
constexpr int NKF = 7;       // reciprocal partials: energy, virial (6)
constexpr int NSH = 31;      // SHAKE partials (SH_* below)
constexpr int KIN_BLOCK = 64;   // threads of a SHAKE block (one molecule each)
constexpr int SF_THREADS = 256; // k-vectors of a structure-factor block

enum {
    SH_AKE = 0, SH_MKE, SH_INTRA, SH_SV, SH_KM = SH_SV + 9, SH_MV = SH_KM + 6, SH_M = SH_MV + 3,
    SH_MC = SH_M + 1, SH_END = SH_MC + 9
};
static_assert(SH_END == NSH, "SHAKE partial layout");

enum : unsigned { E_LIST = 1u, E_KCAP = 2u, E_SHAKE_FAIL = 4u, E_SHAKE_NOCONV = 8u };

// Per-replica constants and block ranges (device copy of the host table).
struct Rep {
    int32_t a0, a1, m0, m1;       // atom and molecule ranges
    int32_t kbase, kcap, nch, pad0;
    int64_t pbase;                // first structure-factor partial (2 doubles each)
    int32_t pb0, pb1;             // pair-kernel blocks
    int32_t fb0, fb1;             // reciprocal-finish blocks
    int32_t sb0, sb1;             // SHAKE blocks
    int32_t hmax, nyb, nzmax, pad1;
    double L, invL, gamma, kcut, e_self, volume;
};

// Per-replica mutable state.
struct RepState {
    double tilt, fnh, fnh_minus;
    double drift[3];
    int32_t K, need, since, apply_drift;
    int32_t pad0, pending, max_row, pad;   // pending: rebuild at the next step
    unsigned long long maxd2;     // bits of the largest |d|^2 of the last step (non-negative doubles order as integers)
    int64_t steps, list_updates;
};

struct Params {
    double dt, temperature, coupl, shake_tol, cutoff, cut_buff, alpha;
    double sigma2_oo, eps4_oo;
    double b2[3];                 // squared constraint lengths H1-O, O-H2, H2-H1
    double mass[3], q[3];         // per site of a molecule
    double pre2;                  // (rlist + 2 r_OH)^2: the list's O-O prefilter
    double two_roh;               // 2 r_OH with a rounding margin
    int32_t max_count, nose_hoover, nb_cap, n;
};

struct Args {
    Params p;
    const Rep* rep;
    RepState* st;
    smd_batch_report* out;
    unsigned* err;
    double* pos;                  // [3n]
    double* vtm;
    double* vel;
    double* disp;
    double* fpair;                // [3n] LJ + real-space forces
    double* frec;                 // [3n] reciprocal-space forces
    double* com;                  // [3 nmol]
    double* mom;                  // [3 nmol]
    int32_t* nb;                  // [n][nb_cap]
    int32_t* nb_cnt;              // [n]
    double4* kvec;                // [sum kcap] (kx, ky, kz, exp(-k^2/4a^2)/k^2)
    double2* kf;                  // [sum kcap] (b C, b S) for the force pass
    double* sf_part;              // structure-factor partials
    double* pair_part;            // [NPAIR][n_pair_blocks]
    double* kf_part;              // [NKF][n_fin_blocks]
    double* sh_part;              // [NSH][n_shake_blocks]
    const int2* pair_blk;         // (replica, first row)
    const int4* sf_blk;           // (replica, k tile, chunk, -)
    const int2* fin_blk;          // (replica, k tile)
    const int2* kf_blk;           // (replica, first atom)
    const int2* sh_blk;           // (replica, first molecule)
    const int2* list_blk;         // (replica, first molecule)
    int32_t n_pair_blk, n_fin_blk, n_sh_blk, sf_chunk;
    int32_t force_rebuild, zero_disp, count_update, kvec_only;
};

// ---- geometry -------------------------------------------------------------

// Minimum image with the Lees-Edwards offset: the y image first (moving x by
// the tilt), then x and z.
__device__ __forceinline__ void le_image(double& dx, double& dy, double& dz, double L, double invL, double tilt) {
    const double ny = rint(dy * invL);
    dy -= ny * L;
    dx -= ny * tilt;
    dx -= L * rint(dx * invL);
    dz -= L * rint(dz * invL);
}

// The reference's apply_pbc loops on a point, with the offset: used for the
// molecular centres, which it keeps in [-L/2, L/2).
__device__ __forceinline__ void le_wrap_point(double& x, double& y, double& z, double L, double tilt) {
    const double h = 0.5 * L;
    if (!(fabs(y) < 2.0 * L && fabs(x) < 2.0 * L && fabs(z) < 2.0 * L)) {
        // far outside (or not a number): one rounding step instead of loops
        const double ny = rint(y / L);
        y -= ny * L;
        x -= ny * tilt;
        x -= L * rint(x / L);
        z -= L * rint(z / L);
        return;
    }
    while (y >= h) { y -= L; x -= tilt; }
    while (y < -h) { y += L; x += tilt; }
    while (x >= h) x -= L;
    while (x < -h) x += L;
    while (z >= h) z -= L;
    while (z < -h) z += L;
}

__device__ __forceinline__ double new_tilt(const Rep& r, double tilt, double dt) {
    double t = tilt + r.gamma * r.L * dt;
    if (t >= 0.5 * r.L) t -= r.L;
    else if (t < -0.5 * r.L) t += r.L;
    return t;
}

// ---- step head: list trigger and the sheared k set -----------------------

// One block per replica.  Decides whether the replica's list is rebuilt this
// step, and enumerates the reciprocal lattice of the current cell with
// |k| < kcut in (h, k, l) order, one of each +-k pair, by an ordered
// compaction of the candidate box h in [0, hmax], k in [-nyb, nyb],
// l in [-lmax, lmax] (a superset of every tilt's candidates).
__global__ __launch_bounds__(256) void k_begin(Args a) {
    const int r = blockIdx.x;
    const Rep rp = a.rep[r];
    RepState& s = a.st[r];
    __shared__ int wtot[256 / WAVE];
    if (threadIdx.x == 0 && !a.kvec_only) {
        int need = a.force_rebuild || s.pending;
        s.pending = 0;
        if (!need) {
            const double rlist = a.p.cutoff + a.p.cut_buff;
            const double slack = a.p.cut_buff - fabs(rp.gamma) * rlist * a.p.dt * (double)s.since;
            const double d2 = __longlong_as_double((long long)s.maxd2);
            need = slack <= 0.0 || d2 > slack * slack / 4.0;
        }
        s.need = need;
        if (need) {
            s.since = 0;
            s.max_row = 0;
            if (a.count_update) s.list_updates++;
        }
        s.maxd2 = 0ull;
    }
    const double tilt = s.tilt;
    const double alphaR4 = 1.0 / (4.0 * (a.p.alpha * a.p.alpha));
    const double kCut2 = rp.kcut * rp.kcut;
    const double v = 2.0 * PI / rp.L;
    const int ny_n = 2 * rp.nyb + 1, nz_n = 2 * rp.nzmax + 1;
    const int total = (rp.hmax + 1) * ny_n * nz_n;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    int count = 0;
    for (int base = 0; base < total; base += blockDim.x) {
        const int c = base + threadIdx.x;
        bool take = false;
        double kx = 0.0, ky = 0.0, kz = 0.0, k2 = 0.0;
        if (c < total) {
            const int nx = c / (ny_n * nz_n);
            const int ny = (c / nz_n) % ny_n - rp.nyb;
            const int nz = c % nz_n - rp.nzmax;
            const bool half = nx > 0 || ny > 0 || (ny == 0 && nz > 0);
            if (half) {
                const double shift = nx * tilt / rp.L;
                kx = nx * v;
                ky = (ny - shift) * v;
                kz = nz * v;
                k2 = kx * kx + ky * ky + kz * kz;
                take = k2 < kCut2;
            }
        }
        const unsigned long long m = __ballot(take);
        if (lane == 0) wtot[w] = __popcll(m);
        __syncthreads();
        int off = count;
        for (int q = 0; q < w; ++q) off += wtot[q];
        int tile = 0;
        for (int q = 0; q < (int)(blockDim.x / WAVE); ++q) tile += wtot[q];
        if (take) {
            const int slot = off + __popcll(m & ((1ull << lane) - 1ull));
            if (slot < rp.kcap) a.kvec[rp.kbase + slot] = make_double4(kx, ky, kz, exp(-k2 * alphaR4) / k2);
        }
        count += tile;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        s.K = min(count, rp.kcap);
        if (count > rp.kcap) atomicOr(a.err, E_KCAP);
    }
}

// ---- pair list: brute force over the replica, one wave per molecule -------

// A block's waves take consecutive molecules of one replica (rows 3m..3m+2);
// the replica's molecules pass through LDS in tiles of 64, lane l testing
// partner molecule l of the tile.  The O-O minimum image first: no site pair
// of two molecules can be within rlist unless their O are within
// rlist + 2 r_OH (pre2), since every site is within r_OH of its O.  Then the
// nine site pairs, each with its own minimum image (as the brute-force list
// does), and an ordered compaction per row: entries are in ascending partner
// index, the order of an atom-by-atom scan.
constexpr int LIST_WAVES = 4;
constexpr int LIST_TILE = WAVE;

__global__ __launch_bounds__(LIST_WAVES * WAVE) void k_list(Args a) {
    __shared__ double tile[LIST_TILE][9];
    const int2 bm = a.list_blk[blockIdx.x];
    const int r = bm.x;
    if (!a.st[r].need) return;   // block-uniform
    const Rep rp = a.rep[r];
    const double tilt = a.st[r].tilt;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    const int mi = bm.y + w;
    const bool live = mi < rp.m1;   // wave-uniform
    const double rl = a.p.cutoff + a.p.cut_buff;
    const double rl2 = rl * rl;
    double p[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    if (live) {
#pragma unroll
        for (int s = 0; s < 3; ++s)
#pragma unroll
            for (int d = 0; d < 3; ++d) p[s][d] = a.pos[9 * mi + 3 * s + d];
        if (a.zero_disp && lane < 9) a.disp[9 * mi + lane] = 0.0;
    }
    int cnt[3] = {0, 0, 0};
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int m0 = rp.m0; m0 < rp.m1; m0 += LIST_TILE) {
        __syncthreads();
        const int nt = min(LIST_TILE, rp.m1 - m0);
        for (int t = threadIdx.x; t < 9 * nt; t += blockDim.x) tile[t / 9][t % 9] = a.pos[9 * m0 + t];
        __syncthreads();
        if (!live) continue;
        const int mj = m0 + lane;
        unsigned bits[3] = {0u, 0u, 0u};   // bit s of bits[q]: site s of mj is within rlist of row site q
        if (lane < nt && mj != mi) {
            const double ox = tile[lane][3] - p[1][0], oy = tile[lane][4] - p[1][1], oz = tile[lane][5] - p[1][2];
            double dx = ox, dy = oy, dz = oz;
            le_image(dx, dy, dz, rp.L, rp.invL, tilt);
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < a.p.pre2) {
                // Within L/2 - 2 r_OH of each other, every site pair of the two
                // molecules is closer than L/2, so its minimum image is the O-O
                // one: shift the partner by that image's lattice vector once
                // instead of searching per site pair.
                const double hc = 0.5 * rp.L - a.p.two_roh;
                const bool same = hc > 0.0 && d2 < hc * hc;
                const double sx = same ? dx - ox : 0.0, sy = same ? dy - oy : 0.0, sz = same ? dz - oz : 0.0;
                double t[3][3];
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    t[u][0] = tile[lane][3 * u] + sx;
                    t[u][1] = tile[lane][3 * u + 1] + sy;
                    t[u][2] = tile[lane][3 * u + 2] + sz;
                }
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int u = 0; u < 3; ++u) {
                        double ex = t[u][0] - p[q][0], ey = t[u][1] - p[q][1], ez = t[u][2] - p[q][2];
                        if (!same) le_image(ex, ey, ez, rp.L, rp.invL, tilt);
                        if (ex * ex + ey * ey + ez * ez < rl2) bits[q] |= 1u << u;
                    }
            }
        }
        // the rows' ordered compaction, skipped by a wave with no site pair
        // inside rlist in this tile
        if (__ballot((bits[0] | bits[1] | bits[2]) != 0u) == 0ull) continue;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const unsigned long long b0 = __ballot(bits[q] & 1u), b1 = __ballot(bits[q] & 2u),
                                     b2 = __ballot(bits[q] & 4u);
            int slot = cnt[q] + __popcll(b0 & lt) + __popcll(b1 & lt) + __popcll(b2 & lt);
            int32_t* out = a.nb + (size_t)(3 * mi + q) * a.p.nb_cap;
#pragma unroll
            for (int s = 0; s < 3; ++s)
                if (bits[q] & (1u << s)) {
                    if (slot < a.p.nb_cap) out[slot] = 3 * mj + s;
                    ++slot;
                }
            cnt[q] += __popcll(b0) + __popcll(b1) + __popcll(b2);
        }
    }
    if (live && lane == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            a.nb_cnt[3 * mi + q] = min(cnt[q], a.p.nb_cap);
            atomicMax(&a.st[r].max_row, cnt[q]);
            if (cnt[q] > a.p.nb_cap) atomicOr(a.err, E_LIST);
        }
    }
}

// ---- pair forces: LJ (O-O) and Ewald real space over full rows ------------

template <typename R>
struct PairMath;

template <>
struct PairMath<double> {
    static __device__ __forceinline__ void ewald(double alpha, double alpha2, double alpha2PI, double r2, double& e,
                                                 double& f) {
        const double r = sqrt(r2);
        e = erfc(alpha * r) / r;
        f = (e + alpha2PI * exp(-alpha2 * r2)) / r2;
    }
};

template <>
struct PairMath<float> {
    static __device__ __forceinline__ void ewald(float alpha, float alpha2, float alpha2PI, float r2, float& e,
                                                 float& f) {
        const float r_1 = __builtin_amdgcn_rsqf(r2);
        const float r = r2 * r_1;
        float ex;
        ewald_erfc_f32(alpha, alpha2, r2, r, r_1, e, ex);
        f = (e + alpha2PI * ex) * (r_1 * r_1);
    }
};

// G lanes per row; a row's lanes take entries sub, sub + G, ...  The force on
// the row's atom is its lanes' FP64 sums added by a shuffle tree; energies
// and the virial come from the pairs with j > i (each pair once), in FP64
// per lane, then block partials.
template <typename R, int G>
__global__ __launch_bounds__(256) void k_pairs(Args a) {
    const int2 bm = a.pair_blk[blockIdx.x];
    const Rep rp = a.rep[bm.x];
    const double tilt = a.st[bm.x].tilt;
    const int row = bm.y + (int)(threadIdx.x / G);
    const int sub = threadIdx.x % G;
    double acc[NPAIR];
#pragma unroll
    for (int k = 0; k < NPAIR; ++k) acc[k] = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const bool live = row < rp.a1;
    if (live) {
        const double xi = a.pos[3 * row], yi = a.pos[3 * row + 1], zi = a.pos[3 * row + 2];
        const bool oi = row % 3 == 1;
        const R qi = (R)a.p.q[row % 3];
        const R cut2 = (R)(a.p.cutoff * a.p.cutoff);
        const R sigma2 = (R)a.p.sigma2_oo, eps4 = (R)a.p.eps4_oo;
        const R alpha = (R)a.p.alpha, alpha2 = (R)(a.p.alpha * a.p.alpha);
        const R alpha2PI = (R)(2.0 * a.p.alpha / SQRT_PI);
        const int cnt = a.nb_cnt[row];
        const int32_t* nb = a.nb + (size_t)row * a.p.nb_cap;
        for (int k = sub; k < cnt; k += G) {
            const int j = nb[k];
            double ddx = a.pos[3 * j] - xi, ddy = a.pos[3 * j + 1] - yi, ddz = a.pos[3 * j + 2] - zi;
            le_image(ddx, ddy, ddz, rp.L, rp.invL, tilt);
            const R dx = (R)ddx, dy = (R)ddy, dz = (R)ddz;
            const R r2 = dx * dx + dy * dy + dz * dz;
            const bool own = j > row;
            R s = (R)0;
            if (oi && j % 3 == 1 && r2 <= cut2) {
                const R rij_2 = (R)1 / r2;
                const R sr2 = sigma2 * rij_2;
                const R sr6 = sr2 * sr2 * sr2;
                const R sr12 = sr6 * sr6;
                const R tmpf = (sr12 + sr12 - sr6) * (R)6 * eps4 * rij_2;
                s -= tmpf;
                if (own) acc[0] += (double)(eps4 * (sr12 - sr6));
            }
            if (r2 < cut2) {
                const R qij = qi * (R)a.p.q[j % 3];
                R e, f;
                PairMath<R>::ewald(alpha, alpha2, alpha2PI, r2, e, f);
                s -= qij * f;
                if (own) acc[1] += (double)(qij * e);
            }
            if (own) {
                const double sd = (double)s;
                acc[2] += sd * (double)(dx * dx); acc[3] += sd * (double)(dx * dy);
                acc[4] += sd * (double)(dx * dz); acc[5] += sd * (double)(dy * dy);
                acc[6] += sd * (double)(dy * dz); acc[7] += sd * (double)(dz * dz);
            }
            fx += (double)(s * dx);
            fy += (double)(s * dy);
            fz += (double)(s * dz);
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        fx += __shfl_xor(fx, o, G);
        fy += __shfl_xor(fy, o, G);
        fz += __shfl_xor(fz, o, G);
    }
    if (live && sub == 0) {
        a.fpair[3 * row] = fx; a.fpair[3 * row + 1] = fy; a.fpair[3 * row + 2] = fz;
    }
    block_store<NPAIR>(acc, a.pair_part, a.n_pair_blk);
}

// ---- reciprocal space -----------------------------------------------------

template <typename R>
__device__ __forceinline__ void phase_sc(double th, R& sn, R& cs);

template <>
__device__ __forceinline__ void phase_sc<double>(double th, double& sn, double& cs) {
    sincos(th, &sn, &cs);
}

// The FP64 phase reduced to [-pi, pi], then the FP32 hardware sine/cosine.
template <>
__device__ __forceinline__ void phase_sc<float>(double th, float& sn, float& cs) {
    const double t = th - (2.0 * PI) * rint(th * (1.0 / (2.0 * PI)));
    const float tf = (float)t;
    sn = __sinf(tf);
    cs = __cosf(tf);
}

// Structure factors: a block takes 256 k-vectors of one replica and a chunk
// of its atoms (staged in LDS), and writes that chunk's (sum q cos, sum q sin)
// per k-vector.
template <typename R>
__global__ __launch_bounds__(SF_THREADS) void k_sfac(Args a) {
    extern __shared__ double4 at[];
    const int4 bm = a.sf_blk[blockIdx.x];
    const Rep rp = a.rep[bm.x];
    const int K = a.st[bm.x].K;
    const int k = bm.y * SF_THREADS + (int)threadIdx.x;
    if (bm.y * SF_THREADS >= K) return;   // block-uniform
    const int c0 = rp.a0 + bm.z * a.sf_chunk;
    const int c1 = min(c0 + a.sf_chunk, rp.a1);
    for (int i = c0 + (int)threadIdx.x; i < c1; i += blockDim.x)
        at[i - c0] = make_double4(a.pos[3 * i], a.pos[3 * i + 1], a.pos[3 * i + 2], a.p.q[i % 3]);
    __syncthreads();
    if (k >= K) return;
    const double4 kv = a.kvec[rp.kbase + k];
    double C = 0.0, S = 0.0;
    for (int i = 0; i < c1 - c0; ++i) {
        const double4 p = at[i];
        const double th = kv.x * p.x + kv.y * p.y + kv.z * p.z;
        R sn, cs;
        phase_sc<R>(th, sn, cs);
        C += (double)((R)p.w * cs);
        S += (double)((R)p.w * sn);
    }
    double* dst = a.sf_part + rp.pbase + 2 * ((int64_t)bm.z * rp.kcap + k);
    dst[0] = C;
    dst[1] = S;
}

// Per k: the chunks' sums in order, the energy and virial terms
// (EwaldForce.cpp:529-698), and (b C, b S) with b = 8 pi / V exp(..)/k^2 for
// the force pass.
__global__ __launch_bounds__(256) void k_kfin(Args a) {
    const int2 bm = a.fin_blk[blockIdx.x];
    const Rep rp = a.rep[bm.x];
    const int K = a.st[bm.x].K;
    const int k = bm.y * 256 + (int)threadIdx.x;
    double v[NKF];
#pragma unroll
    for (int q = 0; q < NKF; ++q) v[q] = 0.0;
    if (k < K) {
        const double4 kv = a.kvec[rp.kbase + k];
        double C = 0.0, S = 0.0;
        for (int c = 0; c < rp.nch; ++c) {
            const double* p = a.sf_part + rp.pbase + 2 * ((int64_t)c * rp.kcap + k);
            C += p[0];
            S += p[1];
        }
        const double alphaR4 = 1.0 / (4.0 * (a.p.alpha * a.p.alpha));
        const double pi4V = 4.0 * PI / rp.volume;
        const double k2 = kv.x * kv.x + kv.y * kv.y + kv.z * kv.z;
        const double te = pi4V * kv.w * (S * S + C * C);
        const double coeff = 2.0 * (alphaR4 + 1.0 / k2);
        v[0] = te;
        v[1] = te - te * coeff * kv.x * kv.x;
        v[2] = -te * coeff * kv.x * kv.y;
        v[3] = -te * coeff * kv.x * kv.z;
        v[4] = te - te * coeff * kv.y * kv.y;
        v[5] = -te * coeff * kv.y * kv.z;
        v[6] = te - te * coeff * kv.z * kv.z;
        const double b = 2.0 * pi4V * kv.w;
        a.kf[rp.kbase + k] = make_double2(b * C, b * S);
    }
    block_store<NKF>(v, a.kf_part, a.n_fin_blk);
}

// Forces: one thread per atom of one replica; the replica's k-vectors pass
// through LDS in tiles.  F_i += q_i k (sin(k.r_i) b C - cos(k.r_i) b S).
template <typename R>
__global__ __launch_bounds__(256) void k_kforce(Args a, int tile) {
    extern __shared__ double kt[];   // tile x (kx, ky, kz, bC, bS)
    const int2 bm = a.kf_blk[blockIdx.x];
    const Rep rp = a.rep[bm.x];
    const int K = a.st[bm.x].K;
    const int i = bm.y + (int)threadIdx.x;
    const bool live = i < rp.a1;
    double x = 0.0, y = 0.0, z = 0.0;
    R q = (R)0;
    if (live) {
        x = a.pos[3 * i]; y = a.pos[3 * i + 1]; z = a.pos[3 * i + 2];
        q = (R)a.p.q[i % 3];
    }
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int t0 = 0; t0 < K; t0 += tile) {
        const int nt = min(tile, K - t0);
        __syncthreads();
        for (int t = threadIdx.x; t < nt; t += blockDim.x) {
            const double4 kv = a.kvec[rp.kbase + t0 + t];
            const double2 f = a.kf[rp.kbase + t0 + t];
            kt[5 * t] = kv.x; kt[5 * t + 1] = kv.y; kt[5 * t + 2] = kv.z;
            kt[5 * t + 3] = f.x; kt[5 * t + 4] = f.y;
        }
        __syncthreads();
        if (live) {
            for (int t = 0; t < nt; ++t) {
                const double kx = kt[5 * t], ky = kt[5 * t + 1], kz = kt[5 * t + 2];
                const double th = kx * x + ky * y + kz * z;
                R sn, cs;
                phase_sc<R>(th, sn, cs);
                const double val = (double)(q * (sn * (R)kt[5 * t + 3] - cs * (R)kt[5 * t + 4]));
                fx += val * kx;
                fy += val * ky;
                fz += val * kz;
            }
        }
    }
    if (live) {
        a.frec[3 * i] = fx; a.frec[3 * i + 1] = fy; a.frec[3 * i + 2] = fz;
    }
}

// ---- SHAKE with SLLOD, kinetics, rewrap, and the replica's epilogue -------

// One thread per molecule.  vtMinus loses the drift the previous epilogue
// found (LFIntegrator.cpp:354-379), the forces are the pair and reciprocal
// sums, the update is LFIntegrator::shake's with the SLLOD terms, then the
// molecule's centre, momentum and kinetic sums (:650-686), the non-affine
// displacement and its largest |d|^2, and the rewrap by centre with the new
// tilt (:897-910).
__global__ __launch_bounds__(KIN_BLOCK) void k_shake(Args a) {
    const int2 bm = a.sh_blk[blockIdx.x];
    const int r = bm.x;
    const Rep rp = a.rep[r];
    RepState& st = a.st[r];
    const int m = bm.y + (int)threadIdx.x;
    double v[NSH];
#pragma unroll
    for (int q = 0; q < NSH; ++q) v[q] = 0.0;
    const double dt = a.p.dt, gamma = rp.gamma;
    const double tilt1 = new_tilt(rp, st.tilt, dt);
    double d2max = 0.0;
    if (m < rp.m1) {
        const int a0 = 3 * m;
        const double fnh = st.fnh;
        const double Px = a.mom[3 * m], Py = a.mom[3 * m + 1], Pz = a.mom[3 * m + 2];
        const double Y = a.com[3 * m + 1];
        double M = 0.0;
#pragma unroll
        for (int s = 0; s < 3; ++s) M += a.p.mass[s];
        double ro[3][3], rn[3][3], vt[3][3], F[3][3], mr[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const int id = a0 + s;
            mr[s] = 1.0 / a.p.mass[s];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                ro[s][d] = a.pos[3 * id + d];
                double vm = a.vtm[3 * id + d];
                if (st.apply_drift) vm -= st.drift[d];
                vt[s][d] = vm;
                F[s][d] = a.fpair[3 * id + d] + a.frec[3 * id + d];
            }
            double vp[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double P = d == 0 ? Px : (d == 1 ? Py : Pz);
                if (a.p.nose_hoover) vp[d] = vt[s][d] + dt * mr[s] * (F[s][d] - (fnh * P) / 3.0);
                else vp[d] = vt[s][d] + dt * mr[s] * F[s][d];
            }
            vp[0] -= dt * gamma * Py / M;
            double adv0 = vp[0] + gamma * Y;
            rn[s][0] = ro[s][0] + dt * adv0;
            rn[s][1] = ro[s][1] + dt * vp[1];
            rn[s][2] = ro[s][2] + dt * vp[2];
        }
        // the intramolecular Ewald correction (EwaldForce.cpp:709-813): with
        // rigid molecules its energy only
        {
            const double alpha = a.p.alpha;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int s0 = b == 2 ? 0 : b, s1 = b == 0 ? 1 : 2;
                const double dx = ro[s1][0] - ro[s0][0], dy = ro[s1][1] - ro[s0][1], dz = ro[s1][2] - ro[s0][2];
                const double rr = sqrt(dx * dx + dy * dy + dz * dz);
                v[SH_INTRA] -= a.p.q[s0] * a.p.q[s1] * erf(alpha * rr) / rr;
            }
        }
        // Gauss-Seidel sweeps (LFIntegrator.cpp:1499-1540) over the cyclic
        // constraints H1-O, O-H2, H2-H1; the bond vectors are short, so the
        // reference's minimum image leaves them as they are.
        double od[3][3], msum[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int nx = b == 2 ? 0 : b + 1;
#pragma unroll
            for (int d = 0; d < 3; ++d) od[b][d] = ro[b][d] - ro[nx][d];
            msum[b] = mr[b] + mr[nx];
        }
        const double tol = a.p.shake_tol, tol_2 = 2.0 * tol;
        double sv[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) sv[q] = 0.0;
        int count = 0;
        bool done = false, failed = false;
        while (!done && count < a.p.max_count && !failed) {
            done = true;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int nx = b == 2 ? 0 : b + 1;
                const double nX = rn[b][0] - rn[nx][0], nY = rn[b][1] - rn[nx][1], nZ = rn[b][2] - rn[nx][2];
                const double diff = a.p.b2[b] - (nX * nX + nY * nY + nZ * nZ);
                if (fabs(diff) > a.p.b2[b] * tol_2) {
                    const double rno = od[b][0] * nX + od[b][1] * nY + od[b][2] * nZ;
                    if (rno < a.p.b2[b] * tol) {
                        failed = true;
                        break;
                    }
                    const double factor = diff / (2 * rno * msum[b]);
                    const double nn[3] = {nX, nY, nZ};
#pragma unroll
                    for (int p = 0; p < 3; ++p)
#pragma unroll
                        for (int q = 0; q < 3; ++q) sv[3 * p + q] += factor * nn[p] * od[b][q];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double fo = factor * od[b][d];
                        rn[b][d] += mr[b] * fo;
                        rn[nx][d] -= mr[nx] * fo;
                    }
                    done = false;
                }
            }
            count++;
        }
        if (failed) atomicOr(a.err, E_SHAKE_FAIL);
        else if (!done) atomicOr(a.err, E_SHAKE_NOCONV);
#pragma unroll
        for (int q = 0; q < 9; ++q) v[SH_SV + q] = sv[q];
        // new positions, displacement, velocities (:1810-1826)
        double P[3] = {0.0, 0.0, 0.0}, Cm[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const int id = a0 + s;
            double vv = 0.0, dd = 0.0;
            const double dy_old = a.disp[3 * id + 1];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double p0 = ro[s][d];
                const double dr = rn[s][d] - p0;
                const double pn = p0 + dr;
                a.pos[3 * id + d] = pn;
                double dn = dr;
                if (d == 0) dn -= dt * gamma * (ro[s][1] - dy_old);
                const double dsp = a.disp[3 * id + d] + dn;
                a.disp[3 * id + d] = dsp;
                dd += dsp * dsp;
                double vp = dr / dt;
                if (d == 0) vp -= gamma * Y;
                const double vel = 0.5 * (vp + vt[s][d]);
                a.vel[3 * id + d] = vel;
                a.vtm[3 * id + d] = vp;
                vv += vel * vel;
                P[d] += a.p.mass[s] * vel;
                Cm[d] += a.p.mass[s] * pn;
                v[SH_MV + d] += a.p.mass[s] * vp;
            }
            v[SH_AKE] += vv * a.p.mass[s];
            v[SH_M] += a.p.mass[s];
            d2max = fmax(d2max, dd);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) Cm[d] /= M;
        v[SH_MKE] += (P[0] * P[0] + P[1] * P[1] + P[2] * P[2]) / M;
        v[SH_KM + 0] += P[0] * P[0] / M; v[SH_KM + 1] += P[0] * P[1] / M; v[SH_KM + 2] += P[0] * P[2] / M;
        v[SH_KM + 3] += P[1] * P[1] / M; v[SH_KM + 4] += P[1] * P[2] / M; v[SH_KM + 5] += P[2] * P[2] / M;
        // sum_i F_i (x) (r_i - R_m) at time t with the centre of time t
        // (Integrator::computePressure's correction, src/Integrator.cpp:215-230)
        {
            const double R0[3] = {a.com[3 * m], a.com[3 * m + 1], a.com[3 * m + 2]};
#pragma unroll
            for (int s = 0; s < 3; ++s)
#pragma unroll
                for (int p = 0; p < 3; ++p)
#pragma unroll
                    for (int q = 0; q < 3; ++q) v[SH_MC + 3 * p + q] += F[s][q] * (ro[s][p] - R0[p]);
        }
        // rewrap by centre with the tilt of t + dt
        double cx = Cm[0], cy = Cm[1], cz = Cm[2];
        le_wrap_point(cx, cy, cz, rp.L, tilt1);
        const double sx = cx - Cm[0], sy = cy - Cm[1], sz = cz - Cm[2];
        if (sx != 0.0 || sy != 0.0 || sz != 0.0) {
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const int id = a0 + s;
                a.pos[3 * id] += sx; a.pos[3 * id + 1] += sy; a.pos[3 * id + 2] += sz;
            }
        }
        a.com[3 * m] = cx; a.com[3 * m + 1] = cy; a.com[3 * m + 2] = cz;
        a.mom[3 * m] = P[0]; a.mom[3 * m + 1] = P[1]; a.mom[3 * m + 2] = P[2];
    }
    // largest |d|^2 of the replica (max is order-independent)
    {
        double d2 = d2max;
        for (int o = WAVE / 2; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, WAVE));
        if (threadIdx.x % WAVE == 0 && d2 > 0.0) atomicMax(&st.maxd2, (unsigned long long)__double_as_longlong(d2));
    }
    block_store<NSH>(v, a.sh_part, a.n_sh_blk);
}

// The replica's epilogue, one block of FINISH_THREADS per replica after
// k_shake (step_kind 0) or k_start (1): its sums in a fixed order (one wave
// per row, lane l adding blocks l, l + 64, ... of the replica in order, then
// a shuffle tree), the report, the friction (:688-709), the drift for the
// next step, the tilt and the list age.
__global__ __launch_bounds__(FINISH_THREADS) void k_finish(Args a, int step_kind) {
    const int r = blockIdx.x;
    __shared__ double tp[NPAIR], tk[NKF], ts[NSH];
    const Rep rp = a.rep[r];
    RepState& st = a.st[r];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int k = w; k < NPAIR + NKF + NSH; k += blockDim.x / WAVE) {
        const double* p;
        int nb;
        double* dst;
        if (k < NPAIR) {
            p = a.pair_part + (size_t)k * a.n_pair_blk + rp.pb0;
            nb = rp.pb1 - rp.pb0;
            dst = tp + k;
        } else if (k < NPAIR + NKF) {
            p = a.kf_part + (size_t)(k - NPAIR) * a.n_fin_blk + rp.fb0;
            nb = rp.fb1 - rp.fb0;
            dst = tk + (k - NPAIR);
        } else {
            p = a.sh_part + (size_t)(k - NPAIR - NKF) * a.n_sh_blk + rp.sb0;
            nb = rp.sb1 - rp.sb0;
            dst = ts + (k - NPAIR - NKF);
        }
        double sum = 0.0;
#pragma unroll 8
        for (int b = lane; b < nb; b += WAVE) sum += p[b];
        sum = wave_sum(sum);
        if (lane == 0) *dst = sum;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    smd_batch_report& o = a.out[r];
    const double dt = a.p.dt;
    o.e_lj = tp[0];
    o.e_real = tp[1];
    o.e_long = tk[0];
    o.e_intra = ts[SH_INTRA];
    o.e_self = rp.e_self;
    o.n_kvec = st.K;
    o.max_row = st.max_row;
    // ensemble virial: -(LJ + real, s d d^T over pairs) + reciprocal + constraint / dt^2
    const int sym[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int q = 0; q < 9; ++q) {
        double w = -tp[2 + sym[q]] + tk[1 + sym[q]];
        if (step_kind == 0) w += ts[SH_SV + q] / (dt * dt);
        o.virial[q] = w;
    }
    // drift of vtMinus for the next step (:354-379)
    const double mt = ts[SH_M];
    double dv[3];
    for (int d = 0; d < 3; ++d) dv[d] = ts[SH_MV + d] / mt;
    st.apply_drift = fabs(dv[0]) > 1e-11 || fabs(dv[1]) > 1e-11 || fabs(dv[2]) > 1e-11;
    for (int d = 0; d < 3; ++d) st.drift[d] = dv[d];
    if (step_kind == 0) {
        o.atom_ke2 = ts[SH_AKE];
        o.mol_ke2 = ts[SH_MKE];
        const double km[6] = {ts[SH_KM], ts[SH_KM + 1], ts[SH_KM + 2], ts[SH_KM + 3], ts[SH_KM + 4], ts[SH_KM + 5]};
        for (int q = 0; q < 9; ++q) o.mol_kin[q] = km[sym[q]];
        if (a.p.nose_hoover) {
            const int nm = rp.m1 - rp.m0;
            const double ndf = (double)(3 * nm - 3);
            const double qq = ndf * BOLTZ * a.p.temperature * a.p.coupl * a.p.coupl;
            const double invqq = 1 / qq;
            const double fnhpl = st.fnh_minus + dt * invqq * (ts[SH_MKE] - ndf * BOLTZ * a.p.temperature);
            st.fnh = 0.5 * (fnhpl + st.fnh_minus);
            st.fnh_minus = fnhpl;
        } else {
            st.fnh = 0.0;
        }
        st.tilt = new_tilt(rp, st.tilt, dt);
        st.since++;
        st.steps++;
    }
    o.tilt = st.tilt;
    o.fnh = st.fnh;
    o.steps = st.steps;
    o.list_updates = st.list_updates;
    if (step_kind == 0)
        for (int q = 0; q < 9; ++q) o.mol_correction[q] = ts[SH_MC + q];
}

// LFIntegrator::start's tail (:1056-1074) for every molecule: vtMinus =
// v - dt/2 F/m, the intramolecular energy, and the sums the epilogue needs
// for the first step's drift.
__global__ __launch_bounds__(KIN_BLOCK) void k_start(Args a) {
    const int2 bm = a.sh_blk[blockIdx.x];
    const int r = bm.x;
    const Rep rp = a.rep[r];
    RepState& st = a.st[r];
    const int m = bm.y + (int)threadIdx.x;
    double v[NSH];
#pragma unroll
    for (int q = 0; q < NSH; ++q) v[q] = 0.0;
    if (m < rp.m1) {
        double ro[3][3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const int id = 3 * m + s;
            const double mass_r = 1.0 / a.p.mass[s];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                ro[s][d] = a.pos[3 * id + d];
                const double F = a.fpair[3 * id + d] + a.frec[3 * id + d];
                const double vm = a.vel[3 * id + d] - 0.5 * a.p.dt * mass_r * F;
                a.vtm[3 * id + d] = vm;
                v[SH_MV + d] += a.p.mass[s] * vm;
            }
            v[SH_M] += a.p.mass[s];
        }
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int s0 = b == 2 ? 0 : b, s1 = b == 0 ? 1 : 2;
            const double dx = ro[s1][0] - ro[s0][0], dy = ro[s1][1] - ro[s0][1], dz = ro[s1][2] - ro[s0][2];
            const double rr = sqrt(dx * dx + dy * dy + dz * dz);
            v[SH_INTRA] -= a.p.q[s0] * a.p.q[s1] * erf(a.p.alpha * rr) / rr;
        }
    }
    block_store<NSH>(v, a.sh_part, a.n_sh_blk);
}

}  // namespace nemd
}  // namespace smd

// ===========================================================================
// host side
// ===========================================================================

struct smd_batch {
    int device = 0;
    hipStream_t stream = nullptr, stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;
    int R = 0, n = 0, nmol = 0;
    std::vector<int> nm;
    std::vector<double> box, gamma, kcut;
    int prec = 0;
    smd::nemd::Params prm{};
    std::vector<smd::nemd::Rep> rep;
    smd_batch_tiles tiles{16, 512, 256, 1, 1};
    // device
    smd::nemd::Rep* d_rep = nullptr;
    smd::nemd::RepState* d_st = nullptr;
    smd_batch_report* d_out = nullptr;
    unsigned* d_err = nullptr;
    double *d_pos = nullptr, *d_vtm = nullptr, *d_vel = nullptr, *d_disp = nullptr;
    double *d_fpair = nullptr, *d_frec = nullptr, *d_com = nullptr, *d_mom = nullptr;
    int32_t *d_nb = nullptr, *d_nb_cnt = nullptr;
    double4* d_kvec = nullptr;
    double2* d_kf = nullptr;
    double *d_sf_part = nullptr, *d_pair_part = nullptr, *d_kf_part = nullptr, *d_sh_part = nullptr;
    size_t sf_part_cap = 0, pair_part_cap = 0;
    int2 *d_pair_blk = nullptr, *d_fin_blk = nullptr, *d_kf_blk = nullptr, *d_sh_blk = nullptr;
    int2* d_list_blk = nullptr;
    int4* d_sf_blk = nullptr;
    size_t pair_blk_cap = 0, sf_blk_cap = 0;
    int n_pair_blk = 0, n_sf_blk = 0, n_fin_blk = 0, n_kf_blk = 0, n_sh_blk = 0, n_list_blk = 0;
    int ktot = 0;
    hipGraphExec_t graph = nullptr;
    bool have_state = false, started = false;
};

namespace smd {
namespace nemd {

static int bfail(smd_batch* b, int code, const char* fmt, ...) {
    if (b) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        b->err = buf;
    }
    return code;
}

#define BHIP(b, expr)                                                                                     \
    do {                                                                                                  \
        hipError_t _e = (expr);                                                                           \
        if (_e != hipSuccess) return bfail((b), SMD_EDEVICE, "%s: %s", #expr, hipGetErrorString(_e));      \
    } while (0)

template <typename T>
static int dalloc(smd_batch* b, T*& p, size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void**)&p, sizeof(T) * count) != hipSuccess)
        return bfail(b, SMD_ENOMEM, "device allocation of %zu bytes", sizeof(T) * count);
    return SMD_OK;
}

static void drop_graph(smd_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
}

static bool lanes_ok(int g) { return g == 1 || g == 2 || g == 4 || g == 8 || g == 16 || g == 32 || g == 64; }

// Block tables for the current tile shapes; every block belongs to one
// replica.  Partial buffers grow as needed.
static int build_tables(smd_batch* b) {
    const int G = b->tiles.pair_lanes, CH = b->tiles.sfac_chunk;
    const int rpb = 256 / G;
    std::vector<int2> pair_blk, fin_blk, kf_blk, sh_blk, list_blk;
    std::vector<int4> sf_blk;
    int64_t pbase = 0;
    for (int r = 0; r < b->R; ++r) {
        Rep& rp = b->rep[r];
        const int na = rp.a1 - rp.a0;
        rp.pb0 = (int)pair_blk.size();
        for (int row = rp.a0; row < rp.a1; row += rpb) pair_blk.push_back(make_int2(r, row));
        rp.pb1 = (int)pair_blk.size();
        rp.nch = (na + CH - 1) / CH;
        rp.pbase = pbase;
        pbase += 2 * (int64_t)rp.nch * rp.kcap;
        const int kt = (rp.kcap + 255) / 256;
        for (int t = 0; t < kt; ++t)
            for (int c = 0; c < rp.nch; ++c) sf_blk.push_back(make_int4(r, t, c, 0));
        rp.fb0 = (int)fin_blk.size();
        for (int t = 0; t < kt; ++t) fin_blk.push_back(make_int2(r, t));
        rp.fb1 = (int)fin_blk.size();
        for (int i = rp.a0; i < rp.a1; i += 256) kf_blk.push_back(make_int2(r, i));
        rp.sb0 = (int)sh_blk.size();
        for (int m = rp.m0; m < rp.m1; m += KIN_BLOCK) sh_blk.push_back(make_int2(r, m));
        rp.sb1 = (int)sh_blk.size();
        for (int m = rp.m0; m < rp.m1; m += LIST_WAVES) list_blk.push_back(make_int2(r, m));
    }
    int rc;
    if ((size_t)pbase > b->sf_part_cap) {
        if ((rc = dalloc(b, b->d_sf_part, (size_t)pbase))) return rc;
        b->sf_part_cap = (size_t)pbase;
    }
    if (pair_blk.size() > b->pair_blk_cap) {
        if ((rc = dalloc(b, b->d_pair_blk, pair_blk.size()))) return rc;
        if ((rc = dalloc(b, b->d_pair_part, NPAIR * pair_blk.size()))) return rc;
        b->pair_blk_cap = pair_blk.size();
    }
    if (sf_blk.size() > b->sf_blk_cap) {
        if ((rc = dalloc(b, b->d_sf_blk, sf_blk.size()))) return rc;
        b->sf_blk_cap = sf_blk.size();
    }
    if (!b->d_fin_blk) {   // these depend on the replicas only
        if ((rc = dalloc(b, b->d_fin_blk, fin_blk.size()))) return rc;
        if ((rc = dalloc(b, b->d_kf_part, NKF * fin_blk.size()))) return rc;
        if ((rc = dalloc(b, b->d_kf_blk, kf_blk.size()))) return rc;
        if ((rc = dalloc(b, b->d_sh_blk, sh_blk.size()))) return rc;
        if ((rc = dalloc(b, b->d_sh_part, NSH * sh_blk.size()))) return rc;
        if ((rc = dalloc(b, b->d_list_blk, list_blk.size()))) return rc;
    }
    b->n_pair_blk = (int)pair_blk.size();
    b->n_sf_blk = (int)sf_blk.size();
    b->n_fin_blk = (int)fin_blk.size();
    b->n_kf_blk = (int)kf_blk.size();
    b->n_sh_blk = (int)sh_blk.size();
    b->n_list_blk = (int)list_blk.size();
    BHIP(b, hipMemcpy(b->d_pair_blk, pair_blk.data(), sizeof(int2) * pair_blk.size(), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_sf_blk, sf_blk.data(), sizeof(int4) * sf_blk.size(), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_fin_blk, fin_blk.data(), sizeof(int2) * fin_blk.size(), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_kf_blk, kf_blk.data(), sizeof(int2) * kf_blk.size(), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_sh_blk, sh_blk.data(), sizeof(int2) * sh_blk.size(), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_list_blk, list_blk.data(), sizeof(int2) * list_blk.size(), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_rep, b->rep.data(), sizeof(Rep) * b->R, hipMemcpyHostToDevice));
    drop_graph(b);
    return SMD_OK;
}

static Args make_args(smd_batch* b) {
    Args a{};
    a.p = b->prm;
    a.rep = b->d_rep;
    a.st = b->d_st;
    a.out = b->d_out;
    a.err = b->d_err;
    a.pos = b->d_pos; a.vtm = b->d_vtm; a.vel = b->d_vel; a.disp = b->d_disp;
    a.fpair = b->d_fpair; a.frec = b->d_frec; a.com = b->d_com; a.mom = b->d_mom;
    a.nb = b->d_nb; a.nb_cnt = b->d_nb_cnt;
    a.kvec = b->d_kvec; a.kf = b->d_kf;
    a.sf_part = b->d_sf_part; a.pair_part = b->d_pair_part; a.kf_part = b->d_kf_part; a.sh_part = b->d_sh_part;
    a.pair_blk = b->d_pair_blk; a.sf_blk = b->d_sf_blk; a.fin_blk = b->d_fin_blk; a.kf_blk = b->d_kf_blk;
    a.sh_blk = b->d_sh_blk;
    a.list_blk = b->d_list_blk;
    a.n_pair_blk = b->n_pair_blk; a.n_fin_blk = b->n_fin_blk; a.n_sh_blk = b->n_sh_blk;
    a.sf_chunk = b->tiles.sfac_chunk;
    return a;
}

template <typename R>
static void launch_pairs(hipStream_t s, const Args& a, int G, int nblk) {
    switch (G) {
    case 1: hipLaunchKernelGGL((k_pairs<R, 1>), dim3(nblk), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_pairs<R, 2>), dim3(nblk), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_pairs<R, 4>), dim3(nblk), dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_pairs<R, 8>), dim3(nblk), dim3(256), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_pairs<R, 16>), dim3(nblk), dim3(256), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_pairs<R, 32>), dim3(nblk), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_pairs<R, 64>), dim3(nblk), dim3(256), 0, s, a); break;
    }
}

enum Pass : unsigned { P_BEGIN = 1, P_LIST = 2, P_PAIRS = 4, P_SFAC = 8, P_KFIN = 16, P_KFORCE = 32 };

static void launch_pass(smd_batch* b, hipStream_t s, const Args& a, unsigned pass) {
    const bool mixed = b->prec == SMD_PREC_MIXED;
    switch (pass) {
    case P_BEGIN: hipLaunchKernelGGL(k_begin, dim3(b->R), dim3(256), 0, s, a); break;
    case P_LIST: hipLaunchKernelGGL(k_list, dim3(b->n_list_blk), dim3(LIST_WAVES * WAVE), 0, s, a); break;
    case P_PAIRS:
        if (mixed) launch_pairs<float>(s, a, b->tiles.pair_lanes, b->n_pair_blk);
        else launch_pairs<double>(s, a, b->tiles.pair_lanes, b->n_pair_blk);
        break;
    case P_SFAC: {
        const size_t lds = sizeof(double4) * b->tiles.sfac_chunk;
        if (mixed) hipLaunchKernelGGL(k_sfac<float>, dim3(b->n_sf_blk), dim3(SF_THREADS), lds, s, a);
        else hipLaunchKernelGGL(k_sfac<double>, dim3(b->n_sf_blk), dim3(SF_THREADS), lds, s, a);
        break;
    }
    case P_KFIN: hipLaunchKernelGGL(k_kfin, dim3(b->n_fin_blk), dim3(256), 0, s, a); break;
    case P_KFORCE: {
        const int tile = b->tiles.kforce_tile;
        const size_t lds = sizeof(double) * 5 * tile;
        if (mixed) hipLaunchKernelGGL(k_kforce<float>, dim3(b->n_kf_blk), dim3(256), lds, s, a, tile);
        else hipLaunchKernelGGL(k_kforce<double>, dim3(b->n_kf_blk), dim3(256), lds, s, a, tile);
        break;
    }
    default: break;
    }
}

// The step: head, then the pair branch (list, pairs) and the reciprocal
// branch (structure factors, per-k finish, forces), on two streams when
// overlap is on, then SHAKE and the epilogues.
static int enqueue_step(smd_batch* b) {
    Args a = make_args(b);
    a.count_update = 1;
    a.zero_disp = 1;
    launch_pass(b, b->stream, a, P_BEGIN);
    hipStream_t s2 = b->stream;
    if (b->tiles.overlap) {
        BHIP(b, hipEventRecord(b->ev_fork, b->stream));
        BHIP(b, hipStreamWaitEvent(b->stream2, b->ev_fork, 0));
        s2 = b->stream2;
    }
    launch_pass(b, b->stream, a, P_LIST);
    launch_pass(b, b->stream, a, P_PAIRS);
    launch_pass(b, s2, a, P_SFAC);
    launch_pass(b, s2, a, P_KFIN);
    launch_pass(b, s2, a, P_KFORCE);
    if (b->tiles.overlap) {
        BHIP(b, hipEventRecord(b->ev_join, b->stream2));
        BHIP(b, hipStreamWaitEvent(b->stream, b->ev_join, 0));
    }
    hipLaunchKernelGGL(k_shake, dim3(b->n_sh_blk), dim3(KIN_BLOCK), 0, b->stream, a);
    hipLaunchKernelGGL(k_finish, dim3(b->R), dim3(FINISH_THREADS), 0, b->stream, a, 0);
    BHIP(b, hipGetLastError());
    return SMD_OK;
}

static int capture(smd_batch* b) {
    drop_graph(b);
    hipGraph_t g = nullptr;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(b);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, SMD_EDEVICE, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return bfail(b, SMD_EDEVICE, "graph instantiate: %s", hipGetErrorString(e));
    }
    return SMD_OK;
}

static int check_err(smd_batch* b) {
    unsigned e = 0;
    BHIP(b, hipMemcpyAsync(&e, b->d_err, sizeof e, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    if (!e) return SMD_OK;
    if (e & E_SHAKE_FAIL) return bfail(b, SMD_ESHAKE, "Shake constraint failure in a replica");
    if (e & E_SHAKE_NOCONV) return bfail(b, SMD_ENOCONV, "SHAKE did not converge in %d sweeps", b->prm.max_count);
    if (e & E_LIST) return bfail(b, SMD_ECAPACITY, "pair-list row capacity %d exceeded", b->prm.nb_cap);
    if (e & E_KCAP) return bfail(b, SMD_ECAPACITY, "k-vector capacity exceeded");
    return bfail(b, SMD_EDEVICE, "device status %u", e);
}

}  // namespace nemd
}  // namespace smd

using namespace smd;
using namespace smd::nemd;

extern "C" int smd_batch_create(smd_batch** out, const smd_batch_params* p, int device) {
    if (!out || !p) return SMD_EINVAL;
    *out = nullptr;
    if (p->n_replicas <= 0 || !p->n_mols || !p->box || !p->shear_rate || !p->kcut) return SMD_EINVAL;
    if (p->prec != SMD_PREC_FP64 && p->prec != SMD_PREC_MIXED) return SMD_EINVAL;
    if (!(p->dt > 0) || !(p->cutoff > 0) || !(p->cut_buff >= 0) || !(p->alpha > 0) || p->max_count <= 0)
        return SMD_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return SMD_EDEVICE;
    smd_batch* b = new smd_batch();
    b->device = device;
    b->R = p->n_replicas;
    b->prec = p->prec;
    const double rl = p->cutoff + p->cut_buff;
    int a = 0, m = 0, kb = 0;
    double dens = 0.0;
    for (int r = 0; r < b->R; ++r) {
        const int nmr = p->n_mols[r];
        const double L = p->box[r];
        if (nmr <= 0 || !(L > 2.0 * rl) || !(p->kcut[r] > 0)) {
            bfail(b, SMD_EINVAL, "replica %d: %d molecules, box %g, kcut %g (the box must exceed 2 (cutoff + cut_buff))",
                  r, nmr, L, p->kcut[r]);
            *out = b;
            return SMD_EINVAL;
        }
        Rep rp{};
        rp.a0 = a; rp.a1 = a + 3 * nmr; rp.m0 = m; rp.m1 = m + nmr;
        rp.L = L; rp.invL = 1.0 / L; rp.gamma = p->shear_rate[r]; rp.kcut = p->kcut[r];
        rp.volume = L * L * L;
        const double nR = p->kcut[r] * L / (2 * PI);
        rp.hmax = (int)nR;
        rp.nzmax = (int)nR;
        rp.nyb = (int)std::ceil(nR + 0.5 * rp.hmax) + 2;
        // lattice points in the half sphere, with room for the surface
        const double half = (4.0 / 3.0) * PI * nR * nR * nR / 2.0;
        rp.kcap = ((int)(half * 1.15 + 4.0 * PI * nR * nR * 0.5) + 64 + 63) / 64 * 64;
        rp.kbase = kb;
        kb += rp.kcap;
        rp.e_self = -(double)nmr * (2 * p->q_h * p->q_h + p->q_o * p->q_o) * p->alpha / SQRT_PI;
        b->rep.push_back(rp);
        b->nm.push_back(nmr);
        dens = std::max(dens, 3.0 * nmr / rp.volume);
        a += 3 * nmr;
        m += nmr;
    }
    b->n = a;
    b->nmol = m;
    b->ktot = kb;
    Params& q = b->prm;
    q.dt = p->dt; q.temperature = p->temperature; q.coupl = p->coupl; q.shake_tol = p->shake_tol;
    q.cutoff = p->cutoff; q.cut_buff = p->cut_buff; q.alpha = p->alpha;
    q.sigma2_oo = p->sigma_o * p->sigma_o;
    q.eps4_oo = 4.0 * p->eps_o;
    q.b2[0] = p->r_oh * p->r_oh; q.b2[1] = p->r_oh * p->r_oh; q.b2[2] = p->r_hh * p->r_hh;
    q.mass[0] = p->mass_h; q.mass[1] = p->mass_o; q.mass[2] = p->mass_h;
    q.q[0] = p->q_h; q.q[1] = p->q_o; q.q[2] = p->q_h;
    {
        q.two_roh = 2.0 * p->r_oh * (1.0 + 1e-6) + 1e-9;
        const double pre = rl + q.two_roh;
        q.pre2 = pre * pre;
    }
    q.max_count = p->max_count;
    q.nose_hoover = p->nose_hoover;
    q.n = a;
    int cap = p->nb_cap;
    if (cap <= 0) cap = (int)(dens * (4.0 / 3.0) * PI * rl * rl * rl * 1.35) + 64;
    q.nb_cap = (cap + 63) / 64 * 64;
    *out = b;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess)
        return bfail(b, SMD_EDEVICE, "stream or event creation");
    int rc;
    const size_t n3 = 3 * (size_t)a;
    if ((rc = dalloc(b, b->d_rep, b->R)) || (rc = dalloc(b, b->d_st, b->R)) || (rc = dalloc(b, b->d_out, b->R)) ||
        (rc = dalloc(b, b->d_err, 1)) || (rc = dalloc(b, b->d_pos, n3)) ||
        (rc = dalloc(b, b->d_vtm, n3)) || (rc = dalloc(b, b->d_vel, n3)) || (rc = dalloc(b, b->d_disp, n3)) ||
        (rc = dalloc(b, b->d_fpair, n3)) || (rc = dalloc(b, b->d_frec, n3)) || (rc = dalloc(b, b->d_com, 3 * (size_t)m)) ||
        (rc = dalloc(b, b->d_mom, 3 * (size_t)m)) || (rc = dalloc(b, b->d_nb, (size_t)a * q.nb_cap)) ||
        (rc = dalloc(b, b->d_nb_cnt, a)) || (rc = dalloc(b, b->d_kvec, kb)) || (rc = dalloc(b, b->d_kf, kb)))
        return rc;
    BHIP(b, hipMemset(b->d_err, 0, sizeof(unsigned)));
    BHIP(b, hipMemset(b->d_out, 0, sizeof(smd_batch_report) * b->R));
    BHIP(b, hipMemset(b->d_nb_cnt, 0, sizeof(int32_t) * a));
    return build_tables(b);
}

extern "C" void smd_batch_destroy(smd_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    void* ptrs[] = {b->d_rep, b->d_st, b->d_out, b->d_err, b->d_pos, b->d_vtm, b->d_vel, b->d_disp,
                    b->d_fpair, b->d_frec, b->d_com, b->d_mom, b->d_nb, b->d_nb_cnt, b->d_kvec, b->d_kf,
                    b->d_sf_part, b->d_pair_part, b->d_kf_part, b->d_sh_part, b->d_pair_blk, b->d_fin_blk,
                    b->d_kf_blk, b->d_sh_blk, b->d_sf_blk, b->d_list_blk};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->stream2) (void)hipStreamDestroy(b->stream2);
    delete b;
}

extern "C" const char* smd_batch_last_error(const smd_batch* b) { return b ? b->err.c_str() : "no batch"; }

extern "C" int smd_batch_set_state(smd_batch* b, const double* pos, const double* vel, const double* tilt) {
    if (!b || !pos || !vel || !tilt) return SMD_EINVAL;
    BHIP(b, hipSetDevice(b->device));
    const size_t n3 = 3 * (size_t)b->n;
    std::vector<double> com(3 * (size_t)b->nmol), zero(n3, 0.0);
    const double* ms = b->prm.mass;
    const double M = ms[0] + ms[1] + ms[2];
    for (int m = 0; m < b->nmol; ++m)
        for (int d = 0; d < 3; ++d) {
            double c = 0.0;
            for (int s = 0; s < 3; ++s) c += ms[s] * pos[3 * (3 * m + s) + d];
            com[3 * m + d] = c / M;
        }
    std::vector<RepState> st(b->R);
    for (int r = 0; r < b->R; ++r) {
        RepState s{};
        s.tilt = tilt[r];
        s.need = 1;
        st[r] = s;
    }
    BHIP(b, hipMemcpy(b->d_pos, pos, sizeof(double) * n3, hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_vel, vel, sizeof(double) * n3, hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_vtm, zero.data(), sizeof(double) * n3, hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_disp, zero.data(), sizeof(double) * n3, hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_com, com.data(), sizeof(double) * com.size(), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_mom, zero.data(), sizeof(double) * com.size(), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_st, st.data(), sizeof(RepState) * b->R, hipMemcpyHostToDevice));
    BHIP(b, hipMemset(b->d_err, 0, sizeof(unsigned)));
    b->have_state = true;
    b->started = false;
    return SMD_OK;
}

extern "C" int smd_batch_start(smd_batch* b) {
    if (!b) return SMD_EINVAL;
    if (!b->have_state) return bfail(b, SMD_ESTATE, "no state set");
    BHIP(b, hipSetDevice(b->device));
    Args a = make_args(b);
    a.force_rebuild = 1;
    a.zero_disp = 1;
    for (unsigned pass : {P_BEGIN, P_LIST, P_PAIRS, P_SFAC, P_KFIN, P_KFORCE}) launch_pass(b, b->stream, a, pass);
    hipLaunchKernelGGL(k_start, dim3(b->n_sh_blk), dim3(KIN_BLOCK), 0, b->stream, a);
    hipLaunchKernelGGL(k_finish, dim3(b->R), dim3(FINISH_THREADS), 0, b->stream, a, 1);
    BHIP(b, hipGetLastError());
    int rc = check_err(b);
    if (rc) return rc;
    b->started = true;
    return SMD_OK;
}

extern "C" int smd_batch_run(smd_batch* b, int64_t nsteps) {
    if (!b || nsteps < 0) return SMD_EINVAL;
    if (!b->started) return bfail(b, SMD_ESTATE, "smd_batch_start has not run");
    BHIP(b, hipSetDevice(b->device));
    if (!b->tiles.use_graph) {
        for (int64_t s = 0; s < nsteps; ++s) {
            int rc = enqueue_step(b);
            if (rc) return rc;
        }
        return check_err(b);
    }
    if (!b->graph) {
        int rc = capture(b);
        if (rc) return rc;
    }
    for (int64_t s = 0; s < nsteps; ++s) BHIP(b, hipGraphLaunch(b->graph, b->stream));
    return check_err(b);
}

extern "C" int smd_batch_get_state(smd_batch* b, double* pos, double* vel, double* vt_minus, double* force,
                                   double* tilt) {
    if (!b) return SMD_EINVAL;
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    const size_t n3 = 3 * (size_t)b->n;
    if (pos) BHIP(b, hipMemcpy(pos, b->d_pos, sizeof(double) * n3, hipMemcpyDeviceToHost));
    if (vel) BHIP(b, hipMemcpy(vel, b->d_vel, sizeof(double) * n3, hipMemcpyDeviceToHost));
    if (vt_minus) BHIP(b, hipMemcpy(vt_minus, b->d_vtm, sizeof(double) * n3, hipMemcpyDeviceToHost));
    if (force) {
        std::vector<double> fr(n3);
        BHIP(b, hipMemcpy(force, b->d_fpair, sizeof(double) * n3, hipMemcpyDeviceToHost));
        BHIP(b, hipMemcpy(fr.data(), b->d_frec, sizeof(double) * n3, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n3; ++k) force[k] += fr[k];
    }
    if (tilt) {
        std::vector<RepState> st(b->R);
        BHIP(b, hipMemcpy(st.data(), b->d_st, sizeof(RepState) * b->R, hipMemcpyDeviceToHost));
        for (int r = 0; r < b->R; ++r) tilt[r] = st[r].tilt;
    }
    return SMD_OK;
}

extern "C" int smd_batch_get_reports(smd_batch* b, smd_batch_report* rep) {
    if (!b || !rep) return SMD_EINVAL;
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipMemcpy(rep, b->d_out, sizeof(smd_batch_report) * b->R, hipMemcpyDeviceToHost));
    return SMD_OK;
}

extern "C" int smd_batch_set_tiles(smd_batch* b, const smd_batch_tiles* t) {
    if (!b || !t) return SMD_EINVAL;
    if (!lanes_ok(t->pair_lanes)) return bfail(b, SMD_EINVAL, "pair_lanes %d: a power of two up to 64", t->pair_lanes);
    if (t->sfac_chunk < 64 || t->sfac_chunk > 2048 || t->sfac_chunk % 64)
        return bfail(b, SMD_EINVAL, "sfac_chunk %d: a multiple of 64 in [64, 2048]", t->sfac_chunk);
    if (t->kforce_tile < 64 || t->kforce_tile > 1536 || t->kforce_tile % 64)
        return bfail(b, SMD_EINVAL, "kforce_tile %d: a multiple of 64 in [64, 1536]", t->kforce_tile);
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    b->tiles = *t;
    b->tiles.overlap = t->overlap ? 1 : 0;
    b->tiles.use_graph = t->use_graph ? 1 : 0;
    return build_tables(b);
}

extern "C" int smd_batch_get_tiles(const smd_batch* b, smd_batch_tiles* t) {
    if (!b || !t) return SMD_EINVAL;
    *t = b->tiles;
    return SMD_OK;
}

__global__ void k_mark_pending(RepState* st, int R) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) st[r].pending = 1;
}

static int time_pass(smd_batch* b, unsigned pass, int reps, double* avg_us) {
    Args a = make_args(b);
    a.zero_disp = 0;
    a.kvec_only = 1;
    hipEvent_t e0, e1;
    BHIP(b, hipEventCreate(&e0));
    BHIP(b, hipEventCreate(&e1));
    launch_pass(b, b->stream, a, pass);   // warm
    BHIP(b, hipEventRecord(e0, b->stream));
    for (int k = 0; k < reps; ++k) launch_pass(b, b->stream, a, pass);
    BHIP(b, hipEventRecord(e1, b->stream));
    BHIP(b, hipEventSynchronize(e1));
    float ms = 0.f;
    BHIP(b, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    BHIP(b, hipGetLastError());
    *avg_us = 1e3 * ms / reps;
    return SMD_OK;
}

extern "C" int smd_batch_time_kernel(smd_batch* b, const char* name, int reps, double* avg_us) {
    if (!b || !name || !avg_us || reps <= 0) return SMD_EINVAL;
    if (!b->started) return bfail(b, SMD_ESTATE, "smd_batch_start has not run");
    BHIP(b, hipSetDevice(b->device));
    const std::string n = name;
    unsigned pass = 0;
    if (n == "list") pass = P_LIST;
    else if (n == "pairs") pass = P_PAIRS;
    else if (n == "kvec") pass = P_BEGIN;
    else if (n == "sfac") pass = P_SFAC;
    else if (n == "kfin") pass = P_KFIN;
    else if (n == "kforce") pass = P_KFORCE;
    else return bfail(b, SMD_EINVAL, "unknown kernel %s", name);
    if (pass == P_LIST) {
        // a rebuild of every replica's list (keeping the displacements and the
        // list age), then a rebuild at the next step so that the trigger's
        // bound holds again
        std::vector<RepState> st(b->R);
        BHIP(b, hipMemcpyAsync(st.data(), b->d_st, sizeof(RepState) * b->R, hipMemcpyDeviceToHost, b->stream));
        BHIP(b, hipStreamSynchronize(b->stream));
        std::vector<RepState> st2 = st;
        for (auto& s : st2) s.need = 1;
        BHIP(b, hipMemcpy(b->d_st, st2.data(), sizeof(RepState) * b->R, hipMemcpyHostToDevice));
        int rc = time_pass(b, P_LIST, reps, avg_us);
        BHIP(b, hipMemcpy(b->d_st, st.data(), sizeof(RepState) * b->R, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_mark_pending, dim3((b->R + 63) / 64), dim3(64), 0, b->stream, b->d_st, b->R);
        BHIP(b, hipStreamSynchronize(b->stream));
        return rc;
    }
    return time_pass(b, pass, reps, avg_us);
}

extern "C" int smd_batch_autotune(smd_batch* b, int reps, double* times_us, int32_t* n_cand) {
    if (!b || reps <= 0) return SMD_EINVAL;
    if (!b->started) return bfail(b, SMD_ESTATE, "smd_batch_start has not run");
    BHIP(b, hipSetDevice(b->device));
    const int lanes[] = {4, 8, 16, 32};
    const int chunks[] = {256, 512, 1024};
    const int ktiles[] = {128, 256, 512};
    const smd_batch_tiles keep = b->tiles;
    smd_batch_tiles best = keep;
    std::vector<double> rows;
    double bt = 1e30;
    int rc;
    for (int g : lanes) {
        smd_batch_tiles t = keep;
        t.pair_lanes = g;
        if ((rc = smd_batch_set_tiles(b, &t))) return rc;
        double us;
        if ((rc = time_pass(b, P_PAIRS, reps, &us))) return rc;
        rows.insert(rows.end(), {(double)g, 0.0, 0.0, us, 0.0, 0.0});
        if (us < bt) { bt = us; best.pair_lanes = g; }
    }
    bt = 1e30;
    for (int c : chunks) {
        smd_batch_tiles t = keep;
        t.pair_lanes = best.pair_lanes;
        t.sfac_chunk = c;
        if ((rc = smd_batch_set_tiles(b, &t))) return rc;
        double us;
        if ((rc = time_pass(b, P_SFAC, reps, &us))) return rc;
        rows.insert(rows.end(), {0.0, (double)c, 0.0, 0.0, us, 0.0});
        if (us < bt) { bt = us; best.sfac_chunk = c; }
    }
    bt = 1e30;
    for (int k : ktiles) {
        smd_batch_tiles t = best;
        t.kforce_tile = k;
        if ((rc = smd_batch_set_tiles(b, &t))) return rc;
        double us;
        if ((rc = time_pass(b, P_KFORCE, reps, &us))) return rc;
        rows.insert(rows.end(), {0.0, 0.0, (double)k, 0.0, 0.0, us});
        if (us < bt) { bt = us; best.kforce_tile = k; }
    }
    if ((rc = smd_batch_set_tiles(b, &best))) return rc;
    // the reciprocal partials of the chosen chunking for the next step's finish
    Args a = make_args(b);
    for (unsigned pass : {P_SFAC, P_KFIN, P_KFORCE}) launch_pass(b, b->stream, a, pass);
    BHIP(b, hipStreamSynchronize(b->stream));
    const int nrow = (int)(rows.size() / 6);
    if (n_cand) {
        if (times_us) std::memcpy(times_us, rows.data(), sizeof(double) * 6 * std::min(nrow, *n_cand));
        *n_cand = nrow;
    }
    return SMD_OK;
}
