// pairlist.hip -- cell grid and half pair list on gfx950.
//
// Reproduces the pair set of Ensemble::set_pairlist/update_pairlist
// (src/Ensemble.cpp:454-525) and CellManager (src/utils/CellManager.h:82-203,
// src/utils/CellManager.cpp:43-314): the same cell grid, the same atom->cell
// arithmetic with clamping, the same 13-cell forward stencil, the same FP64
// distance predicate r^2 < (cutoff+cutBuff)^2 evaluated without contraction,
// and the same row order (in-cell successors, then stencil cells in order,
// atoms in index order).  Rows are stored ELL-style: row i at nb[i*stride].
//
// Rows are written owned-first: the reference's row i fills row i from the
// front, in its order; the same pairs seen from their other atom fill it from
// the back, so every atom's row holds all its partners and the pair kernels
// need no atomics.  Cell grid: one pass per step assigns atoms to cells, scans
// the counts and compacts each cell's atoms in index order; the pairs are then
// found per (cell, neighbour cell) block from LDS-staged positions (see
// k_list_cells).  Without cells, one wavefront per atom tests all atoms.
#include <algorithm>
#include <climits>

#include "smd_internal.h"

namespace smd {

// The 13 forward shifts in CellManager::build_cell_neighbor order
// (src/utils/CellManager.h:161-186).
__constant__ int c_stencil[13][3] = {
    {0, 0, -1}, {1, 0, -1}, {1, 0, 0},  {1, 0, 1},  {-1, 1, -1},
    {0, 1, -1}, {1, 1, -1}, {-1, 1, 0}, {0, 1, 0},  {1, 1, 0},
    {-1, 1, 1}, {0, 1, 1},  {1, 1, 1}};

struct CellGrid {
    int n[3];
    int nxy, ncell;
    double origin[3];
    double rlen[3];
};

__device__ __forceinline__ int cell_coord(double v, int n) {
    // (Int)(delta*rLen) then clamp (CellManager.h:97-119); v is finite for
    // every in-range coordinate, NaN lands in the last cell.
    if (!(v < (double)n)) return n - 1;
    if (v < 0.0) return 0;
    int x = (int)v;
    return x >= n ? n - 1 : x;
}

// Rebuild preamble of the MD step (LFIntegrator::run :306-341): count the
// update, zero the displacements and shift each molecule so its centre lies
// in the box.
struct RebuildPrep {
    int on;
    double* com;
    double* disp;
    double* scalars;   // [2] list updates
};

// Atom -> cell (CellManager::find_cell, src/utils/CellManager.h:97-119) and the
// per-cell counts, aggregated in LDS when the grid fits there.  One thread
// per molecule, so the MD step's molecule rewrap runs in the same pass.
template <bool LDS_HIST>
__global__ __launch_bounds__(128) void k_cell_assign(int nmol, const int32_t* __restrict__ mol_start, double* pos,
                                                      CellGrid g, Box box, RebuildPrep pr,
                                                      int32_t* __restrict__ cell_of, int32_t* __restrict__ cnt,
                                                      const int* gate) {
    if (gate && !*gate) return;
    extern __shared__ int32_t hist[];
    if constexpr (LDS_HIST) {
        for (int k = threadIdx.x; k < g.ncell; k += blockDim.x) hist[k] = 0;
        __syncthreads();
    }
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < nmol) {
        const int s = mol_start[m], e = mol_start[m + 1];
        if (pr.on) {
            if (m == 0) pr.scalars[2] += 1.0;
            for (int i = s; i < e; ++i) pr.disp[3 * i] = pr.disp[3 * i + 1] = pr.disp[3 * i + 2] = 0.0;
            const double c0x = pr.com[3 * m], c0y = pr.com[3 * m + 1], c0z = pr.com[3 * m + 2];
            double cx = c0x, cy = c0y, cz = c0z;
            pbc(cx, cy, cz, box);
            const double dx = cx - c0x, dy = cy - c0y, dz = cz - c0z;
            pr.com[3 * m] = cx; pr.com[3 * m + 1] = cy; pr.com[3 * m + 2] = cz;
            for (int i = s; i < e; ++i) { pos[3 * i] += dx; pos[3 * i + 1] += dy; pos[3 * i + 2] += dz; }
        }
        for (int i = s; i < e; ++i) {
            const double dx = pos[3 * i] - g.origin[0];
            const double dy = pos[3 * i + 1] - g.origin[1];
            const double dz = pos[3 * i + 2] - g.origin[2];
            const int x = cell_coord(dx * g.rlen[0], g.n[0]);
            const int y = cell_coord(dy * g.rlen[1], g.n[1]);
            const int z = cell_coord(dz * g.rlen[2], g.n[2]);
            const int c = z * g.nxy + y * g.n[0] + x;
            cell_of[i] = c;
            if constexpr (LDS_HIST) atomicAdd(&hist[c], 1);
            else atomicAdd(&cnt[c], 1);
        }
    }
    if constexpr (LDS_HIST) {
        __syncthreads();
        for (int k = threadIdx.x; k < g.ncell; k += blockDim.x)
            if (hist[k]) atomicAdd(&cnt[k], hist[k]);
    }
}

// One block per cell: its start (the sum of the counts of lower cells) and
// its atoms compacted in index order (set_atom appends in atom order,
// CellManager.cpp:145-170); it stops once the cell is full.  The counts are
// cleared later, by the counting pass of the list build.
constexpr int FILL_THREADS = 1024;
constexpr int FILL_WAVES = FILL_THREADS / WAVE;

__global__ __launch_bounds__(FILL_THREADS) void k_cell_fill(int n, int ncell, const int32_t* __restrict__ cell_of,
                                                            const int32_t* __restrict__ cnt,
                                                            int32_t* __restrict__ start, int32_t* __restrict__ atoms,
                                                            const int* gate) {
    if (gate && !*gate) return;
    __shared__ int32_t wsum[FILL_WAVES];
    const int c = blockIdx.x;
    if (c >= ncell) return;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    int32_t part = 0;
    for (int k = threadIdx.x; k < c; k += blockDim.x) part += cnt[k];
    part = wave_sum(part);
    if (lane == 0) wsum[w] = part;
    __syncthreads();
    int base = 0;
#pragma unroll
    for (int k = 0; k < FILL_WAVES; ++k) base += wsum[k];
    const int end = base + cnt[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        start[c] = base;
        if (c == ncell - 1) start[ncell] = end;
    }
    const uint64_t below = (1ull << lane) - 1ull;
    for (int a0 = 0; a0 < n && base < end; a0 += FILL_THREADS) {
        const int a = a0 + threadIdx.x;
        const bool hit = (a < n) && (cell_of[a] == c);
        const uint64_t m = __ballot(hit);
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int off = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < FILL_WAVES; ++k) {
            off += k < w ? wsum[k] : 0;
            tot += wsum[k];
        }
        if (hit && base + off + __popcll(m & below) < end) atoms[base + off + __popcll(m & below)] = a;
        base += tot;
        __syncthreads();
    }
}

struct ListArgs {
    int n;
    const double* pos;
    const int32_t* excl_off;
    const int32_t* excl;
    double rlist2;
    Box box;
    int32_t* nb;
    int32_t* nb_cnt;
    int32_t* nb_own;
    int32_t stride;
    unsigned* status;
    uint32_t* mask;    // [27][mask_words][n] hit bits of the cell-pair pass, by cell-order slot
    int mask_words;
};

struct Excl {
    int cnt;
    int lo, hi;        // smallest and largest excluded index
    int ids[8];
    const int32_t* rest;
};

__device__ __forceinline__ bool excluded(const Excl& e, int j) {
    // Ensemble::exclusion_check (src/Ensemble.h:294-305); the first eight
    // entries are compared without a loop so they stay in registers
    if (j < e.lo || j > e.hi) return false;
    bool hit = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) hit = hit || (e.ids[k] == j);
    if (!hit && e.cnt > 8)
        for (int k = 8; k < e.cnt; ++k)
            if (e.rest[k] == j) return true;
    return hit;
}

// Test 64 candidates (one per lane) and append the hits to row i in lane
// order.  Pairs the reference's half list holds in row i ("owned") fill the
// row from the front, in the reference's order; the others fill it from the
// back.
__device__ __forceinline__ void test_and_append(const ListArgs& a, int i, double px, double py, double pz,
                                                const Excl& ex, bool valid, int j, bool owned, bool use_pbc,
                                                int& count, int& n_owned) {
    const int lane = threadIdx.x % WAVE;
    bool hit = false;
    if (valid && !excluded(ex, j)) {
        double dx = a.pos[3 * j] - px;
        double dy = a.pos[3 * j + 1] - py;
        double dz = a.pos[3 * j + 2] - pz;
        if (use_pbc) pbc(dx, dy, dz, a.box);
        hit = len2_exact(dx, dy, dz) < a.rlist2;
    }
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t mo = __ballot(hit && owned), mn = __ballot(hit && !owned);
    const int n_non = count - n_owned;
    if (hit) {
        // the two segments meet only on overflow, which the count reports
        const int slot = owned ? n_owned + __popcll(mo & below) : a.stride - 1 - (n_non + __popcll(mn & below));
        if (slot >= 0 && slot < a.stride && (owned ? slot < a.stride - n_non : slot >= n_owned))
            a.nb[(int64_t)i * a.stride + slot] = j;
    }
    count += __popcll(mo) + __popcll(mn);
    n_owned += __popcll(mo);
}

__device__ __forceinline__ Excl load_excl(const ListArgs& a, int i) {
    Excl e;
    const int o = a.excl_off[i];
    e.cnt = a.excl_off[i + 1] - o;
    e.rest = a.excl + o;
#pragma unroll
    for (int k = 0; k < 8; ++k) e.ids[k] = (k < e.cnt) ? e.rest[k] : -1;
    e.lo = INT_MAX;
    e.hi = -1;
    for (int k = 0; k < e.cnt; ++k) {
        e.lo = min(e.lo, (int)e.rest[k]);
        e.hi = max(e.hi, (int)e.rest[k]);
    }
    return e;
}

// ---- cell-pair list build ----------------------------------------------------
//
// Block (c, s) pairs the atoms of cell c with the atoms of its s-th
// neighbour: s = 0 the cell itself (no PBC, CellManager.cpp:257-279), s =
// 1..13 the forward stencil (the reference's row of the atom in c), s =
// 14..26 the backward stencil (the same pairs seen from the other atom).  The
// neighbour's positions are staged in LDS and each thread owns one atom of c.
// A counting pass stores, per atom, the hits of each segment (self j>i, self
// j<i, forward 1..13, backward 1..13); the writing pass repeats the test and
// writes each hit at its final slot, so row i's front segment is the
// reference's row in its order.

constexpr int LB_THREADS = 256;   // atoms of cell c per pass of a block
constexpr int LB_STAGE = 256;     // neighbour atoms staged in LDS at a time
constexpr int LB_MASK_MAX = 1024; // neighbour-cell atom indices staged for the mask expansion
constexpr int NSEG = 28;          // per-atom hit counters

// Neighbour s of cell c and, per axis, the multiple of L (-1, 0, +1) that
// brings a partner in that cell next to the atom when the stencil wrapped
// around the box.
__device__ __forceinline__ int neighbour_cell(const CellGrid& g, int c, int s, int wrap[3]) {
    wrap[0] = wrap[1] = wrap[2] = 0;
    if (s == 0) return c;
    const int sgn = s <= 13 ? 1 : -1;
    const int k = s <= 13 ? s - 1 : s - 14;
    const int cz = c / g.nxy, cy = (c % g.nxy) / g.n[0], cx = c % g.n[0];
    int xx = cx + sgn * c_stencil[k][0], yy = cy + sgn * c_stencil[k][1], zz = cz + sgn * c_stencil[k][2];
    if (xx >= g.n[0]) { xx = 0; wrap[0] = 1; } else if (xx < 0) { xx = g.n[0] - 1; wrap[0] = -1; }
    if (yy >= g.n[1]) { yy = 0; wrap[1] = 1; } else if (yy < 0) { yy = g.n[1] - 1; wrap[1] = -1; }
    if (zz >= g.n[2]) { zz = 0; wrap[2] = 1; } else if (zz < 0) { zz = g.n[2] - 1; wrap[2] = -1; }
    return zz * g.nxy + yy * g.n[0] + xx;
}

// Tests up to 32 staged candidates against atom i and returns their hit
// bits.  SELF: the cell itself (no PBC; j > i owned).  Otherwise the pair
// vector is shifted by the stencil's wrap: |d + shift| < L/2 on every axis
// is the reference's single PBC step (a value exactly at -L/2 takes the
// exact path), and every hit has the block's ownership.  The first pass only
// counts; the second writes each hit at its slot.
template <bool WRITE, bool SELF>
__device__ __forceinline__ uint32_t test_word(const ListArgs& a, const double4* __restrict__ sj, int nb, int i,
                                              double px, double py, double pz, double shx, double shy, double shz,
                                              const Excl& ex, bool fits, int32_t* row, int& own_at, int& non_at,
                                              bool block_owned = true) {
    uint32_t word = 0;
#pragma unroll 4
    for (int b = 0; b < nb; ++b) {
        const double4 q = sj[b];
        const int j = (int)q.w;
        double dx = q.x - px, dy = q.y - py, dz = q.z - pz;
        if constexpr (!SELF) {
            const double ex_ = dx + shx, ey_ = dy + shy, ez_ = dz + shz;
            const bool inside = (fabs(ex_) < a.box.h[0]) & (fabs(ey_) < a.box.h[1]) & (fabs(ez_) < a.box.h[2]);
            dx = inside ? ex_ : dx;
            dy = inside ? ey_ : dy;
            dz = inside ? ez_ : dz;
            if (!inside) pbc(dx, dy, dz, a.box);
        }
        if (len2_exact(dx, dy, dz) < a.rlist2 && (!SELF || j != i) && !excluded(ex, j)) {
            const bool owned = SELF ? (j > i) : block_owned;
            if constexpr (WRITE) {
                if (fits) {
                    if (owned) row[own_at] = j;
                    else row[a.stride - 1 - non_at] = j;
                }
            } else {
                word |= 1u << b;
            }
            own_at += owned ? 1 : 0;
            non_at += owned ? 0 : 1;
        }
    }
    return word;
}

template <bool WRITE>
__global__ __launch_bounds__(LB_THREADS) void k_list_cells(ListArgs a, CellGrid g, const int32_t* __restrict__ start,
                                                           const int32_t* __restrict__ catoms, int32_t* __restrict__ seg,
                                                           int32_t* cell_cnt, const int* gate) {
    if (gate && !*gate) return;
    __shared__ double4 sj[LB_STAGE];   // x, y, z, atom index
    const int c = blockIdx.x, s = blockIdx.y;
    // the cell counts were last read by k_cell_fill: clear them for the next build
    if constexpr (!WRITE)
        if (c == 0 && s == 0)
            for (int k = threadIdx.x; k < g.ncell; k += blockDim.x) cell_cnt[k] = 0;
    int wrap[3];
    const int nc = neighbour_cell(g, c, s, wrap);
    // A partner across a wrapped face sits about L away: apply_pbc's loop
    // then makes exactly one step, d + shift (shift = +L when the stencil
    // wrapped past the upper face, since d = x_j - x_i is then near -L).
    // Whenever d + shift lands in [-L/2, L/2) it is apply_pbc's result, the
    // same single operation; otherwise the loops run.
    const double shx = wrap[0] * a.box.L[0], shy = wrap[1] * a.box.L[1], shz = wrap[2] * a.box.L[2];
    const int ib = start[c], ie = start[c + 1];
    const int jb = start[nc], je = start[nc + 1];
    const int n = a.n;
    // the counting pass records its hits as bits when the neighbour cell fits
    // the mask; the writing pass then expands the bits instead of testing again
    const bool use_mask = je - jb <= 32 * a.mask_words;
    // the writing pass reads the neighbour cell's atom indices from LDS
    __shared__ int32_t sidx[LB_MASK_MAX];
    const bool lds_idx = WRITE && use_mask && je - jb <= LB_MASK_MAX;
    if (lds_idx) {
        for (int t = threadIdx.x; t < je - jb; t += blockDim.x) sidx[t] = catoms[jb + t];
        __syncthreads();
    }
    for (int i0 = ib; i0 < ie; i0 += LB_THREADS) {
        const int k = i0 + threadIdx.x;
        const bool act = k < ie;
        const int i = act ? catoms[k] : 0;
        double px = 0, py = 0, pz = 0;
        Excl ex{};
        int own_at = 0, non_at = 0;   // counts (first pass) or next slots (second pass)
        bool fits = true;
        // word w of this atom's hits: mrow[w * n] (slot-major, so a wave's
        // stores of one word are contiguous)
        uint32_t* mrow = a.mask + (size_t)s * a.mask_words * n + k;
        if (act) {
            px = a.pos[3 * i]; py = a.pos[3 * i + 1]; pz = a.pos[3 * i + 2];
            ex = load_excl(a, i);
            if constexpr (WRITE) {
                // k_seg_offsets left each segment's first slot in seg
                fits = a.nb_cnt[i] <= a.stride;
                const int at = s == 0 ? 0 : seg[s * n + i];
                own_at = at;
                non_at = at;
            }
        }
        int32_t* row = a.nb + (int64_t)i * a.stride;
        if (WRITE && use_mask) {
            // expand the recorded hits in candidate order
            if (act && fits) {
                const int nw = (je - jb + 31) / 32;
                for (int w = 0; w < nw; ++w) {
                    uint32_t bits = mrow[(size_t)w * n];
                    while (bits) {
                        const int b = __ffs(bits) - 1;
                        bits &= bits - 1u;
                        const int j = lds_idx ? sidx[32 * w + b] : catoms[jb + 32 * w + b];
                        const bool owned = s == 0 ? (j > i) : (s <= 13);
                        if (owned) row[own_at++] = j;
                        else row[a.stride - 1 - non_at++] = j;
                    }
                }
            }
            continue;
        }
        for (int j0 = jb; j0 < je; j0 += LB_STAGE) {
            const int nj = min(LB_STAGE, je - j0);
            __syncthreads();
            for (int t = threadIdx.x; t < nj; t += blockDim.x) {
                const int j = catoms[j0 + t];
                sj[t] = make_double4(a.pos[3 * j], a.pos[3 * j + 1], a.pos[3 * j + 2], (double)j);
            }
            __syncthreads();
            if (act) {
                // candidates in words of 32: word w0/32 of the hit mask is
                // stored once its 32 candidates are tested
                for (int w0 = 0; w0 < nj; w0 += 32) {
                    const int nb = min(32, nj - w0);
                    const uint32_t word = s == 0
                        ? test_word<WRITE, true>(a, sj + w0, nb, i, px, py, pz, 0.0, 0.0, 0.0, ex, fits, row,
                                                 own_at, non_at)
                        : test_word<WRITE, false>(a, sj + w0, nb, i, px, py, pz, shx, shy, shz, ex, fits, row,
                                                  own_at, non_at, s <= 13);
                    if (!WRITE && use_mask) mrow[(size_t)((j0 - jb + w0) >> 5) * n] = word;
                }
            }
        }
        if constexpr (!WRITE) {
            if (act) {
                if (s == 0) { seg[i] = own_at; seg[n + i] = non_at; }
                else seg[(s + 1) * n + i] = s <= 13 ? own_at : non_at;
            }
        }
    }
}

// Between the two passes: per atom, the first slot of each stencil cell's
// segment -- owned hits (same cell j > i, forward cells 1..13) from the front
// of the row in stencil order, the others (same cell j < i, backward cells
// 14..26) from the back -- written over the counts; the row totals.
__global__ __launch_bounds__(256) void k_seg_offsets(ListArgs a, int32_t* __restrict__ seg, const int* gate) {
    if (gate && !*gate) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.n;
    if (i >= n) return;
    int cnt[NSEG];
#pragma unroll
    for (int t = 0; t < NSEG; ++t) cnt[t] = seg[t * n + i];
    // cnt[0], cnt[1]: same cell, owned and not; cnt[s + 1]: stencil cell s
    int own_at = cnt[0], non_at = cnt[1];
    seg[i] = 0;
#pragma unroll
    for (int s = 1; s <= 13; ++s) { seg[s * n + i] = own_at; own_at += cnt[s + 1]; }
#pragma unroll
    for (int s = 14; s <= 26; ++s) { seg[s * n + i] = non_at; non_at += cnt[s + 1]; }
    const int total = own_at + non_at;
    a.nb_cnt[i] = total;
    a.nb_own[i] = own_at;
    if (total > a.stride) atomicOr(a.status, ST_LIST_OVERFLOW);
}

// Ensemble::build_pairlist (src/Ensemble.cpp:492-525): all j != i with the
// min-image predicate; the reference's row i holds the j > i.
__global__ __launch_bounds__(256) void k_pairs_brute(ListArgs a, const int* gate) {
    if (gate && !*gate) return;
    const int i = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (i >= a.n) return;
    const int lane = threadIdx.x % WAVE;
    const double px = a.pos[3 * i], py = a.pos[3 * i + 1], pz = a.pos[3 * i + 2];
    const Excl ex = load_excl(a, i);
    int count = 0, n_owned = 0;
    for (int j0 = 0; j0 < a.n; j0 += WAVE) {
        const int j = j0 + lane;
        test_and_append(a, i, px, py, pz, ex, j < a.n && j != i, j, j > i, true, count, n_owned);
    }
    if (lane == 0) {
        a.nb_cnt[i] = count;
        a.nb_own[i] = n_owned;
        if (count > a.stride) atomicOr(a.status, ST_LIST_OVERFLOW);
    }
}

// ---------------------------------------------------------------------------

static bool cell_setup(double len, const double L[3], int n[3], double rlen[3]) {
    // CellManager::set_cellManager (src/utils/CellManager.cpp:43-84)
    if ((len <= 0) || (len > L[0] / 3.0) || (len > L[1] / 3.0) || (len > L[2] / 3.0)) return false;
    for (int d = 0; d < 3; ++d) {
        n[d] = (int)(L[d] / len);
        const double cl = L[d] / n[d];
        rlen[d] = 1.0 / cl;
    }
    return true;
}

int pairlist_alloc(Ctx* c, int32_t stride) {
    if (stride <= c->stride && c->d_nb) return SMD_OK;
    if (c->d_nb) (void)hipFree(c->d_nb);
    c->d_nb = nullptr;
    const size_t bytes = (size_t)c->n * (size_t)stride * sizeof(int32_t);
    if (hipMalloc(&c->d_nb, bytes ? bytes : 4) != hipSuccess) {
        c->stride = 0;
        return fail(c, SMD_ENOMEM, "pair list allocation of %zu bytes failed", bytes);
    }
    c->stride = stride;
    return SMD_OK;
}

// The molecule rewrap alone, for the brute-force list of the MD step.
__global__ __launch_bounds__(256) void k_rebuild_prep(int nmol, const int32_t* __restrict__ mol_start, double* pos,
                                                      Box box, RebuildPrep pr, const int* gate) {
    if (gate && !*gate) return;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const int s = mol_start[m], e = mol_start[m + 1];
    if (m == 0) pr.scalars[2] += 1.0;
    for (int i = s; i < e; ++i) pr.disp[3 * i] = pr.disp[3 * i + 1] = pr.disp[3 * i + 2] = 0.0;
    const double c0x = pr.com[3 * m], c0y = pr.com[3 * m + 1], c0z = pr.com[3 * m + 2];
    double cx = c0x, cy = c0y, cz = c0z;
    pbc(cx, cy, cz, box);
    const double dx = cx - c0x, dy = cy - c0y, dz = cz - c0z;
    pr.com[3 * m] = cx; pr.com[3 * m + 1] = cy; pr.com[3 * m + 2] = cz;
    for (int i = s; i < e; ++i) { pos[3 * i] += dx; pos[3 * i + 1] += dy; pos[3 * i + 2] += dz; }
}

static int launch_list(Ctx* c, const int* gate, bool md_prep) {
    RebuildPrep pr{};
    pr.on = md_prep ? 1 : 0;
    pr.com = c->d_com;
    pr.disp = c->d_disp;
    pr.scalars = c->d_scalars;
    ListArgs a;
    a.n = c->n;
    a.pos = c->d_pos;
    a.excl_off = c->d_excl_off;
    a.excl = c->d_excl;
    double pc = c->list_cutoff + c->list_buff;   // pairCutOff (src/Ensemble.cpp:456-457)
    pc *= pc;
    a.rlist2 = pc;
    a.box = c->box;
    a.nb = c->d_nb;
    a.nb_cnt = c->d_nb_cnt;
    a.nb_own = c->d_nb_own;
    a.stride = c->stride;
    a.status = c->d_status;
    a.mask = c->d_list_mask;
    a.mask_words = c->list_mask_words;
    const int waves_per_block = 4;
    const dim3 grid((c->n + waves_per_block - 1) / waves_per_block);
    ProfScope ps(c, "pairlist");
    if (c->list_cells) {
        CellGrid g;
        for (int d = 0; d < 3; ++d) {
            g.n[d] = c->ncell[d];
            g.rlen[d] = c->cell_rlen[d];
            g.origin[d] = -c->box.L[d] / 2.0;   // origin0 (CellManager.cpp:74)
        }
        g.nxy = g.n[0] * g.n[1];
        g.ncell = g.nxy * g.n[2];
        const int nc = g.ncell;
        const dim3 ga((c->nmol + 127) / 128);
        if (nc <= 8192)
            hipLaunchKernelGGL((k_cell_assign<true>), ga, dim3(128), sizeof(int32_t) * nc, c->stream, c->nmol,
                               c->d_mol_start, c->d_pos, g, c->box, pr, c->d_cell_of, c->d_cell_cnt, gate);
        else
            hipLaunchKernelGGL((k_cell_assign<false>), ga, dim3(128), 0, c->stream, c->nmol, c->d_mol_start, c->d_pos,
                               g, c->box, pr, c->d_cell_of, c->d_cell_cnt, gate);
        if (c->list_fork) SMD_HIP(c, hipEventRecord(c->list_fork, c->stream));   // positions are final here
        hipLaunchKernelGGL(k_cell_fill, dim3(nc), dim3(FILL_THREADS), 0, c->stream, c->n, nc, c->d_cell_of,
                           c->d_cell_cnt, c->d_cell_start, c->d_cell_atoms, gate);
        const dim3 gl(nc, 27);
        hipLaunchKernelGGL((k_list_cells<false>), gl, dim3(LB_THREADS), 0, c->stream, a, g, c->d_cell_start,
                           c->d_cell_atoms, c->d_list_seg, c->d_cell_cnt, gate);
        hipLaunchKernelGGL(k_seg_offsets, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, a, c->d_list_seg, gate);
        hipLaunchKernelGGL((k_list_cells<true>), gl, dim3(LB_THREADS), 0, c->stream, a, g, c->d_cell_start,
                           c->d_cell_atoms, c->d_list_seg, c->d_cell_cnt, gate);
    } else {
        if (md_prep)
            hipLaunchKernelGGL(k_rebuild_prep, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, c->nmol,
                               c->d_mol_start, c->d_pos, c->box, pr, gate);
        if (c->list_fork) SMD_HIP(c, hipEventRecord(c->list_fork, c->stream));
        hipLaunchKernelGGL(k_pairs_brute, grid, dim3(WAVE * waves_per_block), 0, c->stream, a, gate);
    }
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// Builds the list.  With on_device_flag the kernels are gated on d_flag[0]
// and begin with the MD step's rebuild preamble (used inside the MD step,
// where the Verlet trigger is evaluated on the device); the grid must then
// already be configured.
int pairlist_build(Ctx* c, double cutoff, double cut_buff, int use_cells, bool on_device_flag) {
    if (on_device_flag) return launch_list(c, c->d_flag, true);
    if (!c->box_set) return fail(c, SMD_ESTATE, "box not set");
    c->list_cutoff = cutoff;
    c->list_buff = cut_buff;
    const double len = cutoff + cut_buff;   // cellLen (src/Ensemble.cpp:462)
    int n[3];
    double rl[3];
    c->list_cells = use_cells && cell_setup(len, c->box.L, n, rl);
    if (c->list_cells) {
        for (int d = 0; d < 3; ++d) { c->ncell[d] = n[d]; c->cell_rlen[d] = rl[d]; }
        const int nc = n[0] * n[1] * n[2];
        if (nc + 1 > c->cell_alloc) {
            if (c->d_cell_cnt) (void)hipFree(c->d_cell_cnt);
            if (c->d_cell_start) (void)hipFree(c->d_cell_start);
            SMD_HIP(c, hipMalloc(&c->d_cell_cnt, sizeof(int32_t) * (nc + 1)));
            SMD_HIP(c, hipMalloc(&c->d_cell_start, sizeof(int32_t) * (nc + 1)));
            // the counting pass clears the counts after each use; they start at zero
            SMD_HIP(c, hipMemset(c->d_cell_cnt, 0, sizeof(int32_t) * (nc + 1)));
            c->cell_alloc = nc + 1;
        }
        // hit bits for neighbour cells of up to twice the mean occupancy
        const int words = (2 * c->n / nc + 31) / 32 + 1;
        if (words > c->list_mask_words) {
            if (c->d_list_mask) (void)hipFree(c->d_list_mask);
            c->d_list_mask = nullptr;
            c->list_mask_words = 0;
            const size_t bytes = sizeof(uint32_t) * 27 * (size_t)c->n * words;
            if (hipMalloc(&c->d_list_mask, bytes) != hipSuccess)
                return fail(c, SMD_ENOMEM, "pair-list mask allocation of %zu bytes failed", bytes);
            c->list_mask_words = words;
        }
    }
    // Row capacity: user value, else estimate from the density and grow on
    // overflow (the list is rebuilt until it fits).
    int32_t stride = c->want_stride;
    if (stride <= 0) {
        const double vol = c->box.L[0] * c->box.L[1] * c->box.L[2];
        const double rho = c->n / vol;
        const double half = rho * 4.0 / 3.0 * PI * len * len * len;   // full row
        stride = (int32_t)(2.5 * half) + 64;
        stride = (stride + 63) / 64 * 64;
        if (stride > c->n) stride = ((c->n + 63) / 64) * 64;
        if (stride < 64) stride = 64;
        if (c->stride > stride) stride = c->stride;
    }
    for (int attempt = 0; attempt < 6; ++attempt) {
        int rc = pairlist_alloc(c, stride);
        if (rc) return rc;
        SMD_HIP(c, hipMemsetAsync(c->d_status, 0, sizeof(unsigned), c->stream));
        rc = launch_list(c, nullptr, false);
        if (rc) return rc;
        unsigned st = 0;
        SMD_HIP(c, hipMemcpyAsync(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost, c->stream));
        SMD_HIP(c, hipStreamSynchronize(c->stream));
        if (!(st & ST_LIST_OVERFLOW)) {
            c->list_valid = true;
            return SMD_OK;
        }
        if (c->want_stride > 0) return fail(c, SMD_ECAPACITY, "pair list row capacity %d exceeded", stride);
        // grow to the observed maximum
        std::vector<int32_t> cnt(c->n);
        SMD_HIP(c, hipMemcpy(cnt.data(), c->d_nb_cnt, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost));
        int32_t mx = 0;
        for (int32_t v : cnt) mx = v > mx ? v : mx;
        stride = ((int32_t)(mx * 1.25) + 64 + 63) / 64 * 64;
    }
    return fail(c, SMD_ECAPACITY, "pair list did not fit after resizing");
}

}  // namespace smd

using namespace smd;

extern "C" int smd_build_pairlist(smd_ctx* c, double cutoff, double cut_buff, int use_cells) {
    if (!c) return SMD_EINVAL;
    if (!(cutoff > 0) || !(cut_buff >= 0)) return fail(c, SMD_EINVAL, "bad cutoff/buffer");
    return pairlist_build(c, cutoff, cut_buff, use_cells, false);
}

extern "C" int smd_set_pairlist_capacity(smd_ctx* c, int32_t row_capacity) {
    if (!c || row_capacity < 0) return SMD_EINVAL;
    c->want_stride = row_capacity;
    return SMD_OK;
}

extern "C" int smd_pairlist_info(smd_ctx* c, int64_t* n_pairs, int32_t* max_row, int32_t* used_cells) {
    if (!c) return SMD_EINVAL;
    if (!c->list_valid) return fail(c, SMD_ESTATE, "no pair list");
    std::vector<int32_t> cnt(c->n), own(c->n);
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemcpy(cnt.data(), c->d_nb_cnt, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost));
    SMD_HIP(c, hipMemcpy(own.data(), c->d_nb_own, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost));
    int64_t tot = 0;
    int32_t mx = 0;
    for (int i = 0; i < c->n; ++i) { tot += own[i]; mx = cnt[i] > mx ? cnt[i] : mx; }
    if (n_pairs) *n_pairs = tot;
    if (max_row) *max_row = mx;
    if (used_cells) *used_cells = c->list_cells ? 1 : 0;
    return SMD_OK;
}

// The reference's half list is the front segment of each device row.
extern "C" int smd_get_pairlist(smd_ctx* c, int32_t* offsets, int32_t* nb, int64_t cap) {
    if (!c || !offsets) return SMD_EINVAL;
    if (!c->list_valid) return fail(c, SMD_ESTATE, "no pair list");
    std::vector<int32_t> cnt(c->n), own(c->n);
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemcpy(cnt.data(), c->d_nb_cnt, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost));
    SMD_HIP(c, hipMemcpy(own.data(), c->d_nb_own, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost));
    offsets[0] = 0;
    for (int i = 0; i < c->n; ++i) {
        if (cnt[i] > c->stride) return fail(c, SMD_ECAPACITY, "row %d overflowed", i);
        offsets[i + 1] = offsets[i] + own[i];
    }
    if (offsets[c->n] > cap) return fail(c, SMD_ECAPACITY, "need %d entries, cap %lld", offsets[c->n], (long long)cap);
    if (!nb) return SMD_OK;
    std::vector<int32_t> rows((size_t)c->n * c->stride);
    SMD_HIP(c, hipMemcpy(rows.data(), c->d_nb, rows.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int i = 0; i < c->n; ++i) {
        const int32_t* r = rows.data() + (size_t)i * c->stride;
        std::copy(r, r + own[i], nb + offsets[i]);
    }
    return SMD_OK;
}
