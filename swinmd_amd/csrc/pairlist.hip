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
// Work mapping: one 64-lane wavefront per owner atom.  Lanes test 64
// candidates at a time; hits are compacted with a ballot so the row keeps the
// reference's order without atomics.  The candidate scan is HBM/L2 bound
// (positions are re-read from L2, 24 B per candidate).
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

__global__ void k_zero_i32(int32_t* p, int n, const int* gate) {
    if (gate && !*gate) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_cell_assign(int n, const double* __restrict__ pos, CellGrid g, int32_t* __restrict__ cell_of,
                              int32_t* __restrict__ cnt, const int* gate) {
    if (gate && !*gate) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dx = pos[3 * i] - g.origin[0];
    const double dy = pos[3 * i + 1] - g.origin[1];
    const double dz = pos[3 * i + 2] - g.origin[2];
    const int x = cell_coord(dx * g.rlen[0], g.n[0]);
    const int y = cell_coord(dy * g.rlen[1], g.n[1]);
    const int z = cell_coord(dz * g.rlen[2], g.n[2]);
    const int c = z * g.nxy + y * g.n[0] + x;
    cell_of[i] = c;
    atomicAdd(&cnt[c], 1);
}

// Exclusive scan of the cell counts by one block.
__global__ void k_cell_scan(int ncell, const int32_t* __restrict__ cnt, int32_t* __restrict__ start,
                            const int* gate) {
    if (gate && !*gate) return;
    __shared__ int32_t part[1024];
    const int t = threadIdx.x, T = blockDim.x;
    const int per = (ncell + T - 1) / T;
    const int lo = t * per, hi = min(ncell, lo + per);
    int32_t s = 0;
    for (int k = lo; k < hi; ++k) s += cnt[k];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int32_t run = 0;
        for (int k = 0; k < T; ++k) { int32_t v = part[k]; part[k] = run; run += v; }
        start[ncell] = run;
    }
    __syncthreads();
    int32_t run = part[t];
    for (int k = lo; k < hi; ++k) { start[k] = run; run += cnt[k]; }
}

// One wavefront per cell compacts its atoms in index order
// (set_atom appends in atom order, CellManager.cpp:145-170).
__global__ void k_cell_fill(int n, int ncell, const int32_t* __restrict__ cell_of, const int32_t* __restrict__ start,
                            int32_t* __restrict__ atoms, const int* gate) {
    if (gate && !*gate) return;
    const int c = blockIdx.x;
    if (c >= ncell) return;
    const int lane = threadIdx.x;
    int base = start[c];
    for (int a0 = 0; a0 < n; a0 += WAVE) {
        const int a = a0 + lane;
        const bool hit = (a < n) && (cell_of[a] == c);
        const uint64_t m = __ballot(hit);
        if (hit) atoms[base + __popcll(m & ((1ull << lane) - 1ull))] = a;
        base += __popcll(m);
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
};

struct Excl {
    int cnt;
    int ids[8];
    const int32_t* rest;
};

__device__ __forceinline__ bool excluded(const Excl& e, int j) {
    // Ensemble::exclusion_check (src/Ensemble.h:294-305)
    const int k8 = e.cnt < 8 ? e.cnt : 8;
    for (int k = 0; k < k8; ++k)
        if (e.ids[k] == j) return true;
    for (int k = 8; k < e.cnt; ++k)
        if (e.rest[k] == j) return true;
    return false;
}

// Test 64 candidates (one per lane) and append the hits to row i in lane order.
// Entries the reference's half list holds in row i are tagged with OWNED.
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
    const uint64_t m = __ballot(hit);
    if (hit) {
        const int slot = count + __popcll(m & ((1ull << lane) - 1ull));
        if (slot < a.stride) a.nb[(int64_t)i * a.stride + slot] = owned ? (int32_t)((uint32_t)j | OWNED) : j;
    }
    count += __popcll(m);
    n_owned += __popcll(__ballot(hit && owned));
}

__device__ __forceinline__ Excl load_excl(const ListArgs& a, int i) {
    Excl e;
    const int o = a.excl_off[i];
    e.cnt = a.excl_off[i + 1] - o;
    e.rest = a.excl + o;
#pragma unroll
    for (int k = 0; k < 8; ++k) e.ids[k] = (k < e.cnt) ? e.rest[k] : -1;
    return e;
}

__global__ __launch_bounds__(256) void k_pairs_cells(ListArgs a, CellGrid g, const int32_t* __restrict__ cell_of,
                                                     const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ catoms, const int* gate) {
    if (gate && !*gate) return;
    const int i = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (i >= a.n) return;
    const int lane = threadIdx.x % WAVE;
    const double px = a.pos[3 * i], py = a.pos[3 * i + 1], pz = a.pos[3 * i + 2];
    const Excl ex = load_excl(a, i);
    const int c = cell_of[i];
    int count = 0, n_owned = 0;
    // Same cell: every other member, no PBC (CellManager.cpp:257-279); the
    // reference's row holds the members after i.
    {
        const int s = start[c], e = start[c + 1];
        for (int k0 = s; k0 < e; k0 += WAVE) {
            const int k = k0 + lane;
            const int j = (k < e) ? catoms[k] : -1;
            test_and_append(a, i, px, py, pz, ex, (k < e) && (j != i), j, j > i, false, count, n_owned);
        }
    }
    // The 13 forward cells (owned, reference order), then the 13 backward
    // ones.  With at least 3 cells per edge the 26 neighbours are distinct, so
    // every pair of the reference's list appears once in each of its rows.
    const int cz = c / g.nxy, cy = (c % g.nxy) / g.n[0], cx = c % g.n[0];
    for (int s26 = 0; s26 < 26; ++s26) {
        const int sgn = s26 < 13 ? 1 : -1;
        const int s13 = s26 < 13 ? s26 : s26 - 13;
        int xx = cx + sgn * c_stencil[s13][0], yy = cy + sgn * c_stencil[s13][1], zz = cz + sgn * c_stencil[s13][2];
        if (xx >= g.n[0]) xx = 0; else if (xx < 0) xx = g.n[0] - 1;
        if (yy >= g.n[1]) yy = 0; else if (yy < 0) yy = g.n[1] - 1;
        if (zz >= g.n[2]) zz = 0; else if (zz < 0) zz = g.n[2] - 1;
        const int nc = zz * g.nxy + yy * g.n[0] + xx;
        const int s = start[nc], e = start[nc + 1];
        for (int k0 = s; k0 < e; k0 += WAVE) {
            const int k = k0 + lane;
            const int j = (k < e) ? catoms[k] : -1;
            test_and_append(a, i, px, py, pz, ex, k < e, j, s26 < 13, true, count, n_owned);
        }
    }
    if (lane == 0) {
        a.nb_cnt[i] = count;
        a.nb_own[i] = n_owned;
        if (count > a.stride) atomicOr(a.status, ST_LIST_OVERFLOW);
    }
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

static int launch_list(Ctx* c, const int* gate) {
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
        hipLaunchKernelGGL(k_zero_i32, dim3((nc + 256) / 256), dim3(256), 0, c->stream, c->d_cell_cnt, nc + 1, gate);
        hipLaunchKernelGGL(k_cell_assign, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n, c->d_pos, g,
                           c->d_cell_of, c->d_cell_cnt, gate);
        hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, c->stream, nc, c->d_cell_cnt, c->d_cell_start, gate);
        hipLaunchKernelGGL(k_cell_fill, dim3(nc), dim3(WAVE), 0, c->stream, c->n, nc, c->d_cell_of, c->d_cell_start,
                           c->d_cell_atoms, gate);
        hipLaunchKernelGGL(k_pairs_cells, grid, dim3(WAVE * waves_per_block), 0, c->stream, a, g, c->d_cell_of,
                           c->d_cell_start, c->d_cell_atoms, gate);
    } else {
        hipLaunchKernelGGL(k_pairs_brute, grid, dim3(WAVE * waves_per_block), 0, c->stream, a, gate);
    }
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// Builds the list.  With on_device_flag the kernels are gated on d_flag[0]
// (used inside the MD step, where the Verlet trigger is evaluated on the
// device); the grid must then already be configured.
int pairlist_build(Ctx* c, double cutoff, double cut_buff, int use_cells, bool on_device_flag) {
    if (on_device_flag) return launch_list(c, c->d_flag);
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
            c->cell_alloc = nc + 1;
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
        rc = launch_list(c, nullptr);
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

// The reference's half list is the OWNED subsequence of each device row.
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
        int o = offsets[i];
        for (int k = 0; k < cnt[i]; ++k) {
            const int32_t e = rows[(size_t)i * c->stride + k];
            if (e < 0) nb[o++] = e & INDEX_MASK;
        }
    }
    return SMD_OK;
}
