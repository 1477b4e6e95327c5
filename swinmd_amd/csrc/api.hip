// api.hip -- context lifecycle, state transfer, errors and profiling for the
// C ABI declared in include/swinmd_hip.h.
#include <algorithm>
#include <cstdarg>
#include <cstring>

#include "smd_internal.h"

namespace smd {

int fail(Ctx* c, int code, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

int hip_fail(Ctx* c, hipError_t e, const char* what) {
    return fail(c, SMD_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

int read_terms(Ctx* c, double* dst) {
    SMD_HIP(c, hipMemcpyAsync(c->h_pinned, c->d_terms, sizeof(double) * T_COUNT, hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    std::memcpy(dst, c->h_pinned, sizeof(double) * T_COUNT);
    return SMD_OK;
}

int check_status(Ctx* c) {
    unsigned st = 0;
    SMD_HIP(c, hipMemcpy(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost));
    if (st & ST_LIST_OVERFLOW) return fail(c, SMD_ECAPACITY, "pair list row capacity %d exceeded", c->stride);
    if (st & ST_SHAKE_FAIL) return fail(c, SMD_ESHAKE, "Shake constraint failure");
    if (st & ST_SHAKE_NOCONV) return fail(c, SMD_ENOCONV, "SHAKE not converged");
    if (st & ST_NH_ITER) return fail(c, SMD_ENOCONV, "NH iteration error");
    return SMD_OK;
}

// One wave per value (rows_total's order).
__global__ __launch_bounds__(FINISH_THREADS) void k_red_finish(const double* __restrict__ part, int cap, int nb,
                                                               int nv, double* dst, int split, double* dst2) {
    rows_total(part, cap, nb, nv, dst, split, dst2);
}

int rows_finish(Ctx* c, hipStream_t s, const double* part, int stride, int nb, int nrows, double* dst, int split,
                double* dst2) {
    if (nb <= 0 || nrows <= 0) return SMD_OK;
    hipLaunchKernelGGL(k_red_finish, dim3(1), dim3(FINISH_THREADS), 0, s, part, stride, nb, nrows, dst, split, dst2);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

int red_finish(Ctx* c, int site, unsigned nblocks, int nv, double* dst) {
    if ((int)nblocks > c->red.cap)
        return fail(c, SMD_ECAPACITY, "reduction over %u blocks exceeds the arena's %d", nblocks, c->red.cap);
    if (nblocks == 0) return SMD_OK;
    if (c->red_defer) {
        for (const RedPending& q : c->red_pending)
            if (q.site == site) return fail(c, SMD_ESTATE, "reduction site %d queued twice in one step", site);
        c->red_pending.push_back(RedPending{site, nblocks, nv, dst});
        return SMD_OK;
    }
    hipLaunchKernelGGL(k_red_finish, dim3(1), dim3(FINISH_THREADS), 0, c->stream,
                       c->red.base + (size_t)site * RED_NV * c->red.cap, c->red.cap, (int)nblocks, nv, dst, 1 << 30,
                       (double*)nullptr);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

constexpr int FLUSH_MAX = 24;
struct FlushList {
    const double* part[FLUSH_MAX];
    double* dst[FLUSH_MAX];
    int nb[FLUSH_MAX];
    int nv[FLUSH_MAX];
    int cap;
};

// Block b adds queued finish b: one wave per value, as k_red_finish.
__global__ __launch_bounds__(1024) void k_red_flush(FlushList f) {
    const int b = blockIdx.x;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int k = w; k < f.nv[b]; k += blockDim.x / WAVE) {
        const double* p = f.part[b] + (size_t)k * f.cap;
        double s = 0.0;
#pragma unroll 8
        for (int i = lane; i < f.nb[b]; i += WAVE) s += p[i];
        s = wave_sum(s);
        if (lane == 0) f.dst[b][k] += s;
    }
}

int red_flush(Ctx* c) {
    const size_t n = c->red_pending.size();
    if (n == 0) return SMD_OK;
    if (n > (size_t)FLUSH_MAX) return fail(c, SMD_ESTATE, "%zu queued reductions exceed %d", n, FLUSH_MAX);
    FlushList f{};
    f.cap = c->red.cap;
    for (size_t q = 0; q < n; ++q) {
        const RedPending& r = c->red_pending[q];
        f.part[q] = c->red.base + (size_t)r.site * RED_NV * c->red.cap;
        f.dst[q] = r.dst;
        f.nb[q] = (int)r.nb;
        f.nv[q] = r.nv;
    }
    c->red_pending.clear();
    hipLaunchKernelGGL(k_red_flush, dim3((unsigned)n), dim3(1024), 0, c->stream, f);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

int zero_terms(Ctx* c) {
    SMD_HIP(c, hipMemsetAsync(c->d_terms, 0, sizeof(double) * T_COUNT, c->stream));
    return SMD_OK;
}

static void resolve_prof(Ctx* c) {
    if (c->prof.pending.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto& r : c->prof.pending) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, r.a, r.b);
        auto& e = c->prof.acc[r.fam];
        e.first += ms;
        e.second += 1;
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    c->prof.pending.clear();
}

static void release(Ctx* c) {
    void* ptrs[] = {c->d_type, c->d_mol, c->d_mol_start, c->d_q, c->d_mass, c->d_excl_off, c->d_excl, c->d_pos,
                    c->d_vel, c->d_vtm, c->d_disp, c->d_force, c->d_com, c->d_mom, c->d_bond2, c->d_terms,
                    c->d_status, c->d_flag, c->d_scalars, c->d_cell_of, c->d_cell_cnt, c->d_cell_start,
                    c->d_cell_atoms, c->d_list_seg, c->d_list_mask, c->d_nb, c->d_nb_cnt, c->d_nb_own, c->d_fvirt, c->d_kvec, c->d_kmod2, c->d_expk, c->d_kn, c->d_kgrp, c->d_ksplit, c->d_kf, c->d_e1,
                    c->d_sfac_part, c->d_vdw_a, c->d_vdw_b, c->d_vdw_c, c->d_epos, c->d_fa, c->d_fb,
                    c->d_host, c->d_dipl, c->d_quadl, c->d_polar, c->d_dipg, c->d_quadg, c->d_e0, c->d_mu,
                    c->d_cg, c->d_cg_part, c->d_cg_state, c->d_cg_scal, c->d_frecip, c->d_mp_trq,
                    c->d_mp_terms, c->mp_ks.kvec, c->mp_ks.kmod2, c->mp_ks.expk, c->mp_ks.part, c->mp_ks.kf,
                    c->mp_ks.fpart, c->mp_ks.kstep, c->ind_ks.kstep, c->ind_ks.kvec, c->ind_ks.kmod2, c->ind_ks.expk, c->ind_ks.part,
                    c->ind_ks.kf, c->ind_ks.fpart, c->d_pack, c->d_bd_bonds, c->d_bd_angles,
                    c->d_bd_bond_off, c->d_bd_angle_off, c->d_bd_par, c->d_bd_out, c->d_stress, c->d_real,
                    c->d_it_cnt, c->d_it_j, c->d_it_a, c->d_part, c->d_cg32, c->d_ind_rec32, c->d_ind_fix, c->d_e0_prev, c->d_vrec, c->d_pc_blk, c->d_shell32, c->d_polar_terms, c->d_site32, c->d_site32_ind, c->d_pind32,
                    c->red.base};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    mlp_release(c);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    if (c->h_cg_done) (void)hipHostFree(c->h_cg_done);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->cg_exec) (void)hipGraphExecDestroy(c->cg_exec);
    if (c->pmd_a) (void)hipGraphExecDestroy(c->pmd_a);
    if (c->pmd_c) (void)hipGraphExecDestroy(c->pmd_c);
    if (c->pmd_m) (void)hipGraphExecDestroy(c->pmd_m);
    if (c->pmd_a2) (void)hipGraphExecDestroy(c->pmd_a2);
    if (c->pmd_a3) (void)hipGraphExecDestroy(c->pmd_a3);
    if (c->pmd_m2) (void)hipGraphExecDestroy(c->pmd_m2);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (auto& r : c->prof.pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_cg) (void)hipEventDestroy(c->ev_cg);
    if (c->ev_fork2) (void)hipEventDestroy(c->ev_fork2);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

}  // namespace smd

using namespace smd;

extern "C" int smd_abi_version(void) { return SMD_ABI_VERSION; }

extern "C" const char* smd_strerror(int code) {
    switch (code) {
    case SMD_OK: return "ok";
    case SMD_EINVAL: return "invalid argument";
    case SMD_ENOMEM: return "device allocation failed";
    case SMD_EDEVICE: return "HIP runtime error";
    case SMD_ESTATE: return "call out of order";
    case SMD_ECAPACITY: return "pair list capacity exceeded";
    case SMD_ESHAKE: return "SHAKE constraint failure";
    case SMD_ENOCONV: return "iteration did not converge";
    case SMD_ENOTSUP: return "configuration not supported";
    case SMD_EIO: return "input file missing or malformed";
    default: return "unknown error";
    }
}

extern "C" const char* smd_last_error(const smd_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int smd_ctx_create(smd_ctx** out, const smd_topology* t, int device) {
    if (!out || !t) return SMD_EINVAL;
    *out = nullptr;
    if (t->n_atoms <= 0 || t->n_mols <= 0 || t->n_types <= 0 || !t->atom_type || !t->mol_id || !t->scaled_charge ||
        !t->mass || !t->excl_offsets)
        return SMD_EINVAL;
    const int n = t->n_atoms, nm = t->n_mols;
    // validate: types in range, molecules contiguous and non-decreasing
    std::vector<int32_t> ms(nm + 1, 0);
    for (int i = 0; i < n; ++i) {
        if (t->atom_type[i] < 0 || t->atom_type[i] >= t->n_types) return SMD_EINVAL;
        const int m = t->mol_id[i];
        if (m < 0 || m >= nm) return SMD_EINVAL;
        if (i > 0 && m < t->mol_id[i - 1]) return SMD_EINVAL;
        if (i > 0 && m > t->mol_id[i - 1] + 1) return SMD_EINVAL;
        ms[m + 1]++;
        if (!(t->mass[i] > 0)) return SMD_EINVAL;
    }
    if (t->mol_id[0] != 0 || t->mol_id[n - 1] != nm - 1) return SMD_EINVAL;
    for (int m = 0; m < nm; ++m) ms[m + 1] += ms[m];
    if (t->excl_offsets[0] != 0) return SMD_EINVAL;
    for (int i = 0; i < n; ++i)
        if (t->excl_offsets[i + 1] < t->excl_offsets[i]) return SMD_EINVAL;
    const int nex = t->excl_offsets[n];
    if (nex > 0 && !t->excl_atoms) return SMD_EINVAL;
    for (int k = 0; k < nex; ++k)
        if (t->excl_atoms[k] < 0 || t->excl_atoms[k] >= n) return SMD_EINVAL;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SMD_EDEVICE;
    if (hipSetDevice(device) != hipSuccess) return SMD_EDEVICE;

    smd_ctx* c = new smd_ctx();
    c->device = device;
    c->n = n;
    c->nmol = nm;
    c->ntypes = t->n_types;
    c->h_mol_start = ms;
    c->h_type.assign(t->atom_type, t->atom_type + n);
    c->h_mass.assign(t->mass, t->mass + n);
    c->h_q.assign(t->scaled_charge, t->scaled_charge + n);
    c->h_excl_off.assign(t->excl_offsets, t->excl_offsets + n + 1);
    c->h_excl.assign(t->excl_atoms ? t->excl_atoms : t->excl_offsets, t->excl_atoms ? t->excl_atoms + nex : t->excl_offsets);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SMD_EDEVICE;
    }
    bool ok = true;
    auto A = [&](auto*& p, size_t bytes) {
        if (!ok) return;
        if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) ok = false;
        else if (hipMemset(p, 0, bytes ? bytes : 8) != hipSuccess) ok = false;
    };
    const int nt2 = t->n_types * t->n_types;
    A(c->d_type, sizeof(int32_t) * n);
    A(c->d_mol, sizeof(int32_t) * n);
    A(c->d_mol_start, sizeof(int32_t) * (nm + 1));
    A(c->d_q, sizeof(double) * n);
    A(c->d_mass, sizeof(double) * n);
    A(c->d_excl_off, sizeof(int32_t) * (n + 1));
    A(c->d_excl, sizeof(int32_t) * (nex > 0 ? nex : 1));
    A(c->d_pos, sizeof(double) * 3 * n);
    A(c->d_vel, sizeof(double) * 3 * n);
    A(c->d_vtm, sizeof(double) * 3 * n);
    A(c->d_disp, sizeof(double) * 3 * n);
    A(c->d_real, sizeof(double) * 3 * n);
    A(c->d_force, sizeof(double) * 3 * n);
    A(c->d_com, sizeof(double) * 3 * nm);
    A(c->d_mom, sizeof(double) * 3 * nm);
    A(c->d_bond2, sizeof(double) * 3 * nm);
    A(c->d_terms, sizeof(double) * T_COUNT);
    A(c->d_status, sizeof(unsigned));
    A(c->d_flag, sizeof(int) * 8);
    c->nb_atoms = (n + 255) / 256;
    c->nb_mols = (nm + 255) / 256;
    c->nb_mols64 = (nm + 63) / 64;
    A(c->d_part, sizeof(double) * ((size_t)(4 + 16) * c->nb_atoms + (size_t)KIN_PARTS * c->nb_mols64 +
                                   (size_t)9 * c->nb_mols));
    if (ok) {
        c->part_vom = c->d_part;
        c->part_su = c->part_vom + (size_t)4 * c->nb_atoms;
        c->part_kin = c->part_su + (size_t)16 * c->nb_atoms;
        c->part_sh = c->part_kin + (size_t)KIN_PARTS * c->nb_mols64;
    }
    A(c->d_scalars, sizeof(double) * 8);
    // reduction arena of block_part: up to n/4 + 64 blocks per site (row
    // kernels with 64 lanes per atom), and at least 8,192 (per-k grids)
    c->red.cap = std::max((n + 3) / 4 + 64, 8192);
    A(c->red.base, sizeof(double) * red_doubles(c->red.cap));
    A(c->d_cell_of, sizeof(int32_t) * n);
    A(c->d_cell_atoms, sizeof(int32_t) * n);
    A(c->d_list_seg, sizeof(int32_t) * 28 * n);
    A(c->d_nb_cnt, sizeof(int32_t) * n);
    A(c->d_nb_own, sizeof(int32_t) * n);
    A(c->d_fvirt, sizeof(double) * 3 * n);
    A(c->d_vdw_a, sizeof(double) * nt2);
    A(c->d_vdw_b, sizeof(double) * nt2);
    A(c->d_vdw_c, sizeof(double) * nt2);
    A(c->d_epos, sizeof(double) * 3 * n);
    A(c->d_fa, sizeof(double) * n);
    A(c->d_fb, sizeof(double) * n);
    A(c->d_host, sizeof(int32_t) * n);
    if (ok && hipHostMalloc(&c->h_pinned, sizeof(double) * (T_COUNT + 64)) != hipSuccess) ok = false;
    if (!ok) {
        release(c);
        delete c;
        return SMD_ENOMEM;
    }
    hipError_t e = hipSuccess;
    auto H = [&](void* dst, const void* src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    H(c->d_type, t->atom_type, sizeof(int32_t) * n);
    H(c->d_mol, t->mol_id, sizeof(int32_t) * n);
    H(c->d_mol_start, ms.data(), sizeof(int32_t) * (nm + 1));
    H(c->d_q, t->scaled_charge, sizeof(double) * n);
    H(c->d_mass, t->mass, sizeof(double) * n);
    H(c->d_excl_off, t->excl_offsets, sizeof(int32_t) * (n + 1));
    if (nex > 0) H(c->d_excl, t->excl_atoms, sizeof(int32_t) * nex);
    if (e != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        release(c);
        delete c;
        return SMD_EDEVICE;
    }
    // this device's address of the allocation-free entry point's scratch
    if (static_scratch_init(c) != SMD_OK) {
        release(c);
        delete c;
        return SMD_EDEVICE;
    }
    *out = c;
    return SMD_OK;
}

extern "C" void smd_ctx_destroy(smd_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    delete c;
}

extern "C" int smd_sync(smd_ctx* c) {
    if (!c) return SMD_EINVAL;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return SMD_OK;
}

extern "C" int smd_set_box(smd_ctx* c, const double box[3]) {
    if (!c || !box) return SMD_EINVAL;
    for (int d = 0; d < 3; ++d) {
        if (!(box[d] > 0)) return fail(c, SMD_EINVAL, "box edge %d not positive", d);
        c->box.L[d] = box[d];
        c->box.h[d] = 0.5 * box[d];   // Ensemble::compute_bound_box halfL (src/Ensemble.cpp:431-433)
        c->box.inv[d] = 1.0 / box[d];
    }
    c->box_set = true;
    c->list_valid = false;
    c->ewald_ready = false;
    c->md_ready = false;
    c->ind_last_converged = false;   // the last solve's coupling table is of the old box
    return SMD_OK;
}

static int copy_in(smd_ctx* c, double* dst, const double* src, int64_t n, int64_t want) {
    if (!c || !src || n != want) return c ? fail(c, SMD_EINVAL, "count %lld != %lld", (long long)n, (long long)want) : SMD_EINVAL;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemcpy(dst, src, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    return SMD_OK;
}

static int copy_out(smd_ctx* c, double* dst, const double* src, int64_t n, int64_t want) {
    if (!c || !dst || n != want) return c ? fail(c, SMD_EINVAL, "count %lld != %lld", (long long)n, (long long)want) : SMD_EINVAL;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemcpy(dst, src, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    return SMD_OK;
}

// Ensemble::read_coordinate also starts the unwrapped positions there
// (atoms[i].realPos = position, src/Ensemble.cpp:667-668).
extern "C" int smd_set_positions(smd_ctx* c, const double* x, int64_t n) {
    int rc = copy_in(c, c ? c->d_pos : nullptr, x, n, c ? c->n : 0);
    if (rc == SMD_OK) rc = copy_in(c, c->d_real, x, n, c->n);
    if (rc == SMD_OK) c->ind_last_converged = false;   // the last solve no longer describes this state
    return rc;
}
extern "C" int smd_get_positions(smd_ctx* c, double* x, int64_t n) { return copy_out(c, x, c ? c->d_pos : nullptr, n, c ? c->n : 0); }
extern "C" int smd_set_velocities(smd_ctx* c, const double* v, int64_t n) { return copy_in(c, c ? c->d_vel : nullptr, v, n, c ? c->n : 0); }
extern "C" int smd_get_velocities(smd_ctx* c, double* v, int64_t n) { return copy_out(c, v, c ? c->d_vel : nullptr, n, c ? c->n : 0); }
extern "C" int smd_set_mol_centers(smd_ctx* c, const double* x, int64_t n) { return copy_in(c, c ? c->d_com : nullptr, x, n, c ? c->nmol : 0); }
extern "C" int smd_get_mol_centers(smd_ctx* c, double* x, int64_t n) { return copy_out(c, x, c ? c->d_com : nullptr, n, c ? c->nmol : 0); }
extern "C" int smd_get_forces(smd_ctx* c, double* f, int64_t n) { return copy_out(c, f, c ? c->d_force : nullptr, n, c ? c->n : 0); }

extern "C" int smd_zero_forces(smd_ctx* c) {
    if (!c) return SMD_EINVAL;
    SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    return SMD_OK;
}

extern "C" int smd_profile_enable(smd_ctx* c, int on) {
    if (!c) return SMD_EINVAL;
    resolve_prof(c);
    c->prof.on = on != 0;
    if (on) c->prof.acc.clear();
    return SMD_OK;
}

extern "C" int smd_set_tuning(smd_ctx* c, const char* name, int value) {
    if (!c || !name) return SMD_EINVAL;
    if (std::string(name) == "pair_lanes") {
        if (value != 0 && value != 4 && value != 8 && value != 16 && value != 32 && value != 64)
            return fail(c, SMD_EINVAL, "pair_lanes must be 0 (by precision), 4, 8, 16, 32 or 64 (got %d)", value);
        c->pair_lanes = value;
    } else if (std::string(name) == "pair_packed") {
        if (value != 0 && value != 1) return fail(c, SMD_EINVAL, "pair_packed must be 0 or 1 (got %d)", value);
        c->pair_packed = value;
    } else if (std::string(name) == "md_fused_shake") {
        if (value != 0 && value != 1) return fail(c, SMD_EINVAL, "md_fused_shake must be 0 or 1 (got %d)", value);
        c->md_fused_shake = value;
    } else if (std::string(name) == "cg_lanes") {
        if (value != 8 && value != 16 && value != 32) return fail(c, SMD_EINVAL, "cg_lanes must be 8, 16 or 32 (got %d)", value);
        c->cg_lanes = value;
        if (c->cg_exec) { (void)hipGraphExecDestroy(c->cg_exec); c->cg_exec = nullptr; }
    } else if (std::string(name) == "cg_batch") {
        if (value < 1 || value > 64) return fail(c, SMD_EINVAL, "cg_batch must be 1..64 (got %d)", value);
        c->cg_batch = value;
        if (c->cg_exec) { (void)hipGraphExecDestroy(c->cg_exec); c->cg_exec = nullptr; }
    } else if (std::string(name) == "cg_graph") {
        if (value != 0 && value != 1) return fail(c, SMD_EINVAL, "cg_graph must be 0 or 1 (got %d)", value);
        c->cg_graph = value;
        if (c->cg_exec) { (void)hipGraphExecDestroy(c->cg_exec); c->cg_exec = nullptr; }
    } else if (std::string(name) == "polar_m_sums") {
        if (value != 0 && value != 1) return fail(c, SMD_EINVAL, "polar_m_sums must be 0 or 1 (got %d)", value);
        c->polar_m_sums = value;
    } else if (std::string(name) == "polar_joint_long") {
        if (value != 0 && value != 1) return fail(c, SMD_EINVAL, "polar_joint_long must be 0 or 1 (got %d)", value);
        c->polar_joint = value;
    } else if (std::string(name) == "recip_overlap") {
        if (value != 0 && value != 1) return fail(c, SMD_EINVAL, "recip_overlap must be 0 or 1 (got %d)", value);
        c->recip_overlap = value;
    } else {
        return fail(c, SMD_EINVAL, "unknown tuning parameter '%s'", name);
    }
    if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    if (c->pmd_a) { (void)hipGraphExecDestroy(c->pmd_a); c->pmd_a = nullptr; }
    if (c->pmd_c) { (void)hipGraphExecDestroy(c->pmd_c); c->pmd_c = nullptr; }
    if (c->pmd_m) { (void)hipGraphExecDestroy(c->pmd_m); c->pmd_m = nullptr; }
    if (c->pmd_a2) { (void)hipGraphExecDestroy(c->pmd_a2); c->pmd_a2 = nullptr; }
    if (c->pmd_a3) { (void)hipGraphExecDestroy(c->pmd_a3); c->pmd_a3 = nullptr; }
    if (c->pmd_m2) { (void)hipGraphExecDestroy(c->pmd_m2); c->pmd_m2 = nullptr; }
    return SMD_OK;
}

extern "C" int smd_profile_get(smd_ctx* c, const char* fam, double* total_ms, int64_t* launches) {
    if (!c || !fam) return SMD_EINVAL;
    resolve_prof(c);
    auto it = c->prof.acc.find(fam);
    if (total_ms) *total_ms = it == c->prof.acc.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == c->prof.acc.end() ? 0 : it->second.second;
    return SMD_OK;
}
