// bonded.hip -- intramolecular terms of flexible molecules on gfx950.
//
//   BondForce::compute          src/forces/BondForce.cpp:46-122
//   AngleForce::compute         src/forces/AngleForce.cpp:35-126
//   UreyBradleyForce::compute   src/forces/UreyBradleyForce.cpp:41-101
//
// Integrator::initialise creates the bond and angle terms when the bonds are
// not constrained and the Urey-Bradley term when the config asks for it
// (src/Integrator.cpp:162-192).  Every bond and angle joins atoms of one
// molecule (checked at set-up), so one thread per molecule walks that
// molecule's bonds and angles in topology order and adds into its own atoms'
// forces: no two threads touch one atom, no atomics, and each atom's sum
// over one term is in the reference's order.  The work is a few dozen FP64
// operations per term on data that stays in L2; these kernels are latency-
// bound, tiny next to the pair terms.
#include <algorithm>
#include <vector>

#include "smd_internal.h"

namespace smd {

enum { BD_E = 0, BD_D1, BD_D2, BD_VIR, BD_N = BD_VIR + 9 };   // per term: energy, Lustig, virial

struct BondedArgs {
    int nmol, potential;
    const int32_t* mol_start;
    const int4* items;        // bonds (a1, a2, type, -) or angles (a1, a2, a3, type), grouped by molecule
    const int32_t* off;       // [nmol+1] item range per molecule
    const double* par;        // per type: (r0, k) or (theta0, k)
    const double* pos;
    double* force;
    double* out;              // [BD_N]
    Red red;
    Box box;
    double kl, l0;
};

__device__ __forceinline__ double len3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

// BondForce::compute: harmonic 0.5 k d^2 or the AMOEBA quartic
// k d^2 (1 - 25.5 d + 379.3125 d^2), d = r - r0.
__global__ __launch_bounds__(256) void k_bonds(BondedArgs a) {
    double acc[BD_N];
#pragma unroll
    for (int k = 0; k < BD_N; ++k) acc[k] = 0.0;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < a.nmol) {
        const bool amoeba = a.potential == 2;
        const double c1 = amoeba ? 25.5 : 0.0, c2 = amoeba ? 379.3125 : 0.0;
        for (int b = a.off[m]; b < a.off[m + 1]; ++b) {
            const int4 it = a.items[b];
            const double r0 = a.par[2 * it.z], kf = a.par[2 * it.z + 1];
            double x = a.pos[3 * it.x] - a.pos[3 * it.y], y = a.pos[3 * it.x + 1] - a.pos[3 * it.y + 1],
                   z = a.pos[3 * it.x + 2] - a.pos[3 * it.y + 2];
            pbc(x, y, z, a.box);
            const double r = len3(x, y, z);
            const double diff = r - r0;
            double U, val;
            if (amoeba) {
                const double diff2 = diff * diff;
                const double anh = 1.0 - c1 * diff + c2 * diff2;
                U = kf * diff2 * anh;
                val = 2.0 * kf * diff * anh + kf * diff2 * (-c1 + c2 * 2.0 * diff);
            } else {
                U = 0.5 * kf * diff * diff;
                val = kf * diff;
            }
            const double s = -val / r;
            const double f[3] = {s * x, s * y, s * z};
            const double d[3] = {x, y, z};
            acc[BD_E] += U;
            acc[BD_D1] += val * r;
            acc[BD_D2] += -2.0 * r * val + r * r * kf;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) acc[BD_VIR + 3 * p + q] += f[p] * d[q];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                a.force[3 * it.x + p] += f[p];
                a.force[3 * it.y + p] -= f[p];
            }
        }
    }
    block_part<BD_N>(acc, a.red, RS_BOND);
}

// AngleForce::compute: harmonic 0.5 k dtheta^2 or the AMOEBA sextic in
// degrees, cos(theta) clamped at -0.9999999 as the reference does.
__global__ __launch_bounds__(256) void k_angles(BondedArgs a) {
    double acc[BD_N];
#pragma unroll
    for (int k = 0; k < BD_N; ++k) acc[k] = 0.0;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < a.nmol) {
        const bool amoeba = a.potential == 2;
        for (int g = a.off[m]; g < a.off[m + 1]; ++g) {
            const int4 it = a.items[g];
            const double theta0 = a.par[2 * it.w], kf = a.par[2 * it.w + 1];
            double r12[3], r32[3];
            for (int p = 0; p < 3; ++p) {
                r12[p] = a.pos[3 * it.y + p] - a.pos[3 * it.x + p];
                r32[p] = a.pos[3 * it.y + p] - a.pos[3 * it.z + p];
            }
            pbc(r12[0], r12[1], r12[2], a.box);
            pbc(r32[0], r32[1], r32[2], a.box);
            const double d12r = 1.0 / len3(r12[0], r12[1], r12[2]);
            const double d32r = 1.0 / len3(r32[0], r32[1], r32[2]);
            double cosT = (r12[0] * r32[0] + r12[1] * r32[1] + r12[2] * r32[2]) * d12r * d32r;
            cosT = fmax(cosT, -0.9999999);
            const double sinT = sqrt(1.0 - cosT * cosT);
            const double invSin = 1.0 / sinT;
            double diff = acos(cosT) - theta0;
            double U, v;
            if (amoeba) {
                const double deg = 180 / PI;
                const double c1 = 0.014, c2 = 5.6E-5, c3 = 7.0E-7, c4 = 2.2E-8;
                diff *= deg;
                const double diff2 = diff * diff, diff3 = diff * diff2, diff4 = diff2 * diff2;
                const double anh = 1.0 - c1 * diff + c2 * diff2 - c3 * diff3 + c4 * diff4;
                U = kf * diff2 * anh / (deg * deg);
                v = 2.0 * kf * diff * anh / deg + kf * diff2 * (-c1 + c2 * 2.0 * diff - c3 * 3.0 * diff2 + c4 * 4.0 * diff3) / deg;
            } else {
                U = 0.5 * kf * diff * diff;
                v = kf * diff;
            }
            double f1[3], f3[3];
            for (int p = 0; p < 3; ++p) {
                const double u12 = r12[p] * d12r, u32 = r32[p] * d32r;
                f1[p] = v * d12r * invSin * (u12 * cosT - u32);
                f3[p] = v * d32r * invSin * (u32 * cosT - u12);
            }
            acc[BD_E] += U;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) acc[BD_VIR + 3 * p + q] += f1[p] * r12[q] + f3[p] * r32[q];
            for (int p = 0; p < 3; ++p) {
                a.force[3 * it.x + p] += f1[p];
                a.force[3 * it.y + p] += -(f1[p] + f3[p]);
                a.force[3 * it.z + p] += f3[p];
            }
        }
    }
    block_part<BD_N>(acc, a.red, RS_ANGLE);
}

// UreyBradleyForce::compute: kl (r - l0)^2 between the first and third atom
// of every molecule (H1 and H2 of a water); its virial is accumulated with
// the opposite sign to the bond term's, as the reference does.
__global__ __launch_bounds__(256) void k_urey_bradley(BondedArgs a) {
    double acc[BD_N];
#pragma unroll
    for (int k = 0; k < BD_N; ++k) acc[k] = 0.0;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < a.nmol) {
        const int i1 = a.mol_start[m], i2 = i1 + 2;
        double x = a.pos[3 * i1] - a.pos[3 * i2], y = a.pos[3 * i1 + 1] - a.pos[3 * i2 + 1],
               z = a.pos[3 * i1 + 2] - a.pos[3 * i2 + 2];
        pbc(x, y, z, a.box);
        const double r = len3(x, y, z);
        const double diff = r - a.l0;
        const double U = a.kl * diff * diff;
        const double val = 2.0 * a.kl * diff;
        const double s = -val / r;
        const double f[3] = {s * x, s * y, s * z};
        const double d[3] = {x, y, z};
        acc[BD_E] = U;
        acc[BD_D1] = val * r;
        acc[BD_D2] = -2.0 * r * val + r * r * a.kl;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[BD_VIR + 3 * p + q] = -f[p] * d[q];
        for (int p = 0; p < 3; ++p) {
            a.force[3 * i1 + p] += f[p];
            a.force[3 * i2 + p] -= f[p];
        }
    }
    block_part<BD_N>(acc, a.red, RS_UB);
}

static int bd_alloc(Ctx* c, void** ptr, size_t bytes) {
    if (*ptr) (void)hipFree(*ptr);
    *ptr = nullptr;
    if (bytes == 0) return SMD_OK;
    if (hipMalloc(ptr, bytes) != hipSuccess) return fail(c, SMD_ENOMEM, "bonded-term allocation of %zu bytes", bytes);
    return SMD_OK;
}

// Groups items by the molecule of their first atom (stable: topology order
// within a molecule) after checking that all their atoms are in it.
static int group_by_molecule(Ctx* c, const int32_t* rows, int64_t n, int width, int ntypes, const char* what,
                             std::vector<int4>& items, std::vector<int32_t>& off) {
    std::vector<int32_t> mol_of(c->n);
    for (int m = 0; m < c->nmol; ++m)
        for (int i = c->h_mol_start[m]; i < c->h_mol_start[m + 1]; ++i) mol_of[i] = m;
    off.assign(c->nmol + 1, 0);
    for (int64_t k = 0; k < n; ++k) {
        const int32_t* r = rows + (size_t)k * width;
        for (int j = 0; j + 1 < width; ++j)
            if (r[j] < 0 || r[j] >= c->n) return fail(c, SMD_EINVAL, "%s %lld: atom index %d out of range", what, (long long)k, r[j]);
        if (r[width - 1] < 0 || r[width - 1] >= ntypes)
            return fail(c, SMD_EINVAL, "%s %lld: type %d out of range", what, (long long)k, r[width - 1]);
        for (int j = 1; j + 1 < width; ++j)
            if (mol_of[r[j]] != mol_of[r[0]])
                return fail(c, SMD_ENOTSUP, "%s %lld joins two molecules", what, (long long)k);
        off[mol_of[r[0]] + 1]++;
    }
    for (int m = 0; m < c->nmol; ++m) off[m + 1] += off[m];
    items.assign(n, make_int4(0, 0, 0, 0));
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int64_t k = 0; k < n; ++k) {
        const int32_t* r = rows + (size_t)k * width;
        items[fill[mol_of[r[0]]]++] = make_int4(r[0], r[1], r[2], width > 3 ? r[3] : 0);
    }
    return SMD_OK;
}

static int bonded_args(Ctx* c, BondedArgs& a, int term) {
    a.nmol = c->nmol;
    a.potential = c->bd_potential;
    a.mol_start = c->d_mol_start;
    a.pos = c->d_pos;
    a.force = c->d_force;
    a.box = c->box;
    a.kl = c->bd_kl;
    a.l0 = c->bd_l0;
    a.out = c->d_bd_out + BD_N * term;
    if (term == 0) { a.items = (const int4*)c->d_bd_bonds; a.off = c->d_bd_bond_off; a.par = c->d_bd_par; }
    if (term == 1) { a.items = (const int4*)c->d_bd_angles; a.off = c->d_bd_angle_off; a.par = c->d_bd_par + 2 * c->bd_n_bond_types; }
    return SMD_OK;
}

static int bonded_run(Ctx* c, int term, double* f_accum, smd_bonded_out* out) {
    if (!c->bonded_ready) return fail(c, SMD_ESTATE, "smd_bonded_setup not called");
    if (!c->box_set) return fail(c, SMD_ESTATE, "box not set");
    if (term == 2 && !c->bd_ub_ok) return fail(c, SMD_ENOTSUP, "Urey-Bradley term needs at least 3 atoms per molecule");
    BondedArgs a{};
    bonded_args(c, a, term);
    a.red = c->red;
    if (f_accum) SMD_HIP(c, hipMemsetAsync(c->d_force, 0, sizeof(double) * 3 * c->n, c->stream));
    SMD_HIP(c, hipMemsetAsync(a.out, 0, sizeof(double) * BD_N, c->stream));
    const dim3 grid((c->nmol + 255) / 256);
    if (term == 0) hipLaunchKernelGGL(k_bonds, grid, dim3(256), 0, c->stream, a);
    else if (term == 1) hipLaunchKernelGGL(k_angles, grid, dim3(256), 0, c->stream, a);
    else hipLaunchKernelGGL(k_urey_bradley, grid, dim3(256), 0, c->stream, a);
    SMD_HIP(c, hipGetLastError());
    {
        const int rc = red_finish(c, term == 0 ? RS_BOND : term == 1 ? RS_ANGLE : RS_UB, grid.x, BD_N, a.out);
        if (rc) return rc;
    }
    double h[BD_N];
    SMD_HIP(c, hipMemcpyAsync(h, a.out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    std::vector<double> f;
    if (f_accum) {
        f.resize(3 * (size_t)c->n);
        SMD_HIP(c, hipMemcpyAsync(f.data(), c->d_force, f.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < f.size(); ++k) f_accum[k] += f[k];
    if (out) {
        out->energy = h[BD_E];
        out->deriv = h[BD_D1];
        out->deriv2 = h[BD_D2];
        for (int k = 0; k < 9; ++k) out->virial[k] = h[BD_VIR + k];
    }
    return SMD_OK;
}

}  // namespace smd

using namespace smd;

extern "C" int smd_bonded_setup(smd_ctx* c, const smd_bonded_params* p) {
    if (!c || !p) return SMD_EINVAL;
    if (p->potential != 1 && p->potential != 2)
        return fail(c, SMD_EINVAL, "potential %d: 1 (harmonic) or 2 (AMOEBA) expected", p->potential);
    if (p->n_bonds < 0 || p->n_angles < 0 || p->n_bond_types < 0 || p->n_angle_types < 0 ||
        (p->n_bonds && (!p->bonds || !p->bond_r0 || !p->bond_k)) ||
        (p->n_angles && (!p->angles || !p->angle_theta0 || !p->angle_k)))
        return fail(c, SMD_EINVAL, "bonded-term arrays missing");
    std::vector<int4> bonds, angles;
    std::vector<int32_t> boff, aoff;
    int rc = group_by_molecule(c, p->bonds, p->n_bonds, 3, p->n_bond_types, "bond", bonds, boff);
    if (rc) return rc;
    rc = group_by_molecule(c, p->angles, p->n_angles, 4, p->n_angle_types, "angle", angles, aoff);
    if (rc) return rc;
    std::vector<double> par(2 * (size_t)(p->n_bond_types + p->n_angle_types));
    for (int t = 0; t < p->n_bond_types; ++t) { par[2 * t] = p->bond_r0[t]; par[2 * t + 1] = p->bond_k[t]; }
    for (int t = 0; t < p->n_angle_types; ++t) {
        par[2 * (p->n_bond_types + t)] = p->angle_theta0[t];
        par[2 * (p->n_bond_types + t) + 1] = p->angle_k[t];
    }
    c->bonded_ready = false;
    if ((rc = bd_alloc(c, (void**)&c->d_bd_bonds, sizeof(int4) * std::max<size_t>(1, bonds.size())))) return rc;
    if ((rc = bd_alloc(c, (void**)&c->d_bd_angles, sizeof(int4) * std::max<size_t>(1, angles.size())))) return rc;
    if ((rc = bd_alloc(c, (void**)&c->d_bd_bond_off, sizeof(int32_t) * boff.size()))) return rc;
    if ((rc = bd_alloc(c, (void**)&c->d_bd_angle_off, sizeof(int32_t) * aoff.size()))) return rc;
    if ((rc = bd_alloc(c, (void**)&c->d_bd_par, sizeof(double) * std::max<size_t>(2, par.size())))) return rc;
    if ((rc = bd_alloc(c, (void**)&c->d_bd_out, sizeof(double) * 3 * BD_N))) return rc;
    if (!bonds.empty())
        SMD_HIP(c, hipMemcpy(c->d_bd_bonds, bonds.data(), sizeof(int4) * bonds.size(), hipMemcpyHostToDevice));
    if (!angles.empty())
        SMD_HIP(c, hipMemcpy(c->d_bd_angles, angles.data(), sizeof(int4) * angles.size(), hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_bd_bond_off, boff.data(), sizeof(int32_t) * boff.size(), hipMemcpyHostToDevice));
    SMD_HIP(c, hipMemcpy(c->d_bd_angle_off, aoff.data(), sizeof(int32_t) * aoff.size(), hipMemcpyHostToDevice));
    if (!par.empty()) SMD_HIP(c, hipMemcpy(c->d_bd_par, par.data(), sizeof(double) * par.size(), hipMemcpyHostToDevice));
    c->bd_potential = p->potential;
    c->bd_n_bond_types = p->n_bond_types;
    c->bd_kl = p->kl_ub;
    c->bd_l0 = p->l0_ub;
    c->bd_ub_ok = true;
    for (int m = 0; m < c->nmol; ++m)
        if (c->h_mol_start[m + 1] - c->h_mol_start[m] < 3) c->bd_ub_ok = false;
    c->bonded_ready = true;
    return SMD_OK;
}

extern "C" int smd_bond_compute(smd_ctx* c, double* f_accum, smd_bonded_out* out) {
    if (!c) return SMD_EINVAL;
    return bonded_run(c, 0, f_accum, out);
}

extern "C" int smd_angle_compute(smd_ctx* c, double* f_accum, smd_bonded_out* out) {
    if (!c) return SMD_EINVAL;
    return bonded_run(c, 1, f_accum, out);
}

extern "C" int smd_ub_compute(smd_ctx* c, double* f_accum, smd_bonded_out* out) {
    if (!c) return SMD_EINVAL;
    return bonded_run(c, 2, f_accum, out);
}
