/*
 * oracle.c -- CPU restatement of SwinMD's nonbonded force path (test
 * infrastructure; see oracle.h).  Loop order and floating-point expression
 * order follow the cited reference routines so that results agree with the
 * compiled reference to rounding.  Build: gcc -O2 (no -ffast-math, no FMA
 * target), like the reference's own Makefile.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846;       /* src/NEMD_defs.h:77 */
static const double SQRT_PI = 1.7724538509055160273;   /* src/NEMD_defs.h:78 */
static const double BOLTZ = 0.0083145112;              /* src/NEMD_defs.h:108 */

enum { XX = 0, XY, XZ, YX, YY, YZ, ZX, ZY, ZZ };

/* ---- Vector3 helpers with the reference's operator semantics
 *      (src/utils/Vector3.h:108-193) ---------------------------------------- */
typedef struct { double x, y, z; } v3;

static inline v3 v3_ld(const double *p) { v3 r = {p[0], p[1], p[2]}; return r; }
static inline void v3_st(double *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline v3 v3_sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline v3 v3_add(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline v3 v3_scale(double f, v3 a) { v3 r = {f * a.x, f * a.y, f * a.z}; return r; }
static inline v3 v3_div(v3 a, double f) { v3 r = {a.x / f, a.y / f, a.z / f}; return r; }
static inline double v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double v3_len2(v3 a) { return (a.x * a.x + a.y * a.y + a.z * a.z); }
static inline double v3_len(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline void acc(double *p, v3 a) { p[0] += a.x; p[1] += a.y; p[2] += a.z; }
static inline void dec(double *p, v3 a) { p[0] -= a.x; p[1] -= a.y; p[2] -= a.z; }

/* Ensemble::apply_pbc (src/Ensemble.h:319-334): half-open [-L/2, L/2).  With
 * a Lees-Edwards tilt (the periodic image at +L_y sits at +tilt in x; the
 * sheared replicas, which the reference does not have) the y image is taken
 * first and moves x by the tilt; with tilt 0 every value is the reference's. */
static inline void apply_pbc(v3 *r, const orc_sys *s) {
    const double *box = s->box;
    const double hx = 0.5 * box[0], hy = 0.5 * box[1], hz = 0.5 * box[2];
    while (r->y >= hy) { r->y -= box[1]; r->x -= s->tilt; }
    while (r->y < -hy) { r->y += box[1]; r->x += s->tilt; }
    while (r->x >= hx) r->x -= box[0];
    while (r->x < -hx) r->x += box[0];
    while (r->z >= hz) r->z -= box[2];
    while (r->z < -hz) r->z += box[2];
}

/* Ensemble::exclusion_check (src/Ensemble.h:294-305). */
static inline int excluded(const orc_sys *s, int a, int b) {
    for (int k = s->excl_off[a]; k < s->excl_off[a + 1]; ++k)
        if (s->excl[k] == b) return 1;
    return 0;
}

static inline void vir_add(double *v, v3 f, v3 r) {
    /* virial[XX] += f.x*r.x; virial[XY] += f.y*r.x; ... (LJForce.cpp:206-214) */
    v[XX] += f.x * r.x; v[XY] += f.y * r.x; v[XZ] += f.z * r.x;
    v[YX] += f.x * r.y; v[YY] += f.y * r.y; v[YZ] += f.z * r.y;
    v[ZX] += f.x * r.z; v[ZY] += f.y * r.z; v[ZZ] += f.z * r.z;
}

/* ========================================================================== */
/* (a) cells and pair list                                                    */
/* ========================================================================== */

int orc_cell_setup(double len, const double box[3], int32_t n[3],
                   double cell_len[3], double rcell_len[3]) {
    /* CellManager::set_cellManager (src/utils/CellManager.cpp:43-84) */
    if ((len <= 0) || (len > box[0] / 3.0) || (len > box[1] / 3.0) ||
        (len > box[2] / 3.0))
        return 0;
    for (int d = 0; d < 3; ++d) {
        n[d] = (int32_t)(box[d] / len);
        cell_len[d] = box[d] / n[d];
        rcell_len[d] = 1.0 / cell_len[d];
    }
    return 1;
}

/* The 13 forward neighbour shifts, in the order of
 * CellManager::build_cell_neighbor (src/utils/CellManager.h:161-186). */
static const int STENCIL[13][3] = {
    {0, 0, -1}, {1, 0, -1}, {1, 0, 0},  {1, 0, 1},  {-1, 1, -1},
    {0, 1, -1}, {1, 1, -1}, {-1, 1, 0}, {0, 1, 0},  {1, 1, 0},
    {-1, 1, 1}, {0, 1, 1},  {1, 1, 1}};

static inline int wrap_idx(int v, int n) {
    /* xPlus / xMinus with wrap (CellManager.h:148-159) */
    if (v >= n) return 0;
    if (v < 0) return n - 1;
    return v;
}

typedef struct { int32_t owner, j; } pair_t;

static int64_t scatter_rows(int n, const pair_t *tmp, int64_t np, int32_t *off,
                            int32_t *nb) {
    /* Rows in atom order; within a row, the visit order is kept. */
    memset(off, 0, sizeof(int32_t) * (size_t)(n + 1));
    for (int64_t k = 0; k < np; ++k) off[tmp[k].owner + 1]++;
    for (int i = 0; i < n; ++i) off[i + 1] += off[i];
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) fill[i] = off[i];
    for (int64_t k = 0; k < np; ++k) nb[fill[tmp[k].owner]++] = tmp[k].j;
    free(fill);
    return np;
}

int64_t orc_pairlist(const orc_sys *s, double cutoff, double cut_buff,
                     int use_cells, int32_t *off, int32_t *nb, int64_t cap,
                     int *used_cells) {
    const int n = s->n_atoms;
    /* Ensemble::set_pairlist (src/Ensemble.cpp:454-473) */
    double pair_cut = cutoff + cut_buff;
    pair_cut *= pair_cut;
    const double cell_len_in = cutoff + cut_buff;
    int32_t nc[3];
    double clen[3], rclen[3];
    int cells = use_cells && orc_cell_setup(cell_len_in, s->box, nc, clen, rclen);
    if (used_cells) *used_cells = cells;

    int64_t tcap = cap > 0 ? cap : 0;
    pair_t *tmp = (pair_t *)malloc(sizeof(pair_t) * (size_t)(tcap > 0 ? tcap : 1));
    int64_t np = 0;
#define PUSH(o, jj)                                   \
    do {                                              \
        if (np < tcap) { tmp[np].owner = (o); tmp[np].j = (jj); } \
        ++np;                                         \
    } while (0)

    if (cells) {
        const int ncell = nc[0] * nc[1] * nc[2];
        const int nxy = nc[0] * nc[1];
        const double origin[3] = {-s->box[0] / 2.0, -s->box[1] / 2.0, -s->box[2] / 2.0};
        /* set_atom/find_cell (CellManager.h:91-121, .cpp:145-170): atoms are
         * appended to their cell in index order. */
        int32_t *cell_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        int32_t *cnt = (int32_t *)calloc((size_t)ncell + 1, sizeof(int32_t));
        for (int i = 0; i < n; ++i) {
            v3 d = v3_sub(v3_ld(s->pos + 3 * i), v3_ld(origin));
            int x = (int)(d.x * rclen[0]);
            int y = (int)(d.y * rclen[1]);
            int z = (int)(d.z * rclen[2]);
            if (x >= nc[0]) x = nc[0] - 1; else if (x < 0) x = 0;
            if (y >= nc[1]) y = nc[1] - 1; else if (y < 0) y = 0;
            if (z >= nc[2]) z = nc[2] - 1; else if (z < 0) z = 0;
            cell_of[i] = z * nxy + y * nc[0] + x;
            cnt[cell_of[i] + 1]++;
        }
        for (int c = 0; c < ncell; ++c) cnt[c + 1] += cnt[c];
        int32_t *members = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncell);
        for (int c = 0; c < ncell; ++c) fill[c] = cnt[c];
        for (int i = 0; i < n; ++i) members[fill[cell_of[i]]++] = i;
        free(fill);

        /* CellManager::update_pair_list (src/utils/CellManager.cpp:216-314) */
        for (int c = 0; c < ncell; ++c) {
            const int cz = c / nxy, cy = (c % nxy) / nc[0], cx = c % nc[0];
            int nbr[13];
            for (int k = 0; k < 13; ++k) {
                int xx = wrap_idx(cx + STENCIL[k][0], nc[0]);
                int yy = wrap_idx(cy + STENCIL[k][1], nc[1]);
                int zz = wrap_idx(cz + STENCIL[k][2], nc[2]);
                nbr[k] = zz * nxy + yy * nc[0] + xx;
            }
            for (int a = cnt[c]; a < cnt[c + 1]; ++a) {
                const int i = members[a];
                const v3 pi = v3_ld(s->pos + 3 * i);
                for (int b = a + 1; b < cnt[c + 1]; ++b) {
                    const int j = members[b];
                    if (!excluded(s, i, j)) {
                        v3 d = v3_sub(v3_ld(s->pos + 3 * j), pi);   /* no PBC in-cell */
                        if (v3_len2(d) < pair_cut) PUSH(i, j);
                    }
                }
                for (int k = 0; k < 13; ++k) {
                    const int cc = nbr[k];
                    for (int b = cnt[cc]; b < cnt[cc + 1]; ++b) {
                        const int j = members[b];
                        if (!excluded(s, i, j)) {
                            v3 d = v3_sub(v3_ld(s->pos + 3 * j), pi);
                            apply_pbc(&d, s);
                            if (v3_len2(d) < pair_cut) PUSH(i, j);
                        }
                    }
                }
            }
        }
        free(cell_of); free(cnt); free(members);
    } else {
        /* Ensemble::build_pairlist (src/Ensemble.cpp:492-525) */
        for (int i = 0; i < n - 1; ++i) {
            const v3 pi = v3_ld(s->pos + 3 * i);
            for (int j = i + 1; j < n; ++j) {
                if (!excluded(s, i, j)) {
                    v3 d = v3_sub(v3_ld(s->pos + 3 * j), pi);
                    apply_pbc(&d, s);
                    if (v3_len2(d) < pair_cut) PUSH(i, j);
                }
            }
        }
    }
#undef PUSH
    if (np > cap) { free(tmp); return -np; }
    scatter_rows(n, tmp, np, off, nb);
    free(tmp);
    return np;
}

/* ========================================================================== */
/* (b) van der Waals                                                          */
/* ========================================================================== */

void orc_buff_table(int n_types, const double *type_sigma,
                    const double *type_eps, double *r_ij, double *eps_ij) {
    /* BuffvdWForce::set_buff_param (src/forces/BuffvdWForce.cpp:392-433) */
    for (int i = 0; i < n_types; ++i) {
        const double epsi = type_eps[i], sigmai = type_sigma[i];
        for (int j = 0; j < n_types; ++j) {
            const double epsj = type_eps[j], sigmaj = type_sigma[j];
            const double sigmai2 = sigmai * sigmai, sigmaj2 = sigmaj * sigmaj;
            const double epssqrt = sqrt(epsi) + sqrt(epsj);
            double R, e;
            if (epsi == 0 || epsj == 0) { R = 0.0000001; e = 0; }
            else {
                R = (sigmai2 * sigmai + sigmaj2 * sigmaj) / (sigmai2 + sigmaj2);
                e = (4.0 * epsi * epsj) / (epssqrt * epssqrt);
            }
            r_ij[i * n_types + j] = R;
            eps_ij[i * n_types + j] = e;
        }
    }
}

void orc_lj_lrc(const orc_sys *s, int n_types, const double *sigma2,
                const double *eps4, double cutoff, double *e_lrc,
                double *v_lrc, double *du_lrc, double *d2u_lrc) {
    /* LJForce::long_range_correct (src/forces/LJForce.cpp:260-370).  The
     * reference leaves csix uninitialised (:268); it is started at 0 here. */
    const int n = s->n_atoms;
    int *num = (int *)calloc((size_t)n_types, sizeof(int));
    for (int i = 0; i < n; ++i) num[s->type[i]] += 1;
    const double cutoff3 = cutoff * cutoff * cutoff;
    const double volrc = s->box[0] * s->box[1] * s->box[2];
    const double invvol = 1.0 / volrc;
    double csix = 0.0;
    int npair = 0, nexcl = 0;
    for (int i = 0; i < n_types; ++i)
        for (int j = i; j < n_types; ++j) {
            const double sg2 = sigma2[i * n_types + j], e4 = eps4[i * n_types + j];
            const double sigma6 = sg2 * sg2 * sg2;
            int npair_ij;
            if (i != j) npair_ij = num[i] * num[j];
            else npair_ij = (int)(num[i] * (num[i] - 1) * 0.5);
            csix += npair_ij * e4 * sigma6;
            npair += npair_ij;
        }
    for (int i = 0; i < n; ++i)
        for (int k = s->excl_off[i]; k < s->excl_off[i + 1]; ++k) {
            const int j = s->excl[k];
            if (j < i) continue;
            const int t = s->type[i] * n_types + s->type[j];
            const double sigma6 = sigma2[t] * sigma2[t] * sigma2[t];
            csix -= eps4[t] * sigma6;
            nexcl += 1;
        }
    if (npair - nexcl <= 0) csix = 0;
    else csix /= npair - nexcl;
    *e_lrc = -2.0 * PI * n * n * invvol * csix / (3.0 * cutoff3);
    *du_lrc = 4.0 * PI * n * n * invvol * invvol * csix / (3.0 * cutoff3);
    *d2u_lrc = -4.0 * PI * n * n * invvol * invvol * invvol * csix / cutoff3;
    *v_lrc = -2.0 * (*e_lrc);
    free(num);
}

/* GROMACS-shift Coulomb used by LJ/Mie/Buckingham when computeCoulomb is on
 * (LJForce.cpp:188-201). */
static inline void coulomb_shift(double qij, double r, double r_1, double rij2,
                                 double rij_2, double cut, double rcut_2,
                                 double *elec, double *tmpf) {
    double e = qij * r_1;
    double f = e * rij_2;
    double r_rcut = r / cut;
    double r4_rcut4 = rij2 * rij2 * rcut_2 * rcut_2;
    *elec += e * (1 - 1.666667 * r_rcut + 1.666667 * r4_rcut4 - r_rcut * r4_rcut4);
    *tmpf += f * (1 - 5 * r4_rcut4 + 4 * r_rcut * r4_rcut4);
}

static int vdw_lj_mie(const orc_sys *s, const int32_t *off, const int32_t *nb,
                      const orc_vdw_params *p, double *force, orc_vdw_out *o) {
    const int is_mie = (p->kind == ORC_MIE);
    const double cutOff = p->cutoff, cutOff2 = cutOff * cutOff;
    double switchDist = cutOff2;
    if (p->switch_on) switchDist = p->switch_dist * p->switch_dist;
    double c1 = 0.0, c2 = 0.0, c3 = 0.0;
    if (p->switch_on && (cutOff2 > switchDist)) {
        c1 = 1.0 / (cutOff2 - switchDist);
        c1 = c1 * c1 * c1;
        c2 = cutOff2 - 3.0 * switchDist;
        c3 = 4 * c1;
    }
    double Cn = 0.0;
    if (is_mie) {   /* MieForce ctor (src/forces/MieForce.cpp:32-35) */
        double Cbase = p->n_mie / p->m_mie;
        double Cexp = p->m_mie / (p->n_mie - p->m_mie);
        double Cpow = pow(Cbase, Cexp);
        Cn = (p->n_mie / (p->n_mie - p->m_mie)) * Cpow;
    }
    const double rcut_2 = 1.0 / cutOff2;
    double energy = 0.0, ulus = 0.0, dudv = 0.0, d2udv = 0.0, elecE = 0.0;
    double vir[9] = {0};
    const double voll = s->box[0] * s->box[1] * s->box[2];
    const double vol2 = voll * voll;
    for (int i = 0; i < s->n_atoms; ++i) {
        const v3 rim = v3_ld(s->com + 3 * s->mol[i]);
        const double qi = s->q[i];
        const v3 pi = v3_ld(s->pos + 3 * i);
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int j = nb[k];
            if (s->mol[i] == s->mol[j]) continue;
            v3 rij = v3_sub(v3_ld(s->pos + 3 * j), pi);
            v3 rijm = v3_sub(v3_ld(s->com + 3 * s->mol[j]), rim);
            apply_pbc(&rij, s);
            apply_pbc(&rijm, s);
            const double r = v3_len(rij);
            const double rij2 = r * r;
            if (!(rij2 <= cutOff2)) continue;
            const int t = s->type[i] * p->n_types + s->type[j];
            double tmpE, tmpf, dulj = 0, d2ulj = 0;
            const double r_1 = 1.0 / r;
            const double rij_2 = r_1 * r_1;
            if (!is_mie) {
                /* LJForce::compute (src/forces/LJForce.cpp:122-176) */
                const double sigma2 = p->sigma2[t], eps4 = p->eps4[t];
                const double sr2 = sigma2 * rij_2;
                const double sr6 = sr2 * sr2 * sr2;
                const double sr12 = sr6 * sr6;
                tmpE = eps4 * (sr12 - sr6);
                tmpf = (sr12 + sr12 - sr6) * 6 * eps4 * rij_2;
                const double rijRij = p->use_constraint ? v3_dot(rij, rijm) : rij2;
                if (p->lustig) {
                    dulj = tmpf * rijRij;
                    d2ulj = eps4 * (156.0 * sr12 - 42.0 * sr6) * rijRij * rijRij * rij_2 * rij_2 + 2.0 * dulj;
                }
                if (p->switch_on) {   /* energy shift only (:152-168) */
                    const double sr2cut = sigma2 * rcut_2;
                    const double sr6cut = sr2cut * sr2cut * sr2cut;
                    const double sr12cut = sr6cut * sr6cut;
                    const double Eljcut = eps4 * (sr12cut - sr6cut);
                    tmpE = tmpE - Eljcut;
                }
            } else {
                /* MieForce::compute (src/forces/MieForce.cpp:120-164) */
                const double sigma = sqrt(p->sigma2[t]);
                const double eps = 0.250 * p->eps4[t];
                const double sigrij = sigma * r_1;
                const double srn = pow(sigrij, p->n_mie);
                const double srm = pow(sigrij, p->m_mie);
                tmpE = Cn * eps * (srn - srm);
                tmpf = Cn * eps * (p->n_mie * srn - p->m_mie * srm) * rij_2;
                const double rijRij = v3_dot(rij, rijm);
                if (p->lustig) {
                    dulj = tmpf * rijRij;
                    d2ulj = Cn * eps * (p->n_mie * (p->n_mie + 1) * srn - p->m_mie * (p->m_mie + 1) * srm) *
                                rijRij * rijRij * rij_2 * rij_2 + 2.0 * dulj;
                }
                if (p->switch_on && (rij2 > switchDist)) {
                    const double c = cutOff2 - rij2;
                    const double c4 = c * (2 * rij2 + c2);
                    const double sv = c * c4 * c1;
                    const double dsv = c3 * (c * c - c4);
                    tmpf = tmpf * sv - tmpE * dsv;
                    tmpE *= sv;
                }
            }
            energy += tmpE;
            if (p->lustig) { ulus += tmpE; dudv += dulj; d2udv += d2ulj; }
            if (p->coulomb)
                coulomb_shift(qi * s->q[j], r, r_1, rij2, rij_2, cutOff, rcut_2, &elecE, &tmpf);
            const v3 fij = v3_scale(-tmpf, rij);
            acc(force + 3 * i, fij);
            dec(force + 3 * j, fij);
            vir_add(vir, fij, rij);
        }
    }
    energy += p->e_lrc;
    vir[XX] += p->v_lrc; vir[YY] += p->v_lrc; vir[ZZ] += p->v_lrc;
    o->energy = energy;
    o->lj_energy = energy + elecE;
    memcpy(o->virial, vir, sizeof vir);
    o->ulus = o->dudv = o->d2udv = 0.0;
    if (p->lustig) {
        dudv = -dudv / (3.0 * voll);
        d2udv = d2udv / (9.0 * vol2);
        o->ulus = ulus + p->e_lrc;
        o->dudv = dudv + p->du_lrc;
        o->d2udv = d2udv + p->d2u_lrc;
    }
    return 0;
}

static int vdw_buck(const orc_sys *s, const int32_t *off, const int32_t *nb,
                    const orc_vdw_params *p, double *force, orc_vdw_out *o) {
    /* BuckinghamForce::compute (src/forces/BuckinghamForce.cpp:79-287) with
     * VirtualPositions (:372-398).  Every molecule is treated as CO2 with
     * sites O0, O1, C2 (types 0, 1, 2). */
    const double cutOff = p->cutoff, cutOff2 = cutOff * cutOff;
    double switchDist = cutOff2;
    if (p->switch_on) switchDist = p->switch_dist * p->switch_dist;
    double c1 = 0.0, c2 = 0.0, c3 = 0.0;
    if (p->switch_on && (cutOff2 > switchDist)) {
        c1 = 1.0 / (cutOff2 - switchDist);
        c1 = c1 * c1 * c1;
        c2 = cutOff2 - 3.0 * switchDist;
        c3 = 4 * c1;
    }
    const double rcut_2 = 1.0 / cutOff2;
    const int nm = s->n_mols;
    double *virt = (double *)malloc(sizeof(double) * 6 * (size_t)(nm > 0 ? nm : 1));
    const double ll = p->co2_bond * 0.5;
    const double cc = 0.91 * ll;
    const double bb = ll - cc;
    for (int m = 0; m < nm; ++m) {
        const int a0 = s->mol_start[m];
        const v3 O0 = v3_ld(s->pos + 3 * a0), O1 = v3_ld(s->pos + 3 * (a0 + 1));
        const v3 C2 = v3_ld(s->pos + 3 * (a0 + 2));
        v3_st(virt + 6 * m, v3_add(v3_scale(bb / ll, C2), v3_scale(cc / ll, O0)));
        v3_st(virt + 6 * m + 3, v3_add(v3_scale(bb / ll, C2), v3_scale(cc / ll, O1)));
    }
    double energy = 0.0, ulus = 0.0, dudv = 0.0, d2udv = 0.0, elecE = 0.0;
    double vir[9] = {0};
    const double voll = s->box[0] * s->box[1] * s->box[2];
    const double vol2 = voll * voll;
    for (int i = 0; i < s->n_atoms; ++i) {
        const int mi = s->mol[i];
        const v3 rim = v3_ld(s->com + 3 * mi);
        const double qi = s->q[i];
        v3 ri = v3_ld(s->pos + 3 * i);
        if (s->type[i] == 0) ri = v3_ld(virt + 6 * mi);
        if (s->type[i] == 1) ri = v3_ld(virt + 6 * mi + 3);
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int j = nb[k];
            const int mj = s->mol[j];
            if (mi == mj) continue;
            v3 rj = v3_ld(s->pos + 3 * j);
            if (s->type[j] == 0) rj = v3_ld(virt + 6 * mj);
            if (s->type[j] == 1) rj = v3_ld(virt + 6 * mj + 3);
            v3 rij = v3_sub(ri, rj);
            v3 rijm = v3_sub(v3_ld(s->com + 3 * mj), rim);
            rijm = v3_sub(rijm, rij);
            apply_pbc(&rij, s);
            rijm = v3_add(rijm, rij);
            (void)rijm;
            const double r = v3_len(rij);
            const double rij2 = r * r;
            if (!(rij2 <= cutOff2)) continue;
            const int t = s->type[i] * p->n_types + s->type[j];
            const double Aa = p->buck_a[t], Bb = p->buck_b[t], Cc = p->buck_c[t];
            const double r_1 = 1.0 / r;
            const double rij_2 = r_1 * r_1;
            const double rij6 = rij_2 * rij_2 * rij_2;
            double tmpE = Bb * exp(-Cc * r) - Aa * rij6;
            double tmpf = -Bb * Cc * r_1 * exp(-Cc * r) + 6.0 * Aa * rij6 * rij_2;
            double dulj = 0, d2ulj = 0;
            if (p->lustig) {
                dulj = -Bb * Cc * r * exp(-Cc * r) + 6.0 * Aa * rij6;
                d2ulj = (2 * r + Cc * rij2) * Bb * Cc * exp(-Cc * r) - 54.0 * Aa * rij6;
            }
            if (p->switch_on && (rij2 > switchDist)) {
                const double c = cutOff2 - rij2;
                const double c4 = c * (2 * rij2 + c2);
                const double sv = c * c4 * c1;
                const double dsv = c3 * (c * c - c4);
                tmpf = tmpf * sv - tmpE * dsv;
                tmpE *= sv;
            }
            energy += tmpE;
            if (p->lustig) { ulus += tmpE; dudv += dulj; d2udv += d2ulj; }
            if (p->coulomb)
                coulomb_shift(qi * s->q[j], r, r_1, rij2, rij_2, cutOff, rcut_2, &elecE, &tmpf);
            const v3 fij = v3_scale(-tmpf, rij);
            const int ci = s->mol_start[mi] + 2, cj = s->mol_start[mj] + 2;
            if (s->type[i] == 0 || s->type[i] == 1) {
                acc(force + 3 * i, v3_scale(0.91, fij));
                acc(force + 3 * ci, v3_scale(0.09, fij));
            } else if (s->type[i] == 2) {
                acc(force + 3 * i, fij);
            }
            if (s->type[j] == 0 || s->type[j] == 1) {
                dec(force + 3 * j, v3_scale(0.91, fij));
                dec(force + 3 * cj, v3_scale(0.09, fij));
            } else if (s->type[j] == 2) {
                dec(force + 3 * j, fij);
            }
            vir_add(vir, fij, rij);
        }
    }
    free(virt);
    energy += p->e_lrc;
    vir[XX] += p->v_lrc; vir[YY] += p->v_lrc; vir[ZZ] += p->v_lrc;
    o->energy = energy;
    o->lj_energy = energy + elecE;
    memcpy(o->virial, vir, sizeof vir);
    o->ulus = o->dudv = o->d2udv = 0.0;
    if (p->lustig) {
        dudv = dudv / (3 * voll);
        d2udv = d2udv / (9 * vol2);
        o->ulus = ulus + p->e_lrc;
        o->dudv = dudv + p->du_lrc;
        o->d2udv = d2udv + p->d2u_lrc;
    }
    return 0;
}

static int vdw_buff(const orc_sys *s, const int32_t *off, const int32_t *nb,
                    const orc_vdw_params *p, double *force, orc_vdw_out *o) {
    /* BuffvdWForce::compute (src/forces/BuffvdWForce.cpp:93-343) */
    const int nt = p->n_types;
    double *Rt = (double *)malloc(sizeof(double) * (size_t)(nt * nt));
    double *Et = (double *)malloc(sizeof(double) * (size_t)(nt * nt));
    orc_buff_table(nt, p->type_sigma, p->type_eps, Rt, Et);
    const double cutOff = p->cutoff, cutOff2 = cutOff * cutOff;
    double switchDist = cutOff2;
    if (p->switch_on) switchDist = p->switch_dist * p->switch_dist;
    const double switchd1 = p->switch_dist;
    double c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
    if (p->switch_on && (cutOff2 > switchDist)) {   /* ctor :59-73 */
        const double offcut = cutOff - switchd1;
        const double denom = offcut * offcut * offcut * offcut * offcut;
        c0 = cutOff * cutOff2 * (cutOff2 - 5.0 * cutOff * switchd1 + 10.0 * switchDist) / denom;
        c1 = -30.0 * cutOff2 * switchDist / denom;
        c2 = 30.0 * (cutOff2 * switchd1 + cutOff * switchDist) / denom;
        c3 = -10.0 * (cutOff2 + 4.0 * cutOff * switchd1 + switchDist) / denom;
        c4 = 15.0 * (cutOff + switchd1) / denom;
        c5 = -6.0 / denom;
    }
    const double factor = p->hfactor, n = p->n_buff, m = p->m_buff;
    const double delta = p->delta, gamma = p->gamma;
    /* VirtualPositions (:358-385): H sites pulled towards the molecule's
     * second site (the oxygen). */
    double *virt = (double *)malloc(sizeof(double) * 3 * (size_t)(s->n_atoms > 0 ? s->n_atoms : 1));
    for (int mm = 0; mm < s->n_mols; ++mm) {
        const int oid = s->mol_start[mm] + 1;
        if (s->type[oid] != 1) { free(virt); free(Rt); free(Et); return -1; }
        const v3 po = v3_ld(s->pos + 3 * oid);
        for (int a = s->mol_start[mm]; a < s->mol_start[mm + 1]; ++a)
            if (s->type[a] == 0)
                v3_st(virt + 3 * a, v3_add(v3_scale(factor, v3_ld(s->pos + 3 * a)),
                                           v3_scale(1.0 - factor, po)));
    }
    double energy = 0.0, dudv = 0.0, d2udv = 0.0;
    double vir[9] = {0};
    const double voll = s->box[0] * s->box[1] * s->box[2];
    const double vol2 = voll * voll;
    double r_1 = 0.0;   /* the reference keeps r_1 from the previous pair when eps==0 */
    for (int i = 0; i < s->n_atoms; ++i) {
        const int mi = s->mol[i];
        const v3 rim = v3_ld(s->com + 3 * mi);
        v3 ri = (s->type[i] == 0) ? v3_ld(virt + 3 * i) : v3_ld(s->pos + 3 * i);
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int j = nb[k];
            const int mj = s->mol[j];
            if (mi == mj) { free(virt); free(Rt); free(Et); return -2; }
            v3 rj = (s->type[j] == 0) ? v3_ld(virt + 3 * j) : v3_ld(s->pos + 3 * j);
            v3 rij = v3_sub(ri, rj);
            v3 rijm = v3_sub(rim, v3_ld(s->com + 3 * mj));
            apply_pbc(&rij, s);
            apply_pbc(&rijm, s);
            const double r = v3_len(rij);
            const double rij2 = r * r;
            if (!(rij2 <= cutOff2)) continue;
            const int t = s->type[i] * nt + s->type[j];
            const double sigma = Rt[t], eps = Et[t];
            double tmpE, tmpf;
            if (eps == 0) { tmpE = 0; tmpf = 0; }
            else {
                const double invsigma = 1.0 / sigma;
                const double m_1 = m - 1.0;
                const double rho = r * invsigma;
                const double rho_m = pow(rho, m);
                const double rho_m1 = pow(rho, m_1);
                const double expnm = n - m;
                r_1 = 1.0 / r;
                const double par1 = (1.0 + delta) / (rho + delta);
                const double par2 = (1.0 + gamma) / (rho_m + gamma);
                const double par1_nm = pow(par1, expnm);
                tmpE = eps * par1_nm * (par2 - 2.0);
                tmpf = -eps * expnm * par1_nm * (par2 - 2.0) * invsigma / (rho + delta) -
                       eps * par1_nm * par2 * m * rho_m1 * invsigma / (rho_m + gamma);
            }
            if (p->switch_on && (rij2 > switchDist)) {
                const double rij3 = rij2 * r, rij4 = rij2 * rij2, rij5 = rij2 * rij3;
                const double sv = c5 * rij5 + c4 * rij4 + c3 * rij3 + c2 * rij2 + c1 * r + c0;
                const double dsv = 5.0 * c5 * rij4 + 4.0 * c4 * rij3 + 3.0 * c3 * rij2 + 2.0 * c2 * r + c1;
                tmpf = tmpf * sv + tmpE * dsv;
                tmpE *= sv;
            }
            const double dulj = tmpf * r;
            energy += tmpE;
            dudv += dulj;
            /* d2ulj is 0 in the reference (:218-221) */
            v3 fij = v3_scale(-tmpf, rij);
            fij = v3_scale(r_1, fij);
            const int oi = s->mol_start[mi] + 1, oj = s->mol_start[mj] + 1;
            if (s->type[i] == 0) {
                acc(force + 3 * i, v3_scale(factor, fij));
                acc(force + 3 * oi, v3_scale(1.0 - factor, fij));
            } else acc(force + 3 * i, fij);
            if (s->type[j] == 0) {
                dec(force + 3 * j, v3_scale(factor, fij));
                dec(force + 3 * oj, v3_scale(1.0 - factor, fij));
            } else dec(force + 3 * j, fij);
            v3 nf = {-fij.x, -fij.y, -fij.z};
            vir_add(vir, nf, rij);
            (void)rijm;
        }
    }
    vir[XX] += p->v_lrc; vir[YY] += p->v_lrc; vir[ZZ] += p->v_lrc;
    dudv = dudv / (3.0 * voll);
    o->energy = energy;
    o->lj_energy = energy + p->e_lrc;
    memcpy(o->virial, vir, sizeof vir);
    o->ulus = energy + p->e_lrc;
    o->dudv = dudv + p->du_lrc;
    o->d2udv = 0.0;
    if (p->lustig) {
        d2udv = d2udv / (9.0 * vol2);
        o->d2udv = d2udv + p->d2u_lrc;
    }
    free(virt); free(Rt); free(Et);
    return 0;
}

int orc_vdw(const orc_sys *s, const int32_t *off, const int32_t *nb,
            const orc_vdw_params *p, double *force, orc_vdw_out *out) {
    memset(out, 0, sizeof *out);
    switch (p->kind) {
    case ORC_LJ:
    case ORC_MIE: return vdw_lj_mie(s, off, nb, p, force, out);
    case ORC_BUCK: return vdw_buck(s, off, nb, p, force, out);
    case ORC_BUFF147: return vdw_buff(s, off, nb, p, force, out);
    default: return -3;
    }
}

/* ========================================================================== */
/* (b/c) Ewald for point charges                                              */
/* ========================================================================== */

void orc_ewald_self(const orc_sys *s, double alpha, double volume,
                    double *self_energy, double *charged_energy) {
    /* EwaldForce ctor (src/forces/EwaldForce.cpp:69-79) */
    double qSum = 0.0, q2Sum = 0.0;
    for (int i = 0; i < s->n_atoms; ++i) {
        qSum += s->q[i];
        q2Sum += s->q[i] * s->q[i];
    }
    *self_energy = -q2Sum * alpha / SQRT_PI;
    *charged_energy = 0.0;
    if (fabs(qSum) > 1.0e-5)
        *charged_energy = -PI * qSum * qSum / (2.0 * volume * (alpha * alpha));
}

void orc_ewald_real(const orc_sys *s, const int32_t *off, const int32_t *nb,
                    double alpha, double real_cut, int use_constraint,
                    double *force, orc_ewald_out *o) {
    /* EwaldForce::real_term (src/forces/EwaldForce.cpp:296-527), one rank */
    const double alpha2 = alpha * alpha;
    const double alpha2PI = 2.0 * alpha / SQRT_PI;
    const double realCut2 = real_cut * real_cut;
    double E = 0.0, d1 = 0.0, d2 = 0.0, vir[9] = {0};
    for (int i = 0; i < s->n_atoms; ++i) {
        const v3 ri = v3_ld(s->pos + 3 * i);
        const v3 rim = v3_ld(s->com + 3 * s->mol[i]);
        const double qi = s->q[i];
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int j = nb[k];
            v3 rijm = v3_sub(v3_ld(s->com + 3 * s->mol[j]), rim);
            const double qij = qi * s->q[j];
            v3 rij = v3_sub(v3_ld(s->pos + 3 * j), ri);
            apply_pbc(&rij, s);
            apply_pbc(&rijm, s);
            const double r2 = v3_len2(rij);
            if (r2 < realCut2) {
                const double rijRij = use_constraint ? v3_dot(rij, rijm) : r2;
                const double r = sqrt(r2);
                const double e = erfc(alpha * r) / r;
                const double f = (e + alpha2PI * exp(-alpha2 * r2)) / r2;
                E += qij * e;
                d1 += -qij * f * rijRij;
                d2 += qij * (2.0 * f + 2.0 * alpha2PI * alpha2 * exp(-alpha2 * r2)) * rijRij * rijRij / r2 +
                      2.0 * qij * f * rijRij;
                const v3 fij = v3_scale(-qij * f, rij);
                acc(force + 3 * i, fij);
                dec(force + 3 * j, fij);
                vir_add(vir, fij, rij);
            }
        }
    }
    o->energy = E; o->deriv = d1; o->deriv2 = d2; o->extra = 0.0;
    memcpy(o->virial, vir, sizeof vir);
}

void orc_ewald_intra(const orc_sys *s, double alpha, double volume,
                     int use_constraint, double *force, orc_ewald_out *o) {
    /* EwaldForce::intraMolCorrect (src/forces/EwaldForce.cpp:709-813) */
    const double alpha2 = alpha * alpha;
    const double alpha2PI = 2.0 * alpha / SQRT_PI;
    const double volume2 = volume * volume;
    double E = 0.0, U_i = 0.0, d1 = 0.0, d2 = 0.0;
    for (int m = 0; m < s->n_mols; ++m) {
        for (int a = s->mol_start[m]; a < s->mol_start[m + 1]; ++a) {
            const double qi = s->q[a];
            const v3 ri = v3_ld(s->pos + 3 * a);
            for (int b = a + 1; b < s->mol_start[m + 1]; ++b) {
                if (!excluded(s, a, b)) continue;
                v3 rij = v3_sub(v3_ld(s->pos + 3 * b), ri);
                apply_pbc(&rij, s);
                const double r2 = v3_len2(rij);
                const double r = sqrt(r2);
                const double qij = qi * s->q[b];
                const double e = erf(alpha * r) / r;
                E += qij * e;
                U_i += alpha2PI * qij * exp(-alpha2 * r2);
                if (!use_constraint) {
                    const double dudr = e - alpha2PI * exp(-alpha2 * r2);
                    d1 += qij * dudr / (3.0 * volume);
                    d2 += (-2.0 * qij * dudr + qij * (-2.0 * dudr + 2.0 * alpha2PI * alpha2 * exp(-alpha2 * r2) * r2)) /
                          (9.0 * volume2);
                    const v3 fij = v3_scale(qij * dudr / r2, rij);
                    acc(force + 3 * a, fij);
                    dec(force + 3 * b, fij);
                }
            }
        }
    }
    o->energy = -E; o->extra = -U_i; o->deriv = d1; o->deriv2 = d2;
    memset(o->virial, 0, sizeof o->virial);
}

void orc_ewald_surface(const orc_sys *s, double volume, int use_constraint,
                       double *force, orc_ewald_out *o) {
    /* EwaldForce::surfDipoleCorrect (src/forces/EwaldForce.cpp:816-937) */
    const double volumer = 1.0 / volume;
    const double coeff = 2.0 * PI * volumer / 3.0;
    v3 qr = {0, 0, 0}, qR = {0, 0, 0};
    for (int i = 0; i < s->n_atoms; ++i) {
        qr = v3_add(qr, v3_scale(s->q[i], v3_ld(s->pos + 3 * i)));
        qR = v3_add(qR, v3_scale(s->q[i], v3_ld(s->com + 3 * s->mol[i])));
    }
    o->energy = coeff * v3_len2(qr);
    if (use_constraint) {
        o->deriv = v3_dot(v3_scale(-coeff * volumer, qr), qr) +
                   v3_dot(v3_scale(2.0 * coeff * volumer, qr), qR) / 3.0;
        o->deriv2 = v3_dot(v3_scale(2.0 * coeff * volumer * volumer, qr), qr) -
                    v3_dot(v3_scale(16.0 * coeff * volumer * volumer, qr), qR) / 9.0 +
                    v3_dot(v3_scale(2.0 * coeff * volumer * volumer, qR), qR) / 9.0;
    } else {
        o->deriv = v3_dot(v3_scale(-coeff * volumer, qr), qr) / 3.0;
        o->deriv2 = v3_dot(v3_scale(4.0 * coeff * volumer * volumer, qr), qr) / 9.0;
    }
    qr = v3_scale(coeff, qr);
    double vir[9] = {0};
    for (int i = 0; i < s->n_atoms; ++i) {
        const v3 temp = v3_scale(-s->q[i], qr);
        const v3 ri = v3_ld(s->pos + 3 * i);
        dec(force + 3 * i, v3_scale(2.0 * s->q[i], qr));
        vir_add(vir, temp, ri);
    }
    o->extra = 0.0;
    memcpy(o->virial, vir, sizeof vir);
}

int orc_ewald_kvectors(double kcut, const double box[3], double alpha,
                       double *kvec, double *kmod2, double *expk, int cap) {
    /* EwaldForce::set_up (src/forces/EwaldForce.cpp:198-257) */
    const double kCut2 = kcut * kcut;
    const double alphaR4 = 1.0 / (4.0 * (alpha * alpha));
    const int nxMax = (int)(kcut * box[0] / (2 * PI));
    const int nyMax = (int)(kcut * box[1] / (2 * PI));
    const int nzMax = (int)(kcut * box[2] / (2 * PI));
    const double vx = 2.0 * PI / box[0], vy = 2.0 * PI / box[1], vz = 2.0 * PI / box[2];
    int K = 0;
    for (int nx = 0; nx <= nxMax; nx++) {
        const double kx = nx * vx;
        for (int ny = (nx == 0 ? 0 : -nyMax); ny <= nyMax; ny++) {
            const double ky = ny * vy;
            for (int nz = ((nx == 0 && ny == 0) ? 1 : -nzMax); nz <= nzMax; nz++) {
                const double kz = nz * vz;
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 < kCut2) {
                    if (K < cap) {
                        kvec[3 * K] = kx; kvec[3 * K + 1] = ky; kvec[3 * K + 2] = kz;
                        kmod2[K] = k2;
                        expk[K] = exp(-k2 * alphaR4) / k2;
                    }
                    ++K;
                }
            }
        }
    }
    return K > cap ? -K : K;
}

void orc_ewald_recip(const orc_sys *s, int K, const double *kvec,
                     const double *kmod2, const double *expk, double alpha,
                     double volume, int use_constraint, double *force,
                     orc_ewald_out *o) {
    /* EwaldForce::reciprocal_term (src/forces/EwaldForce.cpp:529-698) */
    const int n = s->n_atoms;
    const double alphaR4 = 1.0 / (4.0 * (alpha * alpha));
    const double volumer = 1.0 / volume;
    const double pi4Volumer = 4.0 * PI * volumer;
    const double pi8Volumer = 2.0 * pi4Volumer;
    double *sc = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    double pe = 0.0, ld = 0.0, ld2 = 0.0, lv[9] = {0};
    for (int j = 0; j < K; ++j) {
        double sinSum = 0.0, cosSum = 0.0;
        const v3 kj = v3_ld(kvec + 3 * j);
        v3 cosrR = {0, 0, 0}, sinrR = {0, 0, 0};
        const double coeff = 2.0 * (alphaR4 + 1.0 / kmod2[j]);
        for (int i = 0; i < n; ++i) {
            const v3 rim = v3_ld(s->com + 3 * s->mol[i]);
            const double qi = s->q[i];
            const v3 ri = v3_ld(s->pos + 3 * i);
            const double a = v3_dot(ri, kj);
            const double sinAqi = qi * sin(a);
            const double cosAqi = qi * cos(a);
            sc[2 * i] = sinAqi;
            sc[2 * i + 1] = cosAqi;
            sinSum += sinAqi;
            cosSum += cosAqi;
            const v3 d = v3_sub(ri, rim);
            cosrR = v3_add(cosrR, v3_scale(cosAqi, d));
            sinrR = v3_add(sinrR, v3_scale(sinAqi, d));
        }
        const double te = pi4Volumer * expk[j] * (sinSum * sinSum + cosSum * cosSum);
        pe += te;
        const double kk = v3_dot(v3_scale(te, kj), kj);
        ld += -te + kk * 2.0 * alphaR4;
        {
            const double t7 = v3_dot(v3_scale(7.0 * te, kj), kj) * 2.0 * alphaR4;
            const double dd = kk * 2.0 * alphaR4;
            const double t2 = v3_dot(v3_scale(dd, kj), kj) * 2.0 * alphaR4;
            ld2 += 4.0 * te - t7 + t2;
        }
        lv[XX] += te - te * coeff * kj.x * kj.x;
        lv[YY] += te - te * coeff * kj.y * kj.y;
        lv[ZZ] += te - te * coeff * kj.z * kj.z;
        lv[XY] -= te * coeff * kj.x * kj.y;
        lv[XZ] -= te * coeff * kj.x * kj.z;
        lv[YZ] -= te * coeff * kj.y * kj.z;
        const double b = pi8Volumer * expk[j];
        const double bkk = v3_dot(v3_scale(b, kj), kj);
        double Ff = 0.0, Ffk2 = 0.0, d2UV1 = 0.0, d2UrR = 0.0;
        for (int i = 0; i < n; ++i) {
            const double val = b * (sc[2 * i] * cosSum - sc[2 * i + 1] * sinSum);
            const double d2u1 = bkk * (sc[2 * i + 1] * cosSum + sc[2 * i] * sinSum);
            const v3 rim = v3_ld(s->com + 3 * s->mol[i]);
            const v3 ri = v3_ld(s->pos + 3 * i);
            const v3 d = v3_sub(ri, rim);
            const v3 w = v3_add(v3_scale(sc[2 * i + 1], cosrR), v3_scale(sc[2 * i], sinrR));
            const double d2u2 = v3_dot(v3_scale(bkk, w), d);
            acc(force + 3 * i, v3_scale(val, kj));
            const double kd = v3_dot(kj, d);
            Ff += val * kd;
            Ffk2 += v3_dot(v3_scale(2.0 * val, kj), kj) * 2.0 * alphaR4 * kd;
            d2UV1 += v3_dot(v3_scale(d2u1, d), d);
            d2UrR += d2u2;
        }
        if (use_constraint) {
            ld += Ff;
            ld2 += -6.0 * Ff + Ffk2 - d2UV1 + d2UrR;
        }
    }
    lv[YX] = lv[XY]; lv[ZX] = lv[XZ]; lv[ZY] = lv[YZ];
    free(sc);
    o->energy = pe; o->deriv = ld; o->deriv2 = ld2; o->extra = 0.0;
    memcpy(o->virial, lv, sizeof lv);
}

/* ========================================================================== */
/* (e) SHAKE and Nose-Hoover                                                  */
/* ========================================================================== */

/* LFIntegrator::shake, NVT branch (src/integrators/LFIntegrator.cpp:1678-1828),
 * with an optional molecular SLLOD shear along x with gradient in y (gamma,
 * 1/ps; no counterpart in the reference).  With gamma = 0 every expression
 * reduces to the reference's.  For gamma != 0: the peculiar unconstrained
 * velocity takes -dt gamma P_y / M in x (P the molecule's peculiar momentum
 * of the previous step, M its mass), the move adds the streaming velocity
 * gamma Y of the molecule's centre (the same for all its atoms, so the bond
 * vectors SHAKE corrects are unchanged), the new peculiar velocity is
 * dr/dt - gamma Y, and the displacement used by the list trigger is the
 * non-affine part: dr - dt gamma (y - d_y) in x, y - d_y being the atom's
 * y when the list was built. */
static int shake_core(const orc_sys *s, double *pos, double *vt_minus,
                      double *velocity, double *displacement, const double *force,
                      const double *bond2, const double *mol_mom, const double *com,
                      double fnh, int nose_hoover, double dt, double tol, int max_count,
                      double gamma, double virial[9]) {
    const double tol_2 = 2 * tol;
    for (int k = 0; k < 9; ++k) virial[k] = 0.0;
    int rc = 0;
    for (int m = 0; m < s->n_mols; ++m) {
        const int a0 = s->mol_start[m];
        const int ns = s->mol_start[m + 1] - a0;
        if (ns != 3) return -3;
        const int nBonds = ns;   /* cyclic water */
        v3 rOld[3], rNew[3], vtPlus[3];
        const v3 P = v3_ld(mol_mom + 3 * m);
        double M = 0.0;
        for (int i = 0; i < ns; ++i) M += s->mass[a0 + i];
        const double Y = com ? com[3 * m + 1] : 0.0;
        for (int i = 0; i < ns; ++i) {
            const int id = a0 + i;
            const double mass_r = 1.0 / s->mass[id];
            rOld[i] = v3_ld(pos + 3 * id);
            const v3 F = v3_ld(force + 3 * id);
            if (nose_hoover) {
                const v3 g = v3_div(v3_scale(fnh, P), (double)ns);
                vtPlus[i] = v3_add(v3_ld(vt_minus + 3 * id), v3_scale(dt * mass_r, v3_sub(F, g)));
            } else {
                vtPlus[i] = v3_add(v3_ld(vt_minus + 3 * id), v3_scale(dt * mass_r, F));
            }
            vtPlus[i].x -= dt * gamma * P.y / M;
            v3 adv = vtPlus[i];
            adv.x += gamma * Y;
            rNew[i] = v3_add(rOld[i], v3_scale(dt, adv));
        }
        int count = 0, done = 0;
        while (!done && count < max_count) {
            done = 1;
            for (int i = 0; i < nBonds; ++i) {
                int next = i + 1;
                if (next == ns) next = 0;
                v3 rijNew = v3_sub(rNew[i], rNew[next]);
                apply_pbc(&rijNew, s);
                const double rijNew2 = v3_len2(rijNew);
                const double diff = bond2[3 * m + i] - rijNew2;
                if (fabs(diff) > (bond2[3 * m + i] * tol_2)) {
                    v3 rijOld = v3_sub(rOld[i], rOld[next]);
                    apply_pbc(&rijOld, s);
                    const double rijNewOld = v3_dot(rijOld, rijNew);
                    if (rijNewOld < bond2[3 * m + i] * tol) return -1;
                    const double mr1 = 1.0 / s->mass[a0 + i], mr2 = 1.0 / s->mass[a0 + next];
                    const double factor = diff / (2 * rijNewOld * (mr1 + mr2));
                    virial[XX] += factor * rijNew.x * rijOld.x;
                    virial[XY] += factor * rijNew.x * rijOld.y;
                    virial[XZ] += factor * rijNew.x * rijOld.z;
                    virial[YX] += factor * rijNew.y * rijOld.x;
                    virial[YY] += factor * rijNew.y * rijOld.y;
                    virial[YZ] += factor * rijNew.y * rijOld.z;
                    virial[ZX] += factor * rijNew.z * rijOld.x;
                    virial[ZY] += factor * rijNew.z * rijOld.y;
                    virial[ZZ] += factor * rijNew.z * rijOld.z;
                    const v3 fo = v3_scale(factor, rijOld);
                    rNew[i] = v3_add(rNew[i], v3_scale(mr1, fo));
                    rNew[next] = v3_sub(rNew[next], v3_scale(mr2, fo));
                    done = 0;
                }
            }
            count++;
        }
        if (!done) { rc = -2; return rc; }
        for (int i = 0; i < ns; ++i) {
            const int id = a0 + i;
            const v3 p0 = v3_ld(pos + 3 * id);
            const v3 dr = v3_sub(rNew[i], p0);
            v3_st(pos + 3 * id, v3_add(p0, dr));
            v3 dn = dr;
            dn.x -= dt * gamma * (p0.y - displacement[3 * id + 1]);
            v3_st(displacement + 3 * id, v3_add(v3_ld(displacement + 3 * id), dn));
            v3 vp = v3_div(dr, dt);
            vp.x -= gamma * Y;
            const v3 vm = v3_ld(vt_minus + 3 * id);
            v3_st(velocity + 3 * id, v3_scale(0.5, v3_add(vp, vm)));
            v3_st(vt_minus + 3 * id, vp);
        }
    }
    for (int k = 0; k < 9; ++k) virial[k] /= (dt * dt);
    return rc;
}

int orc_shake(const orc_sys *s, double *pos, double *vt_minus,
              double *velocity, double *displacement, const double *force,
              const double *bond2, const double *mol_mom, double fnh,
              int nose_hoover, double dt, double tol, int max_count,
              double virial[9]) {
    return shake_core(s, pos, vt_minus, velocity, displacement, force, bond2, mol_mom, NULL, fnh,
                      nose_hoover, dt, tol, max_count, 0.0, virial);
}

void orc_mol_kinetics(const orc_sys *s, const double *pos,
                      const double *velocity, double *com, double *mol_mom,
                      double *atom_ke2, double *mol_ke2) {
    /* LFIntegrator::run (src/integrators/LFIntegrator.cpp:650-686) */
    double ake = 0.0, mke = 0.0;
    for (int m = 0; m < s->n_mols; ++m) {
        v3 P = {0, 0, 0}, C = {0, 0, 0};
        double M = 0.0;
        for (int a = s->mol_start[m]; a < s->mol_start[m + 1]; ++a) M += s->mass[a];
        for (int a = s->mol_start[m]; a < s->mol_start[m + 1]; ++a) {
            const v3 v = v3_ld(velocity + 3 * a);
            P = v3_add(P, v3_scale(s->mass[a], v));
            C = v3_add(C, v3_scale(s->mass[a], v3_ld(pos + 3 * a)));
            ake += v3_dot(v, v) * s->mass[a];
        }
        C = v3_div(C, M);
        mke += v3_dot(P, P) / M;
        v3_st(com + 3 * m, C);
        v3_st(mol_mom + 3 * m, P);
    }
    *atom_ke2 = ake;
    *mol_ke2 = mke;
}

void orc_nh_friction(double ke2, double ndf, double temperature,
                     double coupl, double dt, double *fnh, double *fnh_minus) {
    /* LFIntegrator::run (src/integrators/LFIntegrator.cpp:688-709) */
    const double qq = ndf * BOLTZ * temperature * coupl * coupl;
    const double invqq = 1 / qq;
    const double fnhpl = *fnh_minus + dt * invqq * (ke2 - ndf * BOLTZ * temperature);
    *fnh = 0.5 * (fnhpl + *fnh_minus);
    *fnh_minus = fnhpl;
}

int orc_lf_atoms(int n, double *pos, double *vt_minus, double *velocity,
                 double *displacement, const double *force,
                 const double *mass, double fnh, int nose_hoover, double dt) {
    /* LFIntegrator::run, no constraints, NVT (src/integrators/LFIntegrator.cpp:587-643) */
    for (int i = 0; i < n; ++i) {
        const double mass_r = 1.0 / mass[i];
        const v3 F = v3_ld(force + 3 * i);
        const v3 vm = v3_ld(vt_minus + 3 * i);
        v3 vp;
        if (nose_hoover) {
            double velcheck = 1.0;
            int cnt = 0;
            v3 vel = v3_ld(velocity + 3 * i);
            while ((velcheck > 1E-10) && (cnt < 11)) {
                vp = v3_add(vm, v3_scale(dt, v3_sub(v3_scale(mass_r, F), v3_scale(fnh, vel))));
                const v3 half = v3_scale(0.5, v3_add(vp, vm));
                const v3 vv = v3_sub(half, vel);
                velcheck = v3_len(vv);
                vel = half;
                cnt++;
            }
            if (cnt == 10) return -1;
        } else {
            vp = v3_add(vm, v3_scale(dt * mass_r, F));
        }
        const v3 dr = v3_scale(dt, vp);
        v3_st(pos + 3 * i, v3_add(v3_ld(pos + 3 * i), dr));
        v3_st(displacement + 3 * i, v3_add(v3_ld(displacement + 3 * i), dr));
        v3_st(velocity + 3 * i, v3_scale(0.5, v3_add(vp, vm)));
        v3_st(vt_minus + 3 * i, vp);
    }
    return 0;
}

double orc_nh_chain(double ekin, double temperature, double dt, double ndf,
                    double press_diff, double volume, double wbar,
                    double *vnh, double *gnh, const double *qnh, double *vlog,
                    double *glog) {
    /* NHIntegrator::chain (src/integrators/NHIntegrator.cpp:1814-1912) */
    double scale = 1.0;
    const int nc = 5, ns = 3;
    double w[3];
    w[0] = 1.0 / (2.0 - pow(2.0, (1.0 / 3.0)));
    w[1] = 1.0 - 2.0 * w[0];
    w[2] = w[0];
    const double kT = BOLTZ * temperature;
    const double dtnc = dt / (double)nc;
    const double diff = press_diff;
    const double Nf = ndf;
    const double gn1kt = (Nf + 1.0) * kT;
    const double odnf = 1.0 + 3.0 / Nf;
    double expterm;
    for (int i = 0; i < nc; i++) {
        for (int j = 0; j < ns; j++) {
            const double dts = w[j] * dtnc;
            const double dts2 = 0.5 * dts, dts4 = 0.25 * dts, dts8 = 0.125 * dts;
            gnh[3] = (qnh[2] * vnh[2] * vnh[2] - kT) / qnh[3];
            vnh[3] = vnh[3] + gnh[3] * dts4;
            gnh[2] = (qnh[1] * vnh[1] * vnh[1] - kT) / qnh[2];
            expterm = exp(-vnh[3] * dts8);
            vnh[2] = expterm * (vnh[2] * expterm + gnh[2] * dts4);
            gnh[1] = (qnh[0] * vnh[0] * vnh[0] - kT) / qnh[1];
            expterm = exp(-vnh[2] * dts8);
            vnh[1] = expterm * (vnh[1] * expterm + gnh[1] * dts4);
            gnh[0] = (2.0 * ekin + wbar * (*vlog) * (*vlog) - gn1kt) / qnh[0];
            expterm = exp(-vnh[1] * dts8);
            vnh[0] = expterm * (vnh[0] * expterm + gnh[0] * dts4);
            *glog = (2.0 * ekin * odnf + 3.0 * volume * diff) / wbar;
            expterm = exp(-vnh[0] * dts8);
            *vlog = expterm * ((*vlog) * expterm + (*glog) * dts4);

            expterm = exp(-(vnh[0] + (*vlog) * odnf) * dts2);
            scale *= expterm;
            ekin = ekin * expterm * expterm;

            *glog = (2.0 * ekin * odnf + 3.0 * volume * diff) / wbar;
            expterm = exp(-vnh[0] * dts8);
            *vlog = expterm * ((*vlog) * expterm + (*glog) * dts4);
            gnh[0] = (2.0 * ekin + wbar * (*vlog) * (*vlog) - gn1kt) / qnh[0];
            expterm = exp(-vnh[1] * dts8);
            vnh[0] = expterm * (vnh[0] * expterm + gnh[0] * dts4);
            gnh[1] = (qnh[0] * vnh[0] * vnh[0] - kT) / qnh[1];
            expterm = exp(-vnh[2] * dts8);
            vnh[1] = expterm * (vnh[1] * expterm + gnh[1] * dts4);
            gnh[2] = (qnh[1] * vnh[1] * vnh[1] - kT) / qnh[2];
            expterm = exp(-vnh[3] * dts8);
            vnh[2] = expterm * (vnh[2] * expterm + gnh[2] * dts4);
            gnh[3] = (qnh[2] * vnh[2] * vnh[2] - kT) / qnh[3];
            vnh[3] = vnh[3] + gnh[3] * dts4;
        }
    }
    return scale;
}

/* Ensemble::drand / gauss (src/Ensemble.cpp:1831-1864) */
static double drand_lcg(unsigned int *seed) {
    const int l = 1029, c = 221591, m = 1048576;
    *seed = (*seed * l + c) % m;
    return (double)(*seed) / m;
}

static double gauss12(unsigned int *seed) {
    const double A1 = 3.949846138, A3 = 0.252408784, A5 = 0.076542912;
    const double A7 = 0.008355968, A9 = 0.029899776;
    double sum = 0.0;
    for (int i = 0; i < 12; i++) sum += drand_lcg(seed);
    const double r = (sum - 6.0) / 4.0;
    const double r2 = r * r;
    const double rst = (((A9 * r2 + A7) * r2 + A5) * r2 + A3) * r2 + A1;
    return rst * r;
}

void orc_init_velocity(const orc_sys *s, int per_atom, int ndf,
                       double temperature, unsigned int *seed,
                       double *velocity) {
    /* Ensemble::init_velocity (src/Ensemble.cpp:1752-1827) */
    const int n = per_atom ? s->n_atoms : s->n_mols;
    double *mom = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    double *mmass = NULL;
    if (!per_atom) {
        mmass = (double *)calloc((size_t)(s->n_mols > 0 ? s->n_mols : 1), sizeof(double));
        for (int a = 0; a < s->n_atoms; ++a) mmass[s->mol[a]] += s->mass[a];
    }
    double px = 0.0, py = 0.0, pz = 0.0, tavg = 0.0;
    for (int i = 0; i < n; ++i) {
        mom[3 * i] = gauss12(seed);
        mom[3 * i + 1] = gauss12(seed);
        mom[3 * i + 2] = gauss12(seed);
        px += mom[3 * i]; py += mom[3 * i + 1]; pz += mom[3 * i + 2];
    }
    px = px / n; py = py / n; pz = pz / n;
    for (int i = 0; i < n; ++i) {
        mom[3 * i] -= px; mom[3 * i + 1] -= py; mom[3 * i + 2] -= pz;
        const double l2 = v3_len2(v3_ld(mom + 3 * i));
        tavg += l2 / (per_atom ? s->mass[i] : mmass[i]);
    }
    tavg /= (ndf * BOLTZ);
    const double scale = sqrt(temperature / tavg);
    for (int i = 0; i < n; ++i) {
        const v3 P = v3_scale(scale, v3_ld(mom + 3 * i));
        if (per_atom) {
            v3_st(velocity + 3 * i, v3_div(P, s->mass[i]));
        } else {
            const v3 v = v3_div(P, mmass[i]);
            for (int a = s->mol_start[i]; a < s->mol_start[i + 1]; ++a) v3_st(velocity + 3 * a, v);
        }
    }
    free(mom);
    free(mmass);
}

/* ========================================================================== */
/* one leap-frog NVT step (LFIntegrator::run, src/integrators/LFIntegrator.cpp:278-941) */
/* ========================================================================== */

static void rewrap_molecules(orc_sys *s, double *pos, double *com) {
    /* apply PBC to the molecular centre and shift its atoms (:329-339, :897-910) */
    for (int m = 0; m < s->n_mols; ++m) {
        const v3 c0 = v3_ld(com + 3 * m);
        v3 c = c0;
        apply_pbc(&c, s);
        v3_st(com + 3 * m, c);
        const v3 diff = v3_sub(c, c0);
        for (int a = s->mol_start[m]; a < s->mol_start[m + 1]; ++a)
            v3_st(pos + 3 * a, v3_add(v3_ld(pos + 3 * a), diff));
    }
}

/* Force block of run()/start(): LJ on the list, Ewald real, intramolecular
 * and surface terms, then the reciprocal sum (:382-441, :1001-1055).  vir
 * receives the ensemble virial (each term's virial subtracted, the long
 * virial added). */
static void md_forces(orc_sys *s, const orc_md_params *p, orc_md_state *st, int do_long, double vir[9]) {
    const double volume = s->box[0] * s->box[1] * s->box[2];
    for (int j = 0; j < 3 * s->n_atoms; ++j) st->force[j] = 0.0;
    orc_vdw_params vp;
    memset(&vp, 0, sizeof vp);
    vp.kind = ORC_LJ; vp.n_types = p->n_types; vp.cutoff = p->cutoff;
    vp.use_constraint = p->constraint; vp.sigma2 = p->sigma2; vp.eps4 = p->eps4;
    vp.switch_on = p->switch_on; vp.switch_dist = p->cutoff;
    orc_vdw_out vo;
    orc_vdw(s, st->off, st->nb, &vp, st->force, &vo);
    for (int k = 0; k < 9; ++k) vir[k] = -vo.virial[k];
    st->e_lj = vo.lj_energy;
    st->e_real = st->e_intra = st->e_surf = st->e_self = 0.0;
    if (!p->ewald) return;
    orc_ewald_out eo;
    orc_ewald_real(s, st->off, st->nb, p->alpha, p->cutoff, 1, st->force, &eo);
    st->e_real = eo.energy;
    for (int k = 0; k < 9; ++k) vir[k] -= eo.virial[k];
    orc_ewald_intra(s, p->alpha, volume, 1, st->force, &eo);
    st->e_intra = eo.energy;
    st->e_surf = 0.0;
    if (p->surface) {
        orc_ewald_surface(s, volume, 1, st->force, &eo);
        st->e_surf = eo.energy;
        for (int k = 0; k < 9; ++k) vir[k] -= eo.virial[k];
    }
    if (do_long) {
        orc_ewald_recip(s, st->K, st->kvec, st->kmod2, st->expk, p->alpha, volume, 1, st->force, &eo);
        st->e_long = eo.energy;
        for (int k = 0; k < 9; ++k) vir[k] += eo.virial[k];
    }
    double ce;
    orc_ewald_self(s, p->alpha, volume, &st->e_self, &ce);
    st->e_self += ce;
}

int orc_md_start(const orc_sys *s_in, const orc_md_params *p, orc_md_state *st) {
    /* LFIntegrator::start (src/integrators/LFIntegrator.cpp:991-1174): the
     * pair list must already be built (Ensemble::set_pairlist). */
    orc_sys s = *s_in;
    s.pos = st->pos;
    s.com = st->com;
    double vir[9];
    md_forces(&s, p, st, 1, vir);
    for (int i = 0; i < s.n_atoms; ++i) {
        const double mass_r = 1.0 / s.mass[i];
        for (int d = 0; d < 3; ++d)
            st->vt_minus[3 * i + d] = st->velocity[3 * i + d] - 0.5 * p->dt * mass_r * st->force[3 * i + d];
    }
    for (int k = 0; k < 9; ++k) st->virial[k] = vir[k];
    return 0;
}

int orc_md_step(const orc_sys *s_in, const orc_md_params *p, orc_md_state *st) {
    orc_sys s = *s_in;
    s.pos = st->pos;
    s.com = st->com;
    const int n = s.n_atoms;
    const double drMax = p->cut_buff * p->cut_buff / 4.0;
    int need = 0;
    for (int j = 0; j < n && !need; ++j)
        if (v3_len2(v3_ld(st->displacement + 3 * j)) > drMax) need = 1;
    if (need) {
        for (int j = 0; j < 3 * n; ++j) st->displacement[j] = 0.0;
        rewrap_molecules(&s, st->pos, st->com);
        int used;
        int64_t np = orc_pairlist(&s, p->cutoff, p->cut_buff, p->use_cells, st->off, st->nb, st->cap, &used);
        if (np < 0) return -10;
        st->list_updates++;
    }
    /* remove centre-of-mass drift of vtMinus (:354-379) */
    double vx = 0, vy = 0, vz = 0, mt = 0;
    for (int j = 0; j < n; ++j) {
        vx += st->vt_minus[3 * j] * s.mass[j];
        vy += st->vt_minus[3 * j + 1] * s.mass[j];
        vz += st->vt_minus[3 * j + 2] * s.mass[j];
        mt += s.mass[j];
    }
    vx = vx / mt; vy = vy / mt; vz = vz / mt;
    if (fabs(vx) > 1e-11 || fabs(vy) > 1e-11 || fabs(vz) > 1e-11)
        for (int j = 0; j < n; ++j) {
            st->vt_minus[3 * j] -= vx; st->vt_minus[3 * j + 1] -= vy; st->vt_minus[3 * j + 2] -= vz;
            st->velocity[3 * j] -= vx; st->velocity[3 * j + 1] -= vy; st->velocity[3 * j + 2] -= vz;
        }
    double vir[9];
    md_forces(&s, p, st, (p->long_steps <= 1 || (st->step % p->long_steps) == 0), vir);
    double sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (p->constraint) {
        /* SHAKE with NH molecular friction (:561-586) */
        int rc = orc_shake(&s, st->pos, st->vt_minus, st->velocity, st->displacement, st->force, p->bond2,
                           st->mol_mom, st->fnh, p->nose_hoover, p->dt, p->tol, p->max_count, sv);
        if (rc != 0) return rc;
    } else {
        /* per-atom update with the self-consistent NH loop (:587-643) */
        int rc = orc_lf_atoms(n, st->pos, st->vt_minus, st->velocity, st->displacement, st->force, s.mass,
                              st->fnh, p->nose_hoover, p->dt);
        if (rc != 0) return rc;
    }
    /* centres, momenta, kinetic energy (:650-686) and friction (:688-722);
     * atomNDF is 6 Nmol - 3 with constraints, else 3 N - 3 (src/Integrator.cpp:24-29) */
    orc_mol_kinetics(&s, st->pos, st->velocity, st->com, st->mol_mom, &st->atom_ke2, &st->mol_ke2);
    if (p->nose_hoover) {
        if (p->atom_thermo)
            orc_nh_friction(st->atom_ke2, p->constraint ? (double)(6 * s.n_mols - 3) : (double)(3 * n - 3),
                            p->temperature, p->coupl, p->dt, &st->fnh, &st->fnh_minus);
        else
            orc_nh_friction(st->mol_ke2, (double)(3 * s.n_mols - 3), p->temperature, p->coupl, p->dt,
                            &st->fnh, &st->fnh_minus);
    } else {
        st->fnh = 0.0;
    }
    for (int k = 0; k < 9; ++k) st->virial[k] = vir[k] + sv[k];
    /* per-step rewrap of molecules by centre (:897-910) */
    rewrap_molecules(&s, st->pos, st->com);
    st->step++;
    return 0;
}

/* ========================================================================== */
/* Sheared replicas: molecular SLLOD with Lees-Edwards boundaries.  The       */
/* reference has no shear flow; these functions are this project's own       */
/* definition, built from the reference's step above so that a zero shear    */
/* rate gives the reference's step.                                           */
/* ========================================================================== */

int orc_ewald_kvectors_le(double kcut, const double box[3], double tilt, double alpha,
                          double *kvec, double *kmod2, double *expk, int cap) {
    /* Reciprocal lattice of the sheared cell a = (Lx,0,0), b = (tilt,Ly,0),
     * c = (0,0,Lz): k = 2 pi (h/Lx, (k - h tilt/Lx)/Ly, l/Lz), every k with
     * |k| < kcut, one of each +-k pair (h > 0, or h = 0 and k > 0, or h = k = 0
     * and l > 0), in h, k, l order.  The set is that of the lattice, so it does
     * not change when tilt moves by a multiple of Lx; at tilt = 0 it is
     * EwaldForce::set_up's set in its order (src/forces/EwaldForce.cpp:198-257). */
    const double kCut2 = kcut * kcut;
    const double alphaR4 = 1.0 / (4.0 * (alpha * alpha));
    const int nxMax = (int)(kcut * box[0] / (2 * PI));
    const double nyR = kcut * box[1] / (2 * PI);
    const int nzMax = (int)(kcut * box[2] / (2 * PI));
    const double vx = 2.0 * PI / box[0], vy = 2.0 * PI / box[1], vz = 2.0 * PI / box[2];
    int K = 0;
    for (int nx = 0; nx <= nxMax; nx++) {
        const double kx = nx * vx;
        const double shift = nx * tilt / box[0];
        const int lo = nx == 0 ? 0 : (int)floor(shift - nyR) - 1;
        const int hi = (int)ceil(shift + nyR) + 1;
        for (int ny = lo; ny <= hi; ny++) {
            const double ky = (ny - shift) * vy;
            for (int nz = ((nx == 0 && ny == 0) ? 1 : -nzMax); nz <= nzMax; nz++) {
                const double kz = nz * vz;
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 < kCut2) {
                    if (K < cap) {
                        kvec[3 * K] = kx; kvec[3 * K + 1] = ky; kvec[3 * K + 2] = kz;
                        kmod2[K] = k2;
                        expk[K] = exp(-k2 * alphaR4) / k2;
                    }
                    ++K;
                }
            }
        }
    }
    return K > cap ? -K : K;
}

int orc_sllod_start(const orc_sys *s_in, const orc_md_params *p, orc_md_state *st, double tilt, int kcap) {
    /* orc_md_start at the current tilt: brute-force pair list (built here),
     * sheared k set, first forces, vtMinus = v - dt/2 F/m. */
    orc_sys s = *s_in;
    s.pos = st->pos;
    s.com = st->com;
    s.tilt = tilt;
    if (p->ewald) {
        const int K = orc_ewald_kvectors_le(p->kcut, s.box, tilt, p->alpha, st->kvec, st->kmod2, st->expk, kcap);
        if (K < 0) return -11;
        st->K = K;
    }
    int used;
    if (orc_pairlist(&s, p->cutoff, p->cut_buff, 0, st->off, st->nb, st->cap, &used) < 0) return -10;
    for (int j = 0; j < 3 * s.n_atoms; ++j) st->displacement[j] = 0.0;
    double vir[9];
    md_forces(&s, p, st, 1, vir);
    for (int i = 0; i < s.n_atoms; ++i) {
        const double mass_r = 1.0 / s.mass[i];
        for (int d = 0; d < 3; ++d)
            st->vt_minus[3 * i + d] = st->velocity[3 * i + d] - 0.5 * p->dt * mass_r * st->force[3 * i + d];
    }
    for (int k = 0; k < 9; ++k) st->virial[k] = vir[k];
    return 0;
}

int orc_sllod_step(const orc_sys *s_in, const orc_md_params *p, orc_md_state *st, double gamma,
                   double *tilt, int64_t *since_build, int kcap) {
    /* One step of orc_md_step (rigid water, SHAKE, molecular NH) under shear
     * rate gamma (1/ps): streaming velocity gamma*y along x, Lees-Edwards
     * images (the image across +y sits at x + tilt), tilt advanced by
     * gamma*Ly*dt after the move and kept in [-Lx/2, Lx/2) (a move by Lx maps
     * the lattice onto itself).  Velocities are peculiar.  The pair list is
     * brute force; it is rebuilt when any atom's non-affine displacement d
     * since the build has 2|d| + |gamma| rlist t > cut_buff (t the time since
     * the build): the pair separations change by at most that much, since the
     * affine part of a pair's motion is gamma t times its y separation. */
    orc_sys s = *s_in;
    s.pos = st->pos;
    s.com = st->com;
    s.tilt = *tilt;
    const int n = s.n_atoms;
    if (!p->constraint) return -12;
    const double rlist = p->cutoff + p->cut_buff;
    const double slack = p->cut_buff - fabs(gamma) * rlist * p->dt * (double)(*since_build);
    const double drMax = slack * slack / 4.0;
    int need = slack <= 0.0;
    for (int j = 0; j < n && !need; ++j)
        if (v3_len2(v3_ld(st->displacement + 3 * j)) > drMax) need = 1;
    if (need) {
        for (int j = 0; j < 3 * n; ++j) st->displacement[j] = 0.0;
        rewrap_molecules(&s, st->pos, st->com);
        int used;
        int64_t np = orc_pairlist(&s, p->cutoff, p->cut_buff, 0, st->off, st->nb, st->cap, &used);
        if (np < 0) return -10;
        st->list_updates++;
        *since_build = 0;
    }
    double vx = 0, vy = 0, vz = 0, mt = 0;
    for (int j = 0; j < n; ++j) {
        vx += st->vt_minus[3 * j] * s.mass[j];
        vy += st->vt_minus[3 * j + 1] * s.mass[j];
        vz += st->vt_minus[3 * j + 2] * s.mass[j];
        mt += s.mass[j];
    }
    vx = vx / mt; vy = vy / mt; vz = vz / mt;
    if (fabs(vx) > 1e-11 || fabs(vy) > 1e-11 || fabs(vz) > 1e-11)
        for (int j = 0; j < n; ++j) {
            st->vt_minus[3 * j] -= vx; st->vt_minus[3 * j + 1] -= vy; st->vt_minus[3 * j + 2] -= vz;
            st->velocity[3 * j] -= vx; st->velocity[3 * j + 1] -= vy; st->velocity[3 * j + 2] -= vz;
        }
    if (p->ewald) {
        const int K = orc_ewald_kvectors_le(p->kcut, s.box, s.tilt, p->alpha, st->kvec, st->kmod2, st->expk, kcap);
        if (K < 0) return -11;
        st->K = K;
    }
    double vir[9];
    md_forces(&s, p, st, 1, vir);
    double sv[9];
    int rc = shake_core(&s, st->pos, st->vt_minus, st->velocity, st->displacement, st->force, p->bond2,
                        st->mol_mom, st->com, st->fnh, p->nose_hoover, p->dt, p->tol, p->max_count, gamma, sv);
    if (rc != 0) return rc;
    orc_mol_kinetics(&s, st->pos, st->velocity, st->com, st->mol_mom, &st->atom_ke2, &st->mol_ke2);
    if (p->nose_hoover) {
        if (p->atom_thermo)
            orc_nh_friction(st->atom_ke2, (double)(6 * s.n_mols - 3), p->temperature, p->coupl, p->dt,
                            &st->fnh, &st->fnh_minus);
        else
            orc_nh_friction(st->mol_ke2, (double)(3 * s.n_mols - 3), p->temperature, p->coupl, p->dt,
                            &st->fnh, &st->fnh_minus);
    } else {
        st->fnh = 0.0;
    }
    for (int k = 0; k < 9; ++k) st->virial[k] = vir[k] + sv[k];
    double t = *tilt + gamma * s.box[1] * p->dt;
    if (t >= 0.5 * s.box[0]) t -= s.box[0];
    else if (t < -0.5 * s.box[0]) t += s.box[0];
    *tilt = t;
    s.tilt = t;
    rewrap_molecules(&s, st->pos, st->com);
    st->step++;
    (*since_build)++;
    return 0;
}

/* ========================================================================== */
/* polarizable water: frames, permanent field, induced dipoles                */
/* ========================================================================== */

int orc_local2global(const orc_sys *s, const double *dip_loc, const double *quad_loc,
                     double *dip, double *quad) {
    /* Local2Global::FrameofReference (src/utils/Local2Global.cpp:90-225).
     * Molecule sites are H1, O, H2 (myAtoms[0..2]). */
    const v3 oneX = {1.0, 0.0, 0.0}, oneY = {0.0, 1.0, 0.0};
    for (int m = 0; m < s->n_mols; ++m)
        if (s->mol_start[m + 1] - s->mol_start[m] != 3) return -1;
    for (int m = 0; m < s->n_mols; ++m) {
        const int a0 = s->mol_start[m];
        const v3 h1 = v3_ld(s->pos + 3 * a0), o = v3_ld(s->pos + 3 * (a0 + 1)), h2 = v3_ld(s->pos + 3 * (a0 + 2));
        v3 pp1[3], pp2[3];
        /* H1: axis O - H1, second vector x (y when the axis is close to x) :114-121 */
        pp1[0] = v3_sub(o, h1);
        double val = pp1[0].x / v3_len(pp1[0]);
        pp2[0] = oneX;
        if (fabs(val) > 0.8660) pp2[0] = oneY;
        /* O: sum of the unit vectors to the hydrogens, second vector to H2 :123-129 */
        pp1[1] = v3_sub(h1, o);
        pp2[1] = v3_sub(h2, o);
        pp1[1] = v3_div(pp1[1], v3_len(pp1[1]));
        pp2[1] = v3_div(pp2[1], v3_len(pp2[1]));
        pp1[1] = v3_add(pp1[1], pp2[1]);
        /* H2 :131-137 */
        pp1[2] = v3_sub(o, h2);
        val = pp1[2].x / v3_len(pp1[2]);
        pp2[2] = oneX;
        if (fabs(val) > 0.8660) pp2[2] = oneY;
        for (int i = 0; i < 3; ++i) {
            /* orthonormal u, v, w (:147-160) */
            pp1[i] = v3_div(pp1[i], v3_len(pp1[i]));
            const v3 w = pp1[i];
            const v3 ss = v3_sub(pp2[i], v3_scale(v3_dot(pp2[i], w), w));
            const v3 u = v3_div(ss, v3_len(ss));
            const v3 v = {w.y * u.z - u.y * w.z, u.x * w.z - w.x * u.z, w.x * u.y - u.x * w.y};
            const int id = a0 + i;
            const double dx = dip_loc[3 * id], dy = dip_loc[3 * id + 1], dz = dip_loc[3 * id + 2];
            dip[3 * id] = u.x * dx + v.x * dy + w.x * dz;
            dip[3 * id + 1] = u.y * dx + v.y * dy + w.y * dz;
            dip[3 * id + 2] = u.z * dx + v.z * dy + w.z * dz;
            const double qxx = quad_loc[6 * id], qxy = quad_loc[6 * id + 1], qxz = quad_loc[6 * id + 2];
            const double qyy = quad_loc[6 * id + 3], qyz = quad_loc[6 * id + 4], qzz = quad_loc[6 * id + 5];
            /* R Q R^T with R = [u v w] (:181-205), then the traced form / 3 (:210-215) */
            const double e[3][3] = {{u.x, v.x, w.x}, {u.y, v.y, w.y}, {u.z, v.z, w.z}};
            const int ab[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
            for (int k = 0; k < 6; ++k) {
                const double *ra = e[ab[k][0]], *rb = e[ab[k][1]];
                const double qk = ra[0] * (rb[0] * qxx + rb[1] * qxy + rb[2] * qxz) +
                                  ra[1] * (rb[0] * qxy + rb[1] * qyy + rb[2] * qyz) +
                                  ra[2] * (rb[0] * qxz + rb[1] * qyz + rb[2] * qzz);
                quad[6 * id + k] = qk / 3;
            }
        }
    }
    return 0;
}

int orc_induced_dipoles(const orc_sys *s, const double *dip, const double *quad, const double *polar,
                        const orc_ind_params *p, double *ei0, double *pi, int *iters, double *final_res) {
    /* Induction::InducedDipoles (src/forces/Induction.cpp:181-455) with the
     * dense A[3N][3N], then ConjugateGradient (:2401-2467) or, without
     * minimisation, Pi = alpha E0 (:446-453).  q is the unscaled charge. */
    const int n = s->n_atoms, dim = 3 * n;
    const double polmin = 0.00000001;
    const double cutoffInd2 = p->cutoff_ind * p->cutoff_ind;
    const double a = p->a_damp;
    double *A = (double *)calloc((size_t)dim * dim, sizeof(double));
    if (!A) return -3;
    for (int k = 0; k < dim; ++k) { ei0[k] = 0.0; pi[k] = 0.0; }
    int rc = 0;
    for (int i = 0; i < n; ++i) {
        const int i3 = 3 * i;
        const int molIDi = s->mol[i];
        const v3 ri = v3_ld(s->pos + 3 * i);
        const double qi = s->q[i] / 11.78708976;   /* SQRTCOULOMBCONSTANT, src/NEMD_defs.h:103 */
        const double polarity = polar[i] > polmin ? polar[i] : polmin;
        const v3 dipi = v3_ld(dip + 3 * i);
        const double Qxxi = quad[6 * i], Qxyi = quad[6 * i + 1], Qxzi = quad[6 * i + 2];
        const double Qyyi = quad[6 * i + 3], Qyzi = quad[6 * i + 4], Qzzi = quad[6 * i + 5];
        const double Qyxi = Qxyi, Qzxi = Qxzi, Qzyi = Qyzi;
        A[(size_t)i3 * dim + i3] = 1.0 / polarity;
        A[(size_t)(i3 + 1) * dim + i3 + 1] = 1.0 / polarity;
        A[(size_t)(i3 + 2) * dim + i3 + 2] = 1.0 / polarity;
        for (int j = i + 1; j < n; ++j) {
            const int j3 = 3 * j;
            const int molIDj = s->mol[j];
            const v3 rj = v3_ld(s->pos + 3 * j);
            const double qj = s->q[j] / 11.78708976;
            const v3 dipj = v3_ld(dip + 3 * j);
            const double Qxxj = quad[6 * j], Qxyj = quad[6 * j + 1], Qxzj = quad[6 * j + 2];
            const double Qyyj = quad[6 * j + 3], Qyzj = quad[6 * j + 4], Qzzj = quad[6 * j + 5];
            const double Qyxj = Qxyj, Qzxj = Qxzj, Qzyj = Qyzj;
            v3 rij = v3_sub(ri, rj);
            apply_pbc(&rij, s);
            const double r = v3_len(rij);
            const double rij2 = r * r;
            if (!(rij2 <= cutoffInd2)) continue;
            const double rij_1 = 1.0 / r;
            const double rij_3 = rij_1 * rij_1 * rij_1;
            const double rij_5 = rij_3 * rij_1 * rij_1;
            const double rij_7 = rij_5 * rij_1 * rij_1;
            const double polarj = polar[j] > polmin ? polar[j] : polmin;
            const double u3 = rij2 * r / (sqrt(polarity * polarj));
            const double exp_au3 = exp(-a * u3);
            const double lamb3 = 1.0 - exp_au3;
            const double lamb5 = 1.0 - (1.0 + a * u3) * exp_au3;
            const double lamb7 = 1.0 - (1.0 + a * u3 + 3.0 * a * a * u3 * u3 / 5.0) * exp_au3;
            double T[3][3];
            const double rr[3] = {rij.x, rij.y, rij.z};
            for (int x = 0; x < 3; ++x)
                for (int y = x; y < 3; ++y) {
                    double t = p->damp_tij ? lamb5 * 3.0 * rij_5 * rr[x] * rr[y] : 3.0 * rij_5 * rr[x] * rr[y];
                    if (x == y) t = p->damp_tij ? t - lamb3 * rij_3 : t - rij_3;
                    T[x][y] = t;
                    T[y][x] = t;
                }
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y) {
                    A[(size_t)(i3 + x) * dim + j3 + y] = -T[x][y];
                    A[(size_t)(j3 + x) * dim + i3 + y] = -T[x][y];
                }
            if (excluded(s, i, j)) continue;
            if (molIDi == molIDj) { rc = -2; continue; }   /* the reference stops with ERRORMSG */
            const v3 rji = v3_scale(-1.0, rij);
            v3 RijXmuj, RjiXmui, QiXrji, QjXrij;
            RijXmuj.x = rij.x * rij.x * dipj.x + rij.x * rij.y * dipj.y + rij.x * rij.z * dipj.z;
            RijXmuj.y = rij.y * rij.x * dipj.x + rij.y * rij.y * dipj.y + rij.y * rij.z * dipj.z;
            RijXmuj.z = rij.z * rij.x * dipj.x + rij.z * rij.y * dipj.y + rij.z * rij.z * dipj.z;
            RjiXmui.x = rji.x * rji.x * dipi.x + rji.x * rji.y * dipi.y + rji.x * rji.z * dipi.z;
            RjiXmui.y = rji.y * rji.x * dipi.x + rji.y * rji.y * dipi.y + rji.y * rji.z * dipi.z;
            RjiXmui.z = rji.z * rji.x * dipi.x + rji.z * rji.y * dipi.y + rji.z * rji.z * dipi.z;
            QiXrji.x = Qxxi * rji.x + Qxyi * rji.y + Qxzi * rji.z;
            QiXrji.y = Qyxi * rji.x + Qyyi * rji.y + Qyzi * rji.z;
            QiXrji.z = Qzxi * rji.x + Qzyi * rji.y + Qzzi * rji.z;
            QjXrij.x = Qxxj * rij.x + Qxyj * rij.y + Qxzj * rij.z;
            QjXrij.y = Qyxj * rij.x + Qyyj * rij.y + Qyzj * rij.z;
            QjXrij.z = Qzxj * rij.x + Qzyj * rij.y + Qzzj * rij.z;
            const double QiI = Qxxi + Qyyi + Qzzi;
            const double QjI = Qxxj + Qyyj + Qzzj;
            const double QjRij = Qxxj * rij.x * rij.x + Qxyj * rij.x * rij.y + Qxzj * rij.x * rij.z +
                                 Qyxj * rij.y * rij.x + Qyyj * rij.y * rij.y + Qyzj * rij.y * rij.z +
                                 Qzxj * rij.z * rij.x + Qzyj * rij.z * rij.y + Qzzj * rij.z * rij.z;
            const double QiRji = Qxxi * rij.x * rij.x + Qxyi * rij.x * rij.y + Qxzi * rij.x * rij.z +
                                 Qyxi * rij.y * rij.x + Qyyi * rij.y * rij.y + Qyzi * rij.y * rij.z +
                                 Qzxi * rij.z * rij.x + Qzyi * rij.z * rij.y + Qzzi * rij.z * rij.z;
            /* bi (:423-425) and bj (:430-432), each term scalar * vector, summed left to right */
            const double ri3[3] = {rij.x, rij.y, rij.z}, rj3[3] = {rji.x, rji.y, rji.z};
            const double RM[3] = {RijXmuj.x, RijXmuj.y, RijXmuj.z}, RMi[3] = {RjiXmui.x, RjiXmui.y, RjiXmui.z};
            const double dj[3] = {dipj.x, dipj.y, dipj.z}, di[3] = {dipi.x, dipi.y, dipi.z};
            const double QJ[3] = {QjXrij.x, QjXrij.y, QjXrij.z}, QIv[3] = {QiXrji.x, QiXrji.y, QiXrji.z};
            for (int x = 0; x < 3; ++x) {
                const double bi = qj * lamb3 * rij_3 * ri3[x] + 3.0 * lamb5 * rij_5 * RM[x] - lamb3 * rij_3 * dj[x] +
                                  15.0 * lamb7 * rij_7 * QjRij * ri3[x] - 6.0 * lamb5 * rij_5 * QJ[x] -
                                  3.0 * lamb5 * rij_5 * QjI * ri3[x];
                ei0[i3 + x] += bi;
            }
            for (int x = 0; x < 3; ++x) {
                const double bj = qi * lamb3 * rij_3 * rj3[x] + 3.0 * lamb5 * rij_5 * RMi[x] - lamb3 * rij_3 * di[x] +
                                  15.0 * lamb7 * rij_7 * QiRji * rj3[x] - 6.0 * lamb5 * rij_5 * QIv[x] -
                                  3.0 * lamb5 * rij_5 * QiI * rj3[x];
                ei0[j3 + x] += bj;
            }
        }
    }
    int it = 0;
    double res = 0.0;
    if (p->do_min) {
        /* ConjugateGradient (:2401-2467) */
        const double tolerance = 1.0e-14;
        double *r = (double *)malloc(sizeof(double) * dim);
        double *v = (double *)malloc(sizeof(double) * dim);
        double *z = (double *)malloc(sizeof(double) * dim);
        double c = 0.0, t = 0.0, d = 0.0, dotV = 0.0;
        for (int k = 0; k < dim; ++k) pi[k] = 1.0;
        for (int k = 0; k < dim; ++k) {
            r[k] = ei0[k];
            for (int j = 0; j < dim; ++j) r[k] -= A[(size_t)k * dim + j] * pi[j];
            v[k] = r[k];
            c += r[k] * r[k];
            dotV += v[k] * v[k];
        }
        res = sqrt(c);
        for (int k = 0; k < dim; ++k) {
            if (sqrt(dotV) < tolerance) break;
            double dotVZ = 0.0;
            for (int j = 0; j < dim; ++j) {
                z[j] = 0.0;
                for (int l = 0; l < dim; ++l) z[j] += A[(size_t)j * dim + l] * v[l];
                dotVZ += v[j] * z[j];
            }
            t = c / dotVZ;
            d = 0.0;
            for (int j = 0; j < dim; ++j) {
                pi[j] += t * v[j];
                r[j] -= t * z[j];
                d += r[j] * r[j];
            }
            ++it;
            res = sqrt(d);
            if (sqrt(d) < tolerance) break;
            dotV = 0.0;
            for (int j = 0; j < dim; ++j) {
                v[j] = r[j] + (d / c) * v[j];
                dotV += v[j] * v[j];
            }
            c = d;
        }
        free(r); free(v); free(z);
    } else {
        for (int i = 0; i < n; ++i)
            for (int x = 0; x < 3; ++x) pi[3 * i + x] = polar[i] * ei0[3 * i + x];
    }
    free(A);
    if (iters) *iters = it;
    if (final_res) *final_res = res;
    return rc;
}

/* ---- permanent multipoles under Ewald (src/forces/Multipole.cpp) ---------- */

static const double COULOMBS = 138.93548501;          /* COUlOMBS, src/NEMD_defs.h:104 */
static const double SQRT_COULOMB = 11.78708976;       /* SQRTCOULOMBCONSTANT, src/NEMD_defs.h:103 */

static inline v3 v3_neg(v3 a) { v3 r = {-a.x, -a.y, -a.z}; return r; }
/* cross(v1, v2) (src/utils/Vector3.h:206-210) */
static inline v3 v3_cross(v3 a, v3 b) {
    v3 r = {a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y};
    return r;
}
/* cross(k[3][3], m[3][3], ones): row-wise cross products summed
 * (src/utils/Vector3.h:220-225) */
static inline v3 m3_cross(const double k[3][3], const double m[3][3]) {
    v3 r;
    r.x = k[1][0] * m[2][0] - k[2][0] * m[1][0] + k[1][1] * m[2][1] - k[2][1] * m[1][1] + k[1][2] * m[2][2] -
          k[2][2] * m[1][2];
    r.y = k[2][0] * m[0][0] - k[0][0] * m[2][0] + k[2][1] * m[0][1] - k[0][1] * m[2][1] + k[2][2] * m[0][2] -
          k[0][2] * m[2][2];
    r.z = k[0][0] * m[1][0] - k[1][0] * m[0][0] + k[0][1] * m[1][1] - k[1][1] * m[0][1] + k[0][2] * m[1][2] -
          k[1][2] * m[0][2];
    return r;
}

/* A site's lab-frame multipoles: charge q/sqrt(COULOMB), dipole, the full
 * symmetric quadrupole matrix (Qyx = Qxy, ...) as the reference expands it. */
typedef struct { double q; v3 d; double Q[3][3]; } mp_site;

static inline mp_site mp_load(const orc_sys *s, const double *dip, const double *quad, int i) {
    mp_site m;
    m.q = s->q[i] / SQRT_COULOMB;
    m.d = v3_ld(dip + 3 * i);
    const double *p = quad + 6 * i;
    m.Q[0][0] = p[0]; m.Q[0][1] = p[1]; m.Q[0][2] = p[2];
    m.Q[1][0] = p[1]; m.Q[1][1] = p[3]; m.Q[1][2] = p[4];
    m.Q[2][0] = p[2]; m.Q[2][1] = p[4]; m.Q[2][2] = p[5];
    return m;
}

/* Q.r, row by row, each row summed x, y, z */
static inline v3 m3_mul(const double Q[3][3], v3 r) {
    v3 o = {Q[0][0] * r.x + Q[0][1] * r.y + Q[0][2] * r.z, Q[1][0] * r.x + Q[1][1] * r.y + Q[1][2] * r.z,
            Q[2][0] * r.x + Q[2][1] * r.y + Q[2][2] * r.z};
    return o;
}
/* sum_ab Q_ab r_a r_b, in the reference's term order */
static inline double m3_quadform(const double Q[3][3], v3 r) {
    const double rr[3] = {r.x, r.y, r.z};
    double s = Q[0][0] * rr[0] * rr[0];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            if (a || b) s += Q[a][b] * rr[a] * rr[b];
    return s;
}
/* (Qi Qj) r: (sum_c Qi_ac Qj_cb) r_b summed over b */
static inline v3 m3_mm_mul(const double A[3][3], const double B[3][3], v3 r) {
    const double rr[3] = {r.x, r.y, r.z};
    double o[3];
    for (int a = 0; a < 3; ++a) {
        double s = 0.0;
        for (int b = 0; b < 3; ++b) {
            const double ab = A[a][0] * B[0][b] + A[a][1] * B[1][b] + A[a][2] * B[2][b];
            s = b == 0 ? ab * rr[b] : s + ab * rr[b];
        }
        o[a] = s;
    }
    v3 v = {o[0], o[1], o[2]};
    return v;
}

/* The G_l coefficients of one site pair and, optionally, their gradients and
 * the torque vectors Gt_l on both sites (Multipole::real_term,
 * src/forces/Multipole.cpp:548-617 and :730-877; the same expressions are
 * repeated in intraMolCorrect :1449-1590). */
typedef struct {
    double G[5];
    v3 gG[5];          /* gradG1..gradG4 in [1..4] */
    v3 Gti[5], Gtj[5]; /* torque vectors in [1..4] */
} mp_terms;

static void mp_pair_terms(const mp_site *I, const mp_site *J, v3 rij, int grads, int torques, mp_terms *t) {
    const double qi = I->q, qj = J->q;
    const v3 dipi = I->d, dipj = J->d;
    const double (*Qi)[3] = I->Q, (*Qj)[3] = J->Q;
    const double QiI = Qi[0][0] + Qi[1][1] + Qi[2][2];
    const double QjI = Qj[0][0] + Qj[1][1] + Qj[2][2];
    const double QiRij = m3_quadform(Qi, rij);
    const double QjRij = m3_quadform(Qj, rij);
    double QiQj = Qi[0][0] * Qj[0][0];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            if (a || b) QiQj += Qi[a][b] * Qj[a][b];
    const v3 QiXrij = m3_mul(Qi, rij), QjXrij = m3_mul(Qj, rij);
    const double QimjXrij = dipj.x * QiXrij.x + dipj.y * QiXrij.y + dipj.z * QiXrij.z;
    const double QjmiXrij = dipi.x * QjXrij.x + dipi.y * QjXrij.y + dipi.z * QjXrij.z;
    /* sum over rows a of Qj_a . r weighted by (Qi r)_a: rij.b*Qj[a][b]*(Qi r)_a */
    double QiXrijXQj = 0.0;
    {
        const double rr[3] = {rij.x, rij.y, rij.z};
        const double qr[3] = {QiXrij.x, QiXrij.y, QiXrij.z};
        int first = 1;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                const double term = rr[b] * Qj[a][b] * qr[a];
                QiXrijXQj = first ? term : QiXrijXQj + term;
                first = 0;
            }
    }
    const double mir = v3_dot(dipi, rij), mjr = v3_dot(dipj, rij);
    t->G[0] = qi * qj;
    t->G[1] = v3_dot(v3_scale(qj, dipi), rij) - v3_dot(v3_scale(qi, dipj), rij) - v3_dot(dipi, dipj) + qj * QiI + qi * QjI;
    t->G[2] = qj * QiRij + qi * QjRij - mir * mjr - mjr * QiI + mir * QjI - 2 * QimjXrij + 2 * QjmiXrij + QiI * QjI +
              2 * QiQj;
    t->G[3] = mir * QjRij - mjr * QiRij + QiI * QjRij + QjI * QiRij + 4 * QiXrijXQj;
    t->G[4] = QiRij * QjRij;
    if (grads) {
        const v3 QiXmj = m3_mul(Qi, dipj), QjXmi = m3_mul(Qj, dipi);
        const v3 QiXQjrij = m3_mm_mul(Qi, Qj, rij), QjXQirij = m3_mm_mul(Qj, Qi, rij);
        t->gG[1] = v3_sub(v3_scale(qj, dipi), v3_scale(qi, dipj));
        v3 g = v3_scale(2 * qj, QiXrij);
        g = v3_sub(g, v3_scale(mjr, dipi));
        g = v3_sub(g, v3_scale(mir, dipj));
        g = v3_sub(g, v3_scale(2, QiXmj));
        g = v3_sub(g, v3_scale(QiI, dipj));
        g = v3_add(g, v3_scale(2 * qi, QjXrij));
        g = v3_add(g, v3_scale(2, QjXmi));
        t->gG[2] = v3_add(g, v3_scale(QjI, dipi));
        g = v3_scale(-QiRij, dipj);
        g = v3_sub(g, v3_scale(v3_dot(v3_scale(2, dipj), rij), QiXrij));
        g = v3_add(g, v3_scale(QjRij, dipi));
        g = v3_add(g, v3_scale(v3_dot(v3_scale(2, dipi), rij), QjXrij));
        g = v3_add(g, v3_scale(2 * QiI, QjXrij));
        g = v3_add(g, v3_scale(2 * QjI, QiXrij));
        g = v3_add(g, v3_scale(4, QiXQjrij));
        t->gG[3] = v3_add(g, v3_scale(4, QjXQirij));
        t->gG[4] = v3_add(v3_scale(2 * QjRij, QiXrij), v3_scale(2 * QiRij, QjXrij));
    }
    if (torques) {
        const v3 rji = v3_neg(rij);
        const double rv[3] = {rij.x, rij.y, rij.z}, rjv[3] = {rji.x, rji.y, rji.z};
        const double dj[3] = {dipj.x, dipj.y, dipj.z}, di[3] = {dipi.x, dipi.y, dipi.z};
        const double qjr[3] = {QjXrij.x, QjXrij.y, QjXrij.z}, qir[3] = {QiXrij.x, QiXrij.y, QiXrij.z};
        double Rij[3][3], mjrij[3][3], rijmj[3][3], Qjrijrij[3][3], rijQjrij[3][3];
        double mirji[3][3], rjimi[3][3], Qirjirji[3][3], rjiQirji[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                Rij[a][b] = rv[a] * rv[b];
                mjrij[a][b] = dj[a] * rv[b];
                Qjrijrij[a][b] = qjr[a] * rv[b];
                mirji[a][b] = di[a] * rjv[b];
                Qirjirji[a][b] = qir[a] * rv[b];
            }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                rijmj[a][b] = mjrij[b][a];
                rijQjrij[a][b] = Qjrijrij[b][a];
                rjimi[a][b] = mirji[b][a];
                rjiQirji[a][b] = Qirjirji[b][a];
            }
        const v3 rijXmi = v3_cross(rij, dipi), mjXmi = v3_cross(dipj, dipi), QjrijXmi = v3_cross(QjXrij, dipi);
        const v3 RijXQi = m3_cross(Rij, Qi), mjrijXQi = m3_cross(mjrij, Qi), rijmjXQi = m3_cross(rijmj, Qi);
        const v3 QjXQi = m3_cross(Qj, Qi), rijQjrijXQi = m3_cross(rijQjrij, Qi), QjrijrijXQi = m3_cross(Qjrijrij, Qi);
        const v3 rjiXmj = v3_cross(rji, dipj), QirijXmj = v3_cross(QiXrij, dipj), RjiXQj = m3_cross(Rij, Qj);
        const v3 mirjiXQj = m3_cross(mirji, Qj), rjimiXQj = m3_cross(rjimi, Qj);
        const v3 rjiQirjiXQj = m3_cross(rjiQirji, Qj), QirjirjiXQj = m3_cross(Qirjirji, Qj);
        const double mjr2 = v3_dot(dipj, rij), mir2 = v3_dot(dipi, rji);
        t->Gti[1] = v3_sub(v3_scale(qj, rijXmi), mjXmi);
        v3 g = v3_scale(-mjr2, rijXmi);
        g = v3_add(g, v3_scale(2.0, QjrijXmi));
        g = v3_add(g, v3_scale(QjI, rijXmi));
        g = v3_add(g, v3_scale(2.0 * qj, RijXQi));
        g = v3_sub(g, v3_scale(2.0, mjrijXQi));
        g = v3_sub(g, v3_scale(2.0, rijmjXQi));
        t->Gti[2] = v3_add(g, v3_scale(4.0, QjXQi));
        g = v3_scale(QjRij, rijXmi);
        g = v3_sub(g, v3_scale(2.0 * mjr2, RijXQi));
        g = v3_add(g, v3_scale(2.0 * QjI, RijXQi));
        g = v3_add(g, v3_scale(4.0, rijQjrijXQi));
        t->Gti[3] = v3_add(g, v3_scale(4.0, QjrijrijXQi));
        t->Gti[4] = v3_scale(2.0 * QjRij, RijXQi);
        t->Gtj[1] = v3_add(v3_scale(qi, rjiXmj), mjXmi);
        g = v3_scale(-mir2, rjiXmj);
        g = v3_sub(g, v3_scale(2.0, QirijXmj));
        g = v3_add(g, v3_scale(QiI, rjiXmj));
        g = v3_add(g, v3_scale(2.0 * qi, RjiXQj));
        g = v3_sub(g, v3_scale(2.0, mirjiXQj));
        g = v3_sub(g, v3_scale(2.0, rjimiXQj));
        t->Gtj[2] = v3_sub(g, v3_scale(4.0, QjXQi));
        g = v3_scale(QiRij, rjiXmj);
        g = v3_sub(g, v3_scale(2.0 * mir2, RjiXQj));
        g = v3_add(g, v3_scale(2.0 * QiI, RjiXQj));
        g = v3_add(g, v3_scale(4.0, rjiQirjiXQj));
        t->Gtj[4] = v3_scale(2.0 * QiRij, RjiXQj);
        t->Gtj[3] = v3_add(g, v3_scale(4.0, QirjirjiXQj));
    }
}

/* Multipole::torques2forces (src/forces/Multipole.cpp:1946-2136): the torque
 * on each water site becomes forces on the atoms spanning its frame.  Sites
 * of a molecule are H1, O, H2.  Returns dU/dV before the 1/3V factor. */
static double mp_torques2forces(const orc_sys *s, const double *trq, double *force) {
    const double wrot_f = 0.5;
    double dUdVol = 0.0;
    for (int m = 0; m < s->n_mols; ++m) {
        const int h1 = s->mol_start[m], o = h1 + 1, h2 = h1 + 2;
        const v3 rh1 = v3_ld(s->pos + 3 * h1), ro = v3_ld(s->pos + 3 * o), rh2 = v3_ld(s->pos + 3 * h2);
        for (int side = 0; side < 2; ++side) {
            /* H1 (side 0) and H2 (side 1) share one form (:1988-2022, :2076-2110);
             * the reference handles H1, then O, then H2. */
            if (side == 1) {
                /* O site (:2024-2072) */
                const v3 tau = v3_ld(trq + 3 * o);
                v3 u = v3_sub(ro, rh1), v = v3_sub(ro, rh2);
                v3 w = v3_cross(u, v);
                const double ul = v3_len(u), vl = v3_len(v), wl = v3_len(w);
                u = v3_div(u, ul);
                v = v3_div(v, vl);
                w = v3_div(w, wl);
                const v3 vxu = v3_cross(v, u);
                const double vxuL = v3_len(vxu);
                const v3 wxu = v3_cross(w, u);
                const double wxuL = v3_len(wxu);
                const v3 dUdu = v3_add(v3_div(v3_scale(v3_dot(v3_neg(v), tau), vxu), vxuL * vxuL * ul),
                                       v3_div(v3_scale(wrot_f * v3_dot(v3_neg(w), tau), wxu), wxuL * wxuL * ul));
                const v3 wxv = v3_cross(w, v);
                const double wxvL = v3_len(wxv);
                const v3 dUdv = v3_add(v3_div(v3_scale(v3_dot(v3_neg(u), tau), v3_neg(vxu)), vxuL * vxuL * vl),
                                       v3_div(v3_scale(wrot_f * v3_dot(v3_neg(w), tau), wxv), wxvL * wxvL * vl));
                acc(force + 3 * o, v3_sub(v3_neg(dUdu), dUdv));
                acc(force + 3 * h1, dUdu);
                acc(force + 3 * h2, dUdv);
                dUdVol += v3_dot(dUdu, u) * ul + v3_dot(dUdv, v) * vl;
            }
            const int h = side == 0 ? h1 : h2;
            const v3 rh = side == 0 ? rh1 : rh2;
            const v3 tau = v3_ld(trq + 3 * h);
            v3 u = v3_sub(rh, ro);
            const double ul = v3_len(u);
            v3 v = {1.0, 0.0, 0.0};
            const double val = u.x / ul;
            if (fabs(val) > 0.8666) { v.x = 0.0; v.y = 1.0; }
            const double vl = 1.0;
            v3 w = v3_cross(u, v);
            const double wl = v3_len(w);
            u = v3_div(u, ul);
            v = v3_div(v, vl);
            w = v3_div(w, wl);
            const v3 vxu = v3_cross(v, u);
            const double vxuL = v3_len(vxu);
            const v3 wxu = v3_cross(w, u);
            const double wxuL = v3_len(wxu);
            const v3 dUdu = v3_add(v3_div(v3_scale(v3_dot(v3_neg(v), tau), vxu), vxuL * vxuL * ul),
                                   v3_div(v3_scale(v3_dot(v3_neg(w), tau), wxu), wxuL * wxuL * ul));
            acc(force + 3 * h, v3_neg(dUdu));
            acc(force + 3 * o, dUdu);
            dUdVol += v3_dot(dUdu, u) * ul;
        }
    }
    return dUdVol;
}

static double mp_torque_norms(int n, const double *trq) {
    double t = 0.0;
    for (int i = 0; i < n; ++i) t += v3_len(v3_ld(trq + 3 * i));
    return t;
}

int orc_multipole_short(const orc_sys *s, const int32_t *off, const int32_t *nb, const double *dip,
                        const double *quad, const orc_mp_params *p, double *force, orc_mp_out *out) {
    /* Multipole::compute (src/forces/Multipole.cpp:163-216): self, real space,
     * intramolecular correction and the surface term, each followed by its
     * torques2forces. */
    const int n = s->n_atoms;
    for (int m = 0; m < s->n_mols; ++m)
        if (s->mol_start[m + 1] - s->mol_start[m] != 3) return -1;
    double *trq = (double *)calloc((size_t)3 * n, sizeof(double));
    if (!trq) return -3;
    const double volume = s->box[0] * s->box[1] * s->box[2];
    const double volumer = 1.0 / volume;
    const double alpha = p->alpha, alpha2 = alpha * alpha, alpha2PI = 2.0 * alpha / SQRT_PI;
    const double realCut2 = p->real_cut * p->real_cut;
    memset(out, 0, sizeof *out);

    /* charged-system energy (ctor :96-104) */
    double qSum = 0.0;
    for (int i = 0; i < n; ++i) qSum += s->q[i];
    if (fabs(qSum) > 1.0e-5) out->charged_energy = -PI * qSum * qSum / (2.0 * volume * alpha2);

    /* self_correct (:338-395) */
    for (int i = 0; i < n; ++i) {
        const mp_site I = mp_load(s, dip, quad, i);
        const double QiI = I.Q[0][0] + I.Q[1][1] + I.Q[2][2];
        double QiQi = I.Q[0][0] * I.Q[0][0];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (a || b) QiQi += I.Q[a][b] * I.Q[a][b];
        const double selfterm = I.q * I.q * alpha2PI +
                                (v3_dot(I.d, I.d) - 2 * I.q * QiI) * 2 * alpha2PI * alpha2 / 3 +
                                (QiI * QiI + 2 * QiQi) * 4 * alpha2PI * alpha2 * alpha2 / 5;
        out->self_energy += -0.5 * selfterm * COULOMBS;
    }

    /* real_term (:398-949) over the half list */
    for (int i = 0; i < n; ++i) {
        const v3 ri = v3_ld(s->pos + 3 * i);
        const mp_site I = mp_load(s, dip, quad, i);
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const int j = nb[k];
            const v3 rj = v3_ld(s->pos + 3 * j);
            const mp_site J = mp_load(s, dip, quad, j);
            v3 rij = v3_sub(ri, rj);
            apply_pbc(&rij, s);
            const double r2 = v3_len2(rij);
            if (!(r2 < realCut2)) continue;
            mp_terms t;
            mp_pair_terms(&I, &J, rij, 1, p->torques, &t);
            const double r = sqrt(r2);
            double B[6];
            B[0] = erfc(alpha * r) / r;
            B[1] = (-B[0] - alpha2PI * exp(-alpha2 * r2)) / r2;
            B[2] = (-3 * B[1] + 2 * alpha2PI * alpha2 * exp(-alpha2 * r2)) / r2;
            B[3] = (-5 * B[2] - 4 * alpha2PI * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
            B[4] = (-7 * B[3] + 8 * alpha2PI * alpha2 * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
            B[5] = (-9 * B[4] - 16 * alpha2PI * alpha2 * alpha2 * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
            out->real_energy += (t.G[0] * B[0] + t.G[1] * B[1] + t.G[2] * B[2] + t.G[3] * B[3] + t.G[4] * B[4]) * COULOMBS;
            v3 gsum = v3_scale(B[1], t.gG[1]);
            gsum = v3_add(gsum, v3_scale(B[2], t.gG[2]));
            gsum = v3_add(gsum, v3_scale(B[3], t.gG[3]));
            gsum = v3_add(gsum, v3_scale(B[4], t.gG[4]));
            const double radial = t.G[0] * B[1] + t.G[1] * B[2] + t.G[2] * B[3] + t.G[3] * B[4] + t.G[4] * B[5];
            const v3 fij = v3_sub(v3_neg(gsum), v3_scale(radial, rij));
            out->real_deriv += v3_dot(v3_neg(fij), rij) * COULOMBS;
            acc(force + 3 * i, v3_scale(COULOMBS, fij));
            dec(force + 3 * j, v3_scale(COULOMBS, fij));
            vir_add(out->virial, fij, rij);
            if (p->torques) {
                v3 ti = v3_scale(B[1], t.Gti[1]);
                ti = v3_add(ti, v3_scale(B[2], t.Gti[2]));
                ti = v3_add(ti, v3_scale(B[3], t.Gti[3]));
                ti = v3_add(ti, v3_scale(B[4], t.Gti[4]));
                v3 tj = v3_scale(B[1], t.Gtj[1]);
                tj = v3_add(tj, v3_scale(B[2], t.Gtj[2]));
                tj = v3_add(tj, v3_scale(B[3], t.Gtj[3]));
                tj = v3_add(tj, v3_scale(B[4], t.Gtj[4]));
                acc(trq + 3 * i, v3_scale(COULOMBS, ti));
                acc(trq + 3 * j, v3_scale(COULOMBS, tj));
            }
        }
    }
    if (p->torques) {
        out->tot_torque += mp_torque_norms(n, trq);
        out->real_deriv = out->real_deriv + mp_torques2forces(s, trq, force);
    }

    /* intraMolCorrect (:1290-1750): excluded pairs inside each molecule */
    memset(trq, 0, sizeof(double) * 3 * n);
    {
        double e = 0.0;
        for (int m = 0; m < s->n_mols; ++m) {
            const int a0 = s->mol_start[m], ns = s->mol_start[m + 1] - a0;
            for (int a = 0; a < ns; ++a) {
                const int idi = a0 + a;
                const mp_site I = mp_load(s, dip, quad, idi);
                const v3 ri = v3_ld(s->pos + 3 * idi);
                for (int b = a + 1; b < ns; ++b) {
                    const int idj = a0 + b;
                    if (!excluded(s, idi, idj)) continue;
                    const v3 rj = v3_ld(s->pos + 3 * idj);
                    v3 rij = v3_sub(ri, rj);
                    apply_pbc(&rij, s);
                    const double r2 = v3_len2(rij);
                    const double r = sqrt(r2);
                    const mp_site J = mp_load(s, dip, quad, idj);
                    mp_terms t;
                    mp_pair_terms(&I, &J, rij, !p->use_constraint, !p->use_constraint && p->torques, &t);
                    double D[6];
                    D[0] = erf(alpha * r) / r;
                    D[1] = (-D[0] + alpha2PI * exp(-alpha2 * r2)) / r2;
                    D[2] = (-3 * D[1] - 2 * alpha2PI * alpha2 * exp(-alpha2 * r2)) / r2;
                    D[3] = (-5 * D[2] + 4 * alpha2PI * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
                    D[4] = (-7 * D[3] - 8 * alpha2PI * alpha2 * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
                    D[5] = (-9 * D[4] + 16 * alpha2PI * alpha2 * alpha2 * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
                    e += t.G[0] * D[0] + t.G[1] * D[1] + t.G[2] * D[2] + t.G[3] * D[3] + t.G[4] * D[4];
                    if (!p->use_constraint) {
                        v3 fij = v3_scale(D[1], t.gG[1]);
                        fij = v3_add(fij, v3_scale(D[2], t.gG[2]));
                        fij = v3_add(fij, v3_scale(D[3], t.gG[3]));
                        fij = v3_add(fij, v3_scale(D[4], t.gG[4]));
                        const double radial = t.G[0] * D[1] + t.G[1] * D[2] + t.G[2] * D[3] + t.G[3] * D[4] + t.G[4] * D[5];
                        fij = v3_add(fij, v3_scale(radial, rij));
                        out->mol_deriv += -v3_dot(fij, rij) * COULOMBS / (3.0 * volume);
                        acc(force + 3 * idi, v3_scale(COULOMBS, fij));
                        dec(force + 3 * idj, v3_scale(COULOMBS, fij));
                        if (p->torques) {
                            v3 ti = v3_scale(D[1], t.Gti[1]);
                            ti = v3_add(ti, v3_scale(D[2], t.Gti[2]));
                            ti = v3_add(ti, v3_scale(D[3], t.Gti[3]));
                            ti = v3_add(ti, v3_scale(D[4], t.Gti[4]));
                            v3 tj = v3_scale(D[1], t.Gtj[1]);
                            tj = v3_add(tj, v3_scale(D[2], t.Gtj[2]));
                            tj = v3_add(tj, v3_scale(D[3], t.Gtj[3]));
                            tj = v3_add(tj, v3_scale(D[4], t.Gtj[4]));
                            acc(trq + 3 * idi, v3_scale(-COULOMBS, ti));
                            acc(trq + 3 * idj, v3_scale(-COULOMBS, tj));
                        }
                    }
                }
            }
        }
        out->intra_energy = -e * COULOMBS;
    }
    if (p->torques) {
        out->tot_torque += mp_torque_norms(n, trq);
        out->mol_deriv += mp_torques2forces(s, trq, force) / (3.0 * volume);
    }

    /* surfDipoleCorrect (:1753-1945) */
    if (p->surface) {
        memset(trq, 0, sizeof(double) * 3 * n);
        const double coeff = 2.0 * PI * volumer / 3.0;
        v3 qrSum = {0, 0, 0}, qRsum = {0, 0, 0}, miSum = {0, 0, 0}, qirimi = {0, 0, 0};
        double selfSum = 0.0;
        for (int i = 0; i < n; ++i) {
            const v3 r = v3_ld(s->pos + 3 * i), rim = v3_ld(s->com + 3 * s->mol[i]);
            const mp_site I = mp_load(s, dip, quad, i);
            const double QiI = I.Q[0][0] + I.Q[1][1] + I.Q[2][2];
            qrSum = v3_add(qrSum, v3_scale(I.q, r));
            qRsum = v3_add(qRsum, v3_scale(I.q, rim));
            miSum = v3_add(miSum, I.d);
            qirimi = v3_add(qirimi, v3_add(v3_scale(I.q, r), I.d));
            selfSum += 2 * I.q * QiI - v3_dot(I.d, I.d);
        }
        const double selfterm = v3_dot(qirimi, qirimi) + selfSum;
        out->surf_energy = coeff * selfterm * COULOMBS;
        out->surf_deriv = -out->surf_energy * volumer +
                          v3_dot(v3_scale(2.0 * coeff * volumer, qirimi), qrSum) * COULOMBS / 3.0;
        if (p->use_constraint)
            out->surf_deriv2 = v3_dot(v3_scale(2.0 * coeff * volumer * volumer, qrSum), qrSum) -
                               v3_dot(v3_scale(16.0 * coeff * volumer * volumer, qrSum), qRsum) / 9.0 +
                               v3_dot(v3_scale(2.0 * coeff * volumer * volumer, qRsum), qRsum) / 9.0;
        else
            out->surf_deriv2 = v3_dot(v3_scale(4.0 * coeff * volumer * volumer, qrSum), qrSum) / 9.0;
        for (int i = 0; i < n; ++i) {
            const mp_site I = mp_load(s, dip, quad, i);
            const v3 ri = v3_ld(s->pos + 3 * i);
            const v3 fi = v3_scale(-2.0 * coeff * I.q, qirimi);
            const v3 temp = v3_scale(-s->q[i], qrSum);
            acc(force + 3 * i, v3_scale(COULOMBS, fi));
            vir_add(out->virial, temp, ri);
            if (p->torques) acc(trq + 3 * i, v3_scale(2.0 * coeff * COULOMBS, v3_cross(qirimi, I.d)));
        }
        if (p->torques) {
            out->tot_torque += mp_torque_norms(n, trq);
            out->surf_deriv += mp_torques2forces(s, trq, force) * volumer / 3.0;
        }
    }
    free(trq);
    return 0;
}

int orc_multipole_long(const orc_sys *s, int K, const double *kvec, const double *kmod2, const double *expk,
                       const double *dip, const double *quad, const orc_mp_params *p, double *force,
                       orc_mp_out *out) {
    /* Multipole::compute_long / reciprocal_term (src/forces/Multipole.cpp:218-240,
     * :951-1279). */
    const int n = s->n_atoms;
    for (int m = 0; m < s->n_mols; ++m)
        if (s->mol_start[m + 1] - s->mol_start[m] != 3) return -1;
    const double volume = s->box[0] * s->box[1] * s->box[2];
    const double volumer = 1.0 / volume;
    const double pi4Volumer = 4.0 * PI * volumer, pi8Volumer = 2.0 * pi4Volumer;
    const double alphaR4 = 1.0 / (4.0 * (p->alpha * p->alpha));
    double *buf = (double *)calloc((size_t)19 * n, sizeof(double));
    if (!buf) return -3;
    double *qQcos = buf, *qQsin = buf + n, *mkcos = buf + 2 * n, *mksin = buf + 3 * n, *trq = buf + 4 * n;
    double *kXmsina = buf + 7 * n, *kXmcosa = buf + 10 * n, *KXQcosa = buf + 13 * n, *KXQsina = buf + 16 * n;
    out->long_energy = out->long_deriv = out->long_deriv2 = 0.0;
    memset(out->long_virial, 0, sizeof out->long_virial);
    for (int j = 0; j < K; ++j) {
        const v3 kj = v3_ld(kvec + 3 * j);
        double Sc = 0.0, Ss = 0.0, Ms = 0.0, Mc = 0.0, Vc = 0.0, Vs = 0.0;
        const double coeff = 2.0 * (alphaR4 + 1.0 / kmod2[j]);
        v3 cosrR = {0, 0, 0}, sinrR = {0, 0, 0};
        double Kj[3][3];
        const double kk[3] = {kj.x, kj.y, kj.z};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Kj[a][b] = kk[a] * kk[b];
        for (int i = 0; i < n; ++i) {
            const v3 rim = v3_ld(s->com + 3 * s->mol[i]);
            const mp_site I = mp_load(s, dip, quad, i);
            const v3 ri = v3_ld(s->pos + 3 * i);
            const double a = v3_dot(ri, kj);
            const double QiK = m3_quadform(I.Q, kj);
            const double ca = cos(a), sa = sin(a);
            const double cosAqiQi = (I.q - QiK) * ca;
            const double sinAqiQi = (I.q - QiK) * sa;
            const double mikcos = v3_dot(I.d, kj) * ca;
            const double miksin = v3_dot(I.d, kj) * sa;
            const double qi2Qicos = 2.0 * QiK * ca;
            const double qi2Qisin = 2.0 * QiK * sa;
            qQcos[i] = cosAqiQi;
            qQsin[i] = sinAqiQi;
            mkcos[i] = mikcos;
            mksin[i] = miksin;
            Sc += cosAqiQi;
            Ss += sinAqiQi;
            Mc += mikcos;
            Ms += miksin;
            Vc += qi2Qicos;
            Vs += qi2Qisin;
            cosrR = v3_add(cosrR, v3_scale(cosAqiQi, v3_sub(ri, rim)));
            sinrR = v3_add(sinrR, v3_scale(sinAqiQi, v3_sub(ri, rim)));
            if (p->torques) {
                const v3 kjXmi = v3_cross(kj, I.d);
                const v3 kjkjXQi = m3_cross(Kj, I.Q);
                v3_st(kXmsina + 3 * i, v3_scale(sa, kjXmi));
                v3_st(kXmcosa + 3 * i, v3_scale(ca, kjXmi));
                v3_st(KXQcosa + 3 * i, v3_scale(ca, kjkjXQi));
                v3_st(KXQsina + 3 * i, v3_scale(sa, kjkjXQi));
            }
        }
        const double UqQ = pi4Volumer * expk[j] * (Sc * Sc + Ss * Ss);
        const double UmqQ = pi4Volumer * expk[j] * (Ss * Mc - Sc * Ms);
        const double Umm = pi4Volumer * expk[j] * (Mc * Mc + Ms * Ms);
        const double tmpenergy = (UqQ + 2 * UmqQ + Umm) * COULOMBS;
        out->long_energy += tmpenergy;
        const double dUdVqQ = pi8Volumer * expk[j] * (Vc * Sc + Vs * Ss);
        const double dUdVqQm = pi8Volumer * expk[j] * (Vs * Mc - Vc * Ms);
        const double dUdV3rd = (dUdVqQ - 2.0 * UmqQ + dUdVqQm - 2.0 * Umm) * COULOMBS;
        const double tkk = v3_dot(v3_scale(tmpenergy, kj), kj) * 2.0 * alphaR4;
        out->long_deriv += -tmpenergy + tkk + dUdV3rd;
        out->long_deriv2 += 4.0 * tmpenergy - v3_dot(v3_scale(7.0 * tmpenergy, kj), kj) * 2.0 * alphaR4 +
                            v3_dot(v3_scale(tkk, kj), kj) * 2.0 * alphaR4;
        double *lv = out->long_virial;
        lv[XX] += tmpenergy - tmpenergy * coeff * kj.x * kj.x;
        lv[YY] += tmpenergy - tmpenergy * coeff * kj.y * kj.y;
        lv[ZZ] += tmpenergy - tmpenergy * coeff * kj.z * kj.z;
        lv[XY] -= tmpenergy * coeff * kj.x * kj.y;
        lv[XZ] -= tmpenergy * coeff * kj.x * kj.z;
        lv[YZ] -= tmpenergy * coeff * kj.y * kj.z;
        const double b = pi8Volumer * expk[j];
        double Ff = 0.0, Ffk2 = 0.0, d2UV1 = 0.0, d2UrR = 0.0;
        for (int i = 0; i < n; ++i) {
            const double FqQ = b * (qQsin[i] * Sc - qQcos[i] * Ss);
            const double FmiqQ = b * (mkcos[i] * Sc + mksin[i] * Ss);
            const double FqQmi = -b * (qQcos[i] * Mc + qQsin[i] * Ms);
            const double Fmm = b * (mksin[i] * Mc - mkcos[i] * Ms);
            const double val = COULOMBS * (FqQ + FmiqQ + FqQmi + Fmm);
            const double bkk = v3_dot(v3_scale(b, kj), kj);
            const double d2u1 = bkk * (qQcos[i] * Sc + qQsin[i] * Ss);
            const v3 rim = v3_ld(s->com + 3 * s->mol[i]);
            const v3 ri = v3_ld(s->pos + 3 * i);
            const v3 dr = v3_sub(ri, rim);
            const double d2u2 = v3_dot(v3_scale(bkk, v3_add(v3_scale(qQcos[i], cosrR), v3_scale(qQsin[i], sinrR))), dr);
            acc(force + 3 * i, v3_scale(val, kj));
            Ff += val * v3_dot(kj, dr);
            Ffk2 += v3_dot(v3_scale(2.0 * val, kj), kj) * 2.0 * alphaR4 * v3_dot(kj, dr);
            d2UV1 += v3_dot(v3_scale(d2u1, dr), dr);
            d2UrR += d2u2;
            if (p->torques) {
                const v3 kmc = v3_ld(kXmcosa + 3 * i), kms = v3_ld(kXmsina + 3 * i);
                const v3 kqc = v3_ld(KXQcosa + 3 * i), kqs = v3_ld(KXQsina + 3 * i);
                const v3 taum_qQ = v3_scale(b, v3_sub(v3_scale(Ss, kmc), v3_scale(Sc, kms)));
                const v3 taum_m = v3_scale(b, v3_add(v3_scale(Mc, kmc), v3_scale(Ms, kms)));
                const v3 tauQ_qQ = v3_scale(-2.0 * b, v3_add(v3_scale(Sc, kqc), v3_scale(Ss, kqs)));
                const v3 tauQ_m = v3_scale(2.0 * b, v3_sub(v3_scale(Ms, kqc), v3_scale(Mc, kqs)));
                acc(trq + 3 * i, v3_scale(COULOMBS, v3_add(v3_add(v3_add(taum_qQ, taum_m), tauQ_qQ), tauQ_m)));
            }
        }
        if (p->use_constraint) {
            out->long_deriv += Ff;
            out->long_deriv2 += -6.0 * Ff + Ffk2 - d2UV1 + d2UrR;
        }
    }
    out->long_virial[YX] = out->long_virial[XY];
    out->long_virial[ZX] = out->long_virial[XZ];
    out->long_virial[ZY] = out->long_virial[YZ];
    if (p->torques) {
        out->tot_torque += mp_torque_norms(n, trq);
        out->long_deriv += mp_torques2forces(s, trq, force);
    }
    free(buf);
    return 0;
}

/* ---- induced dipoles in the Ewald sum (src/forces/Induction.cpp) ---------- */

/* cross(Q, a b^T, ones) with the matrices written out, as real_term builds
 * them (:1190-1275) */
static inline v3 m3_cross_qo(const double Q[3][3], v3 a, v3 b) {
    const double av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    double m[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r][c] = av[r] * bv[c];
    return m3_cross(Q, m);
}

/* Induction terms of one site pair (real_term :940-1028 and intraMolCorrect
 * :1880-1935): G_l of site i's permanent multipoles with j's induced dipole
 * (and the reverse), their gradients, and the torque vectors. */
typedef struct {
    double G1i, G2i, G3i, G1j, G2j, G3j;
    v3 gG1i, gG2i, gG3i, gG1j, gG2j, gG3j;
    v3 Gt1i, Gt2i, Gt3i, Gt1j, Gt2j, Gt3j;
} ind_terms;

static void ind_pair_terms(const mp_site *I, const mp_site *J, v3 pi, v3 pj, v3 rij, int grads, int torques,
                           ind_terms *t) {
    const double qi = I->q, qj = J->q;
    const v3 dipi = I->d, dipj = J->d;
    const double (*Qi)[3] = I->Q, (*Qj)[3] = J->Q;
    const v3 rji = v3_neg(rij);
    const double QiI = Qi[0][0] + Qi[1][1] + Qi[2][2];
    const double QjI = Qj[0][0] + Qj[1][1] + Qj[2][2];
    const double QiRij = m3_quadform(Qi, rij);
    const double QjRji = m3_quadform(Qj, rij);
    const v3 QiXrij = m3_mul(Qi, rij), QjXrji = m3_mul(Qj, rji);
    const double QipjXrij = pj.x * QiXrij.x + pj.y * QiXrij.y + pj.z * QiXrij.z;
    const double QjpiXrji = pi.x * QjXrji.x + pi.y * QjXrji.y + pi.z * QjXrji.z;
    t->G1i = v3_dot(v3_scale(-qi, pj), rij) - v3_dot(dipi, pj);
    t->G2i = -v3_dot(dipi, rij) * v3_dot(pj, rij) - 2.0 * QipjXrij - v3_dot(pj, rij) * QiI;
    t->G3i = v3_dot(v3_neg(pj), rij) * QiRij;
    t->G1j = v3_dot(v3_scale(-qj, pi), rji) - v3_dot(dipj, pi);
    t->G2j = -v3_dot(dipj, rji) * v3_dot(pi, rji) - 2.0 * QjpiXrji - v3_dot(pi, rji) * QjI;
    t->G3j = v3_dot(v3_neg(pi), rji) * QjRji;
    if (grads) {
        const v3 QiXpj = m3_mul(Qi, pj), QjXpi = m3_mul(Qj, pi);
        t->gG1i = v3_scale(-qi, pj);
        v3 g = v3_scale(-v3_dot(pj, rij), dipi);
        g = v3_sub(g, v3_scale(v3_dot(dipi, rij), pj));
        g = v3_sub(g, v3_scale(2, QiXpj));
        t->gG2i = v3_sub(g, v3_scale(QiI, pj));
        t->gG3i = v3_sub(v3_scale(-QiRij, pj), v3_scale(v3_dot(v3_scale(2, pj), rij), QiXrij));
        t->gG1j = v3_scale(-qj, pi);
        g = v3_scale(-v3_dot(pi, rji), dipj);
        g = v3_sub(g, v3_scale(v3_dot(dipj, rji), pi));
        g = v3_sub(g, v3_scale(2, QjXpi));
        t->gG2j = v3_sub(g, v3_scale(QjI, pi));
        t->gG3j = v3_sub(v3_scale(-QjRji, pi), v3_scale(v3_dot(v3_scale(2, pi), rji), QjXrji));
    }
    if (torques) {
        double Rij[3][3];
        const double rv[3] = {rij.x, rij.y, rij.z};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Rij[a][b] = rv[a] * rv[b];
        const v3 QiXpjrij = m3_cross_qo(Qi, pj, rij), QiXrijpj = m3_cross_qo(Qi, rij, pj);
        const v3 QiXRij = m3_cross(Qi, Rij);
        const v3 QjXpirji = m3_cross_qo(Qj, pi, rji), QjXrjipi = m3_cross_qo(Qj, rji, pi);
        const v3 QjXRji = m3_cross(Qj, Rij);
        t->Gt1i = v3_cross(dipi, pj);
        t->Gt2i = v3_add(v3_add(v3_scale(v3_dot(pj, rij), v3_cross(dipi, rij)), v3_scale(2.0, QiXpjrij)),
                         v3_scale(2.0, QiXrijpj));
        t->Gt3i = v3_scale(2.0 * v3_dot(pj, rij), QiXRij);
        t->Gt1j = v3_cross(dipj, pi);
        t->Gt2j = v3_add(v3_add(v3_scale(v3_dot(pi, rji), v3_cross(dipj, rji)), v3_scale(2.0, QjXpirji)),
                         v3_scale(2.0, QjXrjipi));
        t->Gt3j = v3_scale(2.0 * v3_dot(pi, rji), QjXRji);
    }
}

int orc_induction_short(const orc_sys *s, const double *dip, const double *quad, const double *polar,
                        const double *pind, const orc_indf_params *p, double *force, orc_indf_out *out) {
    /* Induction::compute after InducedDipoles (src/forces/Induction.cpp:486-565):
     * self_correct (:688-760), real_term (:762-1393), intraMolCorrect
     * (:1760-2186), surfDipoleCorrect (:2188-2399), torques2forces after each
     * (:2557-2744, the same map as Multipole's). */
    const int n = s->n_atoms;
    for (int m = 0; m < s->n_mols; ++m)
        if (s->mol_start[m + 1] - s->mol_start[m] != 3) return -1;
    double *trq = (double *)calloc((size_t)3 * n, sizeof(double));
    if (!trq) return -3;
    const double volume = s->box[0] * s->box[1] * s->box[2], volumer = 1.0 / volume;
    const double alpha = p->alpha, alpha2 = alpha * alpha, alpha2PI = 2.0 * alpha / SQRT_PI;
    const double realCut2 = p->real_cut * p->real_cut, a = p->a_damp;
    memset(out, 0, sizeof *out);

    /* self_correct */
    for (int i = 0; i < n; ++i) {
        const mp_site I = mp_load(s, dip, quad, i);
        const v3 indPi = v3_ld(pind + 3 * i);
        const double selfterm = v3_dot(I.d, indPi) * alpha2PI * alpha2 / 3;
        out->self_energy += -selfterm * COULOMBS;
        if (p->torques) {
            const double term = COULOMBS * (2.0 / 3.0) * alpha2PI * alpha2;
            v3_st(trq + 3 * i, v3_scale(term, v3_cross(I.d, indPi)));
        }
    }
    if (p->torques) {
        out->tot_torque += mp_torque_norms(n, trq);
        out->self_deriv += mp_torques2forces(s, trq, force);
    }

    /* real_term: every pair j > i, minimum image */
    memset(trq, 0, sizeof(double) * 3 * n);
    for (int i = 0; i < n; ++i) {
        const mp_site I = mp_load(s, dip, quad, i);
        const v3 ri = v3_ld(s->pos + 3 * i), indPi = v3_ld(pind + 3 * i);
        for (int j = i + 1; j < n; ++j) {
            const v3 rj = v3_ld(s->pos + 3 * j);
            v3 rij = v3_sub(ri, rj);
            apply_pbc(&rij, s);
            const double r2 = v3_len2(rij);
            if (!(r2 < realCut2)) continue;
            const mp_site J = mp_load(s, dip, quad, j);
            const v3 indPj = v3_ld(pind + 3 * j);
            const v3 rji = v3_neg(rij);
            ind_terms t;
            ind_pair_terms(&I, &J, indPi, indPj, rij, 1, p->torques, &t);
            const double r = sqrt(r2);
            const double u3 = r * r * r / (sqrt(polar[i] * polar[j]));
            double B[6];
            B[0] = erfc(alpha * r) / r;
            B[1] = (-B[0] - alpha2PI * exp(-alpha2 * r2)) / r2;
            B[2] = (-3 * B[1] + 2 * alpha2PI * alpha2 * exp(-alpha2 * r2)) / r2;
            B[3] = (-5 * B[2] - 4 * alpha2PI * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
            B[4] = (-7 * B[3] + 8 * alpha2PI * alpha2 * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
            B[5] = (-9 * B[4] - 16 * alpha2PI * alpha2 * alpha2 * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
            const double expau3 = exp(-a * u3), au3 = a * u3, a2u6 = au3 * au3, a3u9 = a2u6 * au3;
            const double lambda3 = 1.0 - expau3;
            const double lambda5 = 1.0 - (1.0 + au3) * expau3;
            const double lambda7 = 1.0 - (1.0 + au3 + 3.0 * a2u6 / 5.0) * expau3;
            const double gradlam3 = 3.0 * au3 * expau3 / r2;
            const double gradlam5 = 3.0 * a2u6 * expau3 / r2;
            const double gradlam7 = (-3.0 * a2u6 + 9.0 * a3u9) * expau3 / (5.0 * r2);
            const double scale = excluded(s, i, j) ? 0.0 : 1.0;
            double B1d, B2d, B3d, B4d = 0, gradB1 = 0, gradB2 = 0, gradB3 = 0, B2du = 0, gradB1u = 0, gradB2u = 0;
            if (p->damp_amoeba) {
                const double invr = 1.0 / r, invr2 = invr * invr, invr3 = invr2 * invr;
                const double invr5 = 3.0 * invr3 * invr2, invr7 = 5.0 * invr5 * invr2, invr9 = 7.0 * invr7 * invr2;
                B1d = B[1] + (1.0 - scale * lambda3) * invr3;
                B2d = B[2] - (1.0 - scale * lambda5) * invr5;
                B3d = B[3] + (1.0 - scale * lambda7) * invr7;
                gradB1 = B[2] - scale * gradlam3 * invr3 - (1.0 - scale * lambda3) * invr5;
                gradB2 = B[3] + scale * gradlam5 * invr5 + (1.0 - scale * lambda5) * invr7;
                gradB3 = B[4] - scale * gradlam7 * invr7 - (1.0 - scale * lambda7) * invr9;
                B2du = B[2] - (1.0 - lambda5) * invr5;
                gradB1u = B[2] - gradlam3 * invr3 - (1.0 - lambda3) * invr5;
                gradB2u = B[3] + gradlam5 * invr5 + (1.0 - lambda5) * invr7;
            } else {
                B1d = B[1] * lambda3;
                B2d = B[2] * lambda3 + B[1] * gradlam3;
                B3d = B[3] * lambda3 + B[2] * 2.0 * gradlam3 + B[1] * (gradlam3 - 3.0 * gradlam5) / r2;
                B4d = B[4] * lambda3 + B[3] * 3.0 * gradlam3 + B[2] * (3.0 * gradlam3 - 9.0 * gradlam5) / r2 +
                      B[1] * (-gradlam3 - 6.0 * gradlam5 + 15.0 * gradlam7) / (r2 * r2);
            }
            const double Gsum = (t.G1i + t.G1j) * B1d + (t.G2i + t.G2j) * B2d + (t.G3i + t.G3j) * B3d;
            out->real_energy += p->damp_amoeba ? 0.5 * Gsum * COULOMBS : 0.5 * Gsum * scale * COULOMBS;
            const double G1_ip = v3_dot(v3_neg(indPi), indPj);
            const double G2_ip = -v3_dot(indPi, rij) * v3_dot(indPj, rij);
            const v3 gradG2_ip = v3_scale(-v3_dot(indPj, rij), indPi);
            const v3 gradG2_jp = v3_scale(-v3_dot(indPi, rji), indPj);
            v3 fi, fj;
            if (p->damp_amoeba) {
                v3 gi = v3_add(v3_add(v3_scale(B1d, t.gG1i), v3_scale(B2d, t.gG2i)), v3_scale(B3d, t.gG3i));
                v3 gj = v3_add(v3_add(v3_scale(B1d, t.gG1j), v3_scale(B2d, t.gG2j)), v3_scale(B3d, t.gG3j));
                double si = t.G1i * gradB1 + t.G2i * gradB2 + t.G3i * gradB3;
                const double sj = t.G1j * gradB1 + t.G2j * gradB2 + t.G3j * gradB3;
                if (p->do_min) {
                    gi = v3_add(gi, v3_scale(B2du, gradG2_ip));
                    gj = v3_add(gj, v3_scale(B2du, gradG2_jp));
                    si = si + G1_ip * gradB1u + G2_ip * gradB2u;
                }
                fi = v3_sub(v3_neg(gi), v3_scale(si, rij));
                fj = v3_sub(v3_neg(gj), v3_scale(sj, rji));
            } else if (p->do_min) {
                v3 gi = v3_scale(scale, v3_scale(B1d, t.gG1i));
                gi = v3_add(gi, v3_scale(scale, v3_scale(B2d, t.gG2i)));
                gi = v3_add(gi, v3_scale(scale, v3_scale(B3d, t.gG3i)));
                gi = v3_add(gi, v3_scale(B2d, gradG2_ip));
                v3 gj = v3_scale(scale, v3_scale(B1d, t.gG1j));
                gj = v3_add(gj, v3_scale(scale, v3_scale(B2d, t.gG2j)));
                gj = v3_add(gj, v3_scale(scale, v3_scale(B3d, t.gG3j)));
                gj = v3_add(gj, v3_scale(B2d, gradG2_jp));
                const double si = t.G1i * B2d * scale + t.G2i * B3d * scale + t.G3i * B4d * scale + G1_ip * B2d + G2_ip * B3d;
                const double sj = t.G1j * B2d * scale + t.G2j * B3d * scale + t.G3j * B4d * scale;
                fi = v3_sub(v3_neg(gi), v3_scale(si, rij));
                fj = v3_sub(v3_neg(gj), v3_scale(sj, rji));
            } else {
                const v3 gi = v3_add(v3_add(v3_scale(B1d, t.gG1i), v3_scale(B2d, t.gG2i)), v3_scale(B3d, t.gG3i));
                const v3 gj = v3_add(v3_add(v3_scale(B1d, t.gG1j), v3_scale(B2d, t.gG2j)), v3_scale(B3d, t.gG3j));
                const double si = t.G1i * B2d + t.G2i * B3d + t.G3i * B4d;
                const double sj = t.G1j * B2d + t.G2j * B3d + t.G3j * B4d;
                fi = v3_sub(v3_scale(scale, v3_neg(gi)), v3_scale(scale, v3_scale(si, rij)));
                fj = v3_sub(v3_scale(scale, v3_neg(gj)), v3_scale(scale, v3_scale(sj, rji)));
            }
            const v3 fij = v3_sub(fi, fj);
            out->real_deriv += v3_dot(v3_neg(fij), rij) * COULOMBS;
            acc(force + 3 * i, v3_scale(COULOMBS, fij));
            dec(force + 3 * j, v3_scale(COULOMBS, fij));
            vir_add(out->virial, fij, rij);
            if (p->torques) {
                const v3 ti = v3_add(v3_add(v3_scale(B1d, t.Gt1i), v3_scale(B2d, t.Gt2i)), v3_scale(B3d, t.Gt3i));
                const v3 tj = v3_add(v3_add(v3_scale(B1d, t.Gt1j), v3_scale(B2d, t.Gt2j)), v3_scale(B3d, t.Gt3j));
                if (p->damp_amoeba) {
                    acc(trq + 3 * i, v3_scale(COULOMBS, ti));
                    acc(trq + 3 * j, v3_scale(COULOMBS, tj));
                } else {
                    acc(trq + 3 * i, v3_scale(scale, v3_scale(COULOMBS, ti)));
                    acc(trq + 3 * j, v3_scale(scale, v3_scale(COULOMBS, tj)));
                }
            }
        }
    }
    if (p->torques) {
        out->tot_torque += mp_torque_norms(n, trq);
        out->real_deriv += mp_torques2forces(s, trq, force);
    }

    /* intraMolCorrect */
    memset(trq, 0, sizeof(double) * 3 * n);
    {
        double e = 0.0;
        for (int m = 0; m < s->n_mols; ++m) {
            const int a0 = s->mol_start[m], ns = s->mol_start[m + 1] - a0;
            for (int x = 0; x < ns; ++x) {
                const int idi = a0 + x;
                const mp_site I = mp_load(s, dip, quad, idi);
                const v3 indPi = v3_ld(pind + 3 * idi);
                const v3 ri = v3_ld(s->pos + 3 * idi);
                for (int y = x + 1; y < ns; ++y) {
                    const int idj = a0 + y;
                    if (!excluded(s, idi, idj)) continue;
                    const v3 rj = v3_ld(s->pos + 3 * idj);
                    v3 rij = v3_sub(ri, rj);
                    apply_pbc(&rij, s);
                    const double r2 = v3_len2(rij), r = sqrt(r2);
                    const mp_site J = mp_load(s, dip, quad, idj);
                    const v3 indPj = v3_ld(pind + 3 * idj);
                    const v3 rji = v3_neg(rij);
                    ind_terms t;
                    ind_pair_terms(&I, &J, indPi, indPj, rij, !p->use_constraint, !p->use_constraint && p->torques, &t);
                    double D[6];
                    D[0] = erf(alpha * r) / r;
                    D[1] = (-D[0] + alpha2PI * exp(-alpha2 * r2)) / r2;
                    D[2] = (-3 * D[1] - 2 * alpha2PI * alpha2 * exp(-alpha2 * r2)) / r2;
                    D[3] = (-5 * D[2] + 4 * alpha2PI * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
                    D[4] = (-7 * D[3] - 8 * alpha2PI * alpha2 * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
                    D[5] = (-9 * D[4] + 16 * alpha2PI * alpha2 * alpha2 * alpha2 * alpha2 * exp(-alpha2 * r2)) / r2;
                    e += (t.G1i + t.G1j) * D[1] + (t.G2i + t.G2j) * D[2] + (t.G3i + t.G3j) * D[3];
                    if (!p->use_constraint) {
                        v3 fi = v3_add(v3_add(v3_scale(D[1], t.gG1i), v3_scale(D[2], t.gG2i)), v3_scale(D[3], t.gG3i));
                        fi = v3_add(fi, v3_scale(t.G1i * D[2] + t.G2i * D[3] + t.G3i * D[4], rij));
                        v3 fj = v3_add(v3_add(v3_scale(D[1], t.gG1j), v3_scale(D[2], t.gG2j)), v3_scale(D[3], t.gG3j));
                        fj = v3_add(fj, v3_scale(t.G1j * D[2] + t.G2j * D[3] + t.G3j * D[4], rji));
                        const v3 fij = v3_sub(fi, fj);
                        out->mol_deriv += v3_dot(v3_neg(fij), rij) * COULOMBS / (3.0 * volume);
                        acc(force + 3 * idi, v3_scale(COULOMBS, fij));
                        dec(force + 3 * idj, v3_scale(COULOMBS, fij));
                        if (p->torques) {
                            const v3 ti = v3_add(v3_add(v3_scale(D[1], t.Gt1i), v3_scale(D[2], t.Gt2i)), v3_scale(D[3], t.Gt3i));
                            const v3 tj = v3_add(v3_add(v3_scale(D[1], t.Gt1j), v3_scale(D[2], t.Gt2j)), v3_scale(D[3], t.Gt3j));
                            acc(trq + 3 * idi, v3_scale(-COULOMBS, ti));
                            acc(trq + 3 * idj, v3_scale(-COULOMBS, tj));
                        }
                    }
                }
            }
        }
        out->intra_energy = -0.5 * e * COULOMBS;
    }
    if (p->torques) {
        out->tot_torque += mp_torque_norms(n, trq);
        out->mol_deriv += mp_torques2forces(s, trq, force) / (3.0 * volume);
    }

    /* surfDipoleCorrect */
    if (p->surface) {
        memset(trq, 0, sizeof(double) * 3 * n);
        const double coeff = 2.0 * PI * volumer / 3.0;
        v3 qrSum = {0, 0, 0}, qRsum = {0, 0, 0}, miSum = {0, 0, 0}, piSum = {0, 0, 0};
        double selfSum = 0.0, selfpi = 0.0;
        for (int i = 0; i < n; ++i) {
            const v3 r = v3_ld(s->pos + 3 * i), rim = v3_ld(s->com + 3 * s->mol[i]);
            const mp_site I = mp_load(s, dip, quad, i);
            const v3 indPi = v3_ld(pind + 3 * i);
            qrSum = v3_add(qrSum, v3_scale(I.q, r));
            qRsum = v3_add(qRsum, v3_scale(I.q, rim));
            miSum = v3_add(miSum, I.d);
            piSum = v3_add(piSum, indPi);
            selfSum += v3_dot(I.d, indPi);
            selfpi += v3_dot(indPi, indPi);
        }
        const double selfterm = v3_dot(qrSum, piSum) + v3_dot(miSum, piSum) - selfSum;
        out->surf_energy = coeff * selfterm * COULOMBS;
        double dUdV1;
        if (p->do_min)
            dUdV1 = -coeff * volumer * (2.0 * v3_dot(qrSum, piSum) + 2.0 * v3_dot(miSum, piSum) + v3_dot(piSum, piSum) -
                                        2.0 * selfSum - selfpi);
        else
            dUdV1 = -coeff * volumer * (2.0 * v3_dot(qrSum, piSum) + 2.0 * v3_dot(miSum, piSum) - 2.0 * selfSum);
        const double dUdV2 = v3_dot(v3_scale(2.0 * coeff * volumer, qrSum), piSum) / 3.0;
        out->surf_deriv = COULOMBS * (dUdV1 + dUdV2);
        if (p->use_constraint)
            out->surf_deriv2 = v3_dot(v3_scale(2.0 * coeff * volumer * volumer, qrSum), qrSum) -
                               v3_dot(v3_scale(16.0 * coeff * volumer * volumer, qrSum), qRsum) / 9.0 +
                               v3_dot(v3_scale(2.0 * coeff * volumer * volumer, qRsum), qRsum) / 9.0;
        else
            out->surf_deriv2 = v3_dot(v3_scale(4.0 * coeff * volumer * volumer, qrSum), qrSum) / 9.0;
        for (int i = 0; i < n; ++i) {
            const mp_site I = mp_load(s, dip, quad, i);
            const v3 indPi = v3_ld(pind + 3 * i);
            const v3 ri = v3_ld(s->pos + 3 * i);
            const v3 fi = v3_scale(-2.0 * coeff * I.q, piSum);
            const v3 temp = v3_scale(-s->q[i], qrSum);
            acc(force + 3 * i, v3_scale(COULOMBS, fi));
            vir_add(out->virial, temp, ri);
            if (p->torques)
                acc(trq + 3 * i, v3_scale(COULOMBS * 2.0 * coeff, v3_sub(v3_cross(piSum, I.d), v3_cross(indPi, I.d))));
        }
        if (p->torques) {
            out->tot_torque += mp_torque_norms(n, trq);
            out->surf_deriv += mp_torques2forces(s, trq, force) * volumer / 3.0;
        }
    }
    free(trq);
    return 0;
}

int orc_induction_long(const orc_sys *s, int K, const double *kvec, const double *kmod2, const double *expk,
                       const double *dip, const double *quad, const double *pind, const orc_indf_params *p,
                       double *force, orc_indf_out *out) {
    /* Induction::compute_long / reciprocal_term (src/forces/Induction.cpp:567-590,
     * :1395-1758) */
    const int n = s->n_atoms;
    for (int m = 0; m < s->n_mols; ++m)
        if (s->mol_start[m + 1] - s->mol_start[m] != 3) return -1;
    const double volume = s->box[0] * s->box[1] * s->box[2], volumer = 1.0 / volume;
    const double pi4Volumer = 4.0 * PI * volumer, pi8Volumer = 2.0 * pi4Volumer;
    const double alphaR4 = 1.0 / (4.0 * (p->alpha * p->alpha));
    double *buf = (double *)calloc((size_t)21 * n, sizeof(double));
    if (!buf) return -3;
    double *qQcos = buf, *qQsin = buf + n, *mkcos = buf + 2 * n, *mksin = buf + 3 * n, *pkcos = buf + 4 * n,
           *pksin = buf + 5 * n, *trq = buf + 6 * n;
    double *kXmsina = buf + 9 * n, *kXmcosa = buf + 12 * n, *KXQcosa = buf + 15 * n, *KXQsina = buf + 18 * n;
    out->long_energy = out->long_deriv = out->long_deriv2 = 0.0;
    memset(out->long_virial, 0, sizeof out->long_virial);
    for (int j = 0; j < K; ++j) {
        const v3 kj = v3_ld(kvec + 3 * j);
        double Sc = 0, Ss = 0, Ms = 0, Mc = 0, Ps = 0, Pc = 0, Q2s = 0, Q2c = 0;
        const double coeff = 2.0 * (alphaR4 + 1.0 / kmod2[j]);
        v3 cosrR = {0, 0, 0}, sinrR = {0, 0, 0};
        double Kj[3][3];
        const double kk[3] = {kj.x, kj.y, kj.z};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Kj[a][b] = kk[a] * kk[b];
        for (int i = 0; i < n; ++i) {
            const v3 rim = v3_ld(s->com + 3 * s->mol[i]);
            const mp_site I = mp_load(s, dip, quad, i);
            const v3 indPi = v3_ld(pind + 3 * i);
            const v3 ri = v3_ld(s->pos + 3 * i);
            const double rk = v3_dot(ri, kj);
            const double QiK = m3_quadform(I.Q, kj);
            const double cr = cos(rk), sr = sin(rk);
            const double cosAqiQi = (I.q - QiK) * cr, sinAqiQi = (I.q - QiK) * sr;
            const double mikcos = v3_dot(I.d, kj) * cr, miksin = v3_dot(I.d, kj) * sr;
            const double pikcos = v3_dot(indPi, kj) * cr, piksin = v3_dot(indPi, kj) * sr;
            const double Q2sin = 2.0 * QiK * sr, Q2cos = 2.0 * QiK * cr;
            qQcos[i] = cosAqiQi; qQsin[i] = sinAqiQi;
            mkcos[i] = mikcos; mksin[i] = miksin;
            pkcos[i] = pikcos; pksin[i] = piksin;
            Sc += cosAqiQi; Ss += sinAqiQi;
            Mc += mikcos; Ms += miksin;
            Pc += pikcos; Ps += piksin;
            Q2s += Q2sin; Q2c += Q2cos;
            cosrR = v3_add(cosrR, v3_scale(cosAqiQi, v3_sub(ri, rim)));
            sinrR = v3_add(sinrR, v3_scale(sinAqiQi, v3_sub(ri, rim)));
            if (p->torques) {
                const v3 kXmi = v3_cross(kj, I.d);
                const v3 KXQi = m3_cross(Kj, I.Q);
                v3_st(kXmsina + 3 * i, v3_scale(sr, kXmi));
                v3_st(kXmcosa + 3 * i, v3_scale(cr, kXmi));
                v3_st(KXQcosa + 3 * i, v3_scale(cr, KXQi));
                v3_st(KXQsina + 3 * i, v3_scale(sr, KXQi));
            }
        }
        const double UmqQ = pi4Volumer * expk[j] * (Ss * Pc - Sc * Ps);
        const double Ump = pi4Volumer * expk[j] * (Pc * Mc + Ps * Ms);
        const double tmpenergy = (UmqQ + Ump) * COULOMBS;
        out->long_energy += tmpenergy;
        const double dUdVpQ = pi4Volumer * expk[j] * (Q2s * Pc - Q2c * Ps);
        const double Upp = p->do_min ? pi4Volumer * expk[j] * (Pc * Pc + Ps * Ps) : 0.0;
        const double Sk = 2.0 * UmqQ + 2.0 * Ump + Upp;
        const double dUdV1 = -Sk;
        const double dUdV2 = v3_dot(v3_scale(Sk, kj), kj) * 2.0 * alphaR4;
        const double dUdV3 = -2.0 * UmqQ + 2.0 * dUdVpQ - 4.0 * Ump - 2.0 * Upp;
        out->long_deriv += (dUdV1 + dUdV2 + dUdV3) * COULOMBS;
        const double tkk = v3_dot(v3_scale(tmpenergy, kj), kj) * 2.0 * alphaR4;
        out->long_deriv2 += 4.0 * tmpenergy - v3_dot(v3_scale(7.0 * tmpenergy, kj), kj) * 2.0 * alphaR4 +
                            v3_dot(v3_scale(tkk, kj), kj) * 2.0 * alphaR4;
        double *lv = out->long_virial;
        lv[XX] += tmpenergy - tmpenergy * coeff * kj.x * kj.x;
        lv[YY] += tmpenergy - tmpenergy * coeff * kj.y * kj.y;
        lv[ZZ] += tmpenergy - tmpenergy * coeff * kj.z * kj.z;
        lv[XY] -= tmpenergy * coeff * kj.x * kj.y;
        lv[XZ] -= tmpenergy * coeff * kj.x * kj.z;
        lv[YZ] -= tmpenergy * coeff * kj.y * kj.z;
        const double b = pi8Volumer * expk[j];
        double Ff = 0.0, Ffk2 = 0.0, d2UV1 = 0.0, d2UrR = 0.0;
        for (int i = 0; i < n; ++i) {
            const double FpiqQ = b * (pkcos[i] * Sc + pksin[i] * Ss);
            const double FqQpj = -b * (qQcos[i] * Pc + qQsin[i] * Ps);
            const double Fupj = b * (mksin[i] * Pc - mkcos[i] * Ps);
            const double Fpmj = p->do_min ? b * (pksin[i] * (Mc + Pc) - pkcos[i] * (Ms + Ps))
                                          : b * (pksin[i] * Mc - pkcos[i] * Ms);
            const double val = COULOMBS * (FpiqQ + FqQpj + Fpmj + Fupj);
            const double bkk = v3_dot(v3_scale(b, kj), kj);
            const double d2u1 = bkk * (qQcos[i] * Sc + qQsin[i] * Ss);
            const v3 rim = v3_ld(s->com + 3 * s->mol[i]);
            const v3 ri = v3_ld(s->pos + 3 * i);
            const v3 dr = v3_sub(ri, rim);
            const double d2u2 = v3_dot(v3_scale(bkk, v3_add(v3_scale(qQcos[i], cosrR), v3_scale(qQsin[i], sinrR))), dr);
            acc(force + 3 * i, v3_scale(val, kj));
            Ff += val * v3_dot(kj, dr);
            Ffk2 += v3_dot(v3_scale(2.0 * val, kj), kj) * 2.0 * alphaR4 * v3_dot(kj, dr);
            d2UV1 += v3_dot(v3_scale(d2u1, dr), dr);
            d2UrR += d2u2;
            if (p->torques) {
                const v3 taum = v3_scale(b, v3_add(v3_scale(Pc, v3_ld(kXmcosa + 3 * i)), v3_scale(Ps, v3_ld(kXmsina + 3 * i))));
                const v3 tauQ = v3_scale(2.0 * b, v3_sub(v3_scale(Ps, v3_ld(KXQcosa + 3 * i)), v3_scale(Pc, v3_ld(KXQsina + 3 * i))));
                acc(trq + 3 * i, v3_scale(COULOMBS, v3_add(taum, tauQ)));
            }
        }
        if (p->use_constraint) {
            out->long_deriv += Ff;
            out->long_deriv2 += -6.0 * Ff + Ffk2 - d2UV1 + d2UrR;
        }
    }
    out->long_virial[YX] = out->long_virial[XY];
    out->long_virial[ZX] = out->long_virial[XZ];
    out->long_virial[ZY] = out->long_virial[YZ];
    if (p->torques) {
        out->tot_torque += mp_torque_norms(n, trq);
        out->long_deriv += mp_torques2forces(s, trq, force);
    }
    free(buf);
    return 0;
}

/* ========================================================================== */
/* bonded terms of flexible molecules                                         */
/* ========================================================================== */

/* v[3p+q] += f_p r_q, the index order of the bonded terms' virials */
static inline void vir_fr(double *v, v3 f, v3 r, double sign) {
    v[XX] += sign * f.x * r.x; v[XY] += sign * f.x * r.y; v[XZ] += sign * f.x * r.z;
    v[YX] += sign * f.y * r.x; v[YY] += sign * f.y * r.y; v[YZ] += sign * f.y * r.z;
    v[ZX] += sign * f.z * r.x; v[ZY] += sign * f.z * r.y; v[ZZ] += sign * f.z * r.z;
}

int orc_bond(const orc_sys *s, int64_t nb, const int32_t *bonds, const double *r0v, const double *kv,
             int potential, double *force, double out[12]) {
    /* src/forces/BondForce.cpp:46-122 */
    if (potential != 1 && potential != 2) return -1;
    const double c1 = potential == 2 ? 25.5 : 0.0, c2 = potential == 2 ? 379.3125 : 0.0;
    double energy = 0.0, dubond = 0.0, d2ubond = 0.0, vir[9] = {0};
    for (int64_t i = 0; i < nb; ++i) {
        const int a1 = bonds[3 * i], a2 = bonds[3 * i + 1], t = bonds[3 * i + 2];
        const double k = kv[t], r0 = r0v[t];
        v3 r12 = v3_sub(v3_ld(s->pos + 3 * a1), v3_ld(s->pos + 3 * a2));
        apply_pbc(&r12, s);
        const double r = v3_len(r12);
        const double diff = r - r0;
        double U, val;
        if (potential == 1) {
            U = 0.5 * k * diff * diff;
            val = k * diff;
        } else {
            const double diff2 = diff * diff;
            const double anh = 1.0 - c1 * diff + c2 * diff2;
            U = k * diff2 * anh;
            val = 2.0 * k * diff * anh + k * diff2 * (-c1 + c2 * 2.0 * diff);
        }
        const v3 f = v3_div(v3_scale(-val, r12), r);
        energy += U;
        dubond += val * r;
        d2ubond += -2.0 * r * val + r * r * k;
        vir_fr(vir, f, r12, 1.0);
        acc(force + 3 * a1, f);
        dec(force + 3 * a2, f);
    }
    out[0] = energy; out[1] = dubond; out[2] = d2ubond;
    for (int k = 0; k < 9; ++k) out[3 + k] = vir[k];
    return 0;
}

int orc_angle(const orc_sys *s, int64_t na, const int32_t *angles, const double *theta0v, const double *kv,
              int potential, double *force, double out[12]) {
    /* src/forces/AngleForce.cpp:35-126 */
    if (potential != 1 && potential != 2) return -1;
    double energy = 0.0, vir[9] = {0};
    for (int64_t i = 0; i < na; ++i) {
        const int a1 = angles[4 * i], a2 = angles[4 * i + 1], a3 = angles[4 * i + 2], t = angles[4 * i + 3];
        const double theta0 = theta0v[t], k = kv[t];
        v3 r12 = v3_sub(v3_ld(s->pos + 3 * a2), v3_ld(s->pos + 3 * a1));
        apply_pbc(&r12, s);
        const double d12r = 1.0 / v3_len(r12);
        v3 r32 = v3_sub(v3_ld(s->pos + 3 * a2), v3_ld(s->pos + 3 * a3));
        apply_pbc(&r32, s);
        const double d32r = 1.0 / v3_len(r32);
        double cosT = v3_dot(r12, r32) * d12r * d32r;
        cosT = cosT > -0.9999999 ? cosT : -0.9999999;   /* MAX(cosTheta, -0.9999999) */
        const double sinT = sqrt(1.0 - cosT * cosT);
        const double invSin = 1.0 / sinT;
        double diff = acos(cosT) - theta0;
        const v3 u12 = v3_scale(d12r, r12), u32 = v3_scale(d32r, r32);
        double U, v;
        if (potential == 1) {
            U = 0.5 * k * diff * diff;
            v = k * diff;
        } else {
            const double deg = 180 / PI;
            const double c1 = 0.014, c2 = 5.6E-5, c3 = 7.0E-7, c4 = 2.2E-8;
            diff *= deg;
            const double diff2 = diff * diff, diff3 = diff * diff2, diff4 = diff2 * diff2;
            const double anh = 1.0 - c1 * diff + c2 * diff2 - c3 * diff3 + c4 * diff4;
            U = k * diff2 * anh / (deg * deg);
            v = 2.0 * k * diff * anh / deg + k * diff2 * (-c1 + c2 * 2.0 * diff - c3 * 3.0 * diff2 + c4 * 4.0 * diff3) / deg;
        }
        energy += U;
        const v3 f1 = v3_scale(v * d12r * invSin, v3_sub(v3_scale(cosT, u12), u32));
        const v3 f3 = v3_scale(v * d32r * invSin, v3_sub(v3_scale(cosT, u32), u12));
        const v3 f2 = v3_scale(-1.0, v3_add(f1, f3));
        /* virial[XX] += force1.x * r12.x + force3.x * r32.x; ... */
        const double a1v[3] = {f1.x, f1.y, f1.z}, a3v[3] = {f3.x, f3.y, f3.z};
        const double r1v[3] = {r12.x, r12.y, r12.z}, r3v[3] = {r32.x, r32.y, r32.z};
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) vir[3 * p + q] += a1v[p] * r1v[q] + a3v[p] * r3v[q];
        acc(force + 3 * a1, f1);
        acc(force + 3 * a2, f2);
        acc(force + 3 * a3, f3);
    }
    out[0] = energy; out[1] = 0.0; out[2] = 0.0;
    for (int k = 0; k < 9; ++k) out[3 + k] = vir[k];
    return 0;
}

int orc_ub(const orc_sys *s, double kl, double l0, double *force, double out[12]) {
    /* src/forces/UreyBradleyForce.cpp:41-101 */
    double energy = 0.0, dubond = 0.0, d2ubond = 0.0, vir[9] = {0};
    for (int m = 0; m < s->n_mols; ++m) {
        if (s->mol_start[m + 1] - s->mol_start[m] < 3) return -1;
        const int id1 = s->mol_start[m], id2 = s->mol_start[m] + 2;
        v3 r12 = v3_sub(v3_ld(s->pos + 3 * id1), v3_ld(s->pos + 3 * id2));
        apply_pbc(&r12, s);
        const double r = v3_len(r12);
        const double diff = r - l0;
        const double U = kl * diff * diff;
        const double val = 2.0 * kl * diff;
        const v3 f = v3_div(v3_scale(-val, r12), r);
        energy += U;
        dubond += val * r;
        d2ubond += -2.0 * r * val + r * r * kl;
        vir_fr(vir, f, r12, -1.0);
        acc(force + 3 * id1, f);
        dec(force + 3 * id2, f);
    }
    out[0] = energy; out[1] = dubond; out[2] = d2ubond;
    for (int k = 0; k < 9; ++k) out[3 + k] = vir[k];
    return 0;
}

/* ========================================================================== */
/* pressure tensor sums                                                       */
/* ========================================================================== */

void orc_pressure_sums(const orc_sys *s, const double *pos, const double *vel, const double *force,
                       const double *com, double mc[9], double ka[9], double km[9]) {
    /* src/Integrator.cpp:215-261 */
    for (int k = 0; k < 9; ++k) mc[k] = ka[k] = km[k] = 0.0;
    for (int m = 0; m < s->n_mols; ++m) {
        const v3 R = v3_ld(com + 3 * m);
        for (int i = s->mol_start[m]; i < s->mol_start[m + 1]; ++i) {
            const v3 f = v3_ld(force + 3 * i), r = v3_ld(pos + 3 * i);
            mc[XX] += f.x * (r.x - R.x); mc[XY] += f.y * (r.x - R.x); mc[XZ] += f.z * (r.x - R.x);
            mc[YX] += f.x * (r.y - R.y); mc[YY] += f.y * (r.y - R.y); mc[YZ] += f.z * (r.y - R.y);
            mc[ZX] += f.x * (r.z - R.z); mc[ZY] += f.y * (r.z - R.z); mc[ZZ] += f.z * (r.z - R.z);
        }
    }
    for (int i = 0; i < s->n_atoms; ++i) {
        const double m = s->mass[i];
        const v3 p = v3_scale(m, v3_ld(vel + 3 * i));   /* momentum = velocity*mass */
        ka[XX] += p.x * p.x / m; ka[XY] += p.x * p.y / m; ka[XZ] += p.x * p.z / m;
        ka[YY] += p.y * p.y / m; ka[YZ] += p.y * p.z / m; ka[ZZ] += p.z * p.z / m;
    }
    ka[YX] = ka[XY]; ka[ZX] = ka[XZ]; ka[ZY] = ka[YZ];
    for (int m = 0; m < s->n_mols; ++m) {
        v3 P = {0.0, 0.0, 0.0};
        double M = 0.0;
        for (int i = s->mol_start[m]; i < s->mol_start[m + 1]; ++i) {
            P = v3_add(P, v3_scale(s->mass[i], v3_ld(vel + 3 * i)));
            M += s->mass[i];
        }
        km[XX] += P.x * P.x / M; km[XY] += P.x * P.y / M; km[XZ] += P.x * P.z / M;
        km[YY] += P.y * P.y / M; km[YZ] += P.y * P.z / M; km[ZZ] += P.z * P.z / M;
    }
    km[YX] = km[XY]; km[ZX] = km[XZ]; km[ZY] = km[YZ];
}
