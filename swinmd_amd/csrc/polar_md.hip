// polar_md.hip -- the leap-frog step of rigid polarizable water on the device.
//
//   LFIntegrator::run force block, polarizable config
//                                   src/integrators/LFIntegrator.cpp:385-470
//   BuffvdWForce::compute           src/forces/BuffvdWForce.cpp:93-343
//   Multipole::compute/compute_long src/forces/Multipole.cpp:163-240
//   Induction::compute/compute_long src/forces/Induction.cpp:486-590
//   SHAKE, kinetics, NH friction    LFIntegrator.cpp:561-722
//   per-step rewrap                 :897-910
//
// A step runs in parts on the context stream and a second stream:
//   A1 the Verlet trigger and drift partials (cleared forces and terms), the
//      gated list rebuild, the lab frames;
//   M1 (second stream, beside A2) vdW;
//   A2 the solve's field pass and coupling table;
//   M2 (second stream, beside A3 and B) Multipole's short-range kernels,
//      their real space over the coupling table in the mixed form, and (the
//      polar_m_sums tuning switch) the permanent multipoles' structure-factor
//      sums; M1 and M2 write forces and torques, which nothing in A2, A3 and
//      B touches;
//   A3 the solve's start, A x and residual;
//   B  the solve's iterations in captured batches; the host waits for the
//      device's stop flag between batches (the first wait comes after as many
//      batches as the previous step needed);
//   C  (after M2) Induction's short-range kernels and the rest of one k-space
//      pass for both classes with the converged dipoles (the joint pass: both
//      classes on the same k set; its k-space kernels on the second stream),
//      then SHAKE with the kinetic and pressure sums and the NH update.
// Without the second stream (the recip_overlap tuning switch) part A holds
// A1, A2 and A3 and part C also runs vdW and M2's kernels, in order, with the
// structure factors summed in one pass.  Without the joint pass (classes on
// different k sets, or the polar_joint_long switch off), Multipole's kernels
// run in A with their own k-space pass on the second stream, and C holds
// Induction's.  Every part holds no host synchronisation and no memset: each
// is captured once as a hipGraph and replayed.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

#include "smd_internal.h"

namespace smd {

// The polarizable step's term accumulators: Multipole's, then Induction's.
static double* mp_terms(Ctx* c) { return c->d_polar_terms; }
static double* ind_terms(Ctx* c) { return c->d_polar_terms + POLAR_TERMS; }

// Multipole's short range beside the solve: with the joint pass and a second
// stream, part M runs on the second stream between A and C.
static bool split_mp(Ctx* c) { return polar_joint_long(c) && c->recip_overlap && c->stream2; }

// Part A: everything up to the solve's first iteration.
static int part_a(Ctx* c, bool on_device_rebuild) {
    const smd_polar_md_params& p = c->pmd;
    RedDefer defer(c);   // one kernel adds the part's reductions at its end
    int rc = md_head_enqueue(c);
    if (!rc) rc = pairlist_build(c, p.cutoff, p.cut_buff, p.use_cells, on_device_rebuild);
    if (rc) return rc;
    if (polar_joint_long(c)) {
        // vdW and Multipole's short range in part M beside the solve (or in C)
        rc = frames_run(c);
        if (!rc) rc = induce_begin(c, &p.induce, false);
    } else {
        rc = vdw_enqueue(c, &c->pmd_vdw, p.prec);
        if (!rc) rc = frames_run(c);
        if (!rc) rc = multipole_enqueue(c, mp_terms(c));
        if (!rc) rc = induce_begin(c, &p.induce, false);
    }
    if (!rc) rc = red_flush(c);
    return rc;
}

// The split step's part A in three: A1 (head, gated list rebuild, lab
// frames), after which vdW (M1) starts on the second stream; A2 (the solve's
// field pass and coupling table), after which Multipole's short range (M2)
// follows on the second stream; A3 (the solve's start).
static int part_a1(Ctx* c, bool on_device_rebuild) {
    const smd_polar_md_params& p = c->pmd;
    RedDefer defer(c);
    int rc = md_head_enqueue(c);
    if (!rc) rc = pairlist_build(c, p.cutoff, p.cut_buff, p.use_cells, on_device_rebuild);
    if (!rc) rc = frames_run(c);
    if (!rc) rc = red_flush(c);
    return rc;
}

static int part_a2(Ctx* c) {
    RedDefer defer(c);
    int rc = induce_field(c, &c->pmd.induce, false);
    if (!rc) rc = red_flush(c);
    return rc;
}

static int part_a3(Ctx* c) {
    RedDefer defer(c);
    int rc = induce_start(c, &c->pmd.induce);
    if (!rc) rc = red_flush(c);
    return rc;
}

// The mixed real-space kernels walk the solve's coupling table (built this
// step by the mixed field pass) when their cutoff is the table's.
static bool table_real(Ctx* c, double real_cut) {
    const smd_polar_md_params& p = c->pmd;
    return p.prec == SMD_PREC_MIXED && p.induce.do_min && p.induce.prec == SMD_PREC_MIXED &&
           p.induce.cutoff_ind == real_cut;
}

static int part_m1(Ctx* c) {
    RedDefer defer(c);
    int rc = vdw_enqueue(c, &c->pmd_vdw, c->pmd.prec);
    if (!rc) rc = red_flush(c);
    return rc;
}

static int part_m2(Ctx* c) {
    RedDefer defer(c);
    c->mp_real_table = table_real(c, c->mp.real_cut);
    int rc = polar_mp_short_enqueue(c, mp_terms(c), c->polar_m_sums != 0);
    c->mp_real_table = false;
    if (!rc) rc = red_flush(c);
    return rc;
}

// Part C: induction terms (with the joint pass, both classes; Multipole's
// short range unless part M ran it) and the integration tail.
static int part_c(Ctx* c) {
    int rc;
    {
        RedDefer defer(c);
        c->ind_real_table = table_real(c, c->indf.real_cut);
        c->mp_real_table = table_real(c, c->mp.real_cut);
        if (!polar_joint_long(c)) rc = induction_enqueue(c, ind_terms(c));
        else if (split_mp(c)) rc = polar_joint_enqueue(c, mp_terms(c), ind_terms(c), false, c->polar_m_sums != 0);
        else {
            rc = vdw_enqueue(c, &c->pmd_vdw, c->pmd.prec);
            if (!rc) rc = polar_joint_enqueue(c, mp_terms(c), ind_terms(c), true, false);
        }
        c->ind_real_table = false;
        c->mp_real_table = false;
        if (!rc) rc = red_flush(c);
    }
    if (!rc) rc = md_tail_enqueue(c);
    return rc;
}

static int capture(Ctx* c, hipGraphExec_t* exec, int (*fn)(Ctx*)) {
    hipGraph_t g = nullptr;
    SMD_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = fn(c);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture (polarizable step)");
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        *exec = nullptr;
        return hip_fail(c, ei, "hipGraphInstantiate (polarizable step)");
    }
    return SMD_OK;
}

static int part_a_dev(Ctx* c) { return part_a(c, true); }
static int part_a1_dev(Ctx* c) { return part_a1(c, true); }

// Runs the solve's iterations until the device says it stopped; the first
// check comes after `first` batches.  Returns the iteration count.
// With the block preconditioner the stopping test also writes mapped host
// memory (c->h_cg_done, cleared by the host before each step's part A): the
// host spins on it, and on the batches' completion event, instead of
// synchronising the stream and copying the state back, so part C is queued
// as soon as the device stops (behind the batch's remaining, skipping
// kernels).
static int solve(Ctx* c, int first, int* iters, double* res) {
    const smd_induce_params& ip = c->pmd.induce;
    if (!ip.do_min) {
        *iters = 0;
        *res = 0.0;
        return SMD_OK;
    }
    const int max_it = ip.max_iter > 0 ? ip.max_iter : 3 * c->n;
    const bool mapped = c->h_cg_done && ip.precond == 2;
    volatile int* h = c->h_cg_done;
    int launched = 0, done = 0, reason = 0, rc;
    int batch = std::max(1, first);
    while (true) {
        if ((rc = induce_batches(c, &ip, batch))) return rc;
        launched += batch * c->cg_batch;
        if (mapped) {
            SMD_HIP(c, hipEventRecord(c->ev_cg, c->stream));
            bool drained = false;
            while (!h[0] && !drained) {
                const hipError_t q = hipEventQuery(c->ev_cg);
                if (q == hipSuccess) drained = true;
                else if (q != hipErrorNotReady) return hip_fail(c, q, "solve batches");
            }
            if (h[0]) {
                std::atomic_thread_fence(std::memory_order_acquire);
                *iters = h[1];
                reason = h[2];
                *res = reinterpret_cast<volatile double*>(h + 4)[0];
                done = 1;
            }
        }
        if (!done && (rc = induce_poll(c, &done, iters, &reason, res))) return rc;
        if (done) {
            c->ind_last_converged = reason != 0;
            break;
        }
        if (launched >= max_it) return fail(c, SMD_ENOCONV, "induced-dipole solve did not converge in %d iterations", *iters);
        batch = 1;
    }
    return SMD_OK;
}

static void clear_stop_flag(Ctx* c) {
    if (c->h_cg_done) {
        reinterpret_cast<volatile int*>(c->h_cg_done)[0] = 0;
        std::atomic_thread_fence(std::memory_order_seq_cst);
    }
}

static void free_graphs(Ctx* c) {
    if (c->pmd_a) { (void)hipGraphExecDestroy(c->pmd_a); c->pmd_a = nullptr; }
    if (c->pmd_a2) { (void)hipGraphExecDestroy(c->pmd_a2); c->pmd_a2 = nullptr; }
    if (c->pmd_a3) { (void)hipGraphExecDestroy(c->pmd_a3); c->pmd_a3 = nullptr; }
    if (c->pmd_m2) { (void)hipGraphExecDestroy(c->pmd_m2); c->pmd_m2 = nullptr; }
    if (c->pmd_m) { (void)hipGraphExecDestroy(c->pmd_m); c->pmd_m = nullptr; }
    if (c->pmd_c) { (void)hipGraphExecDestroy(c->pmd_c); c->pmd_c = nullptr; }
}

// A part on the second stream (captured there, or launched directly), after
// the work enqueued so far on the step's stream (event `ev`); with `join`,
// ev_join marks its end for part C.
static int launch_side(Ctx* c, hipGraphExec_t* exec, int (*fn)(Ctx*), hipEvent_t ev, bool graph, bool join) {
    SMD_HIP(c, hipEventRecord(ev, c->stream));
    SMD_HIP(c, hipStreamWaitEvent(c->stream2, ev, 0));
    hipStream_t main = c->stream;
    int rc = SMD_OK;
    c->stream = c->stream2;
    if (graph) {
        if (!*exec) rc = capture(c, exec, fn);
        c->stream = main;
        if (rc) return rc;
        SMD_HIP(c, hipGraphLaunch(*exec, c->stream2));
    } else {
        rc = fn(c);
        c->stream = main;
        if (rc) return rc;
    }
    if (join) SMD_HIP(c, hipEventRecord(c->ev_join, c->stream2));
    return SMD_OK;
}

}  // namespace smd

using namespace smd;

extern "C" int smd_polar_md_init(smd_ctx* c, const smd_polar_md_params* p) {
    if (!c || !p) return SMD_EINVAL;
    if (!c->box_set) return fail(c, SMD_ESTATE, "box not set");
    if (!c->multipoles_set) return fail(c, SMD_ESTATE, "multipoles not set (smd_set_multipoles)");
    if (p->prec != SMD_PREC_FP64 && p->prec != SMD_PREC_MIXED) return fail(c, SMD_EINVAL, "bad precision %d", p->prec);
    if (p->vdw.kind != SMD_VDW_BUFF147 && p->vdw.kind != SMD_VDW_LJ)
        return fail(c, SMD_ENOTSUP, "the polarizable step takes LJ or buffered 14-7 vdW");
    for (int m = 0; m < c->nmol; ++m)
        if (c->h_mol_start[m + 1] - c->h_mol_start[m] != 3)
            return fail(c, SMD_ENOTSUP, "the polarizable step is implemented for rigid 3-site water");
    free_graphs(c);
    c->md_ready = false;
    c->pmd_ready = false;
    c->pmd = *p;
    // the step's own copies of the vdW tables
    const int nt = c->ntypes, nt2 = nt * nt;
    auto keep = [](std::vector<double>& v, const double* src, int count) -> const double* {
        if (!src) return nullptr;
        v.assign(src, src + count);
        return v.data();
    };
    c->pmd_vdw = p->vdw;
    c->pmd_vdw.sigma2 = keep(c->pmd_tab[0], p->vdw.sigma2, nt2);
    c->pmd_vdw.eps4 = keep(c->pmd_tab[1], p->vdw.eps4, nt2);
    c->pmd_vdw.type_sigma = keep(c->pmd_tab[2], p->vdw.type_sigma, nt);
    c->pmd_vdw.type_eps = keep(c->pmd_tab[3], p->vdw.type_eps, nt);
    c->pmd_vdw.buck_a = c->pmd_vdw.buck_b = c->pmd_vdw.buck_c = nullptr;
    // the MD head and tail read the step description from c->md
    c->md = smd_md_params{};
    c->md.cutoff = p->cutoff;
    c->md.cut_buff = p->cut_buff;
    c->md.integ = p->integ;
    c->md.use_cells = p->use_cells;
    c->md.lustig = p->vdw.lustig;
    c->md.prec = p->prec;
    c->md.total_steps = p->total_steps;
    c->md.constraint_mode = 2;
    c->md.ewald_on = 0;
    c->mp_prec = p->prec;
    int rc = upload_vdw_tables(c, &c->pmd_vdw);
    if (rc) return rc;
    if (p->prec == SMD_PREC_MIXED && (rc = vdw_packed_alloc(c))) return rc;
    if ((rc = smd_multipole_setup(c, &p->mp, p->mp_kcut, nullptr))) return rc;
    if ((rc = smd_induction_setup(c, &p->indf, p->ind_kcut, nullptr))) return rc;
    if (!c->d_polar_terms) SMD_HIP(c, hipMalloc(&c->d_polar_terms, sizeof(double) * 2 * POLAR_TERMS));
    if (!c->stream2) {
        // the second stream of the step's k-space passes (default priority)
        SMD_HIP(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        SMD_HIP(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        SMD_HIP(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    if (!c->ev_cg) SMD_HIP(c, hipEventCreateWithFlags(&c->ev_cg, hipEventDisableTiming));
    if (!c->ev_fork2) SMD_HIP(c, hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming));
    if (!c->h_cg_done) {
        // the solve's stop flag in mapped (pinned, fine-grained) host memory
        SMD_HIP(c, hipHostMalloc((void**)&c->h_cg_done, 64, hipHostMallocMapped | hipHostMallocCoherent));
        std::memset(c->h_cg_done, 0, 64);
        SMD_HIP(c, hipHostGetDevicePointer((void**)&c->d_cg_done, c->h_cg_done, 0));
    }
    if (!c->d_stress) SMD_HIP(c, hipMalloc(&c->d_stress, sizeof(double) * 12));
    if ((rc = pairlist_build(c, p->cutoff, p->cut_buff, p->use_cells, false))) return rc;
    {
        // headroom for in-step rebuilds: the row capacity is fixed once captured
        int64_t np;
        int32_t mx, uc;
        if ((rc = smd_pairlist_info(c, &np, &mx, &uc))) return rc;
        const int32_t want = ((int32_t)(mx * 1.5) + 64 + 63) / 64 * 64;
        if (want > c->stride) {
            if ((rc = pairlist_alloc(c, want))) return rc;
            if ((rc = pairlist_build(c, p->cutoff, p->cut_buff, p->use_cells, false))) return rc;
        }
    }
    if ((rc = polar_classes_check(c))) return rc;
    if ((rc = smd_compute_mol_centers(c))) return rc;
    // state of the step
    SMD_HIP(c, hipMemsetAsync(c->d_scalars, 0, sizeof(double) * 8, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_stress, 0, sizeof(double) * 12, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_mom, 0, sizeof(double) * 3 * c->nmol, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_disp, 0, sizeof(double) * 3 * c->n, c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_status, 0, sizeof(unsigned), c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_flag, 0, sizeof(int) * 8, c->stream));
    // start(): the forces at t (a full evaluation with the list just built),
    // then vtMinus = v - dt/2 F/m
    clear_stop_flag(c);
    if ((rc = part_a(c, false))) return rc;
    int iters = 0;
    double res = 0.0;
    if ((rc = solve(c, 1, &iters, &res))) return rc;
    if (polar_joint_long(c)) {
        rc = vdw_enqueue(c, &c->pmd_vdw, p->prec);
        c->ind_real_table = table_real(c, c->indf.real_cut);
        c->mp_real_table = table_real(c, c->mp.real_cut);
        if (!rc) rc = polar_joint_enqueue(c, mp_terms(c), ind_terms(c), true, false);
        c->ind_real_table = false;
        c->mp_real_table = false;
    } else {
        rc = induction_enqueue(c, ind_terms(c));
    }
    if (rc) return rc;
    if ((rc = md_start_vtm(c))) return rc;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    SMD_HIP(c, hipMemsetAsync(c->d_disp, 0, sizeof(double) * 3 * c->n, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    if ((rc = check_status(c))) return rc;
    c->pmd_last_iters = iters;
    c->md_step = 0;
    c->pmd_ready = true;
    return SMD_OK;
}

extern "C" int smd_polar_md_run(smd_ctx* c, int64_t nsteps, smd_polar_md_report* out) {
    if (!c || nsteps < 0) return SMD_EINVAL;
    if (!c->pmd_ready) return fail(c, SMD_ESTATE, "smd_polar_md_init not called");
    c->mp_prec = c->pmd.prec;
    const bool graph = c->pmd.use_graph && !c->prof.on;
    int rc;
    const bool split = split_mp(c);
    if (graph && !c->pmd_a) {
        if (split) {
            if ((rc = capture(c, &c->pmd_a, part_a1_dev))) return rc;
            if ((rc = capture(c, &c->pmd_a2, part_a2))) return rc;
            if ((rc = capture(c, &c->pmd_a3, part_a3))) return rc;
        } else if ((rc = capture(c, &c->pmd_a, part_a_dev))) {
            return rc;
        }
        if ((rc = capture(c, &c->pmd_c, part_c))) return rc;
    }
    int64_t total_iters = 0;
    int iters = 0;
    double res = 0.0;
    for (int64_t s = 0; s < nsteps; ++s) {
        clear_stop_flag(c);
        if (split) {
            // M1 beside A2, M2 beside A3 and the solve
            if (graph) SMD_HIP(c, hipGraphLaunch(c->pmd_a, c->stream));
            else if ((rc = part_a1(c, true))) return rc;
            if ((rc = launch_side(c, &c->pmd_m, part_m1, c->ev_fork, graph, false))) return rc;
            if (graph) SMD_HIP(c, hipGraphLaunch(c->pmd_a2, c->stream));
            else if ((rc = part_a2(c))) return rc;
            if ((rc = launch_side(c, &c->pmd_m2, part_m2, c->ev_fork2, graph, true))) return rc;
            if (graph) SMD_HIP(c, hipGraphLaunch(c->pmd_a3, c->stream));
            else if ((rc = part_a3(c))) return rc;
        } else {
            if (graph) SMD_HIP(c, hipGraphLaunch(c->pmd_a, c->stream));
            else if ((rc = part_a(c, true))) return rc;
        }
        // the first check after the batches the previous step needed
        const int first = (c->pmd_last_iters + 1 + c->cg_batch - 1) / c->cg_batch;
        if ((rc = solve(c, first, &iters, &res))) return rc;
        c->pmd_last_iters = iters;
        total_iters += iters;
        if (split) SMD_HIP(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
        if (graph) SMD_HIP(c, hipGraphLaunch(c->pmd_c, c->stream));
        else if ((rc = part_c(c))) return rc;
        c->md_step++;
        c->ind_last_converged = false;   // the solve was of the positions before the move
    }
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    if ((rc = check_status(c))) return rc;
    if (out) {
        double t[T_COUNT], pt[2 * POLAR_TERMS], sc[8];
        if ((rc = read_terms(c, t))) return rc;
        SMD_HIP(c, hipMemcpy(pt, c->d_polar_terms, sizeof pt, hipMemcpyDeviceToHost));
        SMD_HIP(c, hipMemcpy(sc, c->d_scalars, sizeof sc, hipMemcpyDeviceToHost));
        double e[8];
        polar_terms_energies(c, pt, pt + POLAR_TERMS, e);
        *out = smd_polar_md_report{};
        out->steps = (int64_t)sc[3];
        out->list_updates = (int64_t)sc[2];
        out->cg_iterations = total_iters;
        out->last_iterations = iters;
        out->last_residual = res;
        out->e_vdw = t[T_LJ_E] + c->pmd_vdw.e_lrc;
        out->e_mp_real = e[0]; out->e_mp_self = e[1]; out->e_mp_intra = e[2]; out->e_mp_long = e[3];
        out->e_ind_real = e[4]; out->e_ind_self = e[5]; out->e_ind_intra = e[6]; out->e_ind_long = e[7];
        out->atom_ke2 = t[T_KE_ATOM];
        out->mol_ke2 = t[T_KE_MOL];
        out->fnh = sc[0];
    }
    return SMD_OK;
}
