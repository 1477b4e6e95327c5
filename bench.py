#!/usr/bin/env python3
"""Throughput of the rigid-water MD step on one MI355X.

Workload: 4,096 rigid 3-site point-charge waters (12,288 atoms) at 1.0 g/cm^3
in a 4.9669 nm cubic box; LJ on O (sigma 0.3166 nm, eps 0.6502 kJ/mol,
energy-shifted at 0.9 nm), Ewald with alpha 3.5 nm^-1, real cutoff 0.9 nm and
the k-vectors |n|^2 <= 81, cell pair list with a 0.1 nm buffer, SHAKE to
1e-8, Nose-Hoover at 300 K, dt 2 fs: one LFIntegrator::run step per "step"
(src/integrators/LFIntegrator.cpp:278-941).  The state is resident on the GPU
for the whole timed region.

Measured: steps/s of the mixed-precision path (FP32 pair and k-space math,
FP64 accumulation) and of the plain FP64 path from the same equilibrated
coordinates, the relative RMS difference of their total nonbonded forces on
those coordinates, per-kernel averages from HIP events on the library's
stream, and the reference CPU implementation timed on a short sample of the
same system.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s (ns/day) vs plain FP64 GPU path at ≤1e-5 rel RMS force error"
PEAK_FP32_TFLOPS = 157.3     # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6      # MI355X published FP64 vector peak
PEAK_HBM_GBS = 8000.0

MD = dict(cutoff=0.9, cut_buff=0.1, alpha=3.5, dt=0.002, temperature=300.0, coupl=0.1, tol=1e-8)
N2MAX = 81
FAMILIES = ["pairlist", "pairs", "ewald_intra", "recip_sfac", "recip_reduce", "recip_force", "recip_fsum",
            "shake", "kinetics", "shake_kinetics"]

# Algorithmic work of the reference method per unit.  Counting rule: add,
# sub, mul 1; FMA 2; div, sqrt, rsqrt 1; exp 12; erfc 25; sin+cos 20.
#   pairs: per unique pair inside the cutoff, LJ + Ewald real space with the
#     force on both atoms, the energies and a symmetric virial.
FLOP_PER_PAIR = 99
#   reciprocal space per (atom, k): phase k.r 5, sin+cos 20, structure-factor
#     sums 4 (the pass that forms S(k)); force 10 (the pass that uses it).
FLOP_PER_ATOM_K_SFAC = 29
FLOP_PER_ATOM_K_FORCE = 10
#   induced dipoles: per CG iteration, A v streams the solve's coupling table,
#     4 B (partner) + 16 B (two Thole factors) per directed pair inside
#     cutoffInd, and the CG vectors move 11 x 24 B per site (A v reads v and
#     writes z; the step reads x, v, r, z and writes x, r; the direction reads
#     r, v and writes v).
BYTES_PER_DIRECTED_PAIR_TMU = 20
BYTES_PER_SITE_CG = 11 * 24
#   A v alone (mixed solve): 4 B per table entry (the partner index; the
#     Thole factors are recomputed), and per site the FP32 direction and the
#     fixed-point position record (16 + 16), v read and z written in FP64
#     (24 + 24), the polarizability and the row count (8 + 4).
BYTES_PER_ENTRY_APPLY = 4
BYTES_PER_SITE_APPLY = 92


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_md(system, prec, pos, vel, use_graph=True):
    from swinmd_amd import systems as S
    from swinmd_amd.forces import Context, WaterMD
    ctx = Context(system)
    ctx.set_positions(pos)
    ctx.set_velocities(vel)
    md = WaterMD(ctx, cutoff=MD["cutoff"], cut_buff=MD["cut_buff"], alpha=MD["alpha"],
                 kcut=S.kcut_for(N2MAX, system.box[0]), dt=MD["dt"], temperature=MD["temperature"],
                 coupl=MD["coupl"], tol=MD["tol"], prec=prec, use_graph=use_graph)
    md.init()
    return ctx, md


def unique_pairs_in_cutoff(ctx, cutoff):
    """Pairs of the half list closer than the cutoff (minimum image)."""
    off, nb = ctx.pairlist()
    pos = ctx.positions()
    L = np.asarray(ctx.system.box)
    i = np.repeat(np.arange(len(off) - 1), np.diff(off))
    d = pos[nb] - pos[i]
    d -= L * np.round(d / L)
    r2 = np.einsum("ij,ij->i", d, d)
    return int(np.count_nonzero(r2 <= cutoff * cutoff)), int(len(nb))


def time_steps(ctx, md, steps):
    ctx.sync()
    t0 = time.perf_counter()
    r = md.run(steps)
    ctx.sync()
    return time.perf_counter() - t0, r


MPIRUN = "/opt/conda/bin/mpirun"


def _ref_seconds(exe, d, nsteps, ranks):
    cmd = [exe, d, "time", str(nsteps)]
    if ranks > 1:
        cmd = [MPIRUN, "-np", str(ranks)] + cmd
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log(f"cpu baseline ({ranks} ranks) failed: {r.stderr[-400:]}")
        return None
    lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
    return max(json.loads(x)["seconds"] for x in lines)   # slowest rank


def cpu_baseline(system, pos, nsteps, ranks):
    """The reference SwinMD CPU path (its own sources compiled into
    oracle/_ref) on `nsteps` MD steps of the same system, with `ranks` MPI
    ranks (its replicated-data decomposition) and with one."""
    from oracle import refinput
    from swinmd_amd import systems as S
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    if not os.path.exists(MPIRUN):
        ranks = 1
    with tempfile.TemporaryDirectory() as d:
        refinput.write_inputs(d, system, pos=pos, cutoff=MD["cutoff"], cut_buff=MD["cut_buff"], alpha=MD["alpha"],
                              kcut=S.kcut_for(N2MAX, system.box[0]), dt=MD["dt"], nsteps=nsteps,
                              temperature=MD["temperature"], coupl=MD["coupl"], shake_tol=MD["tol"])
        secs = _ref_seconds(exe, d, nsteps, ranks)
        secs1 = _ref_seconds(exe, d, max(1, nsteps // 2), 1) if ranks > 1 else secs
        if secs is None:
            return None
    out = {"value": round(nsteps / secs, 4), "unit": "steps/s", "cores": ranks, "kind": "reference",
           "sample": f"{nsteps} LFIntegrator::run steps of the same 12,288-atom water from the equilibrated "
                     f"coordinates, reference sources compiled with g++ -O2, {ranks} MPI ranks: {secs:.2f} s"}
    if ranks > 1 and secs1:
        out["value_1_core"] = round(max(1, nsteps // 2) / secs1, 4)
    return out


POLAR = dict(cutoff=0.9, cut_buff=0.1, alpha=3.5, cutoff_ind=0.9, a_damp=0.39)
POLAR_FAMILIES = ["pairs", "mp_short", "mp_long", "ind_field", "ind_cg", "ind_short", "ind_long"]


def polar_cpu_baseline(system, reps=1):
    """The reference's polarizable force evaluation (buffered 14-7, Multipole
    and Induction compute + compute_long with its dense-A CG), compiled from
    its sources, on one core: its multi-rank CG (Induction::ParallelCG) reads
    an uninitialised accumulator (src/forces/Induction.cpp:2473), so only the
    serial ConjugateGradient path is timed.  Returns (baseline, forces)."""
    from oracle import refinput
    from swinmd_amd import systems as S
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None, None
    with tempfile.TemporaryDirectory() as d:
        refinput.write_inputs(d, system, cutoff=POLAR["cutoff"], cut_buff=POLAR["cut_buff"],
                              cutoff_ew=POLAR["cutoff"], alpha=POLAR["alpha"], kcut=S.kcut_for(N2MAX, system.box[0]),
                              ewald=False, buff=True, multipole=True, induction=True, torques=True, ind_damp=True,
                              ind_min=True, cutoff_ind=POLAR["cutoff_ind"], a_damp=POLAR["a_damp"], constraint=True)
        r = subprocess.run([exe, d, "time_polar", str(reps)], capture_output=True, text=True, timeout=1200)
        if r.returncode != 0:
            log(f"polarizable cpu baseline failed: {r.stderr[-400:]}")
            return None, None
        secs = [json.loads(x)["seconds"] for x in r.stdout.splitlines() if x.startswith("{")][0]
        forces = np.fromfile(os.path.join(d, "polar_force.bin"), dtype=np.float64).reshape(-1, 3)
    base = {"value": round(reps / secs, 5), "unit": "force evaluations/s", "cores": 1, "kind": "reference",
            "sample": f"{reps} polarizable force evaluation(s) of the same {system.n_atoms}-site water, reference "
                      f"sources compiled with g++ -O2, 1 MPI rank: {secs:.2f} s"}
    return base, forces


def train_learned_guess(ctx, ind, frames, epochs=60, batch=1024):
    """Trains the learned guess on the GPU from converged solves of the given
    frames (positions) of the workload; returns the network and its report."""
    from swinmd_amd.forces import LearnedGuess
    net = LearnedGuess(ctx, seed=1)
    guess = ind.params.guess
    t0 = time.perf_counter()
    for pos in frames:
        ctx.set_positions(pos)
        ctx.set_pairlist(POLAR["cutoff"], POLAR["cut_buff"], True)
        ind.params.guess = 2
        _, conv, _ = ind.induce()
        if not conv:
            raise RuntimeError("CG did not converge on a training frame")
        net.add_frame()
    frames_s = time.perf_counter() - t0
    rep = net.train(epochs=epochs, batch=batch, lr=2e-3, lr_final=2e-5, seed=9)
    ind.params.guess = guess
    rep["frames_seconds"] = round(frames_s, 3)
    rep["frames"] = len(frames)
    return net, rep


# The polarizable-water step (the 8,192-molecule headline workload and its
# 1,024-molecule form): buffered 14-7 on O and H (the reference example's
# parameters; with O alone the model collapses within ~20 steps), multipole
# Ewald with torques,
# Thole-damped induced dipoles solved to 1e-6 D RMS, SHAKE, Nose-Hoover.
POLAR_EQUIL = 400
# algorithmic work per unique pair inside the real-space cutoff (the
# reference's real_term of each class; counting rule as above)
FLOP_PER_PAIR_MP_REAL = 1050    # G_l, their gradients, B_0..B_5, force (650) and torque (400)
FLOP_PER_PAIR_IND_REAL = 500    # both sides' G_l and gradients, Thole-damped B_l, force and torque


def polar_cpu_baseline_md(system, nsteps=1):
    """The reference's LFIntegrator::run on the polarizable model, compiled
    from its sources, on one core (its multi-rank CG, Induction::ParallelCG,
    reads an uninitialised accumulator, src/forces/Induction.cpp:2473): the
    start() evaluation plus `nsteps` steps; returns steps/s of the run."""
    from oracle import refinput
    from swinmd_amd import systems as S
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        refinput.write_inputs(d, system, cutoff=POLAR["cutoff"], cut_buff=POLAR["cut_buff"],
                              cutoff_ew=POLAR["cutoff"], alpha=POLAR["alpha"], kcut=S.kcut_for(N2MAX, system.box[0]),
                              ewald=False, buff=True, multipole=True, induction=True, torques=True, ind_damp=True,
                              ind_min=True, cutoff_ind=POLAR["cutoff_ind"], a_damp=POLAR["a_damp"], constraint=True,
                              dt=MD["dt"], nsteps=nsteps, temperature=MD["temperature"], coupl=MD["coupl"],
                              shake_tol=MD["tol"])
        secs = _ref_seconds(exe, d, nsteps, 1)
    if not secs:
        return None
    return {"value": round(nsteps / secs, 5), "unit": "steps/s", "cores": 1, "kind": "reference",
            "sample": f"{nsteps} LFIntegrator::run step(s) of the polarizable model at 1,024 molecules (3,072 "
                      f"sites, the 1,024-molecule config), reference sources compiled with g++ -O2, 1 MPI rank: "
                      f"{secs:.2f} s",
            "headline_size": "reference method not run at 8,192 molecules: its dense A is (3N)^2 FP64 = 43.5 GB "
                             "rebuilt every step and applied by a dense CG, minutes per step on one core, beyond "
                             "this bench's bounded CPU sample; the 1,024-molecule point is timed instead"}


def bench_polar_md(args):
    """Polarizable water MD steps: the mixed-precision path with the learned
    initial guess against the plain FP64 path (the reference's start, Pi = 1),
    both solving to the same 1e-6 D RMS tolerance with the preconditioned CG,
    from the same equilibrated state."""
    from swinmd_amd import _lib as L
    from swinmd_amd import ensemble
    from swinmd_amd import systems as S
    from swinmd_amd.forces import Context, Induction, LearnedGuess, PolarMD
    nmol = 1024 if args.workload == "polar1024" else 8192
    system = S.polar_water(nmol)
    n = system.n_atoms
    kcut = S.kcut_for(N2MAX, system.box[0])
    tol = 1e-6 * L.DEBYE_E_NM
    use_graph = not args.no_graph
    vel0 = ensemble.init_velocity(system, MD["temperature"], seed=1, per_atom=True)

    def make(pos, vel, prec, guess, net_state=None):
        ctx = Context(system)
        ctx.set_positions(pos)
        ctx.set_velocities(vel)
        ctx.set_pairlist(POLAR["cutoff"], POLAR["cut_buff"], True)
        ctx.set_tuning("cg_batch", args.cg_batch)
        if args.cg_lanes:
            ctx.set_tuning("cg_lanes", args.cg_lanes)
        if args.no_graph:
            ctx.set_tuning("cg_graph", 0)   # direct launches throughout (profilers)
        if args.no_overlap:
            ctx.set_tuning("recip_overlap", 0)
        if args.no_joint:
            ctx.set_tuning("polar_joint_long", 0)
        if args.polar_m_sums is not None:
            ctx.set_tuning("polar_m_sums", args.polar_m_sums)
        ind = Induction(ctx, cutoff_ind=POLAR["cutoff_ind"], a_damp=POLAR["a_damp"], damp_tij=True, do_min=True,
                        tol=tol, precond=args.precond, guess=guess, prec=prec)
        if net_state is not None:
            net = LearnedGuess(ctx, seed=1)
            net.set_params(net_state[0])
            net.set_norm(net_state[1])
        md = PolarMD(ctx, ind, cutoff=POLAR["cutoff"], cut_buff=POLAR["cut_buff"], alpha=POLAR["alpha"], kcut=kcut,
                     dt=MD["dt"], temperature=MD["temperature"], coupl=MD["coupl"], tol=MD["tol"], prec=prec,
                     use_graph=use_graph)
        md.init()
        return ctx, ind, md

    # the lattice start relaxes under the rigid point-charge model (same
    # coordinates, masses and constraints), then under the polarizable model
    # on the mixed path from the extrapolated previous dipoles; three
    # snapshots of that equilibration train the learned guess (on the GPU,
    # converged solves), and every timed path starts from its end
    log(f"point-charge relaxation {args.polar_pre} steps")
    pctx, pmd = build_md(S.point_charge_twin(system), L.PREC_MIXED, system.pos, vel0)
    pmd.run(args.polar_pre)
    pos_pre, vel_pre = pctx.positions(), pctx.velocities()
    pctx.close()
    log(f"equilibrating {args.polar_equil} steps")
    ctx, ind, md = make(pos_pre, vel_pre, L.PREC_MIXED, 4)
    snap_at = [args.polar_equil - 100, args.polar_equil - 50, args.polar_equil]
    done, snaps = 0, []
    for at in snap_at:
        if at > done:
            md.run(at - done)
            done = at
        snaps.append(ctx.positions())
    pos_eq, vel_eq = ctx.positions(), ctx.velocities()
    ctx.close()
    log("training the learned guess on equilibration snapshots")
    tctx = Context(system)
    tctx.set_pairlist(POLAR["cutoff"], POLAR["cut_buff"], True)
    tind = Induction(tctx, cutoff_ind=POLAR["cutoff_ind"], a_damp=POLAR["a_damp"], damp_tij=True, do_min=True,
                     tol=tol, precond=args.precond, guess=2)
    net, train_rep = train_learned_guess(tctx, tind, snaps)
    train_rep["snapshot_steps"] = [max(a, 0) for a in snap_at]
    net_state = (net.params(), net.norm())
    tctx.close()

    paths = [("mixed_learned", L.PREC_MIXED, 3), ("fp64_reference_start", L.PREC_FP64, 0),
             ("fp64_learned", L.PREC_FP64, 3), ("mixed_reference_start", L.PREC_MIXED, 0),
             ("mixed_alpha_e0_start", L.PREC_MIXED, 2), ("mixed_previous_start", L.PREC_MIXED, 1),
             ("mixed_extrapolated_start", L.PREC_MIXED, 4)]
    if args.no_fp64:
        paths = paths[:1]
    res = {}
    for name, prec, guess in paths:
        ctx, ind, md = make(pos_eq, vel_eq, prec, guess, net_state if guess == 3 else None)
        f0 = ctx.forces()
        e0 = md.run(0)
        md.run(args.warmup)
        r0 = md.run(0)
        ctx.sync()
        t0 = time.perf_counter()
        r = md.run(args.steps)
        ctx.sync()
        secs = time.perf_counter() - t0
        res[name] = dict(secs=secs, force=f0, energies=e0, iters=r["cg_iterations"] / args.steps,
                         list_updates=r["list_updates"] - r0["list_updates"])
        log(f"{name}: {args.steps / secs:.2f} steps/s, {res[name]['iters']:.2f} solve iterations per step")
        if name == "mixed_learned":
            # per-family averages (HIP events around each family's launches,
            # direct launches), then the dominant kernels alone
            ctx.profile(True)
            md.run(args.polar_profile_steps)
            fam = {}
            for f in ("pairlist", "pairs", "mp_short", "mp_long", "ind_field", "ind_cg", "ind_short", "ind_long",
                      "shake_kinetics"):
                ms, cnt = ctx.profile_get(f)
                if cnt:
                    fam[f] = round(ms * 1e3 / args.polar_profile_steps, 2)
            ctx.profile(False)
            res[name]["families_us_per_step"] = fam
            iso = {"mp_real": md.time_kernel("mp_real", 50), "ind_real": md.time_kernel("ind_real", 50),
                   "cg_apply": ind.time_apply(100)}
            res[name]["isolated_us"] = iso
            res[name]["pairs_in_cut"], _ = unique_pairs_in_cutoff(ctx, POLAR["cutoff"])
            ind_pairs, _ = unique_pairs_in_cutoff(ctx, POLAR["cutoff_ind"])
            res[name]["ind_pairs"] = ind_pairs + 3 * nmol   # A couples the intramolecular pairs too
        ctx.close()

    head = res["mixed_learned"]
    steps_s = args.steps / head["secs"]
    out = {
        "metric": METRIC, "value": round(steps_s, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * head["secs"] / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 pair and A v math, f64 accumulation and solve",
        "data": f"synthetic: seeded lattice of {nmol} polarizable waters with random orientations (seed 2025), "
                f"reference Maxwell velocities (seed 1), {args.polar_pre} untimed point-charge relaxation steps, "
                f"then {args.polar_equil} untimed polarizable steps; the learned guess trained on 3 snapshots "
                f"of those steps",
        "config": {"workload": f"polarizable water MD step (buffered 14-7 + multipole Ewald + Thole induced dipoles "
                               f"+ SHAKE + Nose-Hoover), {nmol} molecules", "molecules": nmol, "atoms": n,
                   "box_nm": float(system.box[0]), "cutoff_nm": POLAR["cutoff"], "buffer_nm": POLAR["cut_buff"],
                   "alpha_per_nm": POLAR["alpha"], "n2max": N2MAX, "thole_a": POLAR["a_damp"],
                   "scf": ("CG preconditioned by each molecule's 9x9 block of A" if args.precond == 2 else
                           "CG preconditioned by diag(alpha)") + " to an RMS dipole residual of 1e-6 D on every path",
                   "path": "mixed precision + learned initial guess", "dt_ps": MD["dt"],
                   "thermostat": "Nose-Hoover 300 K", "parallelism": "single GPU",
                   "step_launch": "hipGraph replay of the force parts, captured batches of solve iterations"
                   if use_graph else "direct launches"},
        "ns_per_day": round(steps_s * MD["dt"] * 86400.0 / 1000.0, 3),
        "scf_iterations_per_step": {k: round(v["iters"], 2) for k, v in res.items()},
        "list_updates_per_step": round(head["list_updates"] / args.steps, 4),
        "learned_guess_training": train_rep,
    }
    if "fp64_reference_start" in res:
        ref = res["fp64_reference_start"]
        out["fp64_steps_per_s"] = round(args.steps / ref["secs"], 3)
        out["speedup_vs_fp64"] = round(ref["secs"] / head["secs"], 3)
        out["steps_per_s_by_path"] = {k: round(args.steps / v["secs"], 3) for k, v in res.items()}
        d = head["force"] - ref["force"]
        out["force_rel_rms_vs_fp64"] = float(np.sqrt(np.sum(d * d) / np.sum(ref["force"] ** 2)))
        eh, er = head["energies"], ref["energies"]
        keys = ("e_vdw", "e_mp_real", "e_mp_self", "e_mp_intra", "e_mp_long", "e_ind_real", "e_ind_self",
                "e_ind_intra", "e_ind_long")
        out["energy_rel_err_vs_fp64"] = {k: float(abs(eh[k] - er[k]) / max(abs(er[k]), 1e-30)) for k in keys}
        tot_h, tot_r = sum(eh[k] for k in keys), sum(er[k] for k in keys)
        out["potential_energy_rel_err_vs_fp64"] = float(abs(tot_h - tot_r) / abs(tot_r))
        out["virial_rel_err_vs_fp64"] = None   # the polarizable step does not assemble the pressure tensor
    out["families_us_per_step"] = head["families_us_per_step"]
    iso = head["isolated_us"]
    out["kernels_us_isolated"] = {k: round(v, 2) for k, v in iso.items()}
    models = {"mp_real": FLOP_PER_PAIR_MP_REAL * head["pairs_in_cut"],
              "ind_real": FLOP_PER_PAIR_IND_REAL * (head["pairs_in_cut"] + 3 * nmol)}
    top = max(models, key=lambda k: iso[k])
    ach = models[top] / (iso[top] * 1e-6) / 1e12
    out["roofline"] = {"kernel": top, "bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
                       "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 5),
                       "traffic": measured_traffic(f"polar_water{nmol}").get(top), "traffic_unit": "bytes/launch",
                       "flop_per_launch": models[top], "avg_us": round(iso[top], 2),
                       "pairs": head["pairs_in_cut"]}
    alg = BYTES_PER_ENTRY_APPLY * 2 * head["ind_pairs"] + BYTES_PER_SITE_APPLY * n
    out["cg_apply"] = {"us": round(iso["cg_apply"], 2), "algorithmic_bytes": alg,
                       "achieved_GBs": round(alg / (iso["cg_apply"] * 1e-6) / 1e9, 1),
                       "hbm_frac": round(alg / (iso["cg_apply"] * 1e-6) / 1e9 / PEAK_HBM_GBS, 4),
                       "traffic": measured_traffic(f"polar_water{nmol}").get("cg_apply"),
                       "note": "the mixed solve's A v: a 4 B partner index per directed pair, Thole factors "
                               "recomputed in FP32 from fixed-point site records"}
    if not args.no_cpu_baseline:
        log("polarizable cpu baseline (reference, 1 core, 1,024 molecules)")
        out["cpu_baseline"] = polar_cpu_baseline_md(S.polar_water(1024) if nmol != 1024 else system)
    print(json.dumps(out), flush=True)


# Batched sheared replicas: 64 rigid-water replicas of uniform-random size in
# 256..2048 molecules at 1 g/cm^3, shear rates log-spaced 1e-4..1e-2 1/fs,
# Lees-Edwards images, Nose-Hoover 300 K, one graph launch per step.
NEMD = dict(replicas=64, n_min=256, n_max=2048, seed=2026, g_min=0.1, g_max=10.0, rho=33.43, cutoff=0.9,
            cut_buff=0.08, alpha=3.5)
# algorithmic work (counting rule above): LJ + Ewald real per unique pair in
# the cutoff; per (atom, k): phase 5, sin+cos 20 and the two sums 4 in the
# structure-factor pass, phase 5, sin+cos 20 and the force 10 in the force pass
FLOP_PER_PAIR_LJ_EWALD = 99
FLOP_PER_ATOM_K_SFAC = 29
FLOP_PER_ATOM_K_FORCE = 35


def nemd_replicas():
    """The 64 replicas (systems, shear rates in 1/ps) of the batched workload."""
    from swinmd_amd import systems as S
    rng = np.random.default_rng(NEMD["seed"])
    sizes = rng.integers(NEMD["n_min"], NEMD["n_max"] + 1, NEMD["replicas"])
    gam = np.logspace(np.log10(NEMD["g_min"]), np.log10(NEMD["g_max"]), NEMD["replicas"])
    systems = [S.water(int(n), density_mol=NEMD["rho"], seed=1000 + r, subset=True) for r, n in enumerate(sizes)]
    return systems, gam


def nemd_cpu_baseline(systems, gam, velocities):
    """The oracle's sheared-replica step (plain C, one core) on every replica
    once: the CPU time of one batch step."""
    from oracle import oracle as O
    from swinmd_amd import systems as S
    total = 0.0
    for s, g, v in zip(systems, gam, velocities):
        cfg = dict(cutoff=NEMD["cutoff"], cut_buff=NEMD["cut_buff"], alpha=NEMD["alpha"],
                   kcut=S.kcut_for(N2MAX, s.box[0]), dt=MD["dt"], temperature=MD["temperature"], coupl=MD["coupl"])
        t = {}

        def on_step(k, st, t=t):
            t[k] = time.perf_counter()

        t0 = time.perf_counter()
        O.sllod_run(O.Oracle(s), cfg, float(g), 2, velocity=v, on_step=on_step)
        total += t[1] - t[0]
        del t0
    return {"value": round(1.0 / total, 6), "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"one step of each of the {len(systems)} replicas by the oracle's orc_sllod_step (plain C, "
                      f"gcc -O2, one core): {total:.2f} s per batch step"}


def bench_nemd(args):
    """Batched sheared replicas: steps/s of the whole batch, mixed precision
    against FP64, both from the same equilibrated state."""
    from swinmd_amd import _lib as L
    from swinmd_amd import ensemble
    from swinmd_amd.nemd import ShearBatch
    systems, gam = nemd_replicas()
    R = len(systems)
    vels = [ensemble.init_velocity(s, MD["temperature"], seed=r + 1, per_atom=False) for r, s in enumerate(systems)]
    n_mol = sum(s.n_mols for s in systems)

    def make(prec):
        b = ShearBatch(systems, gam, cutoff=NEMD["cutoff"], cut_buff=NEMD["cut_buff"], alpha=NEMD["alpha"],
                       dt=MD["dt"], temperature=MD["temperature"], coupl=MD["coupl"], shake_tol=MD["tol"],
                       prec=prec)
        if args.no_graph:
            b.set_tiles(use_graph=0)
        return b

    log(f"equilibrating {args.nemd_equil} steps ({R} replicas, {n_mol} molecules)")
    b = make(L.PREC_MIXED)
    b.set_state([s.pos for s in systems], vels)
    b.start()
    b.run(args.nemd_equil)
    st = b.state()
    tilt_eq = st["tilt"]
    pos_eq, vel_eq = st["pos"], st["vel"]
    b.close()
    res = {}
    for name, prec in (("mixed", L.PREC_MIXED), ("fp64", L.PREC_FP64)):
        if name == "fp64" and args.no_fp64:
            continue
        b = make(prec)
        b.set_state(pos_eq, vel_eq, tilt_eq)
        b.start()
        f0 = b.state()["force"]
        tune = None
        if not args.no_autotune:
            tiles, table = b.autotune(reps=10)
            tune = {"chosen": tiles, "candidates": table}
            log(f"{name} autotune: {tiles}")
        b.run(args.warmup)
        r0 = b.reports()
        t0 = time.perf_counter()
        b.run(args.steps)
        secs = time.perf_counter() - t0
        r1 = b.reports()
        iso = {k: b.time_kernel(k, 20) for k in ("pairs", "sfac", "kforce", "kfin", "kvec", "list")}
        b.run(1)   # the forced rebuild after the list timing
        res[name] = dict(secs=secs, force=f0, tune=tune, iso=iso,
                         list_updates=sum(a["list_updates"] - c["list_updates"] for a, c in zip(r1, r0)),
                         kvec=sum(a["n_kvec"] for a in r1), reports=r1, state=b.state())
        log(f"{name}: {args.steps / secs:.2f} batch steps/s")
        b.close()
    head = res["mixed"]
    steps_s = args.steps / head["secs"]
    n_atoms = 3 * n_mol
    # unique pairs inside the cutoff and (atom, k) products, from the state
    pairs = 0
    atom_k = 0
    pos_all = head["state"]["pos"]
    off = 0
    for s, rep, tl in zip(systems, head["reports"], head["state"]["tilt"]):
        p = pos_all[off:off + s.n_atoms]
        off += s.n_atoms
        Lb = s.box[0]
        o = p[s.type == 1]
        d = o[None, :, :] - o[:, None, :]
        ny = np.round(d[..., 1] / Lb)
        d[..., 1] -= ny * Lb
        d[..., 0] -= ny * tl
        d[..., 0] -= Lb * np.round(d[..., 0] / Lb)
        d[..., 2] -= Lb * np.round(d[..., 2] / Lb)
        r2 = np.sum(d * d, axis=2)
        n_oo = int((np.count_nonzero(r2 < NEMD["cutoff"] ** 2) - len(o)) // 2)
        pairs += 9 * n_oo   # every site pair of two molecules whose O are in the cutoff, to within a few percent
        atom_k += s.n_atoms * rep["n_kvec"]
    iso = head["iso"]
    models = {"pairs": FLOP_PER_PAIR_LJ_EWALD * pairs, "sfac": FLOP_PER_ATOM_K_SFAC * atom_k,
              "kforce": FLOP_PER_ATOM_K_FORCE * atom_k}
    top = max(models, key=lambda k: iso[k])
    ach = models[top] / (iso[top] * 1e-6) / 1e12
    out = {
        "metric": METRIC, "value": round(steps_s, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * head["secs"] / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 pair and k-space math, f64 accumulation",
        "data": f"synthetic: {R} seeded lattice waters with random orientations, molecular Maxwell velocities, "
                f"{args.nemd_equil} untimed sheared equilibration steps",
        "config": {"workload": f"batched sheared replicas (molecular SLLOD + Lees-Edwards, rigid 3-site water, "
                               f"LJ + Ewald + SHAKE + Nose-Hoover), {R} replicas, {n_mol} molecules",
                   "replicas": R, "molecules": n_mol, "atoms": n_atoms,
                   "sizes": [int(s.n_mols) for s in systems], "shear_rates_per_ps": [round(g, 5) for g in gam],
                   "cutoff_nm": NEMD["cutoff"], "buffer_nm": NEMD["cut_buff"], "alpha_per_nm": NEMD["alpha"],
                   "n2max": N2MAX, "dt_ps": MD["dt"], "thermostat": "Nose-Hoover 300 K on the molecular "
                   "peculiar kinetic energy", "parallelism": "single GPU",
                   "step_launch": "direct launches" if args.no_graph else "one hipGraph launch per batch step"},
        "ns_per_day_per_replica": round(steps_s * MD["dt"] * 86400.0 / 1000.0, 3),
        "molecule_steps_per_s": round(steps_s * n_mol, 1),
        "list_updates_per_replica_step": round(head["list_updates"] / (args.steps * R), 4),
        "autotune": head["tune"],
        "kernels_us_isolated": {k: round(v, 2) for k, v in iso.items()},
    }
    if "fp64" in res:
        ref = res["fp64"]
        out["fp64_steps_per_s"] = round(args.steps / ref["secs"], 3)
        out["speedup_vs_fp64"] = round(ref["secs"] / head["secs"], 3)
        d = head["force"] - ref["force"]
        out["force_rel_rms_vs_fp64"] = float(np.sqrt(np.sum(d * d) / np.sum(ref["force"] ** 2)))
        out["fp64_kernels_us_isolated"] = {k: round(v, 2) for k, v in ref["iso"].items()}
    out["roofline"] = {"kernel": top, "bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
                       "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 5),
                       "traffic": measured_traffic("nemd64").get(top), "traffic_unit": "bytes/launch",
                       "flop_per_launch": models[top], "avg_us": round(iso[top], 2), "pairs": pairs,
                       "atom_k": atom_k}
    if not args.no_cpu_baseline:
        log("batched replicas cpu baseline (oracle, 1 core, one step of every replica)")
        out["cpu_baseline"] = nemd_cpu_baseline(systems, gam, vels)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workload", default="polar8192", choices=["polar8192", "polar1024", "water4096", "nemd64"],
                    help="polar8192: the polarizable-water MD step on 8,192 molecules, mixed precision + "
                         "learned guess against plain FP64 (default); polar1024: the same on 1,024 molecules; "
                         "water4096: the rigid point-charge water MD step; nemd64: 64 sheared replicas "
                         "stepped together")
    ap.add_argument("--nemd-equil", type=int, default=500, help="untimed steps of the batched replicas")
    ap.add_argument("--polar-equil", type=int, default=POLAR_EQUIL,
                    help="untimed equilibration steps of the polarizable workloads")
    ap.add_argument("--polar-profile-steps", type=int, default=10)
    ap.add_argument("--precond", type=int, default=2, choices=[1, 2],
                    help="preconditioner of every polarizable path's solve: 1 diag(alpha), 2 each molecule's 9x9 "
                         "block of A")
    ap.add_argument("--polar-pre", type=int, default=1500,
                    help="point-charge relaxation steps before the polarizable equilibration")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--equil", type=int, default=3000, help="untimed equilibration steps from the lattice start")
    ap.add_argument("--profile-steps", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-ranks", type=int, default=min(16, os.cpu_count() or 1),
                    help="MPI ranks of the reference CPU run (the GPU's share of host cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison run")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch forms instead of timing every combination first")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (counter calibration for profiling runs)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="polarizable workloads: keep the k-space passes on the step's stream")
    ap.add_argument("--no-joint", action="store_true",
                    help="polarizable workloads: a k-space pass per class instead of one joint pass after the solve")
    ap.add_argument("--polar-m-sums", type=int, default=None, choices=[0, 1],
                    help="polarizable workloads: 1 sums the permanent structure factors beside the dipole solve "
                         "(default: the library's)")
    ap.add_argument("--cg-batch", type=int, default=4,
                    help="polarizable workloads: solve iterations per captured batch (the host checks the stop "
                         "flag between batches)")
    ap.add_argument("--cg-lanes", type=int, default=0, choices=[0, 8, 16, 32],
                    help="polarizable workloads: lanes per row of the solve's A v (0: the library's)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each step's kernels directly instead of replaying a captured hipGraph")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only: the MD step runs on a single device and --gpus {args.gpus} "
              f"(WORLD_SIZE={world}) is not supported", file=sys.stderr)
        sys.exit(2)
    from swinmd_amd import _lib as L
    L.load()   # raises if libswinmd_hip.so is missing
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if args.workload == "nemd64":
        bench_nemd(args)
        return
    if args.workload != "water4096":
        bench_polar_md(args)
        return

    from swinmd_amd import ensemble, systems as S

    system = S.c2_water()
    n = system.n_atoms
    vel0 = ensemble.init_velocity(system, MD["temperature"], seed=1, per_atom=True)

    # equilibrate from the lattice start (mixed precision), then start both
    # precisions from the same coordinates and velocities
    log(f"equilibrating {args.equil} steps")
    use_graph = not args.no_graph
    ctx, md = build_md(system, L.PREC_MIXED, system.pos, vel0, use_graph=use_graph)
    md.run(args.equil)
    pos_eq, vel_eq = ctx.positions(), ctx.velocities()
    ctx.close()

    results = {}
    tuning = None
    for prec, name in ((L.PREC_MIXED, "mixed"), (L.PREC_FP64, "fp64")):
        if name == "fp64" and args.no_fp64:
            continue
        ctx, md = build_md(system, prec, pos_eq, vel_eq, use_graph=use_graph)
        f_start = ctx.forces()
        if name == "mixed" and not args.no_autotune:
            # launch forms picked on the live system (untimed, before the warm-up)
            from swinmd_amd.autotune import tune_md_step
            best, table = tune_md_step(md)
            tuning = {"chosen": best, "candidates": [
                {"config": t["config"], "median_steps_per_s": round(t["median_steps_per_s"], 1)} for t in table]}
            log(f"autotune: {best}")
        md.run(args.warmup)
        r0 = md.run(0)
        secs, r = time_steps(ctx, md, args.steps)
        results[name] = dict(secs=secs, force=f_start, list_updates=r["list_updates"] - r0["list_updates"],
                             report=r)
        log(f"{name}: {args.steps / secs:.1f} steps/s, list updates {results[name]['list_updates']}")
        if name == "mixed":
            # per-kernel averages: HIP events around each launch on the
            # library's stream, over a further run of the same step
            ctx.profile(True)
            md.run(args.profile_steps)
            fam = {}
            for f in FAMILIES:
                ms, cnt = ctx.profile_get(f)
                if cnt:
                    fam[f] = dict(us=ms * 1e3 / cnt, launches=cnt, total_ms=ms)
            ctx.profile(False)
            # the modelled kernels alone, launched back to back between two
            # events (what a profiler's kernel trace reports per launch)
            for f in ("pairs", "recip_sfac", "recip_force"):
                if f in fam:
                    fam[f]["isolated_us"] = ctx.time_family(f, reps=200)
            results[name]["families"] = fam
            results[name]["pairs_in_cut"], results[name]["half_list"] = unique_pairs_in_cutoff(ctx, MD["cutoff"])
        ctx.close()

    mix = results["mixed"]
    steps_s = args.steps / mix["secs"]
    ns_day = steps_s * MD["dt"] * 86400.0 / 1000.0
    out = {
        "metric": METRIC, "value": round(steps_s, 2), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * mix["secs"] / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 math, f64 accumulation",
        "data": f"synthetic: jittered 16^3 lattice water with random orientations (seed 2024), reference "
                f"Maxwell velocities (seed 1), {args.equil} untimed equilibration steps",
        "config": {"workload": "rigid 3-site water MD step (LJ + Ewald + SHAKE + Nose-Hoover), 4096 molecules",
                   "molecules": 4096, "atoms": n, "box_nm": float(system.box[0]), "cutoff_nm": MD["cutoff"],
                   "buffer_nm": MD["cut_buff"], "alpha_per_nm": MD["alpha"], "n2max": N2MAX, "dt_ps": MD["dt"],
                   "shake_tol": MD["tol"], "thermostat": "Nose-Hoover 300 K", "parallelism": "single GPU",
                   "step_launch": "hipGraph replay" if use_graph else "direct launches"},
        "ns_per_day": round(ns_day, 2),
        "list_updates_per_step": round(mix["list_updates"] / args.steps, 4),
    }
    if "fp64" in results:
        f64 = results["fp64"]
        out["fp64_steps_per_s"] = round(args.steps / f64["secs"], 2)
        out["speedup_vs_fp64"] = round(f64["secs"] / mix["secs"], 3)
        d = mix["force"] - f64["force"]
        out["force_rel_rms_vs_fp64"] = float(np.sqrt(np.sum(d * d) / np.sum(f64["force"] ** 2)))

    fam = mix["families"]
    out["kernels_us"] = {k: round(v["us"], 2) for k, v in fam.items()}
    if tuning:
        out["autotune"] = tuning
    # roofline of the dominant modelled kernel
    K = count_kvectors(system.box[0])
    models = {}
    if "pairs" in fam:
        models["pairs"] = FLOP_PER_PAIR * mix["pairs_in_cut"]
    if K and "recip_sfac" in fam:
        models["recip_sfac"] = FLOP_PER_ATOM_K_SFAC * n * K
    if K and "recip_force" in fam:
        models["recip_force"] = FLOP_PER_ATOM_K_FORCE * n * K
    out["kernels_us_isolated"] = {k: round(v["isolated_us"], 2) for k, v in fam.items() if "isolated_us" in v}
    if models:
        top = max(models, key=lambda k: fam[k]["isolated_us"])
        us = fam[top]["isolated_us"]
        ach = models[top] / (us * 1e-6) / 1e12
        traffic = measured_traffic().get(top)
        out["roofline"] = {"kernel": top, "bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
                           "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 5), "traffic": traffic,
                           "traffic_unit": "bytes/launch", "flop_per_launch": models[top], "avg_us": round(us, 2)}
        if top == "pairs":
            # the bytes the kernel cannot avoid: the full-row list entries
            # (both directions of each pair), then per atom its position,
            # charge, type, molecule and row counts in, its force out
            alg = 4 * 2 * mix["half_list"] + n * (24 + 8 + 4 + 4 + 8 + 24)
            out["roofline"]["algorithmic_bytes"] = alg
            out["roofline"]["hbm_frac_of_peak"] = round(alg / (us * 1e-6) / 1e9 / PEAK_HBM_GBS, 5)
    if not args.no_cpu_baseline:
        log(f"cpu baseline: {args.cpu_steps} reference steps")
        out["cpu_baseline"] = cpu_baseline(system, pos_eq, args.cpu_steps, args.cpu_ranks)
    print(json.dumps(out), flush=True)


def measured_traffic(workload="water4096"):
    """HBM bytes per launch of the modelled kernels on a workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (lines
    'TRAFFIC <family> <bytes>'); empty when the summary is absent."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "TRAFFIC":
                out[f[1]] = int(f[2])
    return out


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime: a plain
    copy kernel of known traffic (mib read, mib written) against which the
    HBM counters of a profiling run are scaled."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = int(mib) << 20
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0:   # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes} bytes device to device")


def count_kvectors(L):
    """Number of k-vectors of the reference's half space with |k| < kCut."""
    from swinmd_amd import systems as S
    kcut = S.kcut_for(N2MAX, L)
    nmax = int(kcut * L / (2 * math.pi))
    v = 2 * math.pi / L
    K = 0
    for nx in range(0, nmax + 1):
        for ny in range(-nmax if nx else 0, nmax + 1):
            for nz in range(-nmax if (nx or ny) else 1, nmax + 1):
                if (nx * nx + ny * ny + nz * nz) * v * v < kcut * kcut:
                    K += 1
    return K


if __name__ == "__main__":
    main()
