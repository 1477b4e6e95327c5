"""Launch-shape autotuner for the device MD step.

The knobs are smd_set_tuning's switches. Each one picks between kernel
forms whose results agree to rounding; the test suite checks every pair of
forms against each other:
  pair_packed     the mixed LJ + Ewald pairs from packed FP32 site records
                  (1) or from the FP64 positions (0);
  recip_overlap   k-space on a second stream beside the list and pair
                  kernels (1) or on the step's stream (0);
  md_fused_shake  SHAKE and the kinetic and pressure sums as one kernel (1)
                  or two (0);
  pair_lanes      lanes per atom row of the pair kernel (0: 16 for the
                  mixed form, 32 for FP64).
The tuner times the step under every combination on the live system, for
its size, cutoff and k-set on this GPU, and keeps the fastest.  The
trajectory runs on while it measures, so a few hundred steps of the run are
spent on the search.  The reference has no counterpart: its work
decomposition is fixed by MPI.
"""
import itertools
import time

DEFAULT_KNOBS = {"pair_packed": (1, 0), "recip_overlap": (1, 0), "md_fused_shake": (1, 0), "pair_lanes": (0, 8, 32)}


def tune_md_step(md, knobs=None, steps=200, warmup=40, rounds=2):
    """Times `steps` MD steps (after `warmup` untimed ones, which also
    capture the step's graph again) for every combination of `knobs`,
    `rounds` times in alternating order, and leaves the context on the
    combination with the best median rate.  Returns (best, table): best is
    {knob: value}; table lists every combination with its rates in
    steps/s."""
    knobs = dict(knobs or DEFAULT_KNOBS)
    combos = [dict(zip(knobs, values)) for values in itertools.product(*knobs.values())]
    rates = {i: [] for i in range(len(combos))}
    ctx = md.ctx
    for rnd in range(rounds):
        order = range(len(combos)) if rnd % 2 == 0 else reversed(range(len(combos)))
        for i in order:
            for k, v in combos[i].items():
                ctx.set_tuning(k, v)
            md.run(warmup)
            ctx.sync()
            t0 = time.perf_counter()
            md.run(steps)
            ctx.sync()
            rates[i].append(steps / (time.perf_counter() - t0))
    table = []
    for i, cfg in enumerate(combos):
        r = sorted(rates[i])
        table.append({"config": cfg, "median_steps_per_s": r[len(r) // 2], "rates": r})
    best_i = max(range(len(combos)), key=lambda i: table[i]["median_steps_per_s"])
    for k, v in combos[best_i].items():
        ctx.set_tuning(k, v)
    return combos[best_i], table
