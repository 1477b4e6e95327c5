"""The launch-form autotuner's host logic on a stand-in MD object (no GPU):
every combination is timed, the rounds alternate their order, and the
context is left on the combination with the best median rate."""
import itertools

from swinmd_amd.autotune import DEFAULT_KNOBS, tune_md_step


class _Ctx:
    def __init__(self):
        self.knobs = {}
        self.calls = []

    def set_tuning(self, name, value):
        self.knobs[name] = value
        self.calls.append((name, value))

    def sync(self):
        pass


class _Md:
    """Steps cost less with each knob on (pair_packed matters most) and
    with the default lane count."""

    def __init__(self):
        self.ctx = _Ctx()
        # per-step costs large enough that sleep jitter on a busy host
        # cannot reorder the combinations
        self.cost = {"pair_packed": 3e-5, "recip_overlap": 2e-5, "md_fused_shake": 1e-5}
        self.seen = []

    def run(self, n):
        import time
        k = self.ctx.knobs
        self.seen.append(dict(k))
        per = 1e-5 + sum(c for name, c in self.cost.items() if not k.get(name, 1))
        per += {0: 0.0, 8: 5e-6, 32: 8e-6}[k.get("pair_lanes", 0)]
        time.sleep(n * per)


def test_every_combination_is_timed_and_the_fastest_kept():
    md = _Md()
    best, table = tune_md_step(md, steps=400, warmup=2, rounds=2)
    combos = [dict(zip(DEFAULT_KNOBS, v)) for v in itertools.product(*DEFAULT_KNOBS.values())]
    assert [t["config"] for t in table] == combos
    assert all(len(t["rates"]) == 2 for t in table)
    assert best == {"pair_packed": 1, "recip_overlap": 1, "md_fused_shake": 1, "pair_lanes": 0}
    assert md.ctx.knobs == best
    # the slowest stand-in combination (all off) has the lowest median
    slow = min(table, key=lambda t: t["median_steps_per_s"])
    assert slow["config"] == {"pair_packed": 0, "recip_overlap": 0, "md_fused_shake": 0, "pair_lanes": 32}


def test_a_custom_knob_set():
    md = _Md()
    best, table = tune_md_step(md, knobs={"recip_overlap": (0, 1)}, steps=200, warmup=1, rounds=1)
    assert len(table) == 2 and best == {"recip_overlap": 1}
    assert md.ctx.calls[-1] == ("recip_overlap", 1)
