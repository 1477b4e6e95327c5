"""Host-side logic of the package (no GPU): the reference's velocity set-up
and the constants of the FP32 real-space Ewald kernel."""
import os
import re

import numpy as np
import pytest

from swinmd_amd import ensemble, systems as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("per_atom", [True, False])
def test_init_velocity_matches_oracle_bit_for_bit(oracle_mod, per_atom):
    from oracle import oracle as O
    s = S.water(216, density_mol=33.4, seed=3)
    ndf = ensemble.degrees_of_freedom(s, per_atom)
    want = O.Oracle(s).init_velocity(300.0, 7, per_atom=per_atom, ndf=ndf)
    got = ensemble.init_velocity(s, 300.0, 7, per_atom, ndf)
    assert np.array_equal(got, want)


def test_init_velocity_has_zero_momentum_and_target_temperature():
    s = S.water(64, density_mol=33.4, seed=1)
    v = ensemble.init_velocity(s, 250.0, 5, per_atom=True)
    p = (s.mass[:, None] * v).sum(axis=0)
    assert np.abs(p).max() < 1e-10
    ndf = ensemble.degrees_of_freedom(s, True)
    T = (s.mass[:, None] * v * v).sum() / (ndf * ensemble.BOLTZ)
    assert T == pytest.approx(250.0, rel=1e-12)


def test_uniform_stream_is_the_reference_lcg():
    u = ensemble.uniform_stream(1, 3)
    s1 = (1 * 1029 + 221591) % 1048576
    s2 = (s1 * 1029 + 221591) % 1048576
    s3 = (s2 * 1029 + 221591) % 1048576
    assert list(u) == [s1 / 1048576, s2 / 1048576, s3 / 1048576]


def _erfc_constants():
    src = open(os.path.join(ROOT, "swinmd_amd", "csrc", "smd_internal.h")).read()
    p = float(re.search(r"ERFC_P\s*=\s*([0-9.eE+-]+)f", src).group(1))
    body = re.search(r"ERFC_C\[8\]\s*=\s*\{([^}]*)\}", src).group(1)
    c = [float(x.strip().rstrip("f")) for x in body.split(",")]
    return p, c


def test_fp32_erfc_kernel_constants_meet_their_stated_accuracy():
    """erfc(z) = exp(-z^2) t P(t), t = 1/(1+pz).  In FP32, as the kernel
    evaluates it: t P(t) within 4e-7 of erfc(z) exp(z^2) over z in [0, 5.5];
    with exp(-z^2) from an FP32 z^2 (the kernel's r^2), within the 2 z^2 2^-24
    that rounding z^2 allows, on top of that."""
    from scipy.special import erfc, erfcx
    p, c = _erfc_constants()
    assert len(c) == 8
    z = np.linspace(0.0, 5.5, 100001).astype(np.float32)
    t = (np.float32(1) / (np.float32(1) + np.float32(p) * z)).astype(np.float32)
    P = np.full_like(t, np.float32(c[0]))
    for ck in c[1:]:
        P = (P * t + np.float32(ck)).astype(np.float32)
    tp = (t * P).astype(np.float64)
    zz = z.astype(np.float64)
    assert np.max(np.abs(tp - erfcx(zz)) / erfcx(zz)) < 4e-7
    ex = np.exp(-(z * z)).astype(np.float32)
    approx = (t * P * ex).astype(np.float64)
    rel = np.abs(approx - erfc(zz)) / erfc(zz)
    assert np.all(rel < 4e-7 + 2.0 * zz * zz * 2.0 ** -24 + 2e-7)


def test_bench_reads_the_committed_traffic_summaries():
    """bench.py's roofline `traffic` comes from the committed rocprofv3
    summaries; each must carry its machine-readable TRAFFIC lines."""
    import importlib.util
    import os
    from tests.helpers import ROOT
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    md = bench.measured_traffic("water4096")
    assert md.get("pairs", 0) > 20_000_000 and md.get("recip_sfac", 0) > 5_000_000
    polar = bench.measured_traffic("polar_water8192")
    assert polar.get("cg_apply", 0) > 20_000_000
    assert polar.get("field_table32", 0) > 50_000_000
    assert bench.measured_traffic("no_such_workload") == {}
