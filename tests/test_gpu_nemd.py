"""Batched sheared replicas (smd_batch_*) through the C ABI.

Bars: in FP64 every replica follows the oracle's sheared-replica integrator
(orc_sllod_step: molecular SLLOD, Lees-Edwards images, brute-force list with
the non-affine trigger) for 10 steps to 1e-10 in positions and 1e-8 in
velocities, energies and virial, with the same tilt and list rebuilds; a
replica's trajectory does not depend on the other replicas in the batch (bit
for bit); the mixed form stays within 1e-5 relative RMS force error of FP64;
the tile shapes the autotuner picks and the two-branch graph give the
default shapes' trajectory."""
import numpy as np
import pytest

from tests.helpers import rel, rel_rms

pytestmark = pytest.mark.gpu

RHO = 33.43


def _replicas(sizes=(256, 300, 343), seeds=(5, 6, 7)):
    from swinmd_amd import systems as S
    return [S.water(n, density_mol=RHO, seed=sd, subset=True) for n, sd in zip(sizes, seeds)]


def _vel(s, seed):
    from oracle import oracle as O
    return O.Oracle(s).init_velocity(300.0, seed, per_atom=False, ndf=3 * s.n_mols - 3)


CFG = dict(cutoff=0.9, cut_buff=0.08, alpha=3.5, dt=0.002, temperature=300.0, coupl=0.1, atom_thermo=False)


def _batch(systems, gammas, prec=0, **kw):
    from swinmd_amd.nemd import ShearBatch
    return ShearBatch(systems, gammas, cutoff=CFG["cutoff"], cut_buff=CFG["cut_buff"], alpha=CFG["alpha"],
                      dt=CFG["dt"], temperature=CFG["temperature"], coupl=CFG["coupl"], prec=prec, **kw)


def _run_batch(systems, gammas, tilts, vels, nsteps, prec=0, tiles=None, autotune=False):
    b = _batch(systems, gammas, prec)
    if tiles:
        b.set_tiles(**tiles)
    b.set_state([s.pos for s in systems], vels, tilts)
    b.start()
    f0 = b.state()["force"]
    if autotune:
        b.autotune(reps=3)
    b.run(nsteps)
    st = b.state()
    rep = b.reports()
    b.close()
    return f0, st, rep


def test_replicas_follow_the_oracle():
    from oracle import oracle as O
    from swinmd_amd import systems as S
    systems = _replicas()
    gammas = [0.0, 2.0, -5.0]
    L = [s.box[0] for s in systems]
    tilts = [0.0, 0.3 * L[1], -0.45 * L[2]]
    vels = [_vel(s, 11 + k) for k, s in enumerate(systems)]
    n = 10
    print("case nemd 3 replicas fp64 vs oracle", flush=True)
    f0, st, rep = _run_batch(systems, gammas, tilts, vels, n)
    offs = np.concatenate([[0], np.cumsum([s.n_atoms for s in systems])])
    for r, s in enumerate(systems):
        cfg = dict(CFG, kcut=S.kcut_for(81, L[r]))
        want = O.sllod_run(O.Oracle(s), cfg, gammas[r], n, velocity=vels[r], tilt=tilts[r])
        sl = slice(offs[r], offs[r + 1])
        assert rel_rms(st["pos"][sl], want["pos"]) <= 1e-10, r
        assert rel_rms(st["vel"][sl], want["vel"]) <= 1e-8, r
        assert rel_rms(st["vtm"][sl], want["vtm"]) <= 1e-8, r
        assert rel_rms(st["force"][sl], want["force"]) <= 1e-8, r
        assert abs(st["tilt"][r] - want["tilt"]) <= 1e-12 * L[r], r
        g = rep[r]
        assert g["steps"] == n
        assert g["list_updates"] == want["list_updates"], (r, g["list_updates"], want["list_updates"])
        for k in ("e_lj", "e_real", "e_long", "e_intra", "e_self", "atom_ke2", "mol_ke2"):
            assert rel(g[k], want[k]) <= 1e-8, (r, k, g[k], want[k])
        assert abs(g["fnh"] - want["fnh"]) <= 1e-8 * max(1.0, abs(want["fnh"])), r
        assert np.max(np.abs(g["virial"] - want["virial"])) <= 1e-7 * np.max(np.abs(want["virial"])), r
        assert g["n_kvec"] == want["K"], r


def test_a_replica_does_not_depend_on_its_batch():
    systems = _replicas((256, 343), (5, 7))
    vels = [_vel(s, 21 + k) for k, s in enumerate(systems)]
    L = [s.box[0] for s in systems]
    print("case nemd batch independence", flush=True)
    _, both, rb = _run_batch(systems, [3.0, -4.0], [0.1 * L[0], 0.2 * L[1]], vels, 8)
    _, alone, ra = _run_batch(systems[1:], [-4.0], [0.2 * L[1]], vels[1:], 8)
    n0 = systems[0].n_atoms
    for k in ("pos", "vel", "vtm", "force"):
        assert np.array_equal(both[k][n0:], alone[k]), k
    assert both["tilt"][1] == alone["tilt"][0]
    for k in ("e_lj", "e_real", "e_long", "atom_ke2", "fnh"):
        assert rb[1][k] == ra[0][k], k


def test_mixed_precision_stays_on_the_fp64_trajectory():
    systems = _replicas()
    vels = [_vel(s, 31 + k) for k, s in enumerate(systems)]
    gam = [0.5, -1.0, 8.0]
    tilts = [0.0, 0.1, -0.2]
    print("case nemd mixed vs fp64", flush=True)
    f64, s64, r64 = _run_batch(systems, gam, tilts, vels, 10, prec=0)
    f32, s32, r32 = _run_batch(systems, gam, tilts, vels, 10, prec=1)
    assert rel_rms(f32, f64) <= 1e-5
    assert rel_rms(s32["pos"], s64["pos"]) <= 1e-7
    assert rel_rms(s32["force"], s64["force"]) <= 1e-4
    for a, b in zip(r32, r64):
        for k in ("e_lj", "e_real", "e_long"):
            assert rel(a[k], b[k]) <= 1e-5, k


def test_autotuned_tiles_and_two_branch_graph_keep_the_trajectory():
    systems = _replicas()
    vels = [_vel(s, 41 + k) for k, s in enumerate(systems)]
    gam = [1.0, 3.0, -2.0]
    print("case nemd tiles", flush=True)
    _, base, _ = _run_batch(systems, gam, None, vels, 6, tiles={"overlap": 0})
    _, two, _ = _run_batch(systems, gam, None, vels, 6, tiles={"overlap": 1})
    for k in ("pos", "vel", "force"):
        assert np.array_equal(base[k], two[k]), k
    _, tuned, _ = _run_batch(systems, gam, None, vels, 6, autotune=True)
    assert rel_rms(tuned["pos"], base["pos"]) <= 1e-12
    assert rel_rms(tuned["vel"], base["vel"]) <= 1e-9
    other = {"pair_lanes": 4, "sfac_chunk": 256, "kforce_tile": 128}
    _, shaped, _ = _run_batch(systems, gam, None, vels, 6, tiles=other)
    assert rel_rms(shaped["pos"], base["pos"]) <= 1e-12


def test_batch_rejects_what_it_cannot_run():
    from swinmd_amd import _lib as L
    from swinmd_amd import systems as S
    small = S.water(64, density_mol=RHO, seed=1, subset=True)    # box below 2 (cutoff + buffer)
    with pytest.raises(L.SmdError):
        _batch([small], [1.0])
    b = _batch(_replicas((256,), (5,)), [1.0])
    with pytest.raises(L.SmdError):
        b.run(1)                     # no state, no start
    with pytest.raises(L.SmdError):
        b.set_tiles(pair_lanes=3)
    b.close()
