"""HIP path (libswinmd_hip.so through the C ABI) against the reference's golden
vectors and the C oracle.

Bars: the pair set is bit-exact (and rows come out in the reference's order);
FP64 forces, energies and virials within 1e-10 relative (RMS over sites for
forces); the mixed-precision form (FP32 pair math, FP64 accumulation) within
1e-5 relative RMS force error of the FP64 result."""
import numpy as np
import pytest

from tests.helpers import Fixture, pair_set, rel, rel_rms

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _ctx(system, **kw):
    from swinmd_amd.forces import Context
    return Context(system, **kw)


def _close(a, b, tol=TOL, floor=1e-9):
    assert abs(a - b) <= tol * max(abs(b), floor), (a, b)


@pytest.mark.parametrize("name", ["water512_cells_pairlist", "water216_brute_pairlist", "lj864_cells_pairlist"])
def test_pairlist_matches_reference(name):
    fx = Fixture(name)
    ctx = _ctx(fx.system)
    ctx.set_pairlist(fx.cfg["cutoff"], fx.cfg["cut_buff"], True)
    off, nb = ctx.pairlist()
    roff, rnb = fx.pairs()
    _, _, used = ctx.pairlist_info()
    assert used == bool(fx.meta["used_cells"])
    np.testing.assert_array_equal(off, roff)
    np.testing.assert_array_equal(nb, rnb)


def test_pairlist_brute_mode_and_capacity_growth():
    """Forced brute-force path and a too-small row capacity that must grow."""
    fx = Fixture("water512_cells_pairlist")
    ctx = _ctx(fx.system)
    ctx.set_pairlist(fx.cfg["cutoff"], fx.cfg["cut_buff"], False)
    off, nb = ctx.pairlist()
    assert pair_set(off, nb) == pair_set(*fx.pairs())
    ctx2 = _ctx(fx.system)
    ctx2.set_pairlist_capacity(64)
    from swinmd_amd._lib import SmdError, SMD_ECAPACITY
    with pytest.raises(SmdError) as e:
        ctx2.set_pairlist(fx.cfg["cutoff"], fx.cfg["cut_buff"], True)
    assert e.value.code == SMD_ECAPACITY
    ctx2.set_pairlist_capacity(0)
    ctx2.set_pairlist(fx.cfg["cutoff"], fx.cfg["cut_buff"], True)
    off2, nb2 = ctx2.pairlist()
    np.testing.assert_array_equal(nb2, fx.pairs()[1])


def _vdw(fx, kind, prec=0):
    from swinmd_amd import forces as F
    from swinmd_amd.systems import co2_buckingham_table
    c = fx.cfg
    ctx = _ctx(fx.system)
    ctx.set_pairlist(c["cutoff"], c["cut_buff"], True)
    ctx.set_mol_centers(fx.vec("com"))
    kw = dict(cutoff=c["cutoff"], switch_on=c.get("switch_on", False), switch_dist=c.get("switch_dist", 0.0),
              use_constraint=c.get("constraint", True), lustig=c.get("lustig", False),
              coulomb=c.get("coulomb", False), prec=prec)
    cls = {"lj": F.LJForce, "mie": F.MieForce, "buck": F.BuckinghamForce, "buff": F.BuffvdWForce}[kind]
    if kind == "mie":
        kw.update(n_mie=c.get("n_mie", 12.0), m_mie=c.get("m_mie", 6.0))
    if kind == "buck":
        kw.update(buck=co2_buckingham_table(), co2_bond=fx.system.bond_params[0][2])
    force = cls(ctx, **kw)
    f = np.zeros((fx.system.n_atoms, 3))
    out = force.compute(f)
    return f, out


VDW_CASES = [("lj864_forces", "lj"), ("lj864_forces", "mie"), ("lj864_mie147_forces", "mie"),
             ("water216_brute_forces", "lj"), ("water512_cells_forces", "lj"),
             ("amoeba216_buff_forces", "buff"), ("co2_216_buck_forces", "buck")]


@pytest.mark.parametrize("name,kind", VDW_CASES)
def test_vdw_fp64_matches_reference(name, kind):
    fx = Fixture(name)
    f, out = _vdw(fx, kind)
    assert rel_rms(f, fx.vec(kind + "_force")) <= TOL
    _close(out["energy"], fx.meta[kind + "_energy"])
    _close(out["lj_energy"], fx.meta[kind + "_ljEnergy"])
    vref = fx.ref[kind + "_virial"]
    assert np.max(np.abs(out["virial"] - vref)) <= TOL * np.max(np.abs(vref))
    for k in ("dudv", "d2udv"):
        _close(out[k], fx.meta[kind + "_" + k], floor=1.0)


@pytest.mark.parametrize("name,kind", VDW_CASES)
def test_vdw_mixed_precision(name, kind):
    fx = Fixture(name)
    f64, _ = _vdw(fx, kind, prec=0)
    f32, out = _vdw(fx, kind, prec=1)
    assert rel_rms(f32, f64) <= 1e-5
    assert rel(out["energy"], fx.meta[kind + "_energy"]) <= 1e-5


@pytest.mark.parametrize("name", ["water216_brute_forces", "water512_cells_forces"])
@pytest.mark.parametrize("prec", [0, 1])
def test_ewald_matches_reference(name, prec):
    from swinmd_amd.forces import EwaldForce
    fx = Fixture(name)
    c = fx.cfg
    ctx = _ctx(fx.system)
    ctx.set_pairlist(c["cutoff"], c["cut_buff"], True)
    ctx.set_mol_centers(fx.vec("com"))
    ew = EwaldForce(ctx, alpha=c["alpha"], real_cut=c["cutoff"], kcut=c["kcut"], use_constraint=True,
                    surface=bool(c.get("surface")), prec=prec)
    f = np.zeros((fx.system.n_atoms, 3))
    o = ew.compute(f)
    fl = np.zeros_like(f)
    lo = ew.compute_long(fl)
    tol = TOL if prec == 0 else 1e-5
    assert rel_rms(f, fx.vec("ewald_short_force")) <= tol
    assert rel_rms(fl, fx.vec("ewald_long_force")) <= tol
    _close(o["real_energy"], fx.meta["ewald_real"], tol)
    _close(o["real_deriv"], fx.meta["ewald_realDeriv"], tol)
    _close(o["real_deriv2"], fx.meta["ewald_realSecDeriv"], tol)
    _close(o["self_energy"] + o["charged_energy"], fx.meta["ewald_correct"])
    _close(o["intra_energy"], fx.meta["ewald_molCorrect"])
    _close(o["intra_vir_correct"], fx.meta["ewald_intravirCorrect"])
    if c.get("surface"):
        _close(o["surf_energy"], fx.meta["ewald_surfCorrect"])
        _close(o["surf_deriv"], fx.meta["ewald_surfDeriv"])
        _close(o["surf_deriv2"], fx.meta["ewald_surfSecDeriv"])
    vref = fx.ref["ewald_short_virial"]
    assert np.max(np.abs(o["virial"] - vref)) <= tol * np.max(np.abs(vref))
    _close(lo["energy"], fx.meta["ewald_long"], tol)
    _close(lo["deriv"], fx.meta["ewald_longDeriv"], tol)
    _close(lo["deriv2"], fx.meta["ewald_longSecDeriv"], tol)
    lref = fx.ref["ewald_long_virial"]
    assert np.max(np.abs(lo["virial"] - lref)) <= tol * np.max(np.abs(lref))


@pytest.mark.parametrize("name", ["water512_md10", "water216_md20_moltherm"])
def test_md_matches_reference(name):
    """LFIntegrator start() + N steps on the device (SHAKE, NH, list rebuilds)."""
    from oracle import oracle as O
    from swinmd_amd.forces import WaterMD
    fx = Fixture(name)
    c = fx.cfg
    ctx = _ctx(fx.system)
    ctx.set_velocities(fx.vec("init_vel"))
    md = WaterMD(ctx, cutoff=c["cutoff"], cut_buff=c["cut_buff"], alpha=c["alpha"], kcut=c["kcut"], dt=c["dt"],
                 temperature=c["temperature"], coupl=c["coupl"], tol=c["shake_tol"], max_count=c["max_count"],
                 atom_thermo=c.get("atom_thermo", True), use_graph=False, total_steps=int(fx.meta["totsteps"]),
                 lustig=bool(c.get("lustig", False)))
    md.init()
    rep = md.run(int(fx.extra))
    pos, vel = ctx.positions(), ctx.velocities()
    assert rel_rms(pos, fx.vec("end_pos")) <= 1e-11
    assert rel_rms(vel, fx.vec("end_vel")) <= 1e-8
    assert rel(rep["e_lj"], fx.meta["e_lj"]) <= 1e-8
    assert rel(rep["e_real"], fx.meta["e_real"]) <= 1e-8
    assert rel(rep["e_long"], fx.meta["e_long"]) <= 1e-8
    assert rel(0.5 * rep["atom_ke2"], fx.meta["e_atomKin"]) <= 1e-8
    assert rel(0.5 * rep["mol_ke2"], fx.meta["e_molKin"]) <= 1e-8
    o = O.Oracle(fx.system)
    r = O.md_from_cfg(o, c, int(fx.extra))
    assert rep["list_updates"] == r["list_updates"]
    # the last step's computePressure with the shake-virial correction, and the
    # Ensemble::accumulate stress sums over the run
    pr = md.pressure()
    for key in ("virial", "mol_virial", "atom_pressure", "mol_pressure"):
        want = fx.ref["end_" + key]
        assert np.max(np.abs(pr[key] - want)) <= 1e-8 * np.max(np.abs(want)), key
    np.testing.assert_array_equal(rep["virial"], pr["virial"])
    for key, mkey in (("lj_dudv", "ljdUlus"), ("real_deriv", "realDeriv"), ("long_deriv", "longDeriv"),
                      ("mol_deriv", "molDeriv"), ("surf_deriv", "surfDeriv")):
        assert abs(pr[key] - fx.meta[mkey]) <= 1e-8 * max(abs(fx.meta[mkey]), 1e-6), (key, pr[key], fx.meta[mkey])
    assert rel(pr["mol_pt"], fx.meta["molPt"]) <= 1e-8
    assert pr["counter"] == fx.meta["acc_counter"] and pr["samples"] == fx.meta["acc_luscount"]
    st = np.concatenate([pr["atom_stress"], pr["mol_stress"]])
    want = fx.ref["stress_sums"]
    assert np.max(np.abs(st - want)) <= 1e-8 * np.max(np.abs(want))
    # Atom::realPos: advanced with every step, never rewrapped
    assert rel_rms(md.real_positions(), fx.vec("end_real_pos")) <= 1e-11


def test_md_checkpoint_and_restart_continue_the_run(tmp_path):
    """Run 6 steps, write the restart files losslessly (state with positions,
    real positions and velocities; vtMinus; the thermostat friction), continue
    4 steps; a fresh context restarted from the files and run 4 steps lands on
    the same state."""
    from swinmd_amd.forces import WaterMD
    fx = Fixture("water512_md10")
    c = fx.cfg
    kw = dict(cutoff=c["cutoff"], cut_buff=c["cut_buff"], alpha=c["alpha"], kcut=c["kcut"], dt=c["dt"],
              temperature=c["temperature"], coupl=c["coupl"], tol=c["shake_tol"], max_count=c["max_count"],
              use_graph=False)
    ctx = _ctx(fx.system)
    ctx.set_velocities(fx.vec("init_vel"))
    md = WaterMD(ctx, **kw)
    md.init()
    md.run(6)
    nh = md.checkpoint(tmp_path / "state.out", tmp_path / "vtMinus.out", step=6)
    md.run(4)
    want_pos, want_vel, want_real = ctx.positions(), ctx.velocities(), md.real_positions()
    ctx2 = _ctx(fx.system)
    md2 = WaterMD(ctx2, **kw)
    md2.restart(tmp_path / "state.out", tmp_path / "vtMinus.out", thermostat=nh)
    md2.run(4)
    assert rel_rms(ctx2.positions(), want_pos) <= 1e-10
    assert rel_rms(ctx2.velocities(), want_vel) <= 1e-8
    assert rel_rms(md2.real_positions(), want_real) <= 1e-10
    ctx.close()
    ctx2.close()


def test_md_graph_replay_equals_eager():
    """The captured step replays the eager step sequence exactly."""
    from swinmd_amd.forces import WaterMD
    fx = Fixture("water512_md10")
    c = fx.cfg
    outs = []
    for graph in (False, True):
        ctx = _ctx(fx.system)
        ctx.set_velocities(fx.vec("init_vel"))
        md = WaterMD(ctx, cutoff=c["cutoff"], cut_buff=c["cut_buff"], alpha=c["alpha"], kcut=c["kcut"], dt=c["dt"],
                     temperature=c["temperature"], coupl=c["coupl"], use_graph=graph)
        md.init()
        for _ in range(3):           # several replays with host work between them
            md.run(4)
            ctx.sync()
        outs.append(ctx.positions())
    assert rel_rms(outs[1], outs[0]) <= 1e-11


@pytest.mark.parametrize("graph", [False, True])
def test_md_fused_shake_kinetics_equals_two_kernels(graph):
    """SHAKE and the kinetic/pressure sums as one kernel give the two-kernel
    step's trajectory, pressure tensors and thermostat."""
    from swinmd_amd.forces import WaterMD
    fx = Fixture("water512_md10")
    c = fx.cfg
    outs = []
    for fused in (0, 1):
        ctx = _ctx(fx.system)
        ctx.set_tuning("md_fused_shake", fused)
        ctx.set_velocities(fx.vec("init_vel"))
        md = WaterMD(ctx, cutoff=c["cutoff"], cut_buff=c["cut_buff"], alpha=c["alpha"], kcut=c["kcut"], dt=c["dt"],
                     temperature=c["temperature"], coupl=c["coupl"], use_graph=graph)
        md.init()
        rep = md.run(10)
        outs.append((ctx.positions(), ctx.velocities(), md.pressure(), md.thermostat(), rep))
    assert rel_rms(outs[1][0], outs[0][0]) <= 1e-11
    assert rel_rms(outs[1][1], outs[0][1]) <= 1e-9
    for k in ("atom_pressure", "mol_pressure", "virial"):
        assert rel_rms(outs[1][2][k], outs[0][2][k]) <= 1e-9, k
    assert rel(outs[1][3][0], outs[0][3][0]) <= 1e-9
    assert outs[1][4]["list_updates"] == outs[0][4]["list_updates"]


@pytest.mark.parametrize("graph", [False, True])
def test_md_recip_on_second_stream_equals_one_stream(graph):
    """Reciprocal space forked onto the second stream (forces kept apart and
    added in SHAKE) gives the one-stream step's trajectory."""
    from swinmd_amd.forces import WaterMD
    fx = Fixture("water512_md10")
    c = fx.cfg
    outs = []
    for overlap in (0, 1):
        ctx = _ctx(fx.system)
        ctx.set_tuning("recip_overlap", overlap)
        ctx.set_velocities(fx.vec("init_vel"))
        md = WaterMD(ctx, cutoff=c["cutoff"], cut_buff=c["cut_buff"], alpha=c["alpha"], kcut=c["kcut"], dt=c["dt"],
                     temperature=c["temperature"], coupl=c["coupl"], use_graph=graph)
        md.init()
        rep = md.run(10)
        outs.append((ctx.positions(), ctx.forces(), rep))
    assert rel_rms(outs[1][0], outs[0][0]) <= 1e-11
    assert rel_rms(outs[1][1], outs[0][1]) <= 1e-9
    assert rel(outs[1][2]["e_long"], outs[0][2]["e_long"]) <= 1e-9
    assert outs[1][2]["list_updates"] == outs[0][2]["list_updates"]


def test_shake_step_matches_oracle():
    from oracle import oracle as O
    from swinmd_amd.forces import EwaldForce, LFIntegrator, LJForce
    import ctypes as C
    fx = Fixture("water512_md10")
    c = fx.cfg
    s = fx.system
    ctx = _ctx(s)
    ctx.set_pairlist(c["cutoff"], c["cut_buff"], True)
    ctx.compute_mol_centers()
    ctx.zero_forces()
    LJForce(ctx, cutoff=c["cutoff"]).compute()
    ew = EwaldForce(ctx, alpha=c["alpha"], real_cut=c["cutoff"], kcut=c["kcut"])
    ew.compute()
    ew.compute_long()
    F = ctx.forces()
    rng = np.random.default_rng(1)
    vtm = rng.normal(scale=0.5, size=(s.n_atoms, 3))
    mom = rng.normal(scale=0.1, size=(s.n_mols, 3))
    lf = LFIntegrator(ctx, dt=0.002, temperature=300.0, tol=1e-8)
    lf.set_vt_minus(vtm)
    lf.set_mol_momenta(mom)
    vir = lf.shake(fnh=0.37)
    # oracle
    o = O.Oracle(s)
    pos = o.pos.copy()
    vt = vtm.copy()
    vel = np.zeros_like(vt)
    disp = np.zeros_like(vt)
    ovir = np.zeros(9)
    b2 = np.ascontiguousarray(s.constraint_lengths2())
    Fc = np.ascontiguousarray(F)
    momc = np.ascontiguousarray(mom)
    rc = O.lib().orc_shake(C.byref(o._sys), O._p(pos), O._p(vt), O._p(vel), O._p(disp), O._p(Fc), O._p(b2),
                           O._p(momc), 0.37, 1, 0.002, 1e-8, 10000, O._p(ovir))
    assert rc == 0
    assert rel_rms(ctx.positions(), pos) <= 1e-13
    assert rel_rms(lf.vt_minus(), vt) <= 1e-9
    assert rel_rms(ctx.velocities(), vel) <= 1e-9
    assert np.max(np.abs(vir - ovir)) <= 1e-8 * np.max(np.abs(ovir))


def test_nh_chain_matches_oracle():
    import ctypes as C
    from oracle import oracle as O
    from swinmd_amd.forces import Context
    fx = Fixture("water216_md20_moltherm")
    ctx = Context(fx.system)
    v = np.random.default_rng(3).normal(size=(fx.system.n_atoms, 3))
    ctx.set_velocities(v)
    T, dt, ndf, vol = 300.0, 0.002, 3.0 * fx.system.n_atoms - 3, float(np.prod(fx.system.box))
    qnh = np.full(4, 0.0083145112 * T * 0.1 * 0.1)
    qnh[0] *= ndf
    wbar = (ndf + 1.0) * 0.0083145112 * T * 0.5 * 0.5
    state = np.concatenate([[0.1, -0.2, 0.05, 0.0], np.zeros(4), qnh, [0.01, 0.0, wbar]])
    ekin = 1234.5
    sc = C.c_double()
    st = np.ascontiguousarray(state.copy())
    ctx.check(ctx.lib.smd_nh_chain(ctx.h, ekin, T, dt, ndf, 0.3, vol, st.ctypes.data_as(C.c_void_p), C.byref(sc)))
    vnh, gnh = state[:4].copy(), state[4:8].copy()
    vlog, glog = C.c_double(state[12]), C.c_double(state[13])
    scale = O.lib().orc_nh_chain(ekin, T, dt, ndf, 0.3, vol, wbar, O._p(vnh), O._p(gnh), O._p(qnh),
                                 C.byref(vlog), C.byref(glog))
    assert abs(sc.value - scale) <= 1e-13
    np.testing.assert_allclose(st[:4], vnh, rtol=1e-12, atol=1e-15)
    assert abs(st[12] - vlog.value) <= 1e-12 * max(1.0, abs(vlog.value))
    np.testing.assert_allclose(ctx.velocities(), v * scale, rtol=1e-14)


def test_c2_water_forces_against_oracle():
    """The benchmark system (4,096 waters, 12,288 sites) at full size: pair set
    identical to the oracle's and FP64 forces within 1e-10."""
    from oracle import oracle as O
    from swinmd_amd import systems as S
    from swinmd_amd.forces import EwaldForce, LJForce
    s = S.c2_water()
    L = s.box[0]
    cutoff, buff, alpha, kcut = 0.9, 0.1, 3.5, S.kcut_for(81, L)
    ctx = _ctx(s)
    ctx.set_pairlist(cutoff, buff, True)
    ctx.compute_mol_centers()
    o = O.Oracle(s)
    off, nb, used = o.pairlist(cutoff, buff, True)
    goff, gnb = ctx.pairlist()
    assert used
    np.testing.assert_array_equal(goff, off)
    np.testing.assert_array_equal(gnb, nb)
    f = np.zeros((s.n_atoms, 3))
    lj = LJForce(ctx, cutoff=cutoff, lustig=True)
    out = lj.compute(f)
    ew = EwaldForce(ctx, alpha=alpha, real_cut=cutoff, kcut=kcut)
    assert ew.n_kvec == 1535
    eo = ew.compute(f)
    lo = ew.compute_long(f)
    fo, vo = o.vdw(off, nb, "lj", cutoff=cutoff, lustig=True)
    f1, r = o.ewald_real(off, nb, alpha, cutoff)
    f2, it = o.ewald_intra(alpha)
    f3, rc = o.ewald_recip(kcut, alpha)
    assert rel_rms(f, fo + f1 + f2 + f3) <= TOL
    _close(out["energy"], vo["energy"])
    _close(eo["real_energy"], r["energy"])
    _close(lo["energy"], rc["energy"])
    _close(lo["deriv2"], rc["deriv2"])


def test_md_packed_pair_records_match_unpacked():
    """The MD step's mixed LJ + Ewald kernel reading packed FP32 hi/lo site
    records gives the forces and energies of the kernel reading FP64
    positions, to FP32 rounding, and both stay within the mixed form's bar of
    the FP64 forces."""
    from swinmd_amd import systems as S
    from swinmd_amd import _lib as L
    from swinmd_amd.forces import WaterMD
    s = S.water(1000, density_mol=33.4, seed=11, lattice=(10, 10, 10))
    kcut = S.kcut_for(27, s.box[0])
    out = {}
    for name, prec, packed in (("fp64", L.PREC_FP64, 1), ("packed", L.PREC_MIXED, 1), ("plain", L.PREC_MIXED, 0)):
        ctx = _ctx(s)
        ctx.set_tuning("pair_packed", packed)
        md = WaterMD(ctx, cutoff=0.9, cut_buff=0.1, alpha=3.5, kcut=kcut, dt=0.002, temperature=300.0, prec=prec,
                     use_graph=False)
        md.init()
        f0 = ctx.forces()            # start(): the first force evaluation
        rep = md.run(3)
        out[name] = (f0, rep)
    f64, fp, fu = out["fp64"][0], out["packed"][0], out["plain"][0]
    assert rel_rms(fp, fu) <= 2e-6
    assert rel_rms(fp, f64) <= 1e-5 and rel_rms(fu, f64) <= 1e-5
    for k in ("e_lj", "e_real"):
        assert rel(out["packed"][1][k], out["plain"][1][k]) <= 1e-5


@pytest.mark.parametrize("packed", [1, 0])
def test_every_pair_lane_count_runs_and_agrees(packed):
    """Each lane count smd_set_tuning accepts is the one the pair kernel
    runs (packed mixed and general kernels alike): the start forces and
    energies agree across 4..64 lanes to FP64 summation order."""
    from swinmd_amd import systems as S
    from swinmd_amd import _lib as L
    from swinmd_amd.forces import WaterMD
    s = S.water(1000, density_mol=33.4, seed=11, lattice=(10, 10, 10))
    kcut = S.kcut_for(27, s.box[0])
    out = {}
    for lanes in (4, 8, 16, 32, 64):
        print(f"case pair lanes {lanes} packed {packed}", flush=True)
        ctx = _ctx(s)
        ctx.set_tuning("pair_packed", packed)
        ctx.set_tuning("pair_lanes", lanes)
        md = WaterMD(ctx, cutoff=0.9, cut_buff=0.1, alpha=3.5, kcut=kcut, dt=0.002, temperature=300.0,
                     prec=L.PREC_MIXED, use_graph=False)
        md.init()
        out[lanes] = (ctx.forces(), md.run(0))
        ctx.close()
    f16, r16 = out[16]
    for lanes, (f, r) in out.items():
        assert rel_rms(f, f16) <= 1e-12, lanes
        for k in ("e_lj", "e_real"):
            assert rel(r[k], r16[k]) <= 1e-12, (lanes, k)
    with pytest.raises(L.SmdError):
        _ctx(s).set_tuning("pair_lanes", 12)


def test_md_autotuner_picks_a_candidate_and_keeps_the_physics():
    """The launch-shape tuner times every combination of its knobs on the
    live run and leaves the fastest set; the step then gives the same
    trajectory as the default forms from the same state."""
    from swinmd_amd.autotune import DEFAULT_KNOBS, tune_md_step
    from swinmd_amd.forces import WaterMD
    fx = Fixture("water512_md10")
    c = fx.cfg
    kw = dict(cutoff=c["cutoff"], cut_buff=c["cut_buff"], alpha=c["alpha"], kcut=c["kcut"], dt=c["dt"],
              temperature=c["temperature"], coupl=c["coupl"], prec=1)
    ctx = _ctx(fx.system)
    ctx.set_velocities(fx.vec("init_vel"))
    md = WaterMD(ctx, **kw)
    md.init()
    best, table = tune_md_step(md, steps=10, warmup=4, rounds=1)
    print("autotune", best, [(t["config"], round(t["median_steps_per_s"])) for t in table], flush=True)
    combos = int(np.prod([len(v) for v in DEFAULT_KNOBS.values()]))
    assert len(table) == combos and all(t["median_steps_per_s"] > 0 for t in table)
    assert best in [t["config"] for t in table] and set(best) == set(DEFAULT_KNOBS)
    # from the tuned run's state: the chosen forms and the default forms agree
    pos, vel = ctx.positions(), ctx.velocities()
    outs = []
    for cfg in (best, {k: v[0] for k, v in DEFAULT_KNOBS.items()}):
        ctx2 = _ctx(fx.system)
        for k, v in cfg.items():
            ctx2.set_tuning(k, v)
        ctx2.set_positions(pos)
        ctx2.set_velocities(vel)
        md2 = WaterMD(ctx2, **kw)
        md2.init()
        md2.run(10)
        outs.append(ctx2.positions())
        ctx2.close()
    assert rel_rms(outs[0], outs[1]) <= 1e-7   # mixed forms: FP32 rounding differs between pair-kernel forms
    ctx.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_md_trajectory_is_bitwise_reproducible(prec):
    """Two runs of the same 30 steps from the same state on fresh contexts
    give identical bits: the sums that feed the trajectory (drift momentum,
    kinetic energy, SHAKE virial, surface dipole) are fixed-order block
    partials, the forces are written once per atom by one lane group."""
    from swinmd_amd.forces import WaterMD
    fx = Fixture("water512_md10")
    c = fx.cfg
    out = []
    for _ in range(2):
        ctx = _ctx(fx.system)
        ctx.set_velocities(fx.vec("init_vel"))
        md = WaterMD(ctx, cutoff=c["cutoff"], cut_buff=c["cut_buff"], alpha=c["alpha"], kcut=c["kcut"], dt=c["dt"],
                     temperature=c["temperature"], coupl=c["coupl"], tol=c["shake_tol"], max_count=c["max_count"],
                     prec=prec, surface=True)
        md.init()
        rep = md.run(30)
        out.append((ctx.positions(), ctx.velocities(), rep["fnh"], rep["atom_ke2"]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2] and out[0][3] == out[1][3]
