"""Unconstrained Lennard-Jones MD on the device against the compiled
reference: 864 LJ atoms on the fcc lattice at rho* = 0.8442 with the
energy-shifted 2.5 sigma cutoff and the leap-frog Nose-Hoover step of
LFIntegrator::run without constraints (src/integrators/LFIntegrator.cpp:
587-643), 20 steps with Verlet rebuilds; the per-atom update alone
(smd_lf_atoms) and the LJ long-range correction (smd_lj_long_range_correct)
against the oracle, whose own functions the CPU suite pins (the trajectory
fixture, the dispersion-tail integrals)."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import Fixture, rel, rel_rms

pytestmark = pytest.mark.gpu


def _ctx(system):
    from swinmd_amd.forces import Context
    return Context(system)


@pytest.mark.parametrize("use_graph", [False, True])
def test_lj_md_matches_reference(use_graph):
    """start() + 20 steps on the device: positions, velocities, forces,
    energies, list updates, the last step's pressure tensors and the stress
    sums against the compiled reference."""
    from oracle import oracle as O
    from swinmd_amd.forces import WaterMD
    fx = Fixture("lj864_md20_nh")
    c = fx.cfg
    ctx = _ctx(fx.system)
    ctx.set_velocities(fx.vec("init_vel"))
    md = WaterMD(ctx, cutoff=c["cutoff"], cut_buff=c["cut_buff"], dt=c["dt"], temperature=c["temperature"],
                 coupl=c["coupl"], atom_thermo=True, use_graph=use_graph, total_steps=int(fx.meta["totsteps"]),
                 lustig=True, constraint=False, ewald=False, switch_on=True)
    md.init()
    rep = md.run(int(fx.extra))
    assert rel_rms(ctx.positions(), fx.vec("end_pos")) <= 1e-11
    assert rel_rms(ctx.velocities(), fx.vec("end_vel")) <= 1e-9
    assert rel_rms(ctx.forces(), fx.vec("end_force")) <= 1e-9
    assert rel(rep["e_lj"], fx.meta["e_lj"]) <= 1e-10
    assert rel(0.5 * rep["atom_ke2"], fx.meta["e_atomKin"]) <= 1e-10
    assert rel(0.5 * rep["mol_ke2"], fx.meta["e_molKin"]) <= 1e-10
    assert rep["e_real"] == 0.0 and rep["e_long"] == 0.0
    r = O.md_from_cfg(O.Oracle(fx.system), c, int(fx.extra))
    assert rep["list_updates"] == r["list_updates"] and r["list_updates"] > 0
    pr = md.pressure()
    for key in ("virial", "mol_virial", "atom_pressure", "mol_pressure"):
        want = fx.ref["end_" + key]
        assert np.max(np.abs(pr[key] - want)) <= 1e-9 * np.max(np.abs(want)), key
    assert rel(pr["lj_dudv"], fx.meta["ljdUlus"]) <= 1e-9
    assert rel(pr["mol_pt"], fx.meta["molPt"]) <= 1e-9
    st = np.concatenate([pr["atom_stress"], pr["mol_stress"]])
    want = fx.ref["stress_sums"]
    assert np.max(np.abs(st - want)) <= 1e-9 * np.max(np.abs(want))
    assert rel_rms(md.real_positions(), fx.vec("end_real_pos")) <= 1e-11


def test_lf_atoms_matches_oracle():
    """One smd_lf_atoms call (LFIntegrator.cpp:587-643) with a nonzero
    friction on LJ forces against orc_lf_atoms."""
    from oracle import oracle as O
    from swinmd_amd.forces import LFIntegrator, LJForce
    fx = Fixture("lj864_md20_nh")
    s = fx.system
    ctx = _ctx(s)
    ctx.set_pairlist(2.5, 0.1, True)
    ctx.compute_mol_centers()
    ctx.zero_forces()
    LJForce(ctx, cutoff=2.5).compute()
    F = np.ascontiguousarray(ctx.forces())
    rng = np.random.default_rng(4)
    vtm = rng.normal(scale=1.0, size=(s.n_atoms, 3))
    vel = vtm + rng.normal(scale=0.01, size=(s.n_atoms, 3))
    ctx.set_velocities(vel)
    lf = LFIntegrator(ctx, dt=0.005, temperature=86.8, tol=1e-8)
    lf.set_vt_minus(vtm)
    lf.lf_atoms(fnh=0.83)
    o = O.Oracle(s)
    pos, vt, v = o.pos.copy(), vtm.copy(), vel.copy()
    disp = np.zeros_like(vt)
    mass = np.ascontiguousarray(s.mass)
    rc = O.lib().orc_lf_atoms(s.n_atoms, O._p(pos), O._p(vt), O._p(v), O._p(disp), O._p(F), O._p(mass), 0.83, 1,
                              0.005)
    assert rc == 0
    assert rel_rms(ctx.positions(), pos) <= 1e-14
    assert rel_rms(lf.vt_minus(), vt) <= 1e-13
    assert rel_rms(ctx.velocities(), v) <= 1e-13


def test_lj_long_range_correct_matches_oracle():
    """smd_lj_long_range_correct (LJForce::long_range_correct,
    src/forces/LJForce.cpp:260-370) on two LJ types with exclusions against
    orc_lj_lrc."""
    from dataclasses import replace

    from oracle import oracle as O
    from swinmd_amd import systems as S
    from swinmd_amd.forces import lj_long_range_correct
    types = [replace(S.SPCE_TYPES[0], eps=0.0565, sigma=0.2655), S.SPCE_TYPES[1]]
    s = S.water(64, density_mol=33.4, seed=3, types=types)
    ctx = _ctx(s)
    got = lj_long_range_correct(ctx, 0.7)
    want = O.Oracle(s).lj_lrc(0.7)
    for g, w in zip(got, want):
        assert rel(g, w) <= 1e-14
