"""The polarizable-water leap-frog step on the device (smd_polar_md_*) and the
preconditioned induced-dipole solve (smd_induce with precond, FP64 and mixed)
through the C ABI.

Bars: the reference-method step (the reference's CG to |r| < 1e-14) matches
the compiled reference's 10-step trajectory (positions 1e-10, velocities
1e-8, energies 1e-8 relative); the preconditioned solves stop at the RMS
dipole tolerance and their dipoles agree with the reference's to within a
small multiple of it."""
import numpy as np
import pytest

from tests.helpers import Fixture, rel, rel_rms

pytestmark = pytest.mark.gpu


def _ctx(system):
    from swinmd_amd.forces import Context
    return Context(system)


def _polar_md(fx, ctx, use_graph, **induce_kw):
    from swinmd_amd.forces import Induction, PolarMD
    c = fx.cfg
    ind = Induction(ctx, cutoff_ind=c["cutoff_ind"], a_damp=c["a_damp"], damp_tij=c["ind_damp"],
                    do_min=c["ind_min"], **induce_kw)
    md = PolarMD(ctx, ind, cutoff=c["cutoff"], cut_buff=c["cut_buff"], alpha=c["alpha"], kcut=c["kcut"],
                 dt=c["dt"], temperature=c["temperature"], coupl=c["coupl"], tol=c["shake_tol"],
                 max_count=c["max_count"], real_cut=c["cutoff_ew"], vdw="buff", switch_on=c["switch_on"],
                 lustig=c["lustig"], torques=c["torques"], surface=c["surface"], damp_amoeba=c["damp_amoeba"],
                 use_graph=use_graph, total_steps=int(fx.meta["totsteps"]), atom_thermo=c["atom_thermo"])
    return ind, md


@pytest.mark.parametrize("use_graph,joint,overlap,m_sums", [(False, 1, 1, 0), (True, 1, 1, 0), (True, 0, 1, 0),
                                                            (True, 1, 0, 0), (False, 1, 1, 1), (True, 1, 1, 1)])
def test_polar_md_matches_reference(use_graph, joint, overlap, m_sums):
    """start() + 10 steps: buffered 14-7, Multipole, Induction (the
    reference's CG from Pi = 1 to |r| < 1e-14), SHAKE, Nose-Hoover; with
    one k-space pass for both classes after the solve (joint) and with a
    pass per class, with and without the second stream (vdW and Multipole's
    short range beside the solve), with and without the permanent
    multipoles' structure factors beside it."""
    fx = Fixture("amoeba216_md10_polar")
    print(f"case amoeba216 graph {use_graph} joint {joint} overlap {overlap} m_sums {m_sums}", flush=True)
    ctx = _ctx(fx.system)
    ctx.set_tuning("polar_joint_long", joint)
    ctx.set_tuning("recip_overlap", overlap)
    ctx.set_tuning("polar_m_sums", m_sums)
    ctx.set_velocities(fx.vec("init_vel"))
    _, md = _polar_md(fx, ctx, use_graph)
    md.init()
    rep = md.run(int(fx.extra))
    assert rep["steps"] == int(fx.extra)
    assert rel_rms(ctx.positions(), fx.vec("end_pos")) <= 1e-10
    assert rel_rms(ctx.velocities(), fx.vec("end_vel")) <= 1e-8
    assert rel_rms(ctx.forces(), fx.vec("end_force")) <= 1e-8
    assert rel(rep["e_vdw"], fx.meta["e_lj"]) <= 1e-8
    assert rel(rep["e_mp_real"], fx.meta["e_real"]) <= 1e-8
    assert rel(rep["e_mp_long"], fx.meta["e_long"]) <= 1e-8
    assert rel(rep["e_mp_self"], fx.meta["e_correct"]) <= 1e-8
    assert rel(rep["e_mp_intra"], fx.meta["e_molCorrect"]) <= 1e-8
    assert rel(0.5 * rep["atom_ke2"], fx.meta["e_atomKin"]) <= 1e-8
    assert rel(0.5 * rep["mol_ke2"], fx.meta["e_molKin"]) <= 1e-8
    assert rep["cg_iterations"] > 0 and rep["last_residual"] < 1e-14
    assert rel_rms(md_real(ctx), fx.vec("end_real_pos")) <= 1e-10


def test_polar_md_1024_from_the_lattice_start_matches_reference():
    """start() + 3 steps of the 1,024-molecule polarizable water from its
    lattice start (cutoff 0.9 nm, |n|^2 <= 81, cell list), the reference's
    CG, against the compiled reference."""
    fx = Fixture("polar1024_md3")
    print("case polar1024_md3", flush=True)
    ctx = _ctx(fx.system)
    ctx.set_velocities(fx.vec("init_vel"))
    _, md = _polar_md(fx, ctx, True)
    md.init()
    rep = md.run(int(fx.extra))
    assert rep["steps"] == int(fx.extra)
    assert rel_rms(ctx.positions(), fx.vec("end_pos")) <= 1e-10
    assert rel_rms(ctx.velocities(), fx.vec("end_vel")) <= 1e-8
    assert rel_rms(ctx.forces(), fx.vec("end_force")) <= 1e-8
    assert rel(rep["e_vdw"], fx.meta["e_lj"]) <= 1e-8
    assert rel(rep["e_mp_real"], fx.meta["e_real"]) <= 1e-8
    assert rel(rep["e_mp_long"], fx.meta["e_long"]) <= 1e-8
    assert rel(0.5 * rep["atom_ke2"], fx.meta["e_atomKin"]) <= 1e-8


def md_real(ctx):
    r = np.zeros((ctx.n, 3))
    import ctypes as C
    ctx.check(ctx.lib.smd_get_real_positions(ctx.h, r.ctypes.data_as(C.c_void_p), ctx.n))
    return r


@pytest.mark.parametrize("precond", [1, 2])
@pytest.mark.parametrize("prec", [0, 1])
def test_preconditioned_solve_reaches_the_reference_dipoles(prec, precond):
    """CG preconditioned by diag(alpha) (1) or by each molecule's 9x9 block
    of A (2), stopping at an RMS dipole residual of 1e-6 D, FP64 or with FP32
    A v: the dipoles are within a few times the tolerance of the reference's
    (|r| < 1e-14) and the solve needs fewer iterations; a tight tolerance
    reaches them to 1e-9; the block preconditioner needs fewer iterations
    than the diagonal one."""
    from swinmd_amd import _lib as L
    from swinmd_amd.forces import Induction
    fx = Fixture("amoeba216_induction_cg")
    c = fx.cfg
    ctx = _ctx(fx.system)
    ctx.set_pairlist(c["cutoff"], c["cut_buff"], True)
    ind = Induction(ctx, cutoff_ind=c["cutoff_ind"], a_damp=c["a_damp"], damp_tij=c["ind_damp"], do_min=True)
    it_ref, conv, _ = ind.induce()
    mu_ref = ind.results()[3]
    assert conv and rel_rms(mu_ref, fx.vec("ind_pi")) <= 1e-10
    tol = 1e-6 * L.DEBYE_E_NM
    ind2 = Induction(ctx, cutoff_ind=c["cutoff_ind"], a_damp=c["a_damp"], damp_tij=c["ind_damp"], do_min=True,
                     tol=tol, precond=precond, prec=prec)
    it, conv, res = ind2.induce()
    mu = ind2.results()[3]
    assert conv and res < tol
    err = np.sqrt(np.mean(np.sum((mu - mu_ref) ** 2, axis=1)))
    assert err <= 20 * tol, (err, tol)
    assert it < it_ref
    if precond == 2:
        ind1 = Induction(ctx, cutoff_ind=c["cutoff_ind"], a_damp=c["a_damp"], damp_tij=c["ind_damp"], do_min=True,
                         tol=tol, precond=1, prec=prec)
        it1, conv1, _ = ind1.induce()
        print(f"prec {prec}: diagonal {it1}, block {it} iterations", flush=True)
        assert conv1 and it < it1
    if prec == 0:
        ind2.params.tol = 1e-9 * tol
        it_t, conv, _ = ind2.induce()
        assert conv and rel_rms(ind2.results()[3], mu_ref) <= 1e-9
    # every start reaches the same dipoles
    for guess in (1, 2, 4):
        ind2.params.guess = guess
        ind2.params.tol = tol
        it_g, conv, _ = ind2.induce()
        mu_g = ind2.results()[3]
        assert conv and np.sqrt(np.mean(np.sum((mu_g - mu_ref) ** 2, axis=1))) <= 20 * tol


@pytest.mark.parametrize("guess,precond", [(2, 1), (4, 1), (2, 2)])
def test_polar_md_preconditioned_steps_follow_the_reference_method(guess, precond):
    """The step with the preconditioned FP64 solve at 1e-6 D, FP64, from
    alpha E0 (guess 2) or from the previous step's dipoles moved by alpha dE0
    (guess 4), stays on the reference-method trajectory over 10 steps to the
    accuracy the tolerance allows, and needs fewer solve iterations per step."""
    from swinmd_amd import _lib as L
    fx = Fixture("amoeba216_md10_polar")
    ctx = _ctx(fx.system)
    ctx.set_velocities(fx.vec("init_vel"))
    _, md = _polar_md(fx, ctx, True, tol=1e-6 * L.DEBYE_E_NM, precond=precond, guess=guess)
    md.init()
    rep = md.run(int(fx.extra))
    print(f"guess {guess} precond {precond}: {rep['cg_iterations'] / rep['steps']:.2f} iterations per step",
          flush=True)
    assert rel_rms(ctx.positions(), fx.vec("end_pos")) <= 1e-7
    assert rel_rms(ctx.velocities(), fx.vec("end_vel")) <= 1e-4
    assert rel(rep["e_mp_real"], fx.meta["e_real"]) <= 1e-5
    assert rep["cg_iterations"] / rep["steps"] < 40


def test_polar_md_mixed_step_stays_on_the_fp64_trajectory():
    """The mixed form of the step (FP32 pair and k-space math, FP32-table
    solve) against the FP64 form with the same preconditioned solve at
    1e-6 D, 10 steps from the same state."""
    from swinmd_amd import _lib as L
    fx = Fixture("amoeba216_md10_polar")
    out = {}
    for prec in (L.PREC_FP64, L.PREC_MIXED):
        print(f"case polar md prec {prec}", flush=True)
        ctx = _ctx(fx.system)
        ctx.set_velocities(fx.vec("init_vel"))
        _, md = _polar_md(fx, ctx, True, tol=1e-6 * L.DEBYE_E_NM, precond=True, guess=2, prec=prec)
        md.params.prec = prec
        md.init()
        f0 = ctx.forces()
        rep = md.run(10)
        out[prec] = (f0, ctx.positions(), ctx.velocities(), rep)
        ctx.close()
    f64, p64, v64, r64 = out[L.PREC_FP64]
    f32, p32, v32, r32 = out[L.PREC_MIXED]
    assert rel_rms(f32, f64) <= 1e-5
    assert rel_rms(p32, p64) <= 1e-6
    assert rel_rms(v32, v64) <= 1e-2
    assert rel(r32["e_mp_real"], r64["e_mp_real"]) <= 1e-5
    assert rel(r32["e_ind_real"], r64["e_ind_real"]) <= 1e-4
