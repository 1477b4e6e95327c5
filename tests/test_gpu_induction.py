"""Induced dipoles on the GPU (smd_set_multipoles / smd_induce through the C
ABI) against the reference's golden vectors and the C oracle.

Bars: lab-frame multipoles within 1e-12 relative; the permanent field and
the induced dipoles within 1e-10 relative RMS (the CG of both sides converges
to |r| < 1e-14 on the same linear system; the order of the sums differs)."""
import numpy as np
import pytest

from tests.helpers import Fixture, rel_rms

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _ctx(system):
    from swinmd_amd.forces import Context
    return Context(system)


def _induction(fx, **kw):
    from swinmd_amd.forces import Induction
    c = fx.cfg
    ctx = _ctx(fx.system)
    ctx.set_pairlist(c["cutoff"], c["cut_buff"], True)
    ind = Induction(ctx, cutoff_ind=c["cutoff_ind"], a_damp=c["a_damp"], damp_tij=c["ind_damp"],
                    do_min=c["ind_min"], **kw)
    return ctx, ind


@pytest.mark.parametrize("name", ["amoeba216_induction_cg", "amoeba216_induction_direct"])
def test_induction_matches_reference(name, oracle_mod):
    fx = Fixture(name)
    ctx, ind = _induction(fx)
    iters, converged, res = ind.induce()
    dip, quad, e0, mu = ind.results()
    assert rel_rms(dip, fx.vec("ind_dip")) <= 1e-12
    assert rel_rms(quad.ravel(), fx.ref["ind_quad"]) <= 1e-12
    assert rel_rms(e0, fx.vec("ind_ei0")) <= TOL
    assert rel_rms(mu, fx.vec("ind_pi")) <= TOL
    if fx.cfg["ind_min"]:
        o = oracle_mod.Oracle(fx.system)
        d, q = o.local2global()
        _, _, o_iters, _ = o.induced_dipoles(d, q, cutoff_ind=fx.cfg["cutoff_ind"], a_damp=fx.cfg["a_damp"])
        assert converged and res < 1e-14
        assert abs(iters - o_iters) <= 2


def test_induction_guesses_reach_the_same_dipoles():
    """Starting from alpha E0, or from the previous solution, CG reaches the
    same dipoles as from the reference's start, in fewer iterations."""
    fx = Fixture("amoeba216_induction_cg")
    ctx, ind = _induction(fx)
    it0, _, _ = ind.induce()
    mu0 = ind.results()[3]
    ind.params.guess = 2
    it2, conv2, _ = ind.induce()
    mu2 = ind.results()[3]
    ind.params.guess = 1
    it1, conv1, _ = ind.induce()
    mu1 = ind.results()[3]
    assert conv1 and conv2
    assert rel_rms(mu2, mu0) <= 1e-12 and rel_rms(mu1, mu0) <= 1e-12
    assert it2 < it0 and it1 <= 2


def test_induction_512_waters_against_oracle(oracle_mod):
    from swinmd_amd import systems as S
    from swinmd_amd.forces import Induction
    s = S.water(512, density_mol=33.4, seed=31, lattice=(8, 8, 8), r_oh=0.09572, hoh_deg=108.5, types=S.AMOEBA_TYPES)
    ctx = _ctx(s)
    ctx.set_pairlist(0.7, 0.05, True)
    ind = Induction(ctx, cutoff_ind=0.7, a_damp=0.39)
    iters, converged, _ = ind.induce()
    _, _, e0, mu = ind.results()
    o = oracle_mod.Oracle(s)
    d, q = o.local2global()
    oe0, omu, oit, _ = o.induced_dipoles(d, q, cutoff_ind=0.7, a_damp=0.39)
    assert converged
    assert rel_rms(e0, oe0) <= TOL
    assert rel_rms(mu, omu) <= TOL
    assert abs(iters - oit) <= 2


INDF_SHORT = (("real_energy", "indf_real"), ("self_energy", "indf_self"), ("intra_energy", "indf_molCorrect"),
              ("surf_energy", "indf_surf"), ("real_deriv", "indf_realDeriv"), ("mol_deriv", "indf_molDeriv"),
              ("surf_deriv", "indf_surfDeriv"), ("surf_deriv2", "indf_surfSecDeriv"), ("self_deriv", "indf_selfDeriv"))
INDF_LONG = (("energy", "indf_long"), ("deriv", "indf_longDeriv"), ("deriv2", "indf_longSecDeriv"))


@pytest.mark.parametrize("name", ["amoeba216_indforces_cg", "amoeba216_indforces_tinker_flexible",
                                  "amoeba216_indforces_tinker_direct"])
def test_induction_terms_match_reference(name):
    """Induced dipoles solved on the device, then Induction's energy, force and
    torque terms (real space over every pair within cutoffEw, excluded pairs
    included; self, intramolecular, surface; k-space) against the reference.
    Bars as above: 1e-10 relative RMS on forces and virials, 1e-10 relative on
    the energies and volume derivatives, measured on the largest term's scale
    for terms that nearly cancel."""
    from swinmd_amd.forces import Induction
    fx = Fixture(name)
    c = fx.cfg
    print(f"case {name}", flush=True)
    ctx = _ctx(fx.system)
    ctx.set_pairlist(c["cutoff"], c["cut_buff"], True)
    ctx.compute_mol_centers()
    ind = Induction(ctx, cutoff_ind=c["cutoff_ind"], a_damp=c["a_damp"], damp_tij=c["ind_damp"], do_min=c["ind_min"])
    ind.induce()
    assert rel_rms(ind.results()[3], fx.vec("indf_pi")) <= TOL
    ind.setup_terms(alpha=c["alpha"], real_cut=c["cutoff_ew"], kcut=c["kcut"], use_constraint=c["constraint"],
                    surface=c["surface"], torques=c["torques"], damp_amoeba=c["damp_amoeba"])
    n = fx.system.n_atoms
    fs, fl = np.zeros((n, 3)), np.zeros((n, 3))
    out = ind.compute(fs)
    lo = ind.compute_long(fl)
    assert rel_rms(fs, fx.vec("indf_short_force")) <= TOL
    assert rel_rms(fl, fx.vec("indf_long_force")) <= TOL
    assert rel_rms(out["virial"], fx.ref["indf_short_virial"]) <= TOL
    assert rel_rms(lo["virial"], fx.ref["indf_long_virial"]) <= TOL
    scale = max(abs(fx.meta[k]) for _, k in INDF_SHORT) * 1e-6
    for k, mk in INDF_SHORT:
        assert abs(out[k] - fx.meta[mk]) <= TOL * max(abs(fx.meta[mk]), scale), (k, out[k], fx.meta[mk])
    for k, mk in INDF_LONG:
        assert abs(lo[k] - fx.meta[mk]) <= TOL * abs(fx.meta[mk]), (k, lo[k], fx.meta[mk])


def test_time_apply_leaves_the_solution_unchanged():
    from swinmd_amd import _lib as L
    fx = Fixture("amoeba216_induction_cg")
    ctx, ind = _induction(fx)
    with pytest.raises(L.SmdError) as e:
        ind.time_apply(3)                  # no solve yet
    assert e.value.code == L.SMD_ESTATE
    ind.induce()
    mu = ind.results()[3]
    assert ind.time_apply(5) > 0.0
    assert np.array_equal(ind.results()[3], mu)
    ctx.close()
