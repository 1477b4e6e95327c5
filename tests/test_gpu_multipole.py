"""Permanent-multipole Ewald on the GPU (smd_multipole_setup / compute /
compute_long through the C ABI) against the reference's golden vectors.

Bar: the plain FP64 form's 1e-10 relative RMS on forces and virials, 1e-10
relative on energies and volume derivatives (sums run in another order than
the reference's, so results agree to rounding, not bit for bit)."""
import numpy as np
import pytest

from tests.helpers import Fixture, rel, rel_rms

pytestmark = pytest.mark.gpu

TOL = 1e-10

SHORT = (("real_energy", "mp_real"), ("intra_energy", "mp_molCorrect"), ("surf_energy", "mp_surfCorrect"),
         ("real_deriv", "mp_realDeriv"), ("mol_deriv", "mp_molDeriv"), ("surf_deriv", "mp_surfDeriv"),
         ("surf_deriv2", "mp_surfSecDeriv"))
LONG = (("energy", "mp_long"), ("deriv", "mp_longDeriv"), ("deriv2", "mp_longSecDeriv"))


def _close(a, b, scale):
    return abs(a - b) <= TOL * max(abs(b), scale)


@pytest.mark.parametrize("name", ["amoeba216_multipole", "amoeba216_multipole_flexible"])
def test_multipole_matches_reference(name):
    from swinmd_amd.forces import Context, Multipole
    fx = Fixture(name)
    c = fx.cfg
    print(f"case {name}", flush=True)
    ctx = Context(fx.system)
    ctx.set_pairlist(c["cutoff"], c["cut_buff"], True)
    ctx.compute_mol_centers()
    mp = Multipole(ctx, alpha=c["alpha"], real_cut=c["cutoff_ew"], kcut=c["kcut"], use_constraint=c["constraint"],
                   surface=c["surface"], torques=c["torques"])
    n = fx.system.n_atoms
    fs = np.zeros((n, 3))
    out = mp.compute(fs)
    assert rel_rms(fs, fx.vec("mp_short_force")) <= TOL
    assert rel_rms(out["virial"], fx.ref["mp_short_virial"]) <= TOL
    # energies of very different size: compare on the scale of the largest
    scale = max(abs(fx.meta[k]) for _, k in SHORT) * 1e-6
    for k, mk in SHORT:
        assert _close(out[k], fx.meta[mk], scale), (k, out[k], fx.meta[mk])
    assert rel(out["self_energy"] + out["charged_energy"], fx.meta["mp_correct"]) <= TOL
    fl = np.zeros((n, 3))
    lo = mp.compute_long(fl)
    assert rel_rms(fl, fx.vec("mp_long_force")) <= TOL
    assert rel_rms(lo["virial"], fx.ref["mp_long_virial"]) <= TOL
    for k, mk in LONG:
        assert rel(lo[k], fx.meta[mk]) <= TOL, (k, lo[k], fx.meta[mk])
    assert rel(out["tot_torque"] + lo["tot_torque"], fx.meta["mp_totTorque"]) <= 1e-9


def test_multipole_1024_waters_against_oracle(oracle_mod):
    """A larger box (3072 sites, about a thousand k-vectors) against the C
    oracle, with the surface term and torques."""
    from swinmd_amd import systems as S
    from swinmd_amd.forces import Context, Multipole
    s = S.water(1024, density_mol=33.4, seed=37, lattice=(11, 11, 11), r_oh=0.09572, hoh_deg=108.5,
                types=S.AMOEBA_TYPES)
    print("case water1024", flush=True)
    ctx = Context(s)
    ctx.set_pairlist(0.9, 0.1, True)
    ctx.compute_mol_centers()
    kcut = S.kcut_for(81, s.box[0])
    mp = Multipole(ctx, alpha=3.5, real_cut=0.9, kcut=kcut, use_constraint=True, surface=True, torques=True)
    n = s.n_atoms
    fs, fl = np.zeros((n, 3)), np.zeros((n, 3))
    out = mp.compute(fs)
    lo = mp.compute_long(fl)
    o = oracle_mod.Oracle(s)
    dip, quad = o.local2global()
    r = o.multipole(dip, quad, alpha=3.5, real_cut=0.9, cutoff=0.9, cut_buff=0.1, kcut=kcut, use_constraint=True,
                    surface=True, torques=True)
    assert rel_rms(fs, r["short_force"]) <= TOL
    assert rel_rms(fl, r["long_force"]) <= TOL
    assert rel(out["real_energy"], r["real_energy"]) <= TOL
    assert rel(lo["energy"], r["long_energy"]) <= TOL
    assert rel(lo["deriv2"], r["long_deriv2"]) <= TOL
