"""The oracle's sheared-replica integrator (molecular SLLOD with Lees-Edwards
images), which has no counterpart in the reference.  Each function is pinned
by what the mathematics fixes rather than by restating it:

* zero shear rate gives the reference's step (orc_md_step) bit for bit;
* the Lees-Edwards pair set equals a brute-force search over explicit images
  of the sheared lattice;
* the sheared k set is the reciprocal lattice of the sheared cell (from a
  matrix inverse) inside the cutoff, and does not change when the tilt moves
  by a box length;
* the Ewald total in a sheared cell does not depend on the splitting
  parameter, and its forces are the gradient of its energy;
* without a thermostat, the energy changes at the rate SLLOD prescribes,
  dH/dt = gamma (dU/d eps - sum_m P_x P_y / M), where dU/d eps is the energy's
  derivative under the affine shear of the molecular centres (with the tilt),
  taken by finite differences."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from swinmd_amd import systems as S

RHO = 33.43   # molecules / nm^3 (1 g/cm^3)


def _water(n=256, seed=5):
    return S.water(n, density_mol=RHO, seed=seed, subset=True)


def _cfg(L, **kw):
    c = dict(cutoff=0.9, cut_buff=0.08, alpha=3.5, kcut=S.kcut_for(81, L), dt=0.002, temperature=300.0,
             coupl=0.1, use_cells=False, atom_thermo=False, seed=3)
    c.update(kw)
    return c


def _energy(system, pos, com, tilt, alpha, real_cut, kcut, lj=True):
    """Total potential energy (LJ + Ewald real, reciprocal, intramolecular,
    self) of a configuration in the sheared cell."""
    o = O.Oracle(system, pos=pos, com=com)
    o.set_tilt(tilt)
    off, nb, _ = o.pairlist(real_cut, 0.0, use_cells=False)
    e = 0.0
    if lj:
        _, vo = o.vdw(off, nb, "lj", cutoff=real_cut)
        e += vo["lj_energy"]
    _, er = o.ewald_real(off, nb, alpha, real_cut)
    _, ei = o.ewald_intra(alpha)
    _, el = o.ewald_recip(kcut, alpha, kv=o.kvectors_le(kcut, alpha, tilt))
    es, ec = o.ewald_self(alpha)
    return e + er["energy"] + ei["energy"] + el["energy"] + es + ec


def test_zero_shear_is_the_reference_step():
    s = _water()
    cfg = _cfg(s.box[0])
    a = O.md_from_cfg(O.Oracle(s), cfg, 4)
    b = O.sllod_run(O.Oracle(s), cfg, 0.0, 4)
    for k in ("pos", "vel", "vtm", "force", "com"):
        assert np.array_equal(a[k], b[k]), k
    assert a["fnh"] == b["fnh"] and a["e_long"] == b["e_long"]
    assert b["tilt"] == 0.0


@pytest.mark.parametrize("frac", [0.0, 0.31, -0.47, 0.4999])
def test_lees_edwards_pairs_equal_an_explicit_image_search(frac):
    s = _water(256, seed=8)
    L = s.box[0]
    tilt = frac * L
    o = O.Oracle(s)
    o.set_tilt(tilt)
    rl = 0.95
    off, nb, _ = o.pairlist(0.9, 0.05, use_cells=False)
    got = set()
    for i in range(s.n_atoms):
        for j in nb[off[i]:off[i + 1]]:
            got.add((min(i, int(j)), max(i, int(j))))
    # explicit images: r_j - r_i + nx a + ny b + nz c, b = (tilt, L, 0)
    pos = s.pos
    d = pos[None, :, :] - pos[:, None, :]
    best = np.full(d.shape[:2], np.inf)
    for ny in (-1, 0, 1):
        for nx in (-2, -1, 0, 1, 2):
            for nz in (-1, 0, 1):
                sh = np.array([nx * L + ny * tilt, ny * L, nz * L])
                best = np.minimum(best, np.sum((d + sh) ** 2, axis=2))
    ii, jj = np.nonzero(np.triu(best < rl * rl, 1))
    same_mol = s.mol[ii] == s.mol[jj]
    want = {(int(a), int(b)) for a, b, m in zip(ii, jj, same_mol) if not m}
    assert got == want
    assert len(want) > 1000


def _recip_lattice(L, tilt, kcut):
    H = np.array([[L, tilt, 0.0], [0.0, L, 0.0], [0.0, 0.0, L]])   # columns a, b, c
    B = 2 * math.pi * np.linalg.inv(H).T
    m = int(kcut * L / (2 * math.pi)) + 8
    rng = np.arange(-m, m + 1)
    n = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    k = n @ B.T
    keep = (np.sum(k * k, axis=1) < kcut * kcut) & np.any(n != 0, axis=1)
    return k[keep], H


@pytest.mark.parametrize("frac", [0.0, 0.23, -0.5, 0.44])
def test_sheared_kvectors_are_the_reciprocal_lattice(frac):
    o = O.Oracle(_water())
    L = float(o.box[0])
    tilt = frac * L
    kcut = S.kcut_for(81, L)
    kv, k2, ek = o.kvectors_le(kcut, 3.5, tilt)
    full, H = _recip_lattice(L, tilt, kcut)
    # one of each +-k pair: the oracle's set doubled by sign is the lattice set
    both = np.concatenate([kv, -kv])
    key = lambda a: set(map(tuple, np.round(a * 1e6).astype(np.int64)))
    assert key(both) == key(full)
    assert len(kv) * 2 == len(full)
    # k . a, k . b, k . c are 2 pi times integers
    n = kv @ H / (2 * math.pi)
    assert np.max(np.abs(n - np.round(n))) < 1e-9
    np.testing.assert_allclose(k2, np.sum(kv * kv, axis=1), rtol=1e-14)
    np.testing.assert_allclose(ek, np.exp(-k2 / (4 * 3.5 ** 2)) / k2, rtol=1e-14)
    # a tilt one box length away is the same lattice
    kv2, _, _ = o.kvectors_le(kcut, 3.5, tilt - L if frac > 0 else tilt + L)
    assert key(np.concatenate([kv2, -kv2])) == key(full)


def test_sheared_ewald_total_does_not_depend_on_alpha():
    s = _water(256, seed=11)
    L = s.box[0]
    tilt = 0.37 * L
    # both splittings converged: erfc(alpha rc) < 3e-9 and exp(-kcut^2 / 4 alpha^2) < 1e-13
    kcut = 2 * math.pi * 16 / L
    e1 = _energy(s, s.pos, s.com(s.pos), tilt, 4.3, 0.95, kcut, lj=False)
    e2 = _energy(s, s.pos, s.com(s.pos), tilt, 4.6, 0.95, kcut, lj=False)
    assert abs(e1 - e2) <= 3e-6 * abs(e1)
    # and a tilt shifted by a box length gives the same energy
    e3 = _energy(s, s.pos, s.com(s.pos), tilt - L, 4.3, 0.95, kcut, lj=False)
    assert abs(e3 - e1) <= 1e-11 * abs(e1)
    # while a wrong sign of the tilt (the mirror cell) does not
    e4 = _energy(s, s.pos, s.com(s.pos), -tilt, 4.3, 0.95, kcut, lj=False)
    assert abs(e4 - e1) > 1e-3 * abs(e1) or abs(e4 - e1) > 1.0


def test_sheared_reciprocal_forces_are_the_energy_gradient():
    s = _water(256, seed=12)
    L = s.box[0]
    tilt = -0.29 * L
    kcut = S.kcut_for(81, L)
    o = O.Oracle(s)
    o.set_tilt(tilt)
    kv = o.kvectors_le(kcut, 3.5, tilt)
    f, _ = o.ewald_recip(kcut, 3.5, kv=kv)
    h = 1e-6
    for atom in (0, 100, 401):
        for d in range(3):
            e = []
            for sgn in (1, -1):
                p = s.pos.copy()
                p[atom, d] += sgn * h
                oo = O.Oracle(s, pos=p)
                oo.set_tilt(tilt)
                e.append(oo.ewald_recip(kcut, 3.5, kv=kv)[1]["energy"])
            fd = -(e[0] - e[1]) / (2 * h)
            assert abs(fd - f[atom, d]) <= 1e-5 * max(1.0, abs(f[atom, d])), (atom, d)


def _affine_dU(system, pos, com, tilt, cfg, eps=1e-5):
    """dU/d eps under x_m += eps * Y_m for every molecule (rigid) and
    tilt += eps * Ly, by central differences."""
    Ly = system.box[1]
    e = []
    for sgn in (1, -1):
        shift = np.zeros_like(com)
        shift[:, 0] = sgn * eps * com[:, 1]
        p = pos + shift[system.mol]
        e.append(_energy(system, p, com + shift, tilt + sgn * eps * Ly, cfg["alpha"], cfg["cutoff"], cfg["kcut"]))
    return (e[0] - e[1]) / (2 * eps)


@pytest.fixture(scope="module")
def _equilibrated():
    s = _water(256, seed=5)
    cfg = _cfg(s.box[0])
    # a short thermostatted run at zero shear from the lattice start
    return s, cfg, O.sllod_run(O.Oracle(s), cfg, 0.0, 60)


@pytest.mark.parametrize("gamma", [6.0, -6.0])
def test_sllod_energy_balance_without_thermostat(_equilibrated, gamma):
    """With the thermostat off, H = U + K (K from the peculiar velocities of
    the integer steps) changes by the integral of gamma (dU/d eps - K_xy):
    the residual is the size of H's leap-frog fluctuation at zero shear
    (about 10 kJ/mol here), a few percent of the work, for either sign of
    the shear rate (a sign error in the coupling would flip the work)."""
    s, cfg, eq = _equilibrated
    o = O.Oracle(s, pos=eq["pos"], com=eq["com"])
    nve = dict(cfg, nose_hoover=False)
    mass_m = np.add.reduceat(s.mass, s.mol_start[:-1])
    samples = []

    def on_step(k, st):
        kxy = float(np.sum(st["mom"][:, 0] * st["mom"][:, 1] / mass_m))
        samples.append((kxy, _affine_dU(s, st["pos"], st["com"], st["tilt"], cfg)))

    n = 24
    out = O.sllod_run(o, nve, gamma, n, velocity=eq["vel"], tilt=0.0, on_step=on_step)
    rec = out["records"]
    h0 = rec[:, 1] + 0.5 * rec[:, 2]            # U(t_k) + K(t_k), k = 0 .. n-1
    kxy = np.array([a for a, _ in samples])      # at t_k
    due = np.array([b for _, b in samples])      # at t_{k+1}
    rate = gamma * (due[:-1] - kxy[1:])          # at t_j, j = 1 .. n-1
    pred = np.concatenate([[0.0], np.cumsum(0.5 * (rate[1:] + rate[:-1]) * cfg["dt"])])
    dh = h0[1:] - h0[1]
    work = pred[-1]
    assert abs(work) > 300.0, work
    assert np.max(np.abs(dh - pred)) <= 0.04 * abs(work)
