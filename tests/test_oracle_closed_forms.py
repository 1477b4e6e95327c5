"""The C oracle against what the mathematics fixes, independently of the
reference's own outputs: potential minima, the NaCl Madelung constant through
the full Ewald sum, a brute-force full-sphere k-space sum, forces as minus the
energy gradient, and the Nose-Hoover chain against a direct integration of
its equations of motion.  A dropped term, a wrong sign or factor of two, or a
transposed index in the oracle fails one of these."""
import ctypes as C
import math

import numpy as np
import pytest

from swinmd_amd.systems import AtomType, System

COULOMB = 138.93548501          # src/NEMD_defs.h:104, kJ/mol nm / e^2
BOLTZ = 0.0083145112            # src/NEMD_defs.h:108


def _atoms(types, pos, type_ids, box, mol=None):
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    n = len(pos)
    mol = np.arange(n, dtype=np.int32) if mol is None else np.asarray(mol, dtype=np.int32)
    apm = int(np.bincount(mol).max())
    return System(name="t", types=types, mol_type_name="X", atoms_per_mol=apm, pos=pos,
                  type=np.asarray(type_ids, dtype=np.int32), mol=mol, density=n / box ** 3)


def _oracle(s, pos=None):
    from oracle import oracle as O
    return O.Oracle(s, pos=pos)


# ---- 12-6 LJ -------------------------------------------------------------------

def test_lj_pair_minimum_zero_crossing_and_force_direction(oracle_mod):
    sigma, eps = 0.3166, 0.6502
    types = [AtomType(mass=15.9994, charge=0.0, eps=eps, sigma=sigma)]
    for r, e_want, f_want in ((2 ** (1 / 6) * sigma, -eps, 0.0), (sigma, 0.0, 24 * eps / sigma)):
        s = _atoms(types, [[0, 0, 0], [r, 0, 0]], [0, 0], box=3.0)
        o = _oracle(s)
        off, nb, _ = o.pairlist(1.0, 0.1, use_cells=False)
        f, out = o.vdw(off, nb, "lj", cutoff=1.0)
        assert out["lj_energy"] == pytest.approx(e_want, abs=1e-12)
        # force on the atom at +x: +dU/dr pushes it outward when repulsive
        assert f[1, 0] == pytest.approx(f_want, abs=1e-9)
        assert f[0, 0] == pytest.approx(-f_want, abs=1e-9)
        assert np.abs(f[:, 1:]).max() == 0.0


def _lj_cluster(seed=5):
    rng = np.random.default_rng(seed)
    types = [AtomType(mass=15.9994, charge=0.0, eps=0.6502, sigma=0.3166),
             AtomType(mass=12.0, charge=0.0, eps=0.3, sigma=0.35)]
    # points on a jittered lattice so that no pair sits near the cutoff
    grid = np.array([[i, j, k] for i in range(2) for j in range(2) for k in range(2)], dtype=float) * 0.42
    pos = grid + rng.uniform(-0.04, 0.04, grid.shape) + 0.3
    return _atoms(types, pos, [0, 1, 0, 1, 1, 0, 1, 0], box=2.6)


def test_lj_forces_are_minus_energy_gradient(oracle_mod):
    s = _lj_cluster()
    o = _oracle(s)
    off, nb, _ = o.pairlist(1.0, 0.1, use_cells=False)
    f, _ = o.vdw(off, nb, "lj", cutoff=1.0, switch_on=True)
    h = 1e-6
    for a in (0, 3, 6):
        for d in range(3):
            e = []
            for sgn in (1, -1):
                p = s.pos.copy()
                p[a, d] += sgn * h
                op = _oracle(s, pos=p)
                e.append(op.vdw(off, nb, "lj", cutoff=1.0, switch_on=True)[1]["lj_energy"])
            fd = -(e[0] - e[1]) / (2 * h)
            assert f[a, d] == pytest.approx(fd, rel=1e-6, abs=1e-6)


# ---- buffered 14-7 ---------------------------------------------------------------

def test_buffered_14_7_energy_and_force_at_rho_one(oracle_mod):
    """Two waters whose only interaction is O-O.  At r = R (rho = 1) Halgren's
    form gives E = -eps and dE/drho = 7 eps (1/(1+delta) - 1/(1+gamma))."""
    R, eps = 0.3405, 0.46
    types = [AtomType(mass=1.008, charge=0.0, eps=0.0, sigma=0.2),
             AtomType(mass=15.9994, charge=0.0, eps=eps, sigma=R)]
    one = np.array([[0.0816, 0.0577, 0.0], [0.0, 0.0, 0.0], [-0.0816, 0.0577, 0.0]])
    pos = np.vstack([one, one + [R, 0.0, 0.0]])
    s = _atoms(types, pos, [0, 1, 0, 0, 1, 0], box=3.0, mol=[0, 0, 0, 1, 1, 1])
    # bonds and the angle exclude the intramolecular pairs, as in water
    s.bonds = [(0, 1, 0), (1, 2, 0), (3, 4, 0), (4, 5, 0)]
    s.angles = [(0, 1, 2, 0), (3, 4, 5, 0)]
    o = _oracle(s)
    off, nb, _ = o.pairlist(1.0, 0.1, use_cells=False)
    f, out = o.vdw(off, nb, "buff", cutoff=1.0)
    assert out["lj_energy"] == pytest.approx(-eps, abs=1e-12)
    delta, gamma = 0.07, 0.12
    dEdr = 7.0 * eps / R * (1.0 / (1.0 + delta) - 1.0 / (1.0 + gamma))
    # molecule 0 sits at -x of molecule 1: its force is +dE/dr along x
    assert f[0:3, 0].sum() == pytest.approx(dEdr, rel=1e-10)
    assert f[3:6, 0].sum() == pytest.approx(-dEdr, rel=1e-10)
    assert np.abs(f[:, 1:].sum(axis=0)).max() < 1e-12


# ---- Ewald ---------------------------------------------------------------------

def _rock_salt(cells=3, a0=0.282):
    a = 2 * a0
    base = [([0, 0, 0], 0), ([0.5, 0.5, 0], 0), ([0.5, 0, 0.5], 0), ([0, 0.5, 0.5], 0),
            ([0.5, 0, 0], 1), ([0, 0.5, 0], 1), ([0, 0, 0.5], 1), ([0.5, 0.5, 0.5], 1)]
    pos, typ = [], []
    for i in range(cells):
        for j in range(cells):
            for k in range(cells):
                for b, t in base:
                    pos.append((np.array(b) + [i, j, k]) * a)
                    typ.append(t)
    L = cells * a
    pos = np.array(pos) - L / 2
    types = [AtomType(mass=22.99, charge=1.0, eps=0.0, sigma=0.0), AtomType(mass=35.45, charge=-1.0, eps=0.0, sigma=0.0)]
    return _atoms(types, pos, typ, box=L), a0


def test_ewald_total_reproduces_nacl_madelung_constant(oracle_mod):
    s, a0 = _rock_salt()
    o = _oracle(s)
    alpha, rc = 5.0, 0.8
    L = s.box[0]
    kcut = 2 * alpha * math.sqrt(21.0)             # exp(-k^2/4 alpha^2) < 1e-9 beyond it
    off, nb, _ = o.pairlist(rc, 0.0, use_cells=False)
    _, real = o.ewald_real(off, nb, alpha, rc)
    _, rec = o.ewald_recip(kcut, alpha)
    self_e, charged = o.ewald_self(alpha)
    assert charged == 0.0
    e = real["energy"] + rec["energy"] + self_e
    madelung = -2.0 * e * a0 / (s.n_atoms * COULOMB)
    assert madelung == pytest.approx(1.747564594633, abs=2e-6)
    assert L / 2 > rc


def _random_charges(n=24, L=1.6, seed=11):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-L / 2, L / 2, (n, 3))
    q = np.where(np.arange(n) % 2 == 0, 0.8, -0.8)
    types = [AtomType(mass=1.0, charge=0.8, eps=0.0, sigma=0.0), AtomType(mass=1.0, charge=-0.8, eps=0.0, sigma=0.0)]
    return _atoms(types, pos, (q < 0).astype(int), box=L)


def _direct_kspace(pos, q, L, alpha, kcut):
    """(2 pi / V) sum over the full sphere 0 < |k| < kcut of
    exp(-k^2/4a^2)/k^2 |S(k)|^2."""
    nmax = int(kcut * L / (2 * math.pi))
    n = np.array([[i, j, k] for i in range(-nmax, nmax + 1) for j in range(-nmax, nmax + 1)
                  for k in range(-nmax, nmax + 1) if (i, j, k) != (0, 0, 0)], dtype=float)
    kv = n * (2 * math.pi / L)
    k2 = (kv ** 2).sum(1)
    kv, k2 = kv[k2 < kcut * kcut], k2[k2 < kcut * kcut]
    S = (q[None, :] * np.exp(1j * kv @ pos.T)).sum(1)
    V = L ** 3
    return 2 * math.pi / V * np.sum(np.exp(-k2 / (4 * alpha ** 2)) / k2 * np.abs(S) ** 2)


def test_ewald_reciprocal_equals_full_sphere_sum_and_its_gradient(oracle_mod):
    s = _random_charges()
    o = _oracle(s)
    alpha, L = 3.5, s.box[0]
    kcut = 2 * math.pi * math.sqrt(36.5) / L
    f, rec = o.ewald_recip(kcut, alpha)
    e_direct = _direct_kspace(s.pos, s.scaled_charge, L, alpha, kcut)
    assert rec["energy"] == pytest.approx(e_direct, rel=1e-12)
    h = 1e-6
    for a in (0, 7, 13):
        for d in range(3):
            e = []
            for sgn in (1, -1):
                p = s.pos.copy()
                p[a, d] += sgn * h
                e.append(_direct_kspace(p, s.scaled_charge, L, alpha, kcut))
            assert f[a, d] == pytest.approx(-(e[0] - e[1]) / (2 * h), rel=1e-6, abs=1e-5)


def test_ewald_real_forces_are_minus_energy_gradient(oracle_mod):
    s = _random_charges(n=16, L=2.4, seed=3)
    o = _oracle(s)
    off, nb, _ = o.pairlist(1.1, 0.1, use_cells=False)
    f, out = o.ewald_real(off, nb, 3.5, 1.1)
    h = 1e-6
    for a in (1, 8):
        for d in range(3):
            e = []
            for sgn in (1, -1):
                p = s.pos.copy()
                p[a, d] += sgn * h
                e.append(_oracle(s, pos=p).ewald_real(off, nb, 3.5, 1.1)[1]["energy"])
            assert f[a, d] == pytest.approx(-(e[0] - e[1]) / (2 * h), rel=1e-6, abs=1e-5)
    # the real-space energy itself: sum over pairs inside rc of qi qj erfc(a r)/r
    from scipy.special import erfc
    q = s.scaled_charge
    e = 0.0
    for i in range(s.n_atoms):
        for j in range(i + 1, s.n_atoms):
            d = s.pos[j] - s.pos[i]
            d -= s.box[0] * np.round(d / s.box[0])
            r = np.linalg.norm(d)
            if r < 1.1:
                e += q[i] * q[j] * erfc(3.5 * r) / r
    assert out["energy"] == pytest.approx(e, rel=1e-12)


# ---- Nose-Hoover chain -------------------------------------------------------------

def _mtk_rhs(t, y, kT, ndf, Q, W, V, dp):
    K, v0, v1, v2, v3, vlog, lns = y
    odnf = 1.0 + 3.0 / ndf
    return [-2.0 * (v0 + vlog * odnf) * K,
            (2 * K + W * vlog * vlog - (ndf + 1) * kT) / Q[0] - v0 * v1,
            (Q[0] * v0 * v0 - kT) / Q[1] - v1 * v2,
            (Q[1] * v1 * v1 - kT) / Q[2] - v2 * v3,
            (Q[2] * v2 * v2 - kT) / Q[3],
            (2 * K * odnf + 3 * V * dp) / W - vlog * v0,
            -(v0 + vlog * odnf)]


def _chain_vs_ode(dt):
    from scipy.integrate import solve_ivp
    from oracle import oracle as O
    T, ndf, V, dp = 300.0, 600.0, 30.0, 0.4
    kT = BOLTZ * T
    tau, tau_b = 0.05, 0.2
    Q = np.array([ndf * kT * tau ** 2, kT * tau ** 2, kT * tau ** 2, kT * tau ** 2])
    W = (ndf + 3) * kT * tau_b ** 2
    K0 = 1.15 * 0.5 * ndf * kT
    vnh = np.array([3.0, -2.0, 1.5, 0.5])
    gnh = np.zeros(4)
    vlog, glog = C.c_double(0.8), C.c_double(0.0)
    y0 = [K0, *vnh, vlog.value, 0.0]
    p = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    scale = O.lib().orc_nh_chain(K0, T, dt, ndf, dp, V, W, p(vnh), p(gnh), p(Q), C.byref(vlog), C.byref(glog))
    sol = solve_ivp(_mtk_rhs, (0.0, dt / 2), y0, args=(kT, ndf, Q, W, V, dp), rtol=1e-13, atol=1e-15,
                    method="DOP853")
    yT = sol.y[:, -1]
    got = np.array([math.log(scale), *vnh, vlog.value])
    want = np.array([yT[6], *yT[1:5], yT[5]])
    return got, want


def test_nh_chain_integrates_the_mtk_equations_over_half_a_step(oracle_mod):
    """NHIntegrator::chain advances the 4-link chain coupled to the barostat
    by dt/2 with a 5 x 3 Suzuki-Yoshida factorisation.  Against a
    high-accuracy integration of the same ODEs it must agree to the
    factorisation's truncation error (a dropped coupling or a wrong factor
    shows up at ~1e-3), and that error must fall at least 8x when dt halves."""
    got, want = _chain_vs_ode(0.002)
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-3)
    assert err.max() < 1e-6
    got2, want2 = _chain_vs_ode(0.001)
    err2 = np.abs(got2 - want2) / np.maximum(np.abs(want2), 1e-3)
    assert err2.max() < err.max() / 8.0


# ---- induced dipoles -------------------------------------------------------------

def _two_sites(r, q1=0.0, pol2=0.000837):
    types = [AtomType(mass=1.0, charge=q1, eps=0.0, sigma=0.0, polar=1e-12),
             AtomType(mass=1.0, charge=0.0, eps=0.0, sigma=0.0, polar=pol2)]
    return _atoms(types, [[0.0, 0.0, 0.0], [r, 0.0, 0.0]], [0, 1], box=4.0)


def test_point_charge_polarises_a_site_by_alpha_q_over_r2(oracle_mod):
    """Undamped (a large): the field of charge q at distance r is q/r^2 along
    the separation, and the isolated site's dipole is alpha times it."""
    r, q, alpha = 0.3, 0.8, 0.000837
    s = _two_sites(r, q1=q, pol2=alpha)
    o = _oracle(s)
    zero3, zero6 = np.zeros((2, 3)), np.zeros((2, 6))
    e0, pi, _, _ = o.induced_dipoles(zero3, zero6, cutoff_ind=1.0, a_damp=1e9, damp_tij=True, do_min=True)
    assert e0[1, 0] == pytest.approx(q / r ** 2, rel=1e-12)
    assert e0[0] == pytest.approx([0.0, 0.0, 0.0], abs=1e-300)
    # site 0 has alpha = 1e-12: its induced dipole, and its back-reaction, vanish
    assert pi[1, 0] == pytest.approx(alpha * q / r ** 2, rel=1e-6)
    assert abs(pi[1, 1]) + abs(pi[1, 2]) < 1e-12 * abs(pi[1, 0])   # CG stops at |r| < 1e-14


def test_permanent_dipole_field_on_axis_is_2p_over_r3(oracle_mod):
    r, p = 0.35, 0.004
    s = _two_sites(r)
    o = _oracle(s)
    dip = np.array([[p, 0.0, 0.0], [0.0, 0.0, 0.0]])
    e0, _, _, _ = o.induced_dipoles(dip, np.zeros((2, 6)), cutoff_ind=1.0, a_damp=1e9, damp_tij=True,
                                    do_min=False)
    assert e0[1, 0] == pytest.approx(2 * p / r ** 3, rel=1e-12)
    # perpendicular dipole: field -p/r^3
    dip = np.array([[0.0, p, 0.0], [0.0, 0.0, 0.0]])
    e0, _, _, _ = o.induced_dipoles(dip, np.zeros((2, 6)), cutoff_ind=1.0, a_damp=1e9, damp_tij=True,
                                    do_min=False)
    assert e0[1, 1] == pytest.approx(-p / r ** 3, rel=1e-12)


def test_induced_dipoles_solve_the_dense_system(oracle_mod):
    """The CG result solves (diag(1/alpha) - T) mu = E0 with the Thole-damped
    dipole tensor, assembled here independently from its definition."""
    from swinmd_amd import systems as S
    s = S.water(8, density_mol=33.4, seed=4, types=S.AMOEBA_TYPES, r_oh=0.09572, hoh_deg=108.5)
    o = _oracle(s)
    dip, quad = o.local2global()
    a = 0.39
    e0, pi, _, _ = o.induced_dipoles(dip, quad, cutoff_ind=0.8, a_damp=a, damp_tij=True, do_min=True)
    n = s.n_atoms
    pol = o.polarizabilities()
    A = np.zeros((3 * n, 3 * n))
    for i in range(n):
        A[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.eye(3) / pol[i]
        for j in range(i + 1, n):
            d = s.pos[i] - s.pos[j]
            d -= s.box[0] * np.round(d / s.box[0])
            r = np.linalg.norm(d)
            if r > 0.8:
                continue
            u3 = r ** 3 / math.sqrt(pol[i] * pol[j])
            l3 = 1 - math.exp(-a * u3)
            l5 = 1 - (1 + a * u3) * math.exp(-a * u3)
            T = 3 * l5 * np.outer(d, d) / r ** 5 - l3 * np.eye(3) / r ** 3
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = -T
            A[3 * j:3 * j + 3, 3 * i:3 * i + 3] = -T
    mu = np.linalg.solve(A, e0.ravel())
    assert pi.ravel() == pytest.approx(mu, rel=1e-9, abs=1e-15)


def test_multipole_ewald_is_independent_of_alpha(oracle_mod):
    """With both sums converged, the Ewald split parameter cannot change the
    physics: the total multipole energy (real + k-space + self +
    intramolecular correction) and, with the intramolecular forces on, the
    total forces including the torque-derived ones must agree across alpha.
    A wrong coefficient in the B_l or D_l recursions, the structure factors or
    the self term, or a sign error in a torque, moves them apart."""
    from swinmd_amd import systems as S
    s = S.water(64, density_mol=33.4, seed=7, lattice=(4, 4, 4), r_oh=0.09572, hoh_deg=108.5, types=S.AMOEBA_TYPES)
    o = _oracle(s)
    dip, quad = o.local2global()
    rc = 0.49 * s.box[0]
    out = []
    for alpha in (6.5, 7.5):
        r = o.multipole(dip, quad, alpha=alpha, real_cut=rc, cutoff=rc, cut_buff=0.0, kcut=2 * alpha * 4.6,
                        use_constraint=False, surface=False, torques=True, use_cells=False)
        e = r["real_energy"] + r["long_energy"] + r["self_energy"] + r["intra_energy"]
        out.append((e, r["short_force"] + r["long_force"], r["real_energy"], r["long_energy"]))
    (e0, f0, re0, le0), (e1, f1, re1, le1) = out
    assert abs(le1 - le0) > 100.0 and abs(re1 - re0) > 1.0   # the split really moved
    assert e1 == pytest.approx(e0, rel=2e-6)
    assert np.abs(f1 - f0).max() <= 1e-6 * np.abs(f0).max()


@pytest.mark.parametrize("do_min", [False, True])
def test_induction_terms_are_independent_of_alpha(oracle_mod, do_min):
    """Fixed induced dipoles against the permanent multipoles under Ewald, with
    Thole damping switched off (a huge a makes lambda = 1): the total energy and
    forces must not depend on alpha.  The product-rule damping form only; the
    Tinker form subtracts the bare interaction of excluded pairs in real space
    on top of the intramolecular correction, so it is not split-invariant by
    construction and is pinned by the reference's own outputs instead."""
    from swinmd_amd import systems as S
    s = S.water(64, density_mol=33.4, seed=7, lattice=(4, 4, 4), r_oh=0.09572, hoh_deg=108.5, types=S.AMOEBA_TYPES)
    o = _oracle(s)
    dip, quad = o.local2global()
    rc = 0.49 * s.box[0]
    pind = np.random.default_rng(3).normal(scale=0.01, size=(s.n_atoms, 3))
    out = []
    for alpha in (6.5, 7.5):
        r = o.induction_terms(dip, quad, pind, alpha=alpha, real_cut=rc, kcut=2 * alpha * 4.6, a_damp=1e9,
                              use_constraint=False, surface=False, torques=True, damp_amoeba=False, do_min=do_min)
        e = r["real_energy"] + r["long_energy"] + r["self_energy"] + r["intra_energy"]
        out.append((e, r["short_force"] + r["long_force"], r["long_energy"]))
    (e0, f0, l0), (e1, f1, l1) = out
    assert abs(l1 - l0) > 10.0                      # the split really moved
    assert e1 == pytest.approx(e0, rel=1e-6)
    assert np.abs(f1 - f0).max() <= 5e-6 * np.abs(f0).max()


def _flexible_water():
    """The reference's example system (500 flexible waters) with its
    molecules distorted by a few degrees and picometres, so every term is
    away from its minimum."""
    import dataclasses
    import os

    from swinmd_amd import inputs
    from tests.helpers import GOLDEN
    cfg, _, s, _ = inputs.load_inputs(os.path.join(GOLDEN, "ref_test_inputs"))
    rng = np.random.default_rng(11)
    pos = s.pos + rng.normal(scale=0.004, size=s.pos.shape)
    return cfg, dataclasses.replace(s, pos=pos)


@pytest.mark.parametrize("pot", [1, 2])
def test_bonded_forces_are_minus_energy_gradient_and_internal(oracle_mod, pot):
    """Central differences of each bonded energy against its forces, for both
    bending potentials; the forces of each molecule sum to zero and exert no
    net torque (the terms depend on internal coordinates only)."""
    import dataclasses

    cfg, s = _flexible_water()
    terms = {"bond": lambda o: o.bond(pot), "angle": lambda o: o.angle(pot),
             "ub": lambda o: o.urey_bradley(cfg.get_kl_ub(), cfg.get_l0_ub())}
    h = 1e-6
    for name, fn in terms.items():
        r0 = fn(_oracle(s))
        f = r0["force"]
        for atom in (0, 1, 2, 700, 1499):
            for d in range(3):
                pp, pm = s.pos.copy(), s.pos.copy()
                pp[atom, d] += h
                pm[atom, d] -= h
                ep = fn(_oracle(dataclasses.replace(s, pos=pp)))["energy"]
                em = fn(_oracle(dataclasses.replace(s, pos=pm)))["energy"]
                fd = -(ep - em) / (2 * h)
                assert fd == pytest.approx(f[atom, d], rel=1e-5, abs=1e-4 * np.abs(f).max()), (name, atom, d)
        for m in range(0, s.n_mols, 50):
            a = slice(3 * m, 3 * m + 3)
            assert np.abs(f[a].sum(axis=0)).max() < 1e-9 * np.abs(f).max()
            c = s.pos[a].mean(axis=0)
            tq = np.cross(s.pos[a] - c, f[a]).sum(axis=0)
            assert np.abs(tq).max() < 1e-9 * np.abs(f).max() * 0.1


def test_bond_and_ub_volume_derivatives(oracle_mod):
    """Lustig volume derivatives of the stretching terms: with every length
    scaled by (1+e) (positions and box), dE/de = sum r U'(r), the first
    derivative the terms store; for the harmonic bond, d2E/de2 = sum r^2 k =
    deriv2 + 2 deriv by the reference's definition of deriv2."""
    import dataclasses

    cfg, s = _flexible_water()
    e = 1e-5

    def scaled(f):
        return dataclasses.replace(s, pos=s.pos * (1 + f), density=s.density / (1 + f) ** 3)

    for name, fn in (("bond", lambda o: o.bond(1)), ("ub", lambda o: o.urey_bradley(cfg.get_kl_ub(), cfg.get_l0_ub()))):
        r0 = fn(_oracle(s))
        ep, em = fn(_oracle(scaled(e)))["energy"], fn(_oracle(scaled(-e)))["energy"]
        assert (ep - em) / (2 * e) == pytest.approx(r0["deriv"], rel=1e-6), name
        if name == "bond":
            d2 = (ep - 2 * r0["energy"] + em) / (e * e)
            assert d2 == pytest.approx(r0["deriv2"] + 2 * r0["deriv"], rel=1e-4)


# ---- LJ long-range correction -----------------------------------------------------

def test_lj_long_range_correction_is_the_dispersion_tail(oracle_mod):
    """orc_lj_lrc (LJForce::long_range_correct, src/forces/LJForce.cpp:260-370)
    against the tail integrals of the dispersion term -C6/r^6 beyond rc.  The
    reference's own output cannot pin it (its csix starts uninitialised,
    :268), so the integrals are evaluated here by quadrature:
      E(V, rc)  = N^2/(2V) int_rc^inf 4 pi r^2 (-C6/r^6) dr
      eLrc      = E(V0, rc0)
      dUlrc     = dE/dV and d2Ulrc = d2E/dV2 at fixed scaled coordinates,
                  i.e. with rc scaling with the box, rc = rc0 (V/V0)^(1/3)
      vLrc      = each diagonal component of the pair virial sum
                  sum f_i (x) (r_j - r_i) (the reference's convention,
                  LJForce.cpp:211-220), a pair at r adding -F(r) r / 3 with
                  F = -du/dr, over the same tail
    with C6 the mean 4 eps sigma^6 over the non-excluded pairs, counted by
    brute force.  Two LJ types with both epsilons nonzero make the mixed
    pairs and the intramolecular exclusions matter."""
    from dataclasses import replace

    from scipy.integrate import quad

    from swinmd_amd import systems as S

    types = [replace(S.SPCE_TYPES[0], eps=0.0565, sigma=0.2655), S.SPCE_TYPES[1]]
    s = S.water(64, density_mol=33.4, seed=3, types=types)
    o = _oracle(s)
    rc0 = 0.7
    e_lrc, v_lrc, du_lrc, d2u_lrc = o.lj_lrc(rc0)
    n = s.n_atoms
    sig2, eps4 = s.lj_table()
    nt = len(types)
    excl_off, excl = s.exclusions()
    excluded = {(i, int(j)) for i in range(n) for j in excl[excl_off[i]:excl_off[i + 1]]}
    c6 = []
    for i in range(n):
        for j in range(i + 1, n):
            if (i, j) in excluded or (j, i) in excluded:
                continue
            t = s.type[i] * nt + s.type[j]
            c6.append(eps4[t] * sig2[t] ** 3)
    c6 = float(np.mean(c6))
    V0 = s.volume

    def tail(V, integrand):
        rc = rc0 * (V / V0) ** (1.0 / 3.0)
        val, _ = quad(integrand, rc, np.inf, epsabs=0.0, epsrel=1e-13, limit=200)
        return n * n / (2.0 * V) * val

    def energy(V):
        return tail(V, lambda r: 4.0 * math.pi * r * r * (-c6 / r ** 6))

    assert e_lrc == pytest.approx(energy(V0), rel=1e-11)
    h = 1e-3 * V0
    d1 = (energy(V0 - 2 * h) - 8 * energy(V0 - h) + 8 * energy(V0 + h) - energy(V0 + 2 * h)) / (12 * h)
    d2 = (-energy(V0 - 2 * h) + 16 * energy(V0 - h) - 30 * energy(V0) + 16 * energy(V0 + h)
          - energy(V0 + 2 * h)) / (12 * h * h)
    assert du_lrc == pytest.approx(d1, rel=1e-8)
    assert d2u_lrc == pytest.approx(d2, rel=1e-6)
    # F(r) = -du/dr = -6 C6 / r^7; a pair adds -F r / 3 to each diagonal component
    vir = tail(V0, lambda r: 4.0 * math.pi * r * r * (6.0 * c6 / r ** 6) / 3.0)
    assert v_lrc == pytest.approx(vir, rel=1e-11)
    # a plausible slip: the derivative at fixed rc differs by a factor of two
    assert abs(du_lrc - (-energy(V0) / V0)) > 0.1 * abs(du_lrc)
