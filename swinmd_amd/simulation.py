"""The force terms a SimConfiguration asks for and the order they run in.

``ForceField`` mirrors Integrator::initialise (src/Integrator.cpp:111-197),
which creates the terms from the config's flags, and the force half of
LFIntegrator::run (src/integrators/LFIntegrator.cpp:382-557), which evaluates
them: the vdW term, then either Ewald or the permanent multipoles (with
induction), the long-range virial added, then the bond, angle and
Urey-Bradley terms, and the energy sums of the step.  Every term runs on the
GPU through the C ABI; this module only wires them together.

``buff_long_range_correct`` is BuffvdWForce::long_range_correct
(src/forces/BuffvdWForce.cpp:467-601), a host-side integral evaluated once
per volume.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .forces import (AngleForce, BondForce, BuffvdWForce, Context, EwaldForce, Induction, LJForce, Multipole,
                     UreyBradleyForce, lj_long_range_correct)
from .systems import PI

__all__ = ["ForceField", "buff_long_range_correct", "buff_pair_table"]


def buff_pair_table(system):
    """BuffvdWForce::set_buff_param (src/forces/BuffvdWForce.cpp:392-433):
    per type pair the combined Rij (cubic mean of sigma) and eps (HHG)."""
    nt = len(system.types)
    sig = np.zeros((nt, nt))
    eps = np.zeros((nt, nt))
    for i, a in enumerate(system.types):
        for j, b in enumerate(system.types):
            if a.eps == 0 or b.eps == 0:
                sig[i, j], eps[i, j] = 0.0000001, 0.0
            else:
                si2, sj2 = a.sigma * a.sigma, b.sigma * b.sigma
                es = np.sqrt(a.eps) + np.sqrt(b.eps)
                sig[i, j] = (si2 * a.sigma + sj2 * b.sigma) / (si2 + sj2)
                eps[i, j] = (4.0 * a.eps * b.eps) / (es * es)
    return sig, eps


def _buff_pot(r, sigma, eps, n, m, delta, gamma):
    """BuffvdWForce::potFunc (src/forces/BuffvdWForce.cpp:440-465), vectorised over r."""
    if eps == 0:
        z = np.zeros_like(r)
        return z, z
    invs = 1.0 / sigma
    rho = r * invs
    rho_m = rho ** m
    rho_m1 = rho ** (m - 1.0)
    expnm = n - m
    par1 = (1.0 + delta) / (rho + delta)
    par2 = (1.0 + gamma) / (rho_m + gamma)
    p1n = par1 ** expnm
    e = eps * p1n * (par2 - 2.0)
    f = -eps * expnm * p1n * (par2 - 2.0) * invs / (rho + delta) - eps * p1n * par2 * m * rho_m1 * invs / (rho_m + gamma)
    return e, f


def buff_long_range_correct(system, cutoff, *, n=14.0, m=7.0, delta=0.07, gamma=0.12, box=None):
    """(eLrc, vLrc, dUlrc, d2Ulrc): the trapezoidal integral of r^2 U(r) and
    r^3 dU/dr from the cutoff to 5000 nm over 2^20 intervals per type pair,
    averaged over the pairs of the system and scaled as the reference does
    (2 pi N^2 / V and 2 pi N^2 / (3 V^2); vLrc = dUlrc V)."""
    box = system.box if box is None else np.asarray(box)
    vol = float(box[0] * box[1] * box[2])
    invvol = 1.0 / vol
    nt = len(system.types)
    count = np.bincount(system.type, minlength=nt)
    sig, eps = buff_pair_table(system)
    level = 20
    nsteps = 2 ** level - 1
    maxright = 5000.0
    h = (maxright - cutoff) / 2.0 ** level
    r_in = cutoff + np.arange(1, nsteps + 1, dtype=np.float64) * h
    Ulrc = Plrc = 0.0
    npair = 0
    for i in range(nt):
        for j in range(i, nt):
            npair_ij = int(count[i] * count[j]) if i != j else int(count[i] * (count[i] - 1) * 0.5)
            npair += npair_ij
            sumU = sumP = 0.0
            for r, w in ((np.array([cutoff]), 0.5 * h), (r_in, h), (np.array([maxright]), 0.5 * h)):
                e, f = _buff_pot(r, sig[i, j], eps[i, j], n, m, delta, gamma)
                sumU += float(np.sum(w * r * r * e))
                sumP += float(np.sum(w * r * r * r * f))
            Ulrc += npair_ij * sumU
            Plrc += npair_ij * sumP
    if npair <= 0:
        Ulrc = Plrc = 0.0
    else:
        Ulrc /= npair
        Plrc /= npair
    N = system.n_atoms
    e_lrc = 2.0 * PI * N * N * invvol * Ulrc
    du_lrc = 2.0 * PI * N * N * invvol * invvol * Plrc / 3.0
    return e_lrc, du_lrc * vol, du_lrc, 0.0


class ForceField:
    """The terms Integrator::initialise creates for ``cfg`` on ``ctx``, run
    in LFIntegrator::run's order by :meth:`compute`.  The pair list is built
    here as Ensemble::set_pairlist does at start-up (cutoff + cutBuff)."""

    def __init__(self, ctx: Context, cfg, prec=L.PREC_FP64):
        self.ctx = ctx
        self.cfg = cfg
        s = ctx.system
        constrained = cfg.is_constraint_on()
        cutoff = cfg.get_cutoff()
        ctx.set_pairlist(cutoff, cfg.get_cutBuff(), cfg.use_cell_list())
        ctx.compute_mol_centers()
        lustig = cfg.compute_lustig()
        common = dict(cutoff=cutoff, switch_on=cfg.is_switch_on(), switch_dist=cfg.get_switch_distance(),
                      use_constraint=constrained, lustig=lustig, prec=prec)
        if cfg.compute_buffamoeba():
            buff = dict(n_buff=cfg.get_nBuff_param(), m_buff=cfg.get_mBuff_param(), delta=cfg.get_deltaBuff_param(),
                        gamma=cfg.get_gammaBuff_param(), hfactor=cfg.get_hfactor_reduction())
            lrc = (0.0, 0.0, 0.0, 0.0)
            if cfg.use_long_correct():
                lrc = buff_long_range_correct(s, cutoff, n=buff["n_buff"], m=buff["m_buff"], delta=buff["delta"],
                                              gamma=buff["gamma"], box=ctx.box)
            self.vdw = BuffvdWForce(ctx, lrc=lrc, **common, **buff)
        else:
            lrc = lj_long_range_correct(ctx, cutoff) if cfg.use_long_correct() else (0.0, 0.0, 0.0, 0.0)
            self.vdw = LJForce(ctx, lrc=lrc, coulomb=cfg.compute_coulomb(), **common)
        self.ewald = self.multipole = self.induction = None
        elec = dict(alpha=cfg.get_alpha(), real_cut=cfg.get_cutoffEw(), kcut=cfg.get_k_cutoff(),
                    use_constraint=constrained, surface=cfg.compute_surf_correct())
        if cfg.compute_ewald_force():
            self.ewald = EwaldForce(ctx, prec=prec, **elec)
        elif cfg.compute_multipol():
            self.multipole = Multipole(ctx, torques=cfg.compute_torques(), **elec)
            if cfg.compute_induction():
                self.induction = Induction(ctx, cutoff_ind=cfg.get_cutoffInd(), a_damp=cfg.get_adamping(),
                                           damp_tij=cfg.use_damping_induction(),
                                           do_min=cfg.do_induction_minimization())
                self.induction.setup_terms(torques=cfg.compute_torques(), damp_amoeba=cfg.use_damping_indAmoeba(),
                                           **elec)
        elif cfg.compute_wolf():
            raise L.SmdError(L.SMD_ENOTSUP, "WolfForce is not implemented")
        kw = dict(bending_type=cfg.get_bending_type(), kl_ub=cfg.get_kl_ub(), l0_ub=cfg.get_l0_ub())
        self.bond = BondForce(ctx, **kw) if (not constrained and s.bonds) else None
        self.angle = AngleForce(ctx, **kw) if (not constrained and s.angles) else None
        if cfg.compute_harmonic_ub() and constrained:
            # Integrator::initialise stops here (ERRORMSG)
            raise L.SmdError(L.SMD_EINVAL, "a flexible Urey-Bradley term with constrained bonds")
        self.ub = UreyBradleyForce(ctx, **kw) if cfg.compute_harmonic_ub() else None

    def compute(self):
        """Zeroes the forces and evaluates every term into the device
        accumulator (read it with ``ctx.forces()``); returns the step's
        energies under Ensemble::myEnergy's names, totEnergy and inducEnergy
        summed as LFIntegrator::run does (:553-557), and Ensemble::virial
        after the step's terms ("virial", as the reference holds it, and
        "virial_full", with the long-range ZZ component the reference drops)."""
        ctx = self.ctx
        ctx.zero_forces()
        E = dict.fromkeys(("ljEnergy", "realEnergy", "longEnergy", "correctEnergy", "molCorrectEnergy",
                           "surfCorrectEnergy", "realPolEnergy", "longPolEnergy", "correctPolEnergy",
                           "molCorrectPolEnergy", "surfacePolEnergy", "bondEnergy", "angleEnergy", "ubEnergy",
                           "TBEnergy"), 0.0)
        vir = np.zeros(9)
        long_vir = np.zeros(9)     # Ensemble::longVirial
        r = self.vdw.compute()
        E["ljEnergy"] = r["lj_energy"]
        vir -= r["virial"]
        if self.ewald is not None:
            r = self.ewald.compute()
            E["realEnergy"], E["correctEnergy"] = r["real_energy"], r["self_energy"] + r["charged_energy"]
            E["molCorrectEnergy"], E["surfCorrectEnergy"] = r["intra_energy"], r["surf_energy"]
            vir -= r["virial"]
            r = self.ewald.compute_long()
            E["longEnergy"] = r["energy"]
            long_vir = r["virial"].copy()
        elif self.multipole is not None:
            r = self.multipole.compute()
            E["realEnergy"], E["correctEnergy"] = r["real_energy"], r["self_energy"] + r["charged_energy"]
            E["molCorrectEnergy"], E["surfCorrectEnergy"] = r["intra_energy"], r["surf_energy"]
            vir -= r["virial"]
            r = self.multipole.compute_long()
            E["longEnergy"] = r["energy"]
            long_vir = r["virial"].copy()
            if self.induction is not None:
                it, conv, _ = self.induction.induce()
                if not conv:
                    raise L.SmdError(L.SMD_ENOCONV, f"induced dipoles did not converge in {it} iterations")
                r = self.induction.compute()
                E["realPolEnergy"], E["correctPolEnergy"] = r["real_energy"], r["self_energy"]
                E["molCorrectPolEnergy"], E["surfacePolEnergy"] = r["intra_energy"], r["surf_energy"]
                vir -= r["virial"]
                r = self.induction.compute_long()
                E["longPolEnergy"] = r["energy"]
                long_vir += r["virial"]
        # The compute_long methods copy components 0..7 of their long-range
        # virial into Ensemble::longVirial and never ZZ (loops to i < ZZ,
        # EwaldForce.cpp:194, Multipole.cpp:238, Induction.cpp:588), so the
        # virial the reference goes on with has no long-range ZZ part;
        # "virial" reproduces that and "virial_full" keeps it.
        full = vir + long_vir
        vir = vir + np.where(np.arange(9) < 8, long_vir, 0.0)
        for name, term in (("bondEnergy", self.bond), ("angleEnergy", self.angle), ("ubEnergy", self.ub)):
            if term is not None:
                r = term.compute()
                E[name] = r["energy"]
                vir -= r["virial"]
                full -= r["virial"]
        E["inducEnergy"] = (E["realPolEnergy"] + E["longPolEnergy"] + E["correctPolEnergy"]
                            + E["molCorrectPolEnergy"] + E["surfacePolEnergy"])
        E["totEnergy"] = (E["ljEnergy"] + E["realEnergy"] + E["longEnergy"] + E["molCorrectEnergy"]
                          + E["surfCorrectEnergy"] + E["correctEnergy"] + E["angleEnergy"] + E["bondEnergy"]
                          + E["ubEnergy"] + E["inducEnergy"] + E["TBEnergy"])
        E["virial"] = vir
        E["virial_full"] = full
        return E
