"""Host-side mirror of SwinMD's nonbonded plugin interface over the C ABI.

Class names and method meanings follow the reference:

* :class:`Context` holds what ``Ensemble`` holds for the path (topology, box,
  positions, molecule centres, the pair list) on one GPU.
* :class:`LJForce`, :class:`MieForce`, :class:`BuckinghamForce`,
  :class:`BuffvdWForce` mirror ``Force::compute()`` (src/Force.h:59) of
  src/forces/{LJ,Mie,Buckingham,BuffvdW}Force.cpp.
* :class:`EwaldForce` mirrors ``EwaldForce::compute`` / ``compute_long``
  (src/forces/EwaldForce.cpp:128-196).
* :class:`BondForce`, :class:`AngleForce`, :class:`UreyBradleyForce` mirror
  the bonded terms of flexible molecules (src/forces/{Bond,Angle,
  UreyBradley}Force.cpp).
* :class:`LFIntegrator` mirrors the SHAKE + Nose-Hoover leap-frog step of
  src/integrators/LFIntegrator.cpp.

Errors follow the C ABI: a negative status becomes :class:`SmdError` (the
reference would exit through ERRORMSG).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .systems import PI


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _fptr(f_accum):
    """Host force accumulator (n,3) float64 C-contiguous, or NULL."""
    if f_accum is None:
        return None
    assert f_accum.dtype == np.float64 and f_accum.flags.c_contiguous
    return _ptr(f_accum)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


class Context:
    """Device-resident ``Ensemble`` view for one :class:`swinmd_amd.systems.System`."""

    def __init__(self, system, device: int = 0, pos=None, box=None):
        self.lib = L.load()
        s = system
        self.system = s
        self.n = s.n_atoms
        self.nmol = s.n_mols
        self._type = np.ascontiguousarray(s.type, dtype=np.int32)
        self._mol = np.ascontiguousarray(s.mol, dtype=np.int32)
        self._q = _f64(s.scaled_charge)
        self._mass = _f64(s.mass)
        self._eoff, excl = s.exclusions()
        self._excl = excl if excl.size else np.zeros(1, dtype=np.int32)
        t = L.Topology(self.n, self.nmol, len(s.types), _ptr(self._type), _ptr(self._mol), _ptr(self._q),
                       _ptr(self._mass), _ptr(self._eoff), _ptr(self._excl))
        h = C.c_void_p()
        rc = self.lib.smd_ctx_create(C.byref(h), C.byref(t), device)
        if rc != 0:
            raise L.SmdError(rc, f"smd_ctx_create: {self.lib.smd_strerror(rc).decode()}")
        self.h = h
        self.box = _f64(s.box if box is None else box)
        self.check(self.lib.smd_set_box(self.h, _ptr(self.box)))
        self.set_positions(s.pos if pos is None else pos)

    # ---- plumbing -------------------------------------------------------
    def check(self, rc):
        if rc != 0:
            msg = self.lib.smd_last_error(self.h).decode() if self.h else ""
            raise L.SmdError(rc, msg or self.lib.smd_strerror(rc).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.smd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _vec_in(self, fn, a, n):
        a = _f64(a, (n, 3))
        self.check(fn(self.h, _ptr(a), n))

    def _vec_out(self, fn, n):
        out = np.zeros((n, 3))
        self.check(fn(self.h, _ptr(out), n))
        return out

    # ---- state ----------------------------------------------------------
    def set_positions(self, pos):
        self._vec_in(self.lib.smd_set_positions, pos, self.n)

    def positions(self):
        return self._vec_out(self.lib.smd_get_positions, self.n)

    def set_velocities(self, v):
        self._vec_in(self.lib.smd_set_velocities, v, self.n)

    def velocities(self):
        return self._vec_out(self.lib.smd_get_velocities, self.n)

    def set_mol_centers(self, com):
        self._vec_in(self.lib.smd_set_mol_centers, com, self.nmol)

    def mol_centers(self):
        return self._vec_out(self.lib.smd_get_mol_centers, self.nmol)

    def compute_mol_centers(self):
        self.check(self.lib.smd_compute_mol_centers(self.h))

    def zero_forces(self):
        self.check(self.lib.smd_zero_forces(self.h))

    def forces(self):
        return self._vec_out(self.lib.smd_get_forces, self.n)

    # ---- (a) pair list ----------------------------------------------------
    def set_pairlist(self, cutoff, cut_buff, use_cells=True):
        """Ensemble::set_pairlist (src/Ensemble.cpp:454-473)."""
        self.check(self.lib.smd_build_pairlist(self.h, cutoff, cut_buff, int(use_cells)))

    def pairlist_info(self):
        n, mx, uc = C.c_int64(), C.c_int32(), C.c_int32()
        self.check(self.lib.smd_pairlist_info(self.h, C.byref(n), C.byref(mx), C.byref(uc)))
        return n.value, mx.value, bool(uc.value)

    def pairlist(self):
        """Per-atom rows (Atom::myPairList) as CSR: offsets [n+1], neighbours."""
        total, _, _ = self.pairlist_info()
        off = np.zeros(self.n + 1, dtype=np.int32)
        nb = np.zeros(max(total, 1), dtype=np.int32)
        self.check(self.lib.smd_get_pairlist(self.h, _ptr(off), _ptr(nb), nb.size))
        return off, nb[:total]

    def set_pairlist_capacity(self, cap):
        self.check(self.lib.smd_set_pairlist_capacity(self.h, cap))

    # ---- profiling --------------------------------------------------------
    def profile(self, on=True):
        self.check(self.lib.smd_profile_enable(self.h, int(on)))

    def profile_get(self, family):
        ms, n = C.c_double(), C.c_int64()
        self.check(self.lib.smd_profile_get(self.h, family.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def time_family(self, family, reps=50):
        """Average microseconds of one launch of an MD-step kernel family,
        launched back to back (the MD state is left unchanged)."""
        us = C.c_double()
        self.check(self.lib.smd_md_time_family(self.h, family.encode(), int(reps), C.byref(us)))
        return us.value

    def sync(self):
        self.check(self.lib.smd_sync(self.h))

    def set_tuning(self, name, value):
        """Launch-shape knob (e.g. "pair_lanes"); see smd_set_tuning."""
        self.check(self.lib.smd_set_tuning(self.h, name.encode(), int(value)))


class _VdwForce:
    KIND = L.VDW_LJ

    def __init__(self, ctx: Context, *, cutoff, switch_on=False, switch_dist=0.0, use_constraint=True,
                 lustig=False, coulomb=False, lrc=(0.0, 0.0, 0.0, 0.0), prec=L.PREC_FP64, **kind_params):
        self.ctx = ctx
        self.prec = prec
        s = ctx.system
        p = L.VdwParams()
        p.kind = self.KIND
        p.cutoff, p.switch_dist = cutoff, switch_dist
        p.switch_on, p.use_constraint, p.lustig, p.coulomb = int(switch_on), int(use_constraint), int(lustig), int(coulomb)
        self._keep = []
        sig2, eps4 = s.lj_table()
        self._keep += [sig2, eps4]
        p.sigma2, p.eps4 = _ptr(sig2), _ptr(eps4)
        p.n_mie, p.m_mie = kind_params.get("n_mie", 12.0), kind_params.get("m_mie", 6.0)
        if "buck" in kind_params:
            a, b, c = [_f64(x) for x in kind_params["buck"]]
            self._keep += [a, b, c]
            p.buck_a, p.buck_b, p.buck_c = _ptr(a), _ptr(b), _ptr(c)
        p.co2_bond = kind_params.get("co2_bond", 0.0)
        ts = _f64([t.sigma for t in s.types])
        te = _f64([t.eps for t in s.types])
        self._keep += [ts, te]
        p.type_sigma, p.type_eps = _ptr(ts), _ptr(te)
        p.n_buff, p.m_buff = kind_params.get("n_buff", 14.0), kind_params.get("m_buff", 7.0)
        p.delta, p.gamma = kind_params.get("delta", 0.07), kind_params.get("gamma", 0.12)
        p.hfactor = kind_params.get("hfactor", 0.91)
        p.e_lrc, p.v_lrc, p.du_lrc, p.d2u_lrc = lrc
        self.params = p
        self.energy = 0.0
        self.virial = np.zeros(9)

    def compute(self, f_accum=None):
        """Force::compute(): adds forces into f_accum (host, (n,3)) when given,
        else into the device accumulator; returns the reference's outputs."""
        out = L.VdwOut()
        fp = None
        if f_accum is not None:
            assert f_accum.dtype == np.float64 and f_accum.flags.c_contiguous and f_accum.size == 3 * self.ctx.n
            fp = _ptr(f_accum)
        self.ctx.check(self.ctx.lib.smd_vdw_compute(self.ctx.h, C.byref(self.params), self.prec, fp, C.byref(out)))
        self.energy = out.energy
        self.virial = np.array(out.virial[:])
        return {"energy": out.energy, "lj_energy": out.lj_energy, "virial": self.virial.copy(),
                "ulus": out.ulus, "dudv": out.dudv, "d2udv": out.d2udv}

    def get_energy(self):
        return self.energy

    def get_virial(self):
        return self.virial.copy()


class LJForce(_VdwForce):
    KIND = L.VDW_LJ


class MieForce(_VdwForce):
    KIND = L.VDW_MIE


class BuckinghamForce(_VdwForce):
    KIND = L.VDW_BUCK


class BuffvdWForce(_VdwForce):
    KIND = L.VDW_BUFF147


class _BondedTerm:
    """Common set-up of BondForce, AngleForce and UreyBradleyForce: the
    topology's bonds and angles with their type tables (Parameters
    bondParams/angleParams), the bending potential (config bendingType: 1
    harmonic, 2 AMOEBA) and the Urey-Bradley kl, l0 (config klUB, l0UB)."""

    FN = None

    def __init__(self, ctx: Context, *, bending_type=1, kl_ub=0.0, l0_ub=0.0):
        self.ctx = ctx
        key = (int(bending_type), float(kl_ub), float(l0_ub))
        if getattr(ctx, "_bonded_key", None) != key:
            s = ctx.system
            bonds = np.ascontiguousarray(np.array(s.bonds, dtype=np.int32).reshape(-1, 3))
            angles = np.ascontiguousarray(np.array(s.angles, dtype=np.int32).reshape(-1, 4))
            r0 = _f64([b[2] for b in s.bond_params])
            kb = _f64([b[3] for b in s.bond_params])
            th = _f64([a[3] * 2 * PI / 360.0 for a in s.angle_params])   # AngleParam theta0 (Parameters.cpp:186)
            ka = _f64([a[4] for a in s.angle_params])
            keep = [bonds, angles, r0, kb, th, ka]
            p = L.BondedParams(int(bending_type), len(r0), len(th), _ptr(r0), _ptr(kb), _ptr(th), _ptr(ka),
                               len(bonds), len(angles), _ptr(bonds), _ptr(angles), float(kl_ub), float(l0_ub))
            ctx.check(ctx.lib.smd_bonded_setup(ctx.h, C.byref(p)))
            ctx._bonded_key = key
            ctx._bonded_keep = keep
        self.energy = 0.0
        self.virial = np.zeros(9)

    def compute(self, f_accum=None):
        """Force::compute(): adds forces into f_accum (host, (n,3)) when given,
        else into the device accumulator; returns energy, the Lustig
        derivatives and the term's virial."""
        out = L.BondedOut()
        self.ctx.check(getattr(self.ctx.lib, self.FN)(self.ctx.h, _fptr(f_accum), C.byref(out)))
        self.energy = out.energy
        self.virial = np.array(out.virial[:])
        return {"energy": out.energy, "deriv": out.deriv, "deriv2": out.deriv2, "virial": self.virial.copy()}

    def get_energy(self):
        return self.energy

    def get_virial(self):
        return self.virial.copy()


class BondForce(_BondedTerm):
    """BondForce (src/forces/BondForce.cpp:46-122)."""
    FN = "smd_bond_compute"


class AngleForce(_BondedTerm):
    """AngleForce (src/forces/AngleForce.cpp:35-126)."""
    FN = "smd_angle_compute"


class UreyBradleyForce(_BondedTerm):
    """UreyBradleyForce (src/forces/UreyBradleyForce.cpp:41-101)."""
    FN = "smd_ub_compute"


def lj_long_range_correct(ctx: Context, cutoff):
    """LJForce::long_range_correct: (eLrc, vLrc, dUlrc, d2Ulrc)."""
    sig2, eps4 = ctx.system.lj_table()
    out = np.zeros(4)
    ctx.check(ctx.lib.smd_lj_long_range_correct(ctx.h, _ptr(sig2), _ptr(eps4), cutoff, _ptr(out)))
    return tuple(out)


class EwaldForce:
    """EwaldForce (src/forces/EwaldForce.cpp)."""

    def __init__(self, ctx: Context, *, alpha, real_cut, kcut, use_constraint=True, surface=False,
                 prec=L.PREC_FP64):
        self.ctx = ctx
        self.prec = prec
        self.params = L.EwaldParams(alpha, real_cut, kcut, int(use_constraint), int(surface))
        k = C.c_int32()
        ctx.check(ctx.lib.smd_ewald_setup(ctx.h, C.byref(self.params), C.byref(k)))
        self.n_kvec = k.value

    def _fp(self, f_accum):
        if f_accum is None:
            return None
        assert f_accum.dtype == np.float64 and f_accum.flags.c_contiguous and f_accum.size == 3 * self.ctx.n
        return _ptr(f_accum)

    def compute(self, f_accum=None):
        o = L.EwaldOut()
        self.ctx.check(self.ctx.lib.smd_ewald_compute(self.ctx.h, self.prec, self._fp(f_accum), C.byref(o)))
        d = {name: getattr(o, name) for name, _ in L.EwaldOut._fields_ if name != "virial"}
        d["virial"] = np.array(o.virial[:])
        return d

    def compute_long(self, f_accum=None):
        o = L.EwaldLongOut()
        self.ctx.check(self.ctx.lib.smd_ewald_compute_long(self.ctx.h, self.prec, self._fp(f_accum), C.byref(o)))
        return {"energy": o.energy, "deriv": o.deriv, "deriv2": o.deriv2, "virial": np.array(o.virial[:])}


class LFIntegrator:
    """Leap-frog + SHAKE + Nose-Hoover (src/integrators/LFIntegrator.cpp)."""

    def __init__(self, ctx: Context, *, dt, temperature, coupl=0.1, tol=1e-8, max_count=10000,
                 nose_hoover=True, atom_thermo=True):
        self.ctx = ctx
        self.params = L.IntegParams(dt, temperature, coupl, tol, max_count, int(nose_hoover), int(atom_thermo))
        if ctx.system.atoms_per_mol == 3 and ctx.system.bond_params:
            b2 = _f64(ctx.system.constraint_lengths2())
            ctx.check(ctx.lib.smd_set_constraints(ctx.h, _ptr(b2), ctx.nmol))

    def set_vt_minus(self, v):
        self.ctx._vec_in(self.ctx.lib.smd_set_vt_minus, v, self.ctx.n)

    def vt_minus(self):
        return self.ctx._vec_out(self.ctx.lib.smd_get_vt_minus, self.ctx.n)

    def set_mol_momenta(self, p):
        self.ctx._vec_in(self.ctx.lib.smd_set_mol_momenta, p, self.ctx.nmol)

    def shake(self, fnh=0.0):
        """LFIntegrator::shake with the device forces; returns the shake virial."""
        vir = np.zeros(9)
        self.ctx.check(self.ctx.lib.smd_shake(self.ctx.h, C.byref(self.params), fnh, _ptr(vir)))
        return vir

    def lf_atoms(self, fnh=0.0):
        self.ctx.check(self.ctx.lib.smd_lf_atoms(self.ctx.h, C.byref(self.params), fnh))

    def mol_kinetics(self):
        out = np.zeros(2)
        self.ctx.check(self.ctx.lib.smd_mol_kinetics(self.ctx.h, _ptr(out)))
        return out   # (2*atomKinEnergy, 2*molKinEnergy)


class WaterMD:
    """Whole LFIntegrator::run step on the device: rigid point-charge water
    (LJ + Ewald, SHAKE), or with constraint=False and ewald=False unconstrained
    LJ atoms (the per-atom update with the self-consistent Nose-Hoover loop,
    src/integrators/LFIntegrator.cpp:587-643)."""

    def __init__(self, ctx: Context, *, cutoff, cut_buff, dt, temperature, alpha=3.5, kcut=10.0, coupl=0.1,
                 tol=1e-8, max_count=10000, use_cells=True, long_steps=1, lustig=True, prec=L.PREC_FP64,
                 atom_thermo=True, surface=False, use_graph=True, total_steps=0, constraint=True, ewald=True,
                 switch_on=False):
        self.ctx = ctx
        if constraint:
            b2 = _f64(ctx.system.constraint_lengths2())
            ctx.check(ctx.lib.smd_set_constraints(ctx.h, _ptr(b2), ctx.nmol))
        self._sig2, self._eps4 = ctx.system.lj_table()
        p = L.MdParams()
        p.cutoff, p.cut_buff = cutoff, cut_buff
        p.constraint_mode, p.ewald_on, p.switch_on = (2 if constraint else 0), int(ewald), int(switch_on)
        p.ewald = L.EwaldParams(alpha, cutoff, kcut, int(constraint), int(surface))
        p.integ = L.IntegParams(dt, temperature, coupl, tol, max_count, 1, int(atom_thermo))
        p.use_cells, p.long_steps, p.lustig, p.prec = int(use_cells), long_steps, int(lustig), prec
        p.sigma2, p.eps4 = _ptr(self._sig2), _ptr(self._eps4)
        p.use_graph = int(use_graph)
        p.total_steps = int(total_steps)
        self.params = p

    def init(self):
        """LFIntegrator::start: forces at t and vtMinus."""
        self.ctx.check(self.ctx.lib.smd_md_init(self.ctx.h, C.byref(self.params)))

    def run(self, nsteps):
        r = L.MdReport()
        self.ctx.check(self.ctx.lib.smd_md_run(self.ctx.h, nsteps, C.byref(r)))
        d = {name: getattr(r, name) for name, _ in L.MdReport._fields_ if name != "virial"}
        d["virial"] = np.array(r.virial[:])
        return d

    def checkpoint(self, state_path, vt_minus_path=None, step=0, precision=17):
        """Integrator::write_state and NHIntegrator::write_vtMinus for the
        current state (lossless with precision 17); returns the thermostat
        friction (fnh, fnh_minus), which the reference's files do not keep."""
        from .outputs import write_state, write_vt_minus
        ctx = self.ctx
        real = np.zeros((ctx.n, 3))
        ctx.check(ctx.lib.smd_get_real_positions(ctx.h, _ptr(real), ctx.n))
        write_state(state_path, step, ctx.positions(), real, ctx.velocities(), precision)
        if vt_minus_path is not None:
            vtm = np.zeros((ctx.n, 3))
            ctx.check(ctx.lib.smd_get_vt_minus(ctx.h, _ptr(vtm), ctx.n))
            write_vt_minus(vt_minus_path, step, vtm, precision)
        return self.thermostat()

    def restart(self, state_path, vt_minus_path=None, thermostat=None):
        """The restart branch of the Ensemble constructor (state file:
        positions, real positions, velocities, src/Ensemble.cpp:299-317), then
        LFIntegrator::start, then vtMinus from its file when given
        (NHIntegrator::read_vtMinus, :1956-2002) and the thermostat friction
        when given."""
        from .inputs import read_coordinate, read_state
        ctx = self.ctx
        pos, real, vel = read_state(state_path, ctx.n)
        ctx.set_positions(pos)
        ctx.check(ctx.lib.smd_set_real_positions(ctx.h, _ptr(_f64(real)), ctx.n))
        ctx.set_velocities(vel)
        ctx.compute_mol_centers()
        self.init()
        if vt_minus_path is not None:
            vtm = _f64(read_coordinate(vt_minus_path, ctx.n, skip_lines=1))
            ctx.check(ctx.lib.smd_set_vt_minus(ctx.h, _ptr(vtm), ctx.n))
        if thermostat is not None:
            # a lossless continuation: the friction and the molecule momenta
            # SHAKE's friction term reads (Molecule::momenta of the last
            # step, the molecules' summed m v)
            ctx.check(ctx.lib.smd_md_set_thermostat(ctx.h, float(thermostat[0]), float(thermostat[1])))
            s = ctx.system
            mom = np.zeros((ctx.nmol, 3))
            np.add.at(mom, s.mol, s.mass[:, None] * vel)
            ctx.check(ctx.lib.smd_set_mol_momenta(ctx.h, _ptr(_f64(mom)), ctx.nmol))

    def thermostat(self):
        a, b = C.c_double(), C.c_double()
        self.ctx.check(self.ctx.lib.smd_md_get_thermostat(self.ctx.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def real_positions(self):
        r = np.zeros((self.ctx.n, 3))
        self.ctx.check(self.ctx.lib.smd_get_real_positions(self.ctx.h, _ptr(r), self.ctx.n))
        return r

    def pressure(self):
        """The last step's Ensemble::virial, molVirial, atomPressure and
        molPressure (3x3, row-major XX XY XZ ...), myAvLus.P, and the
        Ensemble::accumulate stress sums (pressure, shear, antiShear,
        firstNorm, secondNorm; atom and molecular) with their sample count."""
        o = L.MdPressure()
        self.ctx.check(self.ctx.lib.smd_md_pressure(self.ctx.h, C.byref(o)))
        return {"virial": np.array(o.virial[:]), "mol_virial": np.array(o.mol_virial[:]),
                "atom_pressure": np.array(o.atom_pressure[:]), "mol_pressure": np.array(o.mol_pressure[:]),
                "mol_pt": o.mol_pt, "counter": o.counter, "samples": o.samples,
                "atom_stress": np.array(o.atom_stress[:]), "mol_stress": np.array(o.mol_stress[:]),
                "lj_dudv": o.lj_dudv, "real_deriv": o.real_deriv, "long_deriv": o.long_deriv,
                "mol_deriv": o.mol_deriv, "surf_deriv": o.surf_deriv}


def set_multipoles(ctx: Context):
    """Uploads each site's local-frame dipole, quadrupole and polarizability
    (AtomParam dipx..quadzz, polar) from the system's types."""
    s = ctx.system
    dip = np.ascontiguousarray([s.types[t].dip for t in s.type], dtype=np.float64)
    quad = np.ascontiguousarray([s.types[t].quad for t in s.type], dtype=np.float64)
    pol = np.ascontiguousarray([s.types[t].polar for t in s.type], dtype=np.float64)
    ctx.check(ctx.lib.smd_set_multipoles(ctx.h, _ptr(dip), _ptr(quad), _ptr(pol), ctx.n))


class Multipole:
    """Multipole (src/forces/Multipole.h): permanent charges, dipoles and
    quadrupoles under Ewald with torques, on the device."""

    def __init__(self, ctx: Context, *, alpha, real_cut, kcut, use_constraint=True, surface=False, torques=True,
                 prec=L.PREC_FP64):
        self.ctx = ctx
        self.prec = prec
        set_multipoles(ctx)
        p = L.MpParams(alpha, real_cut, int(use_constraint), int(surface), int(torques))
        nk = C.c_int32()
        ctx.check(ctx.lib.smd_multipole_setup(ctx.h, C.byref(p), kcut, C.byref(nk)))
        self.num_kvectors = nk.value

    def compute(self, f_accum=None):
        """Multipole::compute: returns the energies, Lustig terms and virial;
        forces are added to f_accum [n,3] when given."""
        o = L.MpOut()
        self.ctx.check(self.ctx.lib.smd_multipole_compute(self.ctx.h, self.prec, _fptr(f_accum), C.byref(o)))
        d = {k: getattr(o, k) for k, _ in L.MpOut._fields_ if k != "virial"}
        d["virial"] = np.array(o.virial[:])
        return d

    def compute_long(self, f_accum=None):
        o = L.MpLongOut()
        self.ctx.check(self.ctx.lib.smd_multipole_compute_long(self.ctx.h, self.prec, _fptr(f_accum), C.byref(o)))
        d = {k: getattr(o, k) for k, _ in L.MpLongOut._fields_ if k != "virial"}
        d["virial"] = np.array(o.virial[:])
        return d


class Induction:
    """Induction's induced-dipole solve (src/forces/Induction.cpp:181-455,
    :2401-2467) on the device: frames, permanent field, A mu = E0 by CG."""

    def __init__(self, ctx: Context, *, cutoff_ind, a_damp=0.39, damp_tij=True, do_min=True, tol=0.0,
                 max_iter=0, guess=0, prec=L.PREC_FP64, precond=False):
        """precond=False (0): the reference's CG, stopping at |r|_2 < tol (0:
        its 1e-14).  precond=True (1): CG preconditioned by diag(alpha);
        precond=2: by the inverse of each molecule's 9x9 block of A.  Both
        stop when the RMS dipole residual |alpha_i r_i| over the sites is
        below tol (e nm; the configs' 1e-6 D is 1e-6 * L.DEBYE_E_NM); prec
        MIXED takes these solves with FP32 pair math in A v."""
        self.ctx = ctx
        set_multipoles(ctx)
        pc = int(precond)
        self.params = L.InduceParams(cutoff_ind, a_damp, int(damp_tij), int(do_min), tol, max_iter, guess, prec,
                                     pc, 1 if pc else 0)

    def induce(self):
        """Runs the solve; returns (iterations, converged, residual)."""
        r = L.InduceReport()
        self.ctx.check(self.ctx.lib.smd_induce(self.ctx.h, C.byref(self.params), C.byref(r)))
        return r.iterations, bool(r.converged), r.residual

    terms_prec = L.PREC_FP64

    def setup_terms(self, *, alpha, real_cut, kcut, use_constraint=True, surface=False, torques=True,
                    damp_amoeba=False, prec=L.PREC_FP64):
        """Induction ctor and set_up for the energy, force and torque terms of
        the induced dipoles (src/forces/Induction.cpp:19-158, :592-652); prec
        MIXED evaluates their real-space pair math in FP32."""
        self.terms_prec = prec
        p = L.IndfParams(alpha, real_cut, self.params.a_damp, int(use_constraint), int(surface), int(torques),
                         int(damp_amoeba), int(self.params.do_min))
        nk = C.c_int32()
        self.ctx.check(self.ctx.lib.smd_induction_setup(self.ctx.h, C.byref(p), kcut, C.byref(nk)))
        self.num_kvectors = nk.value

    def compute(self, f_accum=None):
        """Induction::compute after the dipole solve: energies, Lustig terms,
        virial; forces added to f_accum [n,3] when given."""
        o = L.IndfOut()
        self.ctx.check(self.ctx.lib.smd_induction_compute(self.ctx.h, self.terms_prec, _fptr(f_accum), C.byref(o)))
        d = {k: getattr(o, k) for k, _ in L.IndfOut._fields_ if k != "virial"}
        d["virial"] = np.array(o.virial[:])
        return d

    def compute_long(self, f_accum=None):
        o = L.MpLongOut()
        self.ctx.check(self.ctx.lib.smd_induction_compute_long(self.ctx.h, self.terms_prec, _fptr(f_accum),
                                                               C.byref(o)))
        d = {k: getattr(o, k) for k, _ in L.MpLongOut._fields_ if k != "virial"}
        d["virial"] = np.array(o.virial[:])
        return d

    def time_apply(self, reps=50):
        """Average microseconds of one launch of the CG's A v (after a converged induce())."""
        us = C.c_double()
        self.ctx.check(self.ctx.lib.smd_induce_time_apply(self.ctx.h, int(reps), C.byref(us)))
        return us.value

    def results(self):
        """Lab-frame dipoles [n,3], quadrupoles [n,6], permanent field [n,3], induced dipoles [n,3]."""
        n = self.ctx.n
        dip, quad, e0, mu = np.zeros((n, 3)), np.zeros((n, 6)), np.zeros((n, 3)), np.zeros((n, 3))
        self.ctx.check(self.ctx.lib.smd_get_induction(self.ctx.h, _ptr(dip), _ptr(quad), _ptr(e0), _ptr(mu), n))
        return dip, quad, e0, mu


class PolarMD:
    """Whole LFIntegrator::run step for rigid polarizable water on the device
    (smd_polar_md_*): BuffvdWForce (or LJForce), Multipole, Induction with its
    dipole solve, SHAKE and Nose-Hoover (src/integrators/LFIntegrator.cpp:
    385-470, 561-722).  `induce` is an Induction whose solve parameters the
    step uses; constraints are the water's."""

    def __init__(self, ctx: Context, induce: "Induction", *, cutoff, cut_buff, alpha, kcut, dt, temperature,
                 coupl=0.1, tol=1e-8, max_count=10000, real_cut=None, vdw="buff", switch_on=False,
                 switch_dist=0.0, lustig=True, torques=True, surface=False, damp_amoeba=False, prec=L.PREC_FP64,
                 use_graph=True, use_cells=True, total_steps=0, atom_thermo=True):
        self.ctx = ctx
        real_cut = cutoff if real_cut is None else real_cut
        b2 = _f64(ctx.system.constraint_lengths2())
        ctx.check(ctx.lib.smd_set_constraints(ctx.h, _ptr(b2), ctx.nmol))
        cls = BuffvdWForce if vdw == "buff" else LJForce
        self._vdw = cls(ctx, cutoff=cutoff, switch_on=switch_on, switch_dist=switch_dist, use_constraint=True,
                        lustig=lustig, prec=prec)
        p = L.PolarMdParams()
        p.cutoff, p.cut_buff = cutoff, cut_buff
        p.vdw = self._vdw.params
        p.mp = L.MpParams(alpha, real_cut, 1, int(surface), int(torques))
        p.mp_kcut = kcut
        p.indf = L.IndfParams(alpha, real_cut, induce.params.a_damp, 1, int(surface), int(torques),
                              int(damp_amoeba), int(induce.params.do_min))
        p.ind_kcut = kcut
        p.induce = induce.params
        p.integ = L.IntegParams(dt, temperature, coupl, tol, max_count, 1, int(atom_thermo))
        p.use_cells, p.prec, p.use_graph, p.total_steps = int(use_cells), prec, int(use_graph), int(total_steps)
        self.params = p

    def init(self):
        """LFIntegrator::start: forces at t (with a full solve) and vtMinus."""
        self.ctx.check(self.ctx.lib.smd_polar_md_init(self.ctx.h, C.byref(self.params)))

    def run(self, nsteps):
        r = L.PolarMdReport()
        self.ctx.check(self.ctx.lib.smd_polar_md_run(self.ctx.h, int(nsteps), C.byref(r)))
        return {k: getattr(r, k) for k, _ in L.PolarMdReport._fields_}

    def time_kernel(self, name, reps=50):
        """Average microseconds of one launch of "mp_real" or "ind_real"."""
        us = C.c_double()
        self.ctx.check(self.ctx.lib.smd_polar_time_kernel(self.ctx.h, name.encode(), int(reps), C.byref(us)))
        return us.value


class LearnedGuess:
    """The learned initial guess of the induced dipoles (smd_mlp_*): a per-site
    network, trained on the GPU on the dipoles of converged solves, that
    Induction(guess=3) uses as CG's starting point.  The reference's CG starts
    from Pi = 1 (src/forces/Induction.cpp:2404-2410); the solve's tolerance is
    unchanged, so the guess changes the iteration count, not the result."""

    def __init__(self, ctx: Context, seed=1):
        self.ctx = ctx
        ctx.check(ctx.lib.smd_mlp_init(ctx.h, int(seed)))

    def params(self):
        p = np.zeros(L.MLP_NPARAMS, dtype=np.float32)
        self.ctx.check(self.ctx.lib.smd_mlp_get_params(self.ctx.h, _ptr(p), p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).ravel()
        self.ctx.check(self.ctx.lib.smd_mlp_set_params(self.ctx.h, _ptr(p), p.size))

    def norm(self):
        v = np.zeros(L.MLP_NNORM, dtype=np.float32)
        self.ctx.check(self.ctx.lib.smd_mlp_get_norm(self.ctx.h, _ptr(v), v.size))
        return v

    def set_norm(self, v):
        v = np.ascontiguousarray(v, dtype=np.float32).ravel()
        self.ctx.check(self.ctx.lib.smd_mlp_set_norm(self.ctx.h, _ptr(v), v.size))

    def forward(self, x):
        """Standardised outputs [rows, 3] of standardised descriptors x [rows, 32]."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, L.MLP_FEATURES)
        y = np.zeros((x.shape[0], L.MLP_OUTPUTS), dtype=np.float32)
        self.ctx.check(self.ctx.lib.smd_mlp_forward(self.ctx.h, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def loss_grad(self, x, target):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, L.MLP_FEATURES)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(-1, L.MLP_OUTPUTS)
        g = np.zeros(L.MLP_NPARAMS, dtype=np.float32)
        loss = C.c_double()
        self.ctx.check(self.ctx.lib.smd_mlp_loss_grad(self.ctx.h, _ptr(x), _ptr(t), x.shape[0], C.byref(loss),
                                                      _ptr(g)))
        return loss.value, g

    def descriptors(self):
        d = np.zeros((self.ctx.n, L.MLP_FEATURES), dtype=np.float32)
        self.ctx.check(self.ctx.lib.smd_mlp_descriptors(self.ctx.h, _ptr(d), self.ctx.n))
        return d

    def add_frame(self):
        """Appends the last converged solve to the training set; returns its size."""
        k = C.c_int64()
        self.ctx.check(self.ctx.lib.smd_mlp_add_frame(self.ctx.h, C.byref(k)))
        return k.value

    def clear_frames(self):
        self.ctx.check(self.ctx.lib.smd_mlp_clear_frames(self.ctx.h))

    def train(self, epochs=60, batch=1024, lr=2e-3, lr_final=2e-5, seed=7, renormalize=True):
        p = L.MlpTrainParams(epochs, batch, lr, lr_final, seed, int(renormalize))
        r = L.MlpTrainReport()
        self.ctx.check(self.ctx.lib.smd_mlp_train(self.ctx.h, C.byref(p), C.byref(r)))
        return {k: getattr(r, k) for k, _ in L.MlpTrainReport._fields_}

    def state(self):
        """Parameters and normalisation, to carry a trained network to another context."""
        return {"params": self.params(), "norm": self.norm()}

    def load_state(self, st):
        self.set_params(st["params"])
        self.set_norm(st["norm"])

    def save(self, path):
        """Writes the trained network (parameters and normalisation) as .npz."""
        np.savez(path, params=self.params(), norm=self.norm())

    def load(self, path):
        with np.load(path) as st:          # arrays only; nothing is unpickled
            self.load_state({"params": st["params"], "norm": st["norm"]})
