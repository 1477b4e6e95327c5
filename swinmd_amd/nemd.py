"""Batched sheared replicas of rigid 3-site water on the device (the
smd_batch_* C ABI of libswinmd_hip.so).

Each replica is one planar Couette flow: the reference's leap-frog step
(LFIntegrator::run, src/integrators/LFIntegrator.cpp:278-941) for rigid water
with LJ on O and point-charge Ewald, with molecular SLLOD and Lees-Edwards
images added (the reference has no shear flow).  All replicas advance together,
one captured graph launch per step.  There is no CPU path: the library must be
built."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def kcut_for(n2max: float, L_box: float) -> float:
    """k cutoff selecting |n|^2 <= n2max of an unsheared cubic box."""
    return 2 * math.pi * math.sqrt(n2max + 0.5) / L_box


class ShearBatch:
    """Replicas given as rigid-water systems (site order H, O, H; the same
    types in every replica), each with its shear rate (1/ps)."""

    def __init__(self, systems, shear_rates, *, cutoff=0.9, cut_buff=0.08, alpha=3.5, kcut=None, n2max=81,
                 dt=0.002, temperature=300.0, coupl=0.1, shake_tol=1e-8, max_count=10000, nose_hoover=True,
                 prec=L.PREC_FP64, nb_cap=0, device=0):
        self.lib = L.load()
        self.systems = list(systems)
        R = len(self.systems)
        if len(shear_rates) != R:
            raise ValueError("one shear rate per replica")
        s0 = self.systems[0]
        for s in self.systems:
            if s.atoms_per_mol != 3 or s.types[0].mass != s0.types[0].mass or s.types[1].charge != s0.types[1].charge:
                raise ValueError("replicas must be the same 3-site water")
        self.n_mols = np.array([s.n_mols for s in self.systems], dtype=np.int32)
        self.box = np.array([float(s.box[0]) for s in self.systems])
        self.gamma = np.asarray(shear_rates, dtype=np.float64).copy()
        if kcut is None:
            kcut = [kcut_for(n2max, Lb) for Lb in self.box]
        self.kcut = np.asarray(kcut, dtype=np.float64).copy()
        q = s0.scaled_charge
        b2 = s0.constraint_lengths2()
        p = L.BatchParams()
        p.n_replicas = R
        p.n_mols, p.box, p.shear_rate, p.kcut = _ptr(self.n_mols), _ptr(self.box), _ptr(self.gamma), _ptr(self.kcut)
        p.mass_h, p.mass_o = s0.types[0].mass, s0.types[1].mass
        p.q_h, p.q_o = float(q[0]), float(q[1])
        p.sigma_o, p.eps_o = s0.types[1].sigma, s0.types[1].eps
        p.r_oh, p.r_hh = math.sqrt(b2[0]), math.sqrt(b2[2])
        p.cutoff, p.cut_buff, p.alpha = cutoff, cut_buff, alpha
        p.dt, p.temperature, p.coupl, p.shake_tol = dt, temperature, coupl, shake_tol
        p.max_count, p.nose_hoover, p.prec, p.nb_cap = max_count, int(nose_hoover), prec, nb_cap
        self.params = p
        h = C.c_void_p()
        rc = self.lib.smd_batch_create(C.byref(h), C.byref(p), device)
        self.h = h
        self.check(rc)
        self.n = int(3 * self.n_mols.sum())
        self.offsets = np.concatenate([[0], np.cumsum(3 * self.n_mols)])

    def check(self, rc):
        if rc != 0:
            msg = self.lib.smd_batch_last_error(self.h).decode() if self.h else ""
            raise L.SmdError(rc, msg or self.lib.smd_strerror(rc).decode())

    def close(self):
        if self.h:
            self.lib.smd_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ------------------------------------------------------------
    def set_state(self, pos, vel, tilt=None):
        """Positions and velocities [n, 3] of all replicas end to end (or a
        list per replica) and the tilts [R] (nm)."""
        pos = np.ascontiguousarray(np.concatenate(pos) if isinstance(pos, (list, tuple)) else pos, dtype=np.float64)
        vel = np.ascontiguousarray(np.concatenate(vel) if isinstance(vel, (list, tuple)) else vel, dtype=np.float64)
        if pos.size != 3 * self.n or vel.size != 3 * self.n:
            raise ValueError(f"need {self.n} atoms")
        tilt = np.zeros(len(self.systems)) if tilt is None else np.ascontiguousarray(tilt, dtype=np.float64)
        self.check(self.lib.smd_batch_set_state(self.h, _ptr(pos), _ptr(vel), _ptr(tilt)))

    def start(self):
        self.check(self.lib.smd_batch_start(self.h))

    def run(self, nsteps):
        self.check(self.lib.smd_batch_run(self.h, int(nsteps)))

    def state(self):
        out = {k: np.zeros((self.n, 3)) for k in ("pos", "vel", "vtm", "force")}
        tilt = np.zeros(len(self.systems))
        self.check(self.lib.smd_batch_get_state(self.h, _ptr(out["pos"]), _ptr(out["vel"]), _ptr(out["vtm"]),
                                                _ptr(out["force"]), _ptr(tilt)))
        out["tilt"] = tilt
        return out

    def split(self, arr):
        """Per-replica views of an [n, 3] array."""
        return [arr[self.offsets[r]:self.offsets[r + 1]] for r in range(len(self.systems))]

    def reports(self):
        R = len(self.systems)
        rep = (L.BatchReport * R)()
        self.check(self.lib.smd_batch_get_reports(self.h, rep))
        out = []
        for r in rep:
            d = {k: getattr(r, k) for k, _ in L.BatchReport._fields_}
            for k in ("virial", "mol_kin", "mol_correction"):
                d[k] = np.array(d[k][:])
            out.append(d)
        return out

    # ---- tiles --------------------------------------------------------------
    def tiles(self):
        t = L.BatchTiles()
        self.check(self.lib.smd_batch_get_tiles(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in L.BatchTiles._fields_}

    def set_tiles(self, **kw):
        cur = self.tiles()
        cur.update(kw)
        t = L.BatchTiles(**cur)
        self.check(self.lib.smd_batch_set_tiles(self.h, C.byref(t)))

    def autotune(self, reps=20):
        """Times every candidate tile shape on the current state and keeps the
        fastest; returns the chosen tiles and the table of timings (us)."""
        cap = C.c_int32(64)
        buf = np.zeros(6 * 64)
        self.check(self.lib.smd_batch_autotune(self.h, int(reps), _ptr(buf), C.byref(cap)))
        rows = buf[:6 * cap.value].reshape(-1, 6)
        table = []
        for g, ch, kt, tp, ts, tk in rows:
            if g:
                table.append({"pair_lanes": int(g), "us": round(float(tp), 2)})
            elif ch:
                table.append({"sfac_chunk": int(ch), "us": round(float(ts), 2)})
            else:
                table.append({"kforce_tile": int(kt), "us": round(float(tk), 2)})
        return self.tiles(), table

    def time_kernel(self, name, reps=20):
        us = C.c_double()
        self.check(self.lib.smd_batch_time_kernel(self.h, name.encode(), int(reps), C.byref(us)))
        return us.value
