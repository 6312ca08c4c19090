"""Host mirror of the reference's nonbonded entry points over libttc.so.

`Box` holds one device context and exposes the reference's operations by
their reference names, with the reference's argument meaning and error
behaviour (errors carry the reference's exit codes, see include/ttc.h):

    nblist()                 nblist.c:32
    nl_check(lb, ub)         nblist.c:607
    get_nblist(which)        the nlist[k] / nlist_ew[k] arrays
    cnonbondnl()             cnonbond_nblist.c:51   (pair loop + pme_calc)
    pme_calc()               pme_calc.c:18
    cbond()                  cbond.c:65
    forces()                 forces.c:118
    integrate_vverlet(n)     integrate_vverlet.c:36, n steps with the
                             nl_check/nblist trigger of nve_md.c:369-370
    kinet()                  kinet.c:18
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as T


class Box:
    def __init__(self, sysm, ewald_lists: bool = True):
        L = T.lib()
        self.L = L
        self.sys = sysm
        self.n = sysm.n
        p = T.TtcParams(
            n=sysm.n, lx=sysm.box[0], ly=sysm.box[1], lz=sysm.box[2], rc=sysm.rc,
            rclist=sysm.rclist, rclist_ew=sysm.rclist_ew, ewald_lists=int(ewald_lists),
            kappa=sysm.kappa, eps_rf=sysm.eps_rf, e14fac=sysm.e14fac, order=sysm.order,
            gx=sysm.grid[0], gy=sysm.grid[1], gz=sysm.grid[2], ntypes=sysm.ntypes, dt=sysm.dt)
        h = C.c_void_p()
        T.check(L.ttc_create(C.byref(p), C.byref(h)))
        self.h = h
        try:
            self._topology()
            self.set_positions(sysm.pos)
            self.set_velocities(sysm.vel)
        except Exception:
            self.close()
            raise

    def _topology(self):
        s = self.sys
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self._keep = [i32(s.type), f64(s.q), f64(s.mass), f64(s.lj_eps), f64(s.lj_sig),
                      f64(s.lj14_eps), f64(s.lj14_sig), i32(s.xint[:, 0]), i32(s.xint[:, 1]),
                      i32(s.xint[:, 2]), i32(s.bonds[:, 0]), i32(s.bonds[:, 1]), f64(s.bond_r0),
                      f64(s.bond_k)]
        k = self._keep
        T.check(self.L.ttc_set_topology(
            self.h, T.iptr(k[0]), T.dptr(k[1]), T.dptr(k[2]), T.dptr(k[3]), T.dptr(k[4]),
            T.dptr(k[5]), T.dptr(k[6]), len(s.xint), T.iptr(k[7]), T.iptr(k[8]), T.iptr(k[9]),
            len(s.bonds), T.iptr(k[10]), T.iptr(k[11]), T.dptr(k[12]), T.dptr(k[13])))

    def close(self):
        if getattr(self, "h", None):
            self.L.ttc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ----
    def set_positions(self, pos):
        a = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        T.check(self.L.ttc_set_positions(self.h, T.dptr(a), 3))

    def positions(self):
        a = np.zeros((self.n, 3))
        T.check(self.L.ttc_get_positions(self.h, T.dptr(a), 3))
        return a

    def set_velocities(self, vel):
        a = np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        T.check(self.L.ttc_set_velocities(self.h, T.dptr(a), 3))

    def velocities(self):
        a = np.zeros((self.n, 3))
        T.check(self.L.ttc_get_velocities(self.h, T.dptr(a), 3))
        return a

    @property
    def e_self(self) -> float:
        return self.L.ttc_self_energy(self.h)

    # ---- nblist.c ----
    def nblist(self):
        T.check(self.L.ttc_nblist(self.h))

    def nl_check(self, lb: int = 0, ub: int | None = None) -> int:
        f = C.c_int(0)
        T.check(self.L.ttc_nl_check(self.h, lb, self.n if ub is None else ub, C.byref(f)))
        return f.value

    def get_nblist(self, which: int = T.TTC_LIST_NLIST, max_nb: int = 1500):
        cnt = np.zeros(self.n, np.int32)
        lst = np.zeros(self.n * max_nb, np.int32)
        xf = np.zeros(self.n * max_nb, np.int32)
        T.check(self.L.ttc_get_nblist(self.h, which, max_nb, T.iptr(cnt), T.iptr(lst), T.iptr(xf)))
        return cnt, lst.reshape(self.n, max_nb), xf.reshape(self.n, max_nb)

    # ---- forces ----
    def cnonbondnl(self):
        ff = np.zeros((self.n, 3)); en = np.zeros(8); vn = np.zeros(6); vk = np.zeros(6)
        T.check(self.L.ttc_cnonbondnl(self.h, T.dptr(ff), 3, T.dptr(en), T.dptr(vn), T.dptr(vk)))
        return ff, en, vn, vk

    def pme_calc(self):
        ff = np.zeros((self.n, 3)); en = np.zeros(3); vir = np.zeros(6)
        T.check(self.L.ttc_pme_calc(self.h, T.dptr(ff), 3, T.dptr(en), T.dptr(vir)))
        return ff, en, vir

    def pair(self):
        ff = np.zeros((self.n, 3)); e5 = np.zeros(5); vir = np.zeros(6)
        T.check(self.L.ttc_pair(self.h, T.dptr(ff), 3, T.dptr(e5), T.dptr(vir)))
        return ff, e5, vir

    def cbond(self):
        ff = np.zeros((self.n, 3)); eb = np.zeros(1); vir = np.zeros(6)
        T.check(self.L.ttc_cbond(self.h, T.dptr(ff), 3, T.dptr(eb), T.dptr(vir)))
        return ff, float(eb[0]), vir

    def set_angles(self, at):
        """Angle-type terms (system.AngleTerms) on the device."""
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        self._angles = [i32(at.bend), f64(at.bend_eq), f64(at.bend_k), i32(at.urey), f64(at.urey_s0),
                        f64(at.urey_k), i32(at.tors), f64(at.kphi), f64(at.nphi), f64(at.delphi),
                        i32(at.impr), f64(at.impr_k), f64(at.impr_eq)]
        a = self._angles
        T.check(self.L.ttc_set_angles(self.h, len(at.bend), T.iptr(a[0]), T.dptr(a[1]), T.dptr(a[2]),
                                      len(at.urey), T.iptr(a[3]), T.dptr(a[4]), T.dptr(a[5]),
                                      len(at.tors), T.iptr(a[6]), T.dptr(a[7]), T.dptr(a[8]), T.dptr(a[9]),
                                      len(at.impr), T.iptr(a[10]), T.dptr(a[11]), T.dptr(a[12])))

    def angle_kind(self, kind: str, hessian: bool = False):
        """cbend / curey / ctorsion / cimproper (or their CONFIGT variants
        with hessian): forces (n,3), energy, virial (or Hessian sums)."""
        name = {"bend": "ttc_cbend", "urey": "ttc_curey", "tors": "ttc_ctorsion", "impr": "ttc_cimproper"}[kind]
        ff = np.zeros((self.n, 3)); e = np.zeros(1); x = np.zeros(3 if hessian else 6)
        T.check(getattr(self.L, name + ("2" if hessian else ""))(self.h, T.dptr(ff), 3, T.dptr(e), T.dptr(x)))
        return ff, float(e[0]), x

    def drift(self):
        """drift.c: remove the net momentum."""
        T.check(self.L.ttc_drift(self.h))

    def set_drift(self, every: int):
        """sim.drift: momentum removal after every `every`-th MD step."""
        T.check(self.L.ttc_set_drift(self.h, every))

    def bonded_energies(self):
        e = np.zeros(5)
        T.check(self.L.ttc_bonded_energies(self.h, T.dptr(e)))
        return dict(zip(["bond", "bend", "urey", "tors", "impr"], e))

    def forces(self):
        ff = np.zeros((self.n, 3)); en = np.zeros(9)
        T.check(self.L.ttc_forces(self.h, T.dptr(ff), 3, T.dptr(en)))
        keys = ["bond", "nbond", "nbonds", "coulomb", "rewald", "rewalds", "kewald", "sewald", "iewald"]
        return ff, dict(zip(keys, en))

    def get_forces(self):
        """ff[k] of the last evaluation (after a step: integrate_vverlet.c:88)."""
        ff = np.zeros((self.n, 3))
        T.check(self.L.ttc_get_forces(self.h, T.dptr(ff), 3))
        return ff

    def set_capacity_limits(self, row_limit: int = 0, window_limit: int = 0):
        T.check(self.L.ttc_set_capacity_limits(self.h, row_limit, window_limit))

    def window_stats(self):
        m = C.c_int(0); cap = C.c_int(0); h = C.c_int(0)
        T.check(self.L.ttc_window_stats(self.h, C.byref(m), C.byref(cap), C.byref(h)))
        return m.value, cap.value, bool(h.value)

    def kinet(self):
        k = np.zeros(1); ken = np.zeros(6)
        T.check(self.L.ttc_kinet(self.h, T.dptr(k), T.dptr(ken)))
        return float(k[0]), ken

    # ---- CONFIGT: cnonbond_nblist2.c, cbond2.c, configtemp.c ----
    def cnonbondnl2(self):
        ff = np.zeros((self.n, 3)); en = np.zeros(8); h = np.zeros(3)
        T.check(self.L.ttc_cnonbondnl2(self.h, T.dptr(ff), 3, T.dptr(en), T.dptr(h)))
        return ff, en, h

    def cbond2(self):
        ff = np.zeros((self.n, 3)); eb = np.zeros(1); h = np.zeros(3)
        T.check(self.L.ttc_cbond2(self.h, T.dptr(ff), 3, T.dptr(eb), T.dptr(h)))
        return ff, float(eb[0]), h

    def configtemp(self):
        num = np.zeros(1); den = np.zeros(1); t = np.zeros(1); h = np.zeros(3)
        T.check(self.L.ttc_configtemp(self.h, T.dptr(num), T.dptr(den), T.dptr(t), T.dptr(h)))
        return float(t[0]), float(num[0]), float(den[0]), h

    # ---- integrate_vverlet.c ----
    def integrate_vverlet(self, nsteps: int, vscal: float = 1.0, energy_every: int = 0):
        rows = max(1, nsteps // energy_every) if energy_every else 1
        buf = np.zeros((rows, 9))
        nr = C.c_int(0)
        nreb = C.c_int(0)
        T.check(self.L.ttc_md_vverlet(self.h, nsteps, vscal, energy_every, T.dptr(buf), rows,
                                      C.byref(nr), C.byref(nreb)))
        return buf[:nr.value], nreb.value

    # ---- force_short.c / force_long.c / integrate_mts.c ----
    def force_short(self):
        ff = np.zeros((self.n, 3)); eb = np.zeros(1)
        T.check(self.L.ttc_force_short(self.h, T.dptr(ff), 3, T.dptr(eb)))
        return ff, float(eb[0])

    def force_long(self):
        ff = np.zeros((self.n, 3)); en = np.zeros(8)
        T.check(self.L.ttc_force_long(self.h, T.dptr(ff), 3, T.dptr(en)))
        return ff, en

    def integrate_mts(self, nouter: int, ninner: int, dtlong: float, vscal: float = 1.0, energy_every: int = 0):
        rows = max(1, nouter // energy_every) if energy_every else 1
        buf = np.zeros((rows, 9)); nr = C.c_int(0); nreb = C.c_int(0)
        T.check(self.L.ttc_md_mts(self.h, nouter, ninner, dtlong, vscal, energy_every, T.dptr(buf), rows,
                                  C.byref(nr), C.byref(nreb)))
        return buf[:nr.value], nreb.value

    # ---- integrate_nhc.c / nvt_md.c ----
    def set_nhc(self, T_K: float, Q: float, M: int = 2, Nc: int = 1, Nys: int = 1, w=None, nfree=None):
        w = np.ascontiguousarray(np.ones(Nys) if w is None else w, dtype=np.float64)
        nfree = 3.0 * self.n - 3.0 if nfree is None else nfree
        T.check(self.L.ttc_set_nhc(self.h, T_K, Q, M, Nc, Nys, T.dptr(w), nfree))
        self._nhc_M = M

    def nhc_state(self):
        z = np.zeros(3 * self._nhc_M)
        T.check(self.L.ttc_get_nhc(self.h, T.dptr(z)))
        return z

    def integrate_nhc(self, nsteps: int, energy_every: int = 0):
        rows = max(1, nsteps // energy_every) if energy_every else 1
        buf = np.zeros((rows, 9)); nr = C.c_int(0); nreb = C.c_int(0)
        T.check(self.L.ttc_md_nhc(self.h, nsteps, energy_every, T.dptr(buf), rows, C.byref(nr), C.byref(nreb)))
        return buf[:nr.value], nreb.value

    # ---- enbnd_mc.c / nblist_pivot.c ----
    def enbnd_mc(self, lb: int, ub: int):
        en = np.zeros(8)
        T.check(self.L.ttc_enbnd_mc(self.h, lb, ub, T.dptr(en)))
        return en

    def nblist_pivot(self, ub: int):
        T.check(self.L.ttc_nblist_pivot(self.h, ub))

    def timing(self, which: int) -> float:
        t = np.zeros(1)
        T.check(self.L.ttc_timing(self.h, which, T.dptr(t)))
        return float(t[0])

    def tune(self):
        """Pick the pair kernel's waves per group by timing (ttc_tune)."""
        w = C.c_int(0); t = np.zeros(3)
        T.check(self.L.ttc_tune(self.h, C.byref(w), T.dptr(t)))
        return w.value, t

    def tune_spread(self):
        """Pick the spread kernel's sites per pass by timing (ttc_tune_spread)."""
        ch = C.c_int(0); t = np.zeros(3)
        T.check(self.L.ttc_tune_spread(self.h, C.byref(ch), T.dptr(t)))
        return ch.value, t

    def tune_cached(self, path):
        """Both tuners behind a cache file (ttc_tune_cached): returns
        (pair waves, spread chunk, taken from the cache)."""
        w = C.c_int(0); ch = C.c_int(0); hit = C.c_int(0)
        T.check(self.L.ttc_tune_cached(self.h, os.fsencode(path), C.byref(w), C.byref(ch), C.byref(hit)))
        return w.value, ch.value, bool(hit.value)

    def list_stats(self):
        e = C.c_int64(0); m = C.c_int(0); cap = C.c_int(0)
        T.check(self.L.ttc_list_stats(self.h, C.byref(e), C.byref(m), C.byref(cap)))
        return e.value, m.value, cap.value


ROW_KEYS = ["step", "kinet", "bond", "nbond", "nbonds", "coulomb", "kewald", "rewald", "iewald"]
