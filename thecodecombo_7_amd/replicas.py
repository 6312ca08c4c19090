"""Batched temperature replica exchange over libttc.so (include/ttc.h,
ttc_batch_*): many uneven-size replicas, each a device context running the
Nose-Hoover chain NVT step of nvt_md.c:528-531, advanced together, with the
exchange rule of swap_box.c:7-11 applied between neighbouring temperatures
every `exchange_every` steps (nvt_mc.c:143-162 alternates even/odd pairs).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as T
from .nonbond import Box

KB = 1.380662e-23


class ReplicaBatch:
    def __init__(self, systems, temps, tau: float = 0.1, M: int = 2, seed: int = 1, angles=None):
        """angles: optional system.AngleTerms per replica (None entries: none)."""
        L = T.lib()
        self.L = L
        self.boxes = []
        self.h = None
        try:
            for r, (s, t) in enumerate(zip(systems, temps)):
                b = Box(s)
                if angles is not None and angles[r] is not None:
                    b.set_angles(angles[r])
                nfree = 3.0 * s.n - 3.0
                b.set_nhc(float(t), nfree * KB * float(t) * 1e-4 * tau * tau, M=M, nfree=nfree)
                b.nblist()
                self.boxes.append(b)
            self.nrep = len(self.boxes)
            arr = (C.c_void_p * self.nrep)(*[b.h.value for b in self.boxes])
            h = C.c_void_p()
            T.check(L.ttc_batch_create(arr, self.nrep, seed, C.byref(h)))
            self.h = h
        except Exception:
            self.close()
            raise
        self.sites = np.array([b.n for b in self.boxes])

    def close(self):
        if getattr(self, "h", None):
            self.L.ttc_batch_destroy(self.h)
            self.h = None
        for b in getattr(self, "boxes", []):
            b.close()
        self.boxes = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, mode: int, steps_per_launch: int):
        T.check(self.L.ttc_batch_set_launch(self.h, mode, steps_per_launch))
        self.spl = steps_per_launch

    def autotune(self, trial_steps: int = 50):
        m = C.c_int(0); s = C.c_int(0); t = np.zeros(12)
        T.check(self.L.ttc_batch_autotune(self.h, trial_steps, C.byref(m), C.byref(s), T.dptr(t)))
        self.spl = s.value
        return m.value, s.value, t

    def md_nhc(self, nsteps: int):
        pot = np.zeros(self.nrep); reb = np.zeros(self.nrep, np.int32)
        T.check(self.L.ttc_batch_md_nhc(self.h, nsteps, T.dptr(pot), T.iptr(reb)))
        return pot, reb

    def exchange(self, potens, parity: int) -> int:
        p = np.ascontiguousarray(potens, dtype=np.float64)
        acc = C.c_int(0)
        T.check(self.L.ttc_batch_exchange(self.h, T.dptr(p), parity, C.byref(acc)))
        return acc.value

    def state(self):
        order = np.zeros(self.nrep, np.int32); temps = np.zeros(self.nrep); ms = np.zeros(1)
        att = C.c_longlong(0); acc = C.c_longlong(0)
        T.check(self.L.ttc_batch_state(self.h, T.iptr(order), T.dptr(temps), T.dptr(ms), C.byref(att),
                                       C.byref(acc)))
        return order, temps, float(ms[0]), att.value, acc.value

    def run(self, nsteps: int, exchange_every: int = 500):
        """nsteps NHC steps of every replica with an exchange attempt every
        exchange_every steps (alternating even/odd temperature pairs)."""
        done = 0
        parity = 0
        ms = 0.0
        rebuilds = np.zeros(self.nrep, np.int64)
        while done < nsteps:
            k = min(exchange_every, nsteps - done)
            pot, reb = self.md_nhc(k)
            ms += self.state()[2] * k
            rebuilds += reb
            done += k
            if k == exchange_every:
                self.exchange(pot, parity)
                parity ^= 1
        order, temps, _, att, acc = self.state()
        return {"ms_per_step": ms / max(1, nsteps), "attempts": att, "accepts": acc,
                "temps": temps, "rebuilds": rebuilds}
