"""Host mirror of the batched Monte Carlo / density-of-states path.

`Walkers` holds one device batch of C-alpha Go-model chains and exposes the
reference's MC operations:

    energy()                      full Go energy of every walker (egolik/enbnd)
    delta_energy(w, lb, ub, new)  forces_mc(lb,ub,1) - forces_mc(lb,ub,0)
    run(ntrials, mode, lnf)       pivot / displacement trials with Wang-Landau
                                  acceptance (pivot.c:236-253); mode 1 adds the
                                  learned first stage (delayed acceptance)
    snapshot() / restore()        device copy of the sampling state
    flat_histogram(w)             the flatness test of ctdos.c:268
    train_surrogate(...)          Adam on the logged (move, dE) pairs
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as T
from . import system

NF = 52
HID = 64
NSTAT = 7   # trials, accepted, stage-1 rejects, exact evaluations, outside window, pivots, accepted pivots


class McParams(C.Structure):
    _fields_ = [("nwalk", C.c_int), ("maxn", C.c_int), ("nbins", C.c_int), ("ebin", C.c_double),
                ("kbond", C.c_double), ("r0bond", C.c_double), ("eps", C.c_double), ("sig_rep", C.c_double),
                ("pivot_frac", C.c_double), ("theta_max", C.c_double), ("disp_max", C.c_double),
                ("seed", C.c_uint64)]


def _declare(L):
    if getattr(L, "_mc_declared", False):
        return
    dp, ip, fp, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_float), C.c_void_p
    i64p = C.POINTER(C.c_int64)
    L.ttc_mc_create.argtypes = [C.POINTER(McParams), C.POINTER(vp)]
    L.ttc_mc_destroy.argtypes = [vp]
    L.ttc_mc_set_walker.argtypes = [vp, C.c_int, C.c_int, dp, C.c_int, ip, ip, dp, C.c_double]
    L.ttc_mc_get_walker.argtypes = [vp, C.c_int, dp, dp]
    L.ttc_mc_get_streams.argtypes = [vp, C.c_int, C.POINTER(C.c_uint64)]
    L.ttc_mc_set_streams.argtypes = [vp, C.c_int, C.POINTER(C.c_uint64)]
    L.ttc_mc_energy.argtypes = [vp, dp]
    L.ttc_mc_delta_energy.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
    L.ttc_mc_run.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int]
    L.ttc_mc_checkpoint.argtypes = [vp, C.c_int]
    L.ttc_mc_get_stats.argtypes = [vp, i64p]
    L.ttc_mc_reset_stats.argtypes = [vp]
    L.ttc_mc_get_dos.argtypes = [vp, C.c_int, dp, i64p]
    L.ttc_mc_set_dos.argtypes = [vp, C.c_int, dp]
    L.ttc_mc_reset_hist.argtypes = [vp]
    L.ttc_mc_swap_windows.argtypes = [vp, C.c_int, C.c_double, ip]
    L.ttc_mc_merge_windows.argtypes = [vp, C.c_int, C.c_double]
    L.ttc_mc_get_log.argtypes = [vp, C.c_int, fp, dp, ip]
    L.ttc_mc_clear_log.argtypes = [vp]
    L.ttc_mc_last_time.argtypes = [vp, dp]
    L.ttc_mlp_nparams.restype = C.c_int
    L.ttc_mc_set_mlp.argtypes = [vp, fp]
    L.ttc_mc_get_mlp.argtypes = [vp, fp]
    L.ttc_mlp_forward.argtypes = [vp, C.c_int, fp, fp]
    L.ttc_mlp_loss_grad.argtypes = [vp, C.c_int, fp, fp, dp, fp]
    L.ttc_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_double, dp]
    L._mc_declared = True


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


class Walkers:
    def __init__(self, chains, *, nbins=512, ebin=0.5 * system.KCAL, seed=2024, pivot_frac=0.5,
                 theta_max=0.25, disp_max=0.6, window_below=10.0, emin=None, kbond=system.GO_KBOND,
                 sig_rep=system.GO_SIG_REP, eps=system.GO_EPS):
        """emin: the windows' lower edges (scalar or one per walker); by
        default each window starts window_below under the walker's energy."""
        L = T.lib()
        _declare(L)
        self.L = L
        self.chains = chains
        self.nwalk = len(chains)
        self.maxn = max(c.n for c in chains)
        self.nbins = nbins
        self.ebin = ebin
        self.p = McParams(nwalk=self.nwalk, maxn=self.maxn, nbins=nbins, ebin=ebin, kbond=kbond,
                          r0bond=system.GO_BOND, eps=eps, sig_rep=sig_rep,
                          pivot_frac=pivot_frac, theta_max=theta_max, disp_max=disp_max, seed=seed)
        h = C.c_void_p()
        T.check(L.ttc_mc_create(C.byref(self.p), C.byref(h)))
        self.h = h
        if emin is not None:
            self.emin = np.broadcast_to(np.asarray(emin, dtype=np.float64), (self.nwalk,)).copy()
        else:
            self.emin = np.zeros(self.nwalk)
            for w, c in enumerate(chains):
                self.set_walker(w, c.pos, c, None)
            e = self.energy()
            self.emin = e - window_below
        for w, c in enumerate(chains):
            self.set_walker(w, c.pos, c, self.emin[w])

    def set_walker(self, w, pos, chain, emin):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        na = np.ascontiguousarray(chain.contacts[:, 0], dtype=np.int32)
        nb = np.ascontiguousarray(chain.contacts[:, 1], dtype=np.int32)
        r0 = np.ascontiguousarray(chain.r0, dtype=np.float64)
        T.check(self.L.ttc_mc_set_walker(self.h, w, chain.n, T.dptr(pos), len(na), T.iptr(na), T.iptr(nb),
                                         T.dptr(r0), 0.0 if emin is None else float(emin)))

    def close(self):
        if getattr(self, "h", None):
            self.L.ttc_mc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def streams(self, w):
        """The 64 xoshiro256** states of walker w, (64, 4) uint64."""
        s = np.zeros((64, 4), dtype=np.uint64)
        T.check(self.L.ttc_mc_get_streams(self.h, w, s.ctypes.data_as(C.POINTER(C.c_uint64))))
        return s

    def set_streams(self, w, s):
        s = np.ascontiguousarray(s, dtype=np.uint64).reshape(64, 4)
        T.check(self.L.ttc_mc_set_streams(self.h, w, s.ctypes.data_as(C.POINTER(C.c_uint64))))

    def energy(self):
        e = np.zeros(self.nwalk)
        T.check(self.L.ttc_mc_energy(self.h, T.dptr(e)))
        return e

    def walker(self, w):
        n = self.chains[w].n
        pos = np.zeros((n, 3)); e = np.zeros(1)
        T.check(self.L.ttc_mc_get_walker(self.h, w, T.dptr(pos), T.dptr(e)))
        return pos, float(e[0])

    def delta_energy(self, w, lb, ub, newpos):
        a = np.ascontiguousarray(newpos, dtype=np.float64).reshape(-1)
        d = np.zeros(1)
        T.check(self.L.ttc_mc_delta_energy(self.h, w, lb, ub, T.dptr(a), T.dptr(d)))
        return float(d[0])

    def run(self, ntrials, mode=0, lnf=0.0, log=False):
        T.check(self.L.ttc_mc_run(self.h, ntrials, mode, lnf, int(log)))
        t = np.zeros(1)
        T.check(self.L.ttc_mc_last_time(self.h, T.dptr(t)))
        return float(t[0])

    def snapshot(self):
        """Take a device copy of the sampling state (one slot per batch)."""
        T.check(self.L.ttc_mc_checkpoint(self.h, 0))
        return True

    def restore(self, _snap=True):
        """Put the state of the last snapshot() back."""
        T.check(self.L.ttc_mc_checkpoint(self.h, 1))

    def stats(self):
        s = np.zeros(self.nwalk * NSTAT, np.int64)
        T.check(self.L.ttc_mc_get_stats(self.h, s.ctypes.data_as(C.POINTER(C.c_int64))))
        return s.reshape(self.nwalk, NSTAT)

    def reset_stats(self):
        T.check(self.L.ttc_mc_reset_stats(self.h))

    def dos(self, w):
        lng = np.zeros(self.nbins); hist = np.zeros(self.nbins, np.int64)
        T.check(self.L.ttc_mc_get_dos(self.h, w, T.dptr(lng), hist.ctypes.data_as(C.POINTER(C.c_int64))))
        return lng, hist

    def set_dos(self, w, lng):
        a = np.ascontiguousarray(lng, dtype=np.float64)
        T.check(self.L.ttc_mc_set_dos(self.h, w, T.dptr(a)))

    def reset_hist(self):
        T.check(self.L.ttc_mc_reset_hist(self.h))

    def swap_windows(self, parity, lnf=0.0):
        """swap_mpi.c: configuration exchange between neighbouring energy
        windows (pairs (w, w+1), w = parity mod 2); returns the per-walker
        accepted flags."""
        acc = np.zeros(self.nwalk, np.int32)
        T.check(self.L.ttc_mc_swap_windows(self.h, int(parity), float(lnf), T.iptr(acc)))
        return acc

    def merge_windows(self, w, lnf):
        """swap_mpi.c merge_gofe of windows w (left) and w + 1 (right)."""
        T.check(self.L.ttc_mc_merge_windows(self.h, int(w), float(lnf)))

    def flat_histogram(self, w, tol=0.8):
        """ctdos.c:268 flat_histogram: every visited bin at least tol x mean."""
        _, h = self.dos(w)
        v = h[h > 0]
        return len(v) > 0 and v.min() >= tol * v.mean()

    def log(self, max_rows=1 << 20):
        cnt = C.c_int(0)
        x = np.zeros((max_rows, NF), np.float32)
        de = np.zeros(max_rows)
        _, xp = _f(x)
        T.check(self.L.ttc_mc_get_log(self.h, max_rows, x.ctypes.data_as(C.POINTER(C.c_float)), T.dptr(de),
                                      C.byref(cnt)))
        k = min(cnt.value, max_rows)
        return x[:k], de[:k]

    def clear_log(self):
        T.check(self.L.ttc_mc_clear_log(self.h))

    # ---- surrogate network ----
    @staticmethod
    def nparams():
        L = T.lib()
        _declare(L)
        return L.ttc_mlp_nparams()

    def set_mlp(self, params):
        a, p = _f(params)
        T.check(self.L.ttc_mc_set_mlp(self.h, p))

    def get_mlp(self):
        a = np.zeros(self.nparams(), np.float32)
        T.check(self.L.ttc_mc_get_mlp(self.h, a.ctypes.data_as(C.POINTER(C.c_float))))
        return a

    def mlp_forward(self, x):
        x, xp = _f(x)
        out = np.zeros(len(x), np.float32)
        T.check(self.L.ttc_mlp_forward(self.h, len(x), xp, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def mlp_loss_grad(self, x, y):
        x, xp = _f(x)
        y, yp = _f(y)
        g = np.zeros(self.nparams(), np.float32)
        loss = np.zeros(1)
        T.check(self.L.ttc_mlp_loss_grad(self.h, len(x), xp, yp, T.dptr(loss), g.ctypes.data_as(C.POINTER(C.c_float))))
        return float(loss[0]), g

    def train_surrogate(self, nepoch=2, lr=1e-3, escale=2.5):
        loss = np.zeros(1)
        T.check(self.L.ttc_mlp_train(self.h, nepoch, 32, lr, escale, T.dptr(loss)))
        return float(loss[0])
