"""ctypes binding of the CPU oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and there only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class OrcSystem(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("rc", C.c_double), ("rclist", C.c_double), ("rclist_ew", C.c_double),
        ("kappa", C.c_double), ("eps_rf", C.c_double), ("e14fac", C.c_double),
        ("order", C.c_int), ("gx", C.c_int), ("gy", C.c_int), ("gz", C.c_int),
        ("ntypes", C.c_int), ("type", C.POINTER(C.c_int)), ("q", C.POINTER(C.c_double)),
        ("lj_eps", C.POINTER(C.c_double)), ("lj_sig", C.POINTER(C.c_double)),
        ("lj14_eps", C.POINTER(C.c_double)), ("lj14_sig", C.POINTER(C.c_double)),
    ]


class OrcExcl(C.Structure):
    _fields_ = [("n", C.c_int), ("off", C.POINTER(C.c_int)), ("partner", C.POINTER(C.c_int)),
                ("flag", C.POINTER(C.c_int))]


class OrcPairOut(C.Structure):
    _fields_ = [("enonbond", C.c_double), ("enonbonds", C.c_double), ("eewald_real", C.c_double),
                ("eewald_reals", C.c_double), ("eewald_intra", C.c_double), ("vir", C.c_double * 6),
                ("hes", C.c_double * 3)]


class Ran2State(C.Structure):
    _fields_ = [("idum", C.c_long), ("idum2", C.c_long), ("iy", C.c_long), ("iv", C.c_long * 32)]


class WlParams(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("kbond", "r0bond", "eps", "sig_rep", "pivot_frac", "theta_max",
                                          "disp_max", "emin", "ebin")] + [("nbins", C.c_int)]


def build(fast: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _load(name: str):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    return C.CDLL(path)


_lib = None
_fast = None


def lib(fast: bool = False):
    global _lib, _fast
    if fast:
        if _fast is None:
            _fast = _load("liboracle_fast.so")
            _declare(_fast)
        return _fast
    if _lib is None:
        _lib = _load("liboracle.so")
        _declare(_lib)
    return _lib


def _declare(L):
    P = C.POINTER
    L.orc_erfc.restype = C.c_double; L.orc_erfc.argtypes = [C.c_double]
    L.orc_ewald_self.restype = C.c_double
    L.orc_ewald_self.argtypes = [C.c_int, _dp, C.c_double, C.c_double]
    L.orc_check_neutral.restype = C.c_int; L.orc_check_neutral.argtypes = [C.c_int, _dp]
    L.orc_fillspline.argtypes = [C.c_double, C.c_int, _dp, _dp]
    L.orc_dftmod.argtypes = [_dp, _dp, C.c_int]
    L.orc_pme_bspline.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    L.orc_fft3d.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_dft3d_naive.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_excl_build.restype = C.c_int
    L.orc_excl_build.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip, P(OrcExcl)]
    L.orc_excl_free.argtypes = [P(OrcExcl)]
    L.orc_nblist.restype = C.c_int
    L.orc_nblist.argtypes = [P(OrcSystem), _dp, P(OrcExcl), C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip]
    L.orc_nblist_cells.restype = C.c_int
    L.orc_nblist_cells.argtypes = [P(OrcSystem), _dp, P(OrcExcl), C.c_int, _ip, _ip, _ip]
    L.orc_nl_check.restype = C.c_int
    L.orc_nl_check.argtypes = [P(OrcSystem), _dp, _dp, C.c_int, C.c_int]
    L.orc_pme_calc.argtypes = [P(OrcSystem), _dp, C.c_double, _dp, _dp, _dp, C.c_void_p]
    L.orc_cnonbondnl.argtypes = [P(OrcSystem), _dp, _ip, _ip, _ip, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    L.orc_cbond.restype = C.c_double
    L.orc_cbond.argtypes = [P(OrcSystem), _dp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
    L.orc_pbc_all.argtypes = [P(OrcSystem), _dp, C.c_int]
    L.orc_drift.argtypes = [C.c_int, _dp, _dp]
    L.orc_cbend.restype = C.c_double
    L.orc_cbend.argtypes = [P(OrcSystem), _dp, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
    L.orc_curey.restype = C.c_double
    L.orc_curey.argtypes = [P(OrcSystem), _dp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
    L.orc_ctorsion.restype = C.c_double
    L.orc_ctorsion.argtypes = [P(OrcSystem), _dp, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    L.orc_cimproper.restype = C.c_double
    L.orc_cimproper.argtypes = [P(OrcSystem), _dp, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
    for nm, args in (("orc_cbend2", [P(OrcSystem), _dp, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp]),
                     ("orc_curey2", [P(OrcSystem), _dp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]),
                     ("orc_ctorsion2", [P(OrcSystem), _dp, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]),
                     ("orc_cimproper2", [P(OrcSystem), _dp, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp])):
        getattr(L, nm).restype = C.c_double
        getattr(L, nm).argtypes = args
    L.orc_kinet.restype = C.c_double
    L.orc_kinet.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    L.orc_vverlet_a.argtypes = [P(OrcSystem), C.c_int, _dp, C.c_double, C.c_double, _dp, C.c_void_p, _dp, _dp]
    L.orc_vverlet_b.argtypes = [C.c_int, _dp, C.c_double, C.c_double, _dp, _dp, _dp]
    L.orc_nhc.restype = C.c_double
    L.orc_nhc.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                          C.c_int, C.c_int, _dp, _dp, C.c_double]
    L.orc_go_energy.restype = C.c_double
    L.orc_go_energy.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_int, _ip, _ip, _dp,
                                C.c_double, C.c_double, C.c_double]
    L.orc_go_range_energy.restype = C.c_double
    L.orc_go_range_energy.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_double, C.c_double, _u8,
                                      _dp, C.c_double, C.c_double, C.c_double]
    L.orc_ran2_seed.argtypes = [P(Ran2State), C.c_long]
    L.orc_ran2.restype = C.c_double; L.orc_ran2.argtypes = [P(Ran2State)]
    L.orc_fillspline2.argtypes = [C.c_double, C.c_int, _dp, _dp, _dp]
    L.orc_pair.argtypes = [P(OrcSystem), _dp, _ip, _ip, _ip, C.c_int, _dp, P(OrcPairOut)]
    L.orc_pme_hessian.argtypes = [P(OrcSystem), _dp, _dp, C.c_void_p]
    L.orc_cbond2.restype = C.c_double
    L.orc_cbond2.argtypes = [P(OrcSystem), _dp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
    L.orc_cnonbondnl2.argtypes = [P(OrcSystem), _dp, _ip, _ip, _ip, C.c_int, C.c_double, _dp, _dp, _dp]
    L.orc_splitmix64.restype = C.c_uint64; L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
    L.orc_xoshiro_next.restype = C.c_uint64; L.orc_xoshiro_next.argtypes = [C.POINTER(C.c_uint64)]
    L.orc_xoshiro_uniform.restype = C.c_double; L.orc_xoshiro_uniform.argtypes = [C.POINTER(C.c_uint64)]
    L.orc_dos_interp.restype = C.c_double
    L.orc_dos_interp.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_int, C.c_double]
    L.orc_wl_run.restype = C.c_int
    L.orc_wl_run.argtypes = [C.POINTER(WlParams), C.c_int, _dp, C.POINTER(C.c_double), _u8, _dp, _dp,
                             C.POINTER(C.c_longlong),
                             C.POINTER(C.c_uint64), C.c_int, C.c_double, C.POINTER(C.c_longlong)]
    L.orc_mlp_nparams.restype = C.c_int; L.orc_mlp_nparams.argtypes = [C.c_int, C.c_int]
    L.orc_mlp_forward.restype = C.c_double
    L.orc_mlp_forward.argtypes = [C.c_int, C.c_int, _dp, _dp]
    L.orc_mlp_loss_grad.restype = C.c_double
    L.orc_mlp_loss_grad.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, _dp]


class Oracle:
    """Holds an orc_system built from a ChainSystem-like object and keeps the
    numpy arrays it points into alive."""

    def __init__(self, sysm, fast: bool = False):
        self.L = lib(fast)
        self.s = sysm
        self._keep = []

        def dp(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        t = np.ascontiguousarray(sysm.type, dtype=np.int32)
        self._keep.append(t)
        self.cs = OrcSystem(
            n=sysm.n, lx=sysm.box[0], ly=sysm.box[1], lz=sysm.box[2], rc=sysm.rc,
            rclist=sysm.rclist, rclist_ew=sysm.rclist_ew, kappa=sysm.kappa, eps_rf=sysm.eps_rf,
            e14fac=sysm.e14fac, order=sysm.order, gx=sysm.grid[0], gy=sysm.grid[1], gz=sysm.grid[2],
            ntypes=sysm.ntypes, type=t.ctypes.data_as(C.POINTER(C.c_int)), q=dp(sysm.q),
            lj_eps=dp(sysm.lj_eps), lj_sig=dp(sysm.lj_sig), lj14_eps=dp(sysm.lj14_eps),
            lj14_sig=dp(sysm.lj14_sig))
        x = np.ascontiguousarray(sysm.xint, dtype=np.int32)
        self.ex = OrcExcl()
        xa, xb, xf = (np.ascontiguousarray(x[:, i]) for i in range(3))
        if self.L.orc_excl_build(sysm.n, len(x), xa, xb, xf, C.byref(self.ex)) != 0:
            raise ValueError("bad exclusion list")
        self.e_self = self.L.orc_ewald_self(sysm.n, np.ascontiguousarray(sysm.q, dtype=np.float64),
                                            sysm.kappa, sysm.eps_rf)

    def __del__(self):
        try:
            self.L.orc_excl_free(C.byref(self.ex))
        except Exception:
            pass

    def nblist(self, pos, max_nb=1500, use_ew=True, cells=False):
        n = self.s.n
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        cnt = np.zeros(n, np.int32); lst = np.zeros(n * max_nb, np.int32); xf = np.zeros(n * max_nb, np.int32)
        if cells:
            rc = self.L.orc_nblist_cells(C.byref(self.cs), pos, C.byref(self.ex), max_nb, cnt, lst, xf)
            return rc, cnt, lst.reshape(n, max_nb), xf.reshape(n, max_nb)
        ce = np.zeros(n, np.int32); le = np.zeros(n * max_nb, np.int32); xe = np.zeros(n * max_nb, np.int32)
        rc = self.L.orc_nblist(C.byref(self.cs), pos, C.byref(self.ex), int(use_ew), max_nb,
                               cnt, lst, xf, ce, le, xe)
        return (rc, cnt, lst.reshape(n, max_nb), xf.reshape(n, max_nb),
                ce, le.reshape(n, max_nb), xe.reshape(n, max_nb))

    def cnonbondnl(self, pos, cnt, lst, xf):
        n = self.s.n
        max_nb = lst.shape[1]
        ff = np.zeros(n * 3)
        en = np.zeros(8); vn = np.zeros(6); vk = np.zeros(6)
        self.L.orc_cnonbondnl(C.byref(self.cs), np.ascontiguousarray(pos, dtype=np.float64),
                              np.ascontiguousarray(cnt, dtype=np.int32),
                              np.ascontiguousarray(lst, dtype=np.int32).ravel(),
                              np.ascontiguousarray(xf, dtype=np.int32).ravel(), max_nb,
                              self.e_self, ff, en, vn, vk)
        return ff.reshape(n, 3), en, vn, vk

    def pme(self, pos, want_grid=False):
        n = self.s.n
        ff = np.zeros(n * 3); en = np.zeros(3); vir = np.zeros(6)
        g = None
        if want_grid:
            g = np.zeros(2 * int(np.prod(self.s.grid)))
            self.L.orc_pme_calc(C.byref(self.cs), np.ascontiguousarray(pos, dtype=np.float64),
                                self.e_self, ff, en, vir, g.ctypes.data_as(C.c_void_p))
            g = g[0::2].reshape(self.s.grid)
        else:
            self.L.orc_pme_calc(C.byref(self.cs), np.ascontiguousarray(pos, dtype=np.float64),
                                self.e_self, ff, en, vir, None)
        return ff.reshape(n, 3), en, vir, g

    def bonds(self, pos):
        n = self.s.n
        ff = np.zeros(n * 3); vir = np.zeros(6)
        b = np.ascontiguousarray(self.s.bonds, dtype=np.int32)
        e = self.L.orc_cbond(C.byref(self.cs), np.ascontiguousarray(pos, dtype=np.float64), len(b),
                             np.ascontiguousarray(b[:, 0]), np.ascontiguousarray(b[:, 1]),
                             np.ascontiguousarray(self.s.bond_r0, dtype=np.float64),
                             np.ascontiguousarray(self.s.bond_k, dtype=np.float64), ff, vir)
        return e, ff.reshape(n, 3), vir

    def forces(self, pos, max_nb=1500, lists=None):
        """forces.c:118-219 for a system with bonds only: returns (ff, en dict)."""
        if lists is None:
            rc, cnt, lst, xf, *_ = self.nblist(pos, max_nb)
            if rc:
                raise RuntimeError(f"neighbour list overflow (reference exit {rc})")
        else:
            cnt, lst, xf = lists
        ff, en, vn, vk = self.cnonbondnl(pos, cnt, lst, xf)
        eb, fb, _ = self.bonds(pos)
        return ff + fb, {"bond": eb, "nbond": en[0], "nbonds": en[1], "coulomb": en[2],
                         "rewalds": en[3], "kewald": en[4], "sewald": en[5], "iewald": en[6],
                         "rewald": en[7], "slots": en, "vir_nbond": vn, "vir_kspace": vk}


    # ---- CONFIGT (cnonbond_nblist2.c, cbond2.c) ----
    def angles(self, pos, at, which=("bend", "urey", "tors", "impr"), hessian=False):
        """Angle-type terms of system.AngleTerms `at` at pos: forces (n,3)
        summed over the chosen kinds, energies and virials per kind (with
        hessian: the CONFIGT variants, Hessian diagonal sums per kind)."""
        sfx = "2" if hessian else ""
        L = self.L
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        ff = np.zeros(self.s.n * 3)
        en, vir = {}, {}
        col = lambda a, k: np.ascontiguousarray(a[:, k], dtype=np.int32)
        for w in which:
            v = np.zeros(3 if hessian else 6)
            if w == "bend":
                e = getattr(L, 'orc_cbend' + sfx)(C.byref(self.cs), pos, len(at.bend), col(at.bend, 0), col(at.bend, 1),
                                col(at.bend, 2), np.ascontiguousarray(at.bend_eq), np.ascontiguousarray(at.bend_k), ff, v)
            elif w == "urey":
                e = getattr(L, 'orc_curey' + sfx)(C.byref(self.cs), pos, len(at.urey), col(at.urey, 0), col(at.urey, 1),
                                np.ascontiguousarray(at.urey_s0), np.ascontiguousarray(at.urey_k), ff, v)
            elif w == "tors":
                e = getattr(L, 'orc_ctorsion' + sfx)(C.byref(self.cs), pos, len(at.tors), col(at.tors, 0), col(at.tors, 1),
                                   col(at.tors, 2), col(at.tors, 3), np.ascontiguousarray(at.kphi).ravel(),
                                   np.ascontiguousarray(at.nphi).ravel(), np.ascontiguousarray(at.delphi).ravel(), ff, v)
            else:
                e = getattr(L, 'orc_cimproper' + sfx)(C.byref(self.cs), pos, len(at.impr), col(at.impr, 0), col(at.impr, 1),
                                    col(at.impr, 2), col(at.impr, 3), np.ascontiguousarray(at.impr_k),
                                    np.ascontiguousarray(at.impr_eq), ff, v)
            en[w] = e
            vir[w] = v
        return ff.reshape(-1, 3), en, vir

    def pair(self, pos, cnt, lst, xf):
        """The pair loop alone: forces and the orc_pair_out fields."""
        n = self.s.n
        ff = np.zeros(n * 3)
        po = OrcPairOut()
        self.L.orc_pair(C.byref(self.cs), np.ascontiguousarray(pos, dtype=np.float64).ravel(),
                        np.ascontiguousarray(cnt, dtype=np.int32), np.ascontiguousarray(lst, dtype=np.int32).ravel(),
                        np.ascontiguousarray(xf, dtype=np.int32).ravel(), lst.shape[1], ff, C.byref(po))
        return ff.reshape(n, 3), po

    def pme_hessian(self, pos, per_site=False):
        h = np.zeros(3)
        ps = np.zeros(3 * self.s.n) if per_site else None
        self.L.orc_pme_hessian(C.byref(self.cs), np.ascontiguousarray(pos, dtype=np.float64).ravel(), h,
                               ps.ctypes.data_as(C.c_void_p) if per_site else None)
        return (h, ps.reshape(-1, 3)) if per_site else h

    def bonds2(self, pos):
        n = self.s.n
        ff = np.zeros(n * 3); h = np.zeros(3)
        b = np.ascontiguousarray(self.s.bonds, dtype=np.int32)
        e = self.L.orc_cbond2(C.byref(self.cs), np.ascontiguousarray(pos, dtype=np.float64).ravel(), len(b),
                              np.ascontiguousarray(b[:, 0]), np.ascontiguousarray(b[:, 1]),
                              np.ascontiguousarray(self.s.bond_r0, dtype=np.float64),
                              np.ascontiguousarray(self.s.bond_k, dtype=np.float64), ff, h)
        return e, ff.reshape(n, 3), h

    def cnonbondnl2(self, pos, cnt, lst, xf):
        n = self.s.n
        ff = np.zeros(n * 3); en = np.zeros(8); h = np.zeros(3)
        self.L.orc_cnonbondnl2(C.byref(self.cs), np.ascontiguousarray(pos, dtype=np.float64).ravel(),
                               np.ascontiguousarray(cnt, dtype=np.int32),
                               np.ascontiguousarray(lst, dtype=np.int32).ravel(),
                               np.ascontiguousarray(xf, dtype=np.int32).ravel(), lst.shape[1], self.e_self, ff, en, h)
        return ff.reshape(n, 3), en, h


# --------------------------------------------------------------------------
# oracle/_ref: the reference's own routines compiled from its sources
# --------------------------------------------------------------------------

REF_SO = os.path.join(HERE, "_ref", "libttcref.so")


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    L = C.CDLL(REF_SO)
    L.ref_setup.restype = C.c_int
    L.ref_setup.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                            C.c_double, C.c_double, C.c_double, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp,
                            C.c_int, _ip, _ip, _ip, C.c_int]
    L.ref_nblist.argtypes = [_dp, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip]
    L.ref_nl_check.restype = C.c_int
    L.ref_nl_check.argtypes = [_dp, C.c_int, C.c_int]
    L.ref_ewald_setup.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.ref_cnonbondnl.argtypes = [_dp, _dp, _dp, _dp]
    L.ref_ewald_kspace.argtypes = [_dp, _dp, _dp]
    L.ref_erfc.restype = C.c_double; L.ref_erfc.argtypes = [C.c_double]
    L.ref_self_energy.restype = C.c_double
    L.ref_ran2_seq.argtypes = [C.c_long, C.c_int, _dp]
    L.ref_vverlet.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
    L.ref_pbc_all.argtypes = [_dp]
    L.ref_kinet.restype = C.c_double
    L.ref_kinet.argtypes = [_dp, _dp, _dp, _dp]
    L.ref_bonded_box.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    L.ref_drift.argtypes = [_dp, _dp]
    L.ref_cbend.restype = C.c_double
    L.ref_cbend.argtypes = [_dp, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
    L.ref_curey.restype = C.c_double
    L.ref_curey.argtypes = [_dp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
    L.ref_ctorsion.restype = C.c_double
    L.ref_ctorsion.argtypes = [_dp, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    L.ref_cimproper.restype = C.c_double
    L.ref_cimproper.argtypes = [_dp, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
    return L


REF_DOS_SO = os.path.join(HERE, "_ref", "libttcref_dos.so")


def ref_dos_lib():
    """The reference's own dos_interp (dos_interpolate.c, DOS build)."""
    L = C.CDLL(REF_DOS_SO)
    L.ref_dos_interp.restype = C.c_double
    L.ref_dos_interp.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_int, C.c_double]
    return L


def _u64(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def splitmix_seq(seed: int, count: int) -> np.ndarray:
    """count consecutive splitmix64 outputs from state seed."""
    L = lib()
    x = C.c_uint64(seed)
    return np.array([L.orc_splitmix64(C.byref(x)) for _ in range(count)], dtype=np.uint64)


def xoshiro_advance(state: np.ndarray, draws: int) -> np.ndarray:
    """The (..., 4) xoshiro256** states after `draws` outputs each."""
    L = lib()
    s = np.ascontiguousarray(state, dtype=np.uint64).reshape(-1, 4).copy()
    for row in s:
        for _ in range(draws):
            L.orc_xoshiro_next(_u64(row))
    return s.reshape(state.shape)


def go_tables(n, contacts, r0):
    """n x n native mask and native distances of a Go chain."""
    nat = np.zeros((n, n), np.uint8)
    r0m = np.zeros((n, n))
    for (a, b), r in zip(contacts, r0):
        nat[a, b] = nat[b, a] = 1
        r0m[a, b] = r0m[b, a] = r
    return nat, r0m


def wl_run(params: WlParams, pos, ecur, nat, r0m, lng, hist, streams, ntrials, lnf, stats, fast=False):
    """orc_wl_run on numpy arrays, all updated in place; returns the new ecur.
    fast: the build with the reference's own compiler flags (CPU baseline)."""
    L = lib(fast)
    e = C.c_double(ecur)
    for a, dt in ((pos, np.float64), (nat, np.uint8), (r0m, np.float64), (lng, np.float64), (hist, np.int64),
                  (streams, np.uint64), (stats, np.int64)):
        assert a.dtype == dt and a.flags.c_contiguous
    assert lng.size == params.nbins and hist.size == params.nbins and streams.size == 256 and stats.size == 7
    n = pos.size // 3
    assert nat.size == n * n and r0m.size == n * n
    rc = L.orc_wl_run(C.byref(params), n, pos, C.byref(e), nat, r0m, lng,
                      hist.ctypes.data_as(C.POINTER(C.c_longlong)), _u64(streams), ntrials, lnf,
                      stats.ctypes.data_as(C.POINTER(C.c_longlong)))
    assert rc == 0
    return e.value


REF_CT_SO = os.path.join(HERE, "_ref", "libttcref_ct.so")


def ref_ct_lib():
    """The reference's own CONFIGT bonded terms (cbend2.c, curey2.c,
    ctorsion2.c, cimproper2.c)."""
    L = C.CDLL(REF_CT_SO)
    L.ref_ct_box.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    L.ref_cbend2.restype = C.c_double
    L.ref_cbend2.argtypes = [_dp, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
    L.ref_curey2.restype = C.c_double
    L.ref_curey2.argtypes = [_dp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
    L.ref_ctorsion2.restype = C.c_double
    L.ref_ctorsion2.argtypes = [_dp, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    L.ref_cimproper2.restype = C.c_double
    L.ref_cimproper2.argtypes = [_dp, C.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
    return L
