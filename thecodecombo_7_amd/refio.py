"""Reference-format inputs and trajectories.

read_inputs(input_dir)  reads a reference input directory (simul.input,
                        mol.input, the parameter file, simul.psf,
                        simul<k>.crd) through the native readers of libttc.so
                        (include/ttc_io.h) into a ChainSystem + AngleTerms, so
                        that a reference-format system drives the device path
                        (nonbond.Box, the MD loops, the replica batch).
RefInputs.write(dir)    writes a system in those formats (the inverse, used to
                        export synthetic systems and to build test inputs).
write_trr(path, ...)    appends one .trr frame as write_trr.c writes it.

Units and conventions are the reference's (read_atom.c, read_topology.c,
setlj.c): parameter files hold CHARMM values (kcal/mol, degrees, LJ eps
negative and Rmin/2); the readers convert to kJ/mol and radians.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .system import KCAL, AngleTerms, ChainSystem

WILD = 100          # wildcard atom type of torsion / improper rows


class IoCounts(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("nsite", "ncomp", "nbond", "nbend", "nurey", "ntors", "nimpr", "nexcl",
                                       "n14", "nxint", "ntype")]


class IoSim(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("lx", "ly", "lz", "rc", "rclist", "rc_ew", "rclist_ew", "kappa",
                                          "eps_rf", "dt", "temp", "nhc_q")] + \
               [("order", C.c_int), ("grid", C.c_int * 3), ("nhc_m", C.c_int), ("respa", C.c_int),
                ("nboxes", C.c_int), ("mode", C.c_int), ("intmode", C.c_int)]


_bound = False


def _io():
    global _bound
    L = _lib.lib()
    if not _bound:
        dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
        L.ttc_io_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp)]
        L.ttc_io_free.argtypes = [vp]
        L.ttc_io_last_error.restype = C.c_char_p
        L.ttc_io_get_counts.argtypes = [vp, C.POINTER(IoCounts)]
        L.ttc_io_get_sim.argtypes = [vp, C.POINTER(IoSim)]
        L.ttc_io_get_sites.argtypes = [vp, dp, dp, dp, dp, ip, ip, ip, C.c_char_p]
        L.ttc_io_get_types.argtypes = [vp, ip, dp, dp, dp, dp]
        L.ttc_io_get_bonds.argtypes = [vp, ip, dp, dp]
        L.ttc_io_get_bends.argtypes = [vp, ip, dp, dp]
        L.ttc_io_get_ureys.argtypes = [vp, ip, dp, dp]
        L.ttc_io_get_torsions.argtypes = [vp, ip, dp, dp, dp]
        L.ttc_io_get_impropers.argtypes = [vp, ip, dp, dp]
        L.ttc_io_get_exclusions.argtypes = [vp, ip]
        L.ttc_io_get_14.argtypes = [vp, ip]
        L.ttc_io_get_xint.argtypes = [vp, ip, ip, ip]
        L.ttc_io_write_trr.argtypes = [C.c_char_p, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, C.c_double,
                                       C.c_double, C.c_double, C.c_long, C.c_double]
        _bound = True
    return L


class IoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ttc_io error {code}: {msg}")
        self.code = code


def _check(code: int):
    if code:
        raise IoError(code, _io().ttc_io_last_error().decode())


def _dp(a):
    return None if a is None else _lib.dptr(a)


def _ip(a):
    return None if a is None else _lib.iptr(a)


def read_raw(input_dir: str, ibox: int = 0) -> dict:
    """Every array the native readers return for box ibox, as numpy arrays
    (the reference's units and orders; see include/ttc_io.h)."""
    L = _io()
    h = C.c_void_p()
    _check(L.ttc_io_read(os.fsencode(input_dir), ibox, C.byref(h)))
    try:
        c = IoCounts()
        m = IoSim()
        _check(L.ttc_io_get_counts(h, C.byref(c)))
        _check(L.ttc_io_get_sim(h, C.byref(m)))
        n, T = c.nsite, c.ntype
        out = {"counts": {k: getattr(c, k) for k, _ in IoCounts._fields_},
               "sim": {k: (list(getattr(m, k)) if k == "grid" else getattr(m, k)) for k, _ in IoSim._fields_}}
        d = lambda *s: np.zeros(s, np.float64)
        i = lambda *s: np.zeros(s, np.int32)
        xn, x, q, mass = d(n, 3), d(n, 3), d(n), d(n)
        atomid, typ, res = i(n), i(n), i(n)
        names = C.create_string_buffer(5 * n + 1)
        _check(L.ttc_io_get_sites(h, _dp(xn), _dp(x), _dp(q), _dp(mass), _ip(atomid), _ip(typ), _ip(res), names))
        out.update(pos_nopbc=xn, pos=x, q=q, mass=mass, atomid=atomid, type=typ, resid=res,
                   names=[names.raw[5 * k:5 * k + 5].split(b"\0")[0].decode() for k in range(n)])
        tid, e, s, e14, s14 = i(T), d(T, T), d(T, T), d(T, T), d(T, T)
        _check(L.ttc_io_get_types(h, _ip(tid), _dp(e), _dp(s), _dp(e14), _dp(s14)))
        out.update(type_atomid=tid, lj_eps=e, lj_sig=s, lj14_eps=e14, lj14_sig=s14)
        ab, r0, k = i(c.nbond, 2), d(c.nbond), d(c.nbond)
        _check(L.ttc_io_get_bonds(h, _ip(ab), _dp(r0), _dp(k)))
        out.update(bonds=ab, bond_r0=r0, bond_k=k)
        abc, eq, kb = i(c.nbend, 3), d(c.nbend), d(c.nbend)
        _check(L.ttc_io_get_bends(h, _ip(abc), _dp(eq), _dp(kb)))
        out.update(bends=abc, bend_eq=eq, bend_k=kb)
        ub, s0, ku = i(c.nurey, 2), d(c.nurey), d(c.nurey)
        _check(L.ttc_io_get_ureys(h, _ip(ub), _dp(s0), _dp(ku)))
        out.update(ureys=ub, urey_s0=s0, urey_k=ku)
        tq, kp, npf, dlt = i(c.ntors, 4), d(c.ntors, 6), d(c.ntors, 6), d(c.ntors, 6)
        _check(L.ttc_io_get_torsions(h, _ip(tq), _dp(kp), _dp(npf), _dp(dlt)))
        out.update(tors=tq, kphi=kp, nphi=npf, delphi=dlt)
        iq, ki, ei = i(c.nimpr, 4), d(c.nimpr), d(c.nimpr)
        _check(L.ttc_io_get_impropers(h, _ip(iq), _dp(ki), _dp(ei)))
        out.update(impr=iq, impr_k=ki, impr_eq=ei)
        ex = i(c.nexcl, 2)
        _check(L.ttc_io_get_exclusions(h, _ip(ex)))
        p14 = i(c.n14, 4)
        _check(L.ttc_io_get_14(h, _ip(p14)))
        xa, xb, xf = i(c.nxint), i(c.nxint), i(c.nxint)
        _check(L.ttc_io_get_xint(h, _ip(xa), _ip(xb), _ip(xf)))
        out.update(exclusions=ex, in14=p14, xint=np.column_stack([xa, xb, xf]).astype(np.int32))
        return out
    finally:
        L.ttc_io_free(h)


def read_inputs(input_dir: str, ibox: int = 0, *, e14fac: float = 0.4, seed: int = 0, name: str = ""):
    """A ChainSystem (+ AngleTerms when the psf has angle-type terms) from a
    reference input directory.  Velocities are drawn at the box temperature
    (vinit.c's Maxwell distribution, centre-of-mass momentum removed) from
    seed; the reference's E14FAC (0.4 in a non-BGO build) scales 1-4
    charges."""
    r = read_raw(input_dir, ibox)
    m, n = r["sim"], r["counts"]["nsite"]
    box = np.array([m["lx"], m["ly"], m["lz"]])
    rng = np.random.default_rng(seed)
    temp = m["temp"]
    kT_over_m = 1.380662e-23 * temp / (r["mass"] * 1.6605655e-27) * 1e-4
    vel = rng.standard_normal((n, 3)) * np.sqrt(kT_over_m)[:, None]
    vel -= (r["mass"][:, None] * vel).sum(0) / r["mass"].sum()
    rc = m["rc"]
    s = ChainSystem(
        n=n, box=box, pos=np.ascontiguousarray(r["pos"]), vel=vel, mass=r["mass"], q=r["q"],
        type=r["type"].astype(np.int32), ntypes=len(r["type_atomid"]), lj_eps=r["lj_eps"], lj_sig=r["lj_sig"],
        lj14_eps=r["lj14_eps"], lj14_sig=r["lj14_sig"], bonds=r["bonds"], bond_r0=r["bond_r0"], bond_k=r["bond_k"],
        xint=r["xint"], rc=rc, rclist=m["rclist"], rclist_ew=m["rclist_ew"] or m["rclist"], kappa=m["kappa"],
        eps_rf=m["eps_rf"] or 1.0, e14fac=e14fac, order=m["order"], grid=tuple(m["grid"]), dt=m["dt"],
        temperature=temp, name=name or os.path.basename(os.path.normpath(input_dir)),
        meta={"atomid": r["atomid"], "type_atomid": r["type_atomid"], "names": r["names"], "resid": r["resid"]})
    angles = None
    # angle-type terms whose force constants are all zero (a topology listed
    # only for its exclusions and 1-4 pairs) contribute nothing: left out
    if np.any(r["bend_k"]) or np.any(r["urey_k"]) or np.any(r["kphi"]) or np.any(r["impr_k"]):
        angles = AngleTerms(
            bend=r["bends"], bend_eq=r["bend_eq"], bend_k=r["bend_k"], urey=r["ureys"], urey_s0=r["urey_s0"],
            urey_k=r["urey_k"], tors=r["tors"], kphi=r["kphi"], nphi=r["nphi"], delphi=r["delphi"],
            impr=r["impr"], impr_k=r["impr_k"], impr_eq=r["impr_eq"])
    return s, angles


def write_trr(path: str, pos: np.ndarray, box, *, vel: np.ndarray | None = None, force: np.ndarray | None = None,
              step: int = 0, time: float = 0.0) -> None:
    """Append one frame (write_trr.c layout) of n x 3 arrays in A, A/ps and
    kJ/mol/A; the file holds each value divided by 10 (nm for positions and
    the box), rounded to float and stored as a big-endian double, as
    write_trr.c:122-170 writes them."""
    pos = np.ascontiguousarray(pos, np.float64)
    vel = None if vel is None else np.ascontiguousarray(vel, np.float64)
    force = None if force is None else np.ascontiguousarray(force, np.float64)
    _check(_io().ttc_io_write_trr(os.fsencode(path), len(pos), _dp(pos), 3, _dp(vel), 3, _dp(force), 3,
                                  float(box[0]), float(box[1]), float(box[2]), int(step), float(time)))


def read_trr(path: str):
    """Frames of a .trr file written as write_trr.c writes it: a list of dicts
    with step, time, box (3,) and x/v/f (n x 3, in the file's units) where
    present."""
    with open(path, "rb") as fh:
        raw = fh.read()
    out, p = [], 0
    while p < len(raw):
        head = np.frombuffer(raw, ">i4", 3, p)
        if head[0] != 1993:
            raise ValueError(f"{path}: bad magic at byte {p}")
        # magic, version, title length, the title's bytes, then 13 sizes
        p += 12 + int(head[2])
        hdr = np.concatenate([head, np.frombuffer(raw, ">i4", 13, p)])
        p += 52
        t, lam = np.frombuffer(raw, ">f8", 2, p)
        p += 16
        bx = np.frombuffer(raw, ">f8", 9, p).reshape(3, 3)
        p += int(hdr[5])
        n = int(hdr[13])
        fr = {"step": int(hdr[14]), "time": float(t), "box": np.diag(bx).copy()}
        for key, size in (("x", hdr[10]), ("v", hdr[11]), ("f", hdr[12])):
            if size:
                fr[key] = np.frombuffer(raw, ">f8", 3 * n, p).reshape(n, 3).astype(np.float64)
                p += int(size)
        out.append(fr)
    return out


# --------------------------------------------------------------------------
# writing reference-format inputs
# --------------------------------------------------------------------------

@dataclass
class RefInputs:
    """A system in the reference's input vocabulary.

    atom types: ids, 4-character names, masses, LJ eps (kcal/mol, CHARMM
    sign: negative) and Rmin/2 (A), and the 1-4 pair; parameter rows by
    type name (bonds: k [kcal/mol/A^2], req; bends: k, theta0 [deg];
    torsions: k, n, delta [deg]; impropers: k, 0, psi0 [deg]; Urey-Bradley:
    k, S0; NBFIX: eps, sig).  Sites: per-component molecule counts and
    sizes, residue numbers and names, site names, atom ids, charges,
    coordinates; bonded lists with 0-based site indices; explicit
    exclusions as (a, b) pairs."""
    box: tuple
    type_id: list
    type_name: list
    type_mass: list
    type_eps: list
    type_rmin2: list
    type_eps14: list
    type_rmin2_14: list
    bond_rows: list = field(default_factory=list)
    bend_rows: list = field(default_factory=list)
    tors_rows: list = field(default_factory=list)
    impr_rows: list = field(default_factory=list)
    urey_rows: list = field(default_factory=list)
    nbfix_rows: list = field(default_factory=list)
    comp_nmol: list = field(default_factory=list)
    comp_nsite: list = field(default_factory=list)
    comp_nres: list = field(default_factory=list)
    nsolute: int = 1
    resno: np.ndarray = None
    resname: list = None
    site_name: list = None
    atomid: np.ndarray = None
    q: np.ndarray = None
    pos: np.ndarray = None
    bonds: np.ndarray = None
    bends: np.ndarray = None
    tors: np.ndarray = None
    impr: np.ndarray = None
    exclusions: np.ndarray = None
    sim: dict = field(default_factory=dict)
    param_name: str = "ttc.param"

    # -- simul.input ------------------------------------------------------
    def _sim_positional(self) -> str:
        g = self.sim
        lx, ly, lz = self.box
        head = [f"* simul.input written by thecodecombo_7_amd.refio ({k})" for k in range(12)]
        lines = head + [
            g.get("title", "synthetic chains") + " !",
            f"Parameter File Name={self.param_name}",
            f"{g.get('mode', 1)}\t\t! sim.ID", f"{g.get('intmode', 1)}\t\t! sim.ID2",
            "0\t\t! cyc_eq", "1000\t\t! cyc_pr", "0\t\t! cyc_hybmd", "0\t\t! cyc_swap", "100\t\t! blockd",
            "100\t100\t! blockc blockt", "100\t\t! drift", f"{g.get('dt', 0.001)!r}\t\t! dt [ps]",
            "1\t\t! nsteps", f"{g['rc']!r} {g['rclist']!r}\t! rc rclist",
            f"{g.get('rc_ew', g['rc'])!r} {g.get('rclist_ew', g['rclist'])!r}\t! rc_ew rclist_ew",
            "1\t\t! nlist", "1\t\t! NB", f"{g.get('eps_rf', 1.0)!r}\t\t! epsRF", "! box 0",
            "0\t\t! box number", f"{lx!r} {ly!r} {lz!r}\t! box size",
            f"{g.get('temp', 300.0)!r}\t\t! T", "0.0\t\t! dtemp", "0.2\t\t! Ttau",
            f"{g.get('nhc_m', 2)}\t\t! NHC M", f"{g.get('nhc_q', 1.0e-18)!r}\t\t! NHC Q",
            "0.0 0.0 0.0\t! P", "0.0 0.0 0.0\t! dpress", "0.5 0.5 0.5\t! Ptau", "1e-06 1e-06 1e-06\t! Pcomp",
            "7.813381e-22\t! Wg", f"{g.get('respa', 0)}\t\t! xRESPA", f"{g['kappa']!r}\t\t! kappa",
            "6\t6\t6\t! kmax", f"{g.get('order', 4)}\t\t! spline order",
            "{} {} {}\t! grid".format(*g["grid"]), f"{g.get('eps_rf', 1.0)!r}\t\t! epsRF", ""]
        return "\n".join(lines)

    def _sim_keywords(self) -> str:
        g = self.sim
        lx, ly, lz = self.box
        return "\n".join([
            "# simul.input (keyword form) written by thecodecombo_7_amd.refio",
            f"sim_paramfile {self.param_name}", f"sim_mode {g.get('mode', 1)}",
            f"sim_intmode {g.get('intmode', 1)}", f"sim_timestep {g.get('dt', 0.001)!r}",
            f"sim_cutoff {g['rc']!r}", f"sim_cutoff_nblist {g['rclist']!r}",
            f"sim_cutoff_ewald {g.get('rc_ew', g['rc'])!r}",
            f"sim_cutoff_ewald_nblist {g.get('rclist_ew', g['rclist'])!r}", "sim_nboxes 1",
            "BOX_NUMBER 0", f"box_size_xyz {lx!r} {ly!r} {lz!r}", f"box_temp {g.get('temp', 300.0)!r}",
            f"box_nh_tstat_n {g.get('nhc_m', 2)}", f"box_nh_tstat_mass {g.get('nhc_q', 1.0e-18)!r}",
            f"box_respa {g.get('respa', 0)}", f"box_kappa {g['kappa']!r}",
            f"box_spme_spline_order {g.get('order', 4)}", "box_grid_pts_xyz {} {} {}".format(*g["grid"]),
            f"box_dielectric {g.get('eps_rf', 1.0)!r}", ""])

    # -- parameter file -----------------------------------------------------
    def _param(self) -> str:
        out = ["* parameter file written by thecodecombo_7_amd.refio"]
        nt = len(self.type_id)
        out.append(f"{nt}\t! atom types")
        for i in range(nt):
            out.append(f"{self.type_id[i]} {self.type_name[i]} {self.type_mass[i]!r}")
        out.append(f"{len(self.bond_rows)}\t! bonds")
        out += [f"{a} {b} {k!r} {r!r}" for a, b, k, r in self.bond_rows]
        out.append(f"{len(self.bend_rows)}\t! bends")
        out += [f"{a} {b} {c} {k!r} {t!r}" for a, b, c, k, t in self.bend_rows]
        out.append(f"{len(self.tors_rows)}\t! torsions")
        out += [f"{a} {b} {c} {d} {k!r} {n!r} {p!r}" for a, b, c, d, k, n, p in self.tors_rows]
        out.append("NONBONDED eps Rmin/2")
        out += [f"{self.type_name[i]} {self.type_eps[i]!r} {self.type_rmin2[i]!r}" for i in range(nt)]
        out.append(f"{len(self.impr_rows)}\t! impropers")
        out += [f"{a} {b} {c} {d} {k!r} 0 {p!r}" for a, b, c, d, k, p in self.impr_rows]
        out.append("NONBONDED 1-4 eps Rmin/2")
        out += [f"{self.type_name[i]} {self.type_eps14[i]!r} {self.type_rmin2_14[i]!r}" for i in range(nt - 1)]
        out.append(f"{len(self.urey_rows)}\t! Urey-Bradley")
        out += [f"{a} {b} {c} {k!r} {s!r}" for a, b, c, k, s in self.urey_rows]
        if self.nbfix_rows:
            out.append("NBFIX")
            out.append(f"{len(self.nbfix_rows)}\t! pairs")
            out += [f"{a} {b} {e!r} {s!r}" for a, b, e, s in self.nbfix_rows]
        out.append("END")
        return "\n".join(out) + "\n"

    # -- mol.input ------------------------------------------------------------
    def _mol(self) -> str:
        nc = len(self.comp_nmol)
        out = ["* mol.input written by thecodecombo_7_amd.refio", "*",
               f"{nc}\t! number of molecule types", f"{sum(self.comp_nmol)}\t! total molecules", "! molecules"]
        out += [f"{v}\t! type {i}" for i, v in enumerate(self.comp_nmol)]
        out.append("! sites per molecule")
        out += [f"{v}\t! type {i}" for i, v in enumerate(self.comp_nsite)]
        out.append("! residues per molecule")
        out += [f"{v}\t! type {i}" for i, v in enumerate(self.comp_nres)]
        out.append("! fancy moves")
        out += [f"{self.nsolute}\t! Nsolute", "1\t! SITE1", "2\t! SITE2", "3\t! SITE3", ""]
        return "\n".join(out)

    # -- psf --------------------------------------------------------------------
    @staticmethod
    def _records(arr, per_line: int) -> list:
        if len(arr) == 0:
            return []
        flat = [" ".join(f"{v + 1:7d}" for v in rec) for rec in np.asarray(arr).reshape(len(arr), -1)]
        return ["  ".join(flat[i:i + per_line]) for i in range(0, len(flat), per_line)]

    def _psf(self) -> str:
        out = ["PSF", "", "       1 !NTITLE", " * written by thecodecombo_7_amd.refio", ""]
        k = 0
        for cc in range(len(self.comp_nmol)):
            ns = self.comp_nmol[cc] * self.comp_nsite[cc]
            out.append(f"{ns:8d} !NATOM")
            for _ in range(ns):
                out.append(f"{k + 1:8d} C{cc:<4d} {int(self.resno[k]):5d} {self.resname[k]:<4s} {self.site_name[k]:<4s}"
                           f" {int(self.atomid[k]):4d} {float(self.q[k])!r:>22s}   0.0   0")
                k += 1
            out.append("")
        for arr, per, tag in ((self.bonds, 4, "!NBOND: bonds"), (self.bends, 3, "!NTHETA: angles"),
                              (self.tors, 2, "!NPHI: dihedrals"), (self.impr, 2, "!NIMPHI: impropers")):
            arr = np.zeros((0, 2), np.int32) if arr is None else arr
            out.append(f"{len(arr):8d} {tag}")
            out += self._records(arr, per)
            out.append("")
        out += ["       0 !NDON: donors", "", "       0 !NACC: acceptors", ""]
        ex = np.zeros((0, 2), np.int32) if self.exclusions is None else np.asarray(self.exclusions)
        n = len(self.atomid)
        order = np.lexsort((np.arange(len(ex)), ex[:, 0])) if len(ex) else np.zeros(0, np.int64)
        ex = ex[order]
        out.append(f"{len(ex):8d} !NNB")
        vals = [f"{b + 1:8d}" for b in ex[:, 1]]
        out += ["".join(vals[i:i + 8]) for i in range(0, len(vals), 8)]
        bounds = np.searchsorted(ex[:, 0], np.arange(n), side="right") if len(ex) else np.zeros(n, np.int64)
        bv = [f"{int(b):8d}" for b in bounds]
        out += ["".join(bv[i:i + 8]) for i in range(0, len(bv), 8)]
        out.append("")
        return "\n".join(out) + "\n"

    def _crd(self) -> str:
        out = [f"{len(self.atomid)}"]
        for i, (x, y, z) in enumerate(self.pos):
            out.append(f"{i + 1:10d} {int(self.resno[i]):9d}  {self.resname[i]:<8s} {self.site_name[i]:<8s}"
                       f" {float(x)!r:>24s} {float(y)!r:>24s} {float(z)!r:>24s}  SEG {int(self.resno[i])}  0.0")
        return "\n".join(out) + "\n"

    def write(self, input_dir: str, *, keyword_sim: bool = False) -> str:
        os.makedirs(input_dir, exist_ok=True)
        files = {"simul.input": self._sim_keywords() if keyword_sim else self._sim_positional(),
                 "mol.input": self._mol(), self.param_name: self._param(), "simul.psf": self._psf(),
                 "simul0.crd": self._crd()}
        for name, text in files.items():
            with open(os.path.join(input_dir, name), "w") as f:
                f.write(text)
        return input_dir


def from_chain_system(s: ChainSystem, angles: AngleTerms | None = None, *, ncomp: int = 1,
                      torsion_terms: int = 3, exclusions: np.ndarray | None = None,
                      nbfix: list | None = None, seed: int = 11) -> RefInputs:
    """RefInputs for a linear-chain ChainSystem: one atom type per LJ type
    (ids 1..ntypes, names T01..; the wildcard X = 100 last), LJ parameters
    recovered from the pair tables' diagonals, bond rows per type pair from
    the bonds.  With angles, bends / Urey-Bradley / torsions / impropers of
    the same site tuples are written with parameters drawn per type tuple
    (seeded), since the parameter file can only hold type-based values;
    torsions get up to torsion_terms exact Fourier rows per type quad, and a
    wildcard row X-b-c-X covers the central pairs that have no exact row.
    Chains become ncomp molecule types of equal counts."""
    rng = np.random.default_rng(seed)
    T = s.ntypes
    ids = list(range(1, T + 1)) + [WILD]
    names = [f"T{t + 1:02d}" for t in range(T)] + ["X"]
    diag = np.arange(T)
    masses = [float(s.mass[np.flatnonzero(s.type == t)[0]]) if np.any(s.type == t) else 14.027 for t in range(T)]
    eps = [float(v) for v in -s.lj_eps[diag, diag] / KCAL] + [0.0]
    rmin2 = [float(v) for v in s.lj_sig[diag, diag] / 2.0] + [0.0]
    eps14 = [float(v) for v in -s.lj14_eps[diag, diag] / KCAL] + [0.0]
    rmin2_14 = [float(v) for v in s.lj14_sig[diag, diag] / 2.0] + [0.0]
    tn = lambda t: names[int(t)]
    # bonds: one row per type pair present
    bond_rows, seen = [], {}
    for (a, b), r0, k in zip(s.bonds, s.bond_r0, s.bond_k):
        key = tuple(sorted((int(s.type[a]), int(s.type[b]))))
        if key in seen:
            if seen[key] != (k, r0):
                raise ValueError("bond parameters differ within one type pair: not expressible by type")
            continue
        seen[key] = (k, r0)
        bond_rows.append((tn(key[0]), tn(key[1]), float(k / KCAL), float(r0)))
    # chains: runs of consecutive bonds i -> i+1
    nb = len(s.bonds)
    starts = [0] + [k for k in range(1, nb) if s.bonds[k, 0] != s.bonds[k - 1, 1]] + [nb]
    chains = [[int(s.bonds[c0, 0])] + [int(x) for x in s.bonds[c0:c1, 1]] for c0, c1 in zip(starts[:-1], starts[1:])]
    nchain = len(chains)
    nsite = len(chains[0])
    if any(len(c) != nsite for c in chains) or nchain * nsite != s.n:
        raise ValueError("from_chain_system needs equal linear chains covering every site")
    bend_rows, tors_rows, impr_rows, urey_rows = [], [], [], []
    bends = tors = impr = None
    if angles is None:
        # bonds-only chemistry: the chains' bends are still listed, since the
        # reference derives the 1-3 exclusions and the 1-4 pairs from them,
        # with zero force constants
        bends = np.array([c[j:j + 3] for c in chains for j in range(len(c) - 2)], np.int32).reshape(-1, 3)
        done = set()
        for a3, b3, c3 in bends:
            ta, tb, tc = int(s.type[a3]), int(s.type[b3]), int(s.type[c3])
            key = (min(ta, tc), tb, max(ta, tc))
            if key not in done:
                done.add(key)
                bend_rows.append((tn(key[0]), tn(tb), tn(key[2]), 0.0, 112.0))
    else:
        bends, tors, impr = angles.bend, angles.tors, angles.impr
        done = set()
        for a, b, c in bends:
            ta, tb, tc = int(s.type[a]), int(s.type[b]), int(s.type[c])
            key = (min(ta, tc), tb, max(ta, tc))
            if key in done:
                continue
            done.add(key)
            bend_rows.append((tn(key[0]), tn(tb), tn(key[2]), float(rng.uniform(40, 70)), float(rng.uniform(105, 118))))
            urey_rows.append((tn(key[0]), tn(tb), tn(key[2]), float(rng.uniform(5, 15)), float(rng.uniform(2.4, 2.7))))
        done = set()
        central = set()
        for a, b, c, d in tors:
            t = tuple(int(s.type[x]) for x in (a, b, c, d))
            key = min(t, t[::-1])
            if key in done:
                continue
            done.add(key)
            if rng.uniform() < 0.75:      # exact rows for most quads ...
                for m in range(1, 1 + int(rng.integers(1, torsion_terms + 1))):
                    tors_rows.append((tn(key[0]), tn(key[1]), tn(key[2]), tn(key[3]),
                                      float(rng.uniform(0.05, 0.3)), float(m), float(rng.choice([0.0, 180.0]))))
            else:                          # ... a wildcard row for the rest
                cb = tuple(sorted((key[1], key[2])))
                if cb not in central:
                    central.add(cb)
                    tors_rows.append(("X", tn(cb[0]), tn(cb[1]), "X", float(rng.uniform(0.05, 0.3)), 3.0, 0.0))
        # every central pair gets a wildcard row at the end, so each torsion
        # matches something (earlier wildcard rows shadow later exact rows of
        # the same central pair: the reference's order rule)
        for q in tors:
            cb = tuple(sorted((int(s.type[q[1]]), int(s.type[q[2]]))))
            if cb not in central:
                central.add(cb)
                tors_rows.append(("X", tn(cb[0]), tn(cb[1]), "X", float(rng.uniform(0.05, 0.3)), 2.0, 180.0))
        done = set()
        for a, b, c, d in impr:
            t = tuple(int(s.type[x]) for x in (a, b, c, d))
            key = min(t, t[::-1])
            if key in done:
                continue
            done.add(key)
            impr_rows.append((tn(key[0]), tn(key[1]), tn(key[2]), tn(key[3]), float(rng.uniform(1, 3)),
                              float(rng.uniform(-10, 10))))
    if nbfix:
        nbfix_rows = [(tn(a), tn(b), float(e), float(sg)) for a, b, e, sg in nbfix]
    else:
        nbfix_rows = []
    if nchain % ncomp:
        raise ValueError("the chain count must split evenly into the molecule types")
    per = nchain // ncomp
    resno = np.repeat(np.arange(1, nchain + 1), nsite)
    site_name = [f"C{j + 1}" for j in range(nsite)] * nchain
    order = np.concatenate(chains)
    if not np.array_equal(order, np.arange(s.n)):
        raise ValueError("from_chain_system needs chains stored site by site in order")
    return RefInputs(
        box=tuple(float(v) for v in s.box), type_id=ids, type_name=names, type_mass=masses + [0.0],
        type_eps=eps, type_rmin2=rmin2, type_eps14=eps14, type_rmin2_14=rmin2_14, bond_rows=bond_rows,
        bend_rows=bend_rows, tors_rows=tors_rows, impr_rows=impr_rows, urey_rows=urey_rows, nbfix_rows=nbfix_rows,
        comp_nmol=[per] * ncomp, comp_nsite=[nsite] * ncomp, comp_nres=[1] * ncomp, nsolute=ncomp,
        resno=resno, resname=["CHN"] * s.n, site_name=site_name, atomid=(s.type + 1).astype(np.int32), q=s.q,
        pos=s.pos, bonds=s.bonds, bends=bends, tors=tors, impr=impr, exclusions=exclusions,
        sim={"rc": float(s.rc), "rclist": float(s.rclist), "rc_ew": float(s.rc), "rclist_ew": float(s.rclist_ew),
             "kappa": float(s.kappa), "order": int(s.order), "grid": [int(g) for g in s.grid], "dt": float(s.dt),
             "temp": float(s.temperature), "eps_rf": float(s.eps_rf)})


__all__ = ["read_inputs", "read_raw", "write_trr", "read_trr", "RefInputs", "from_chain_system", "WILD"]
