"""Runs the reference's own readers (oracle/_ref/libttcref_io*.so, compiled
from read_sim.c, read_atom.c, read_config.c, read_topology.c, setlj.c,
pbc_init.c and write_trr.c) and dumps what they leave in the reference's
globals, for comparison with the native readers.

The reference's readers keep state in globals that a second read would
accumulate into, so each read runs in its own process:

    python tests/ref_io_dump.py read <case_dir> <out.npz> [kw]
    python tests/ref_io_dump.py trr <out_dir> <in.npz> [kw]

Test infrastructure only.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def lib(kw: bool):
    return C.CDLL(os.path.join(REF, "libttcref_io_kw.so" if kw else "libttcref_io.so"))


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def dump_read(case_dir: str, out: str, kw: bool) -> None:
    L = lib(kw)
    rc = L.ref_io_read(os.fsencode(case_dir))
    if rc:
        raise SystemExit(f"ref_io_read failed: {rc}")
    c = np.zeros(10, np.int32)
    L.ref_io_counts(ip(c))
    n, nc, nb, nbe, nu, nt, ni, ne, n14, nx = (int(v) for v in c)
    sd, si = np.zeros(12), np.zeros(9, np.int32)
    L.ref_io_sim(dp(sd), ip(si))
    xn, x, q, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n)
    atomid, res = np.zeros(n, np.int32), np.zeros(n, np.int32)
    names = C.create_string_buffer(5 * n + 1)
    L.ref_io_sites(dp(xn), dp(x), dp(q), dp(mass), ip(atomid), ip(res), names)
    ids = np.unique(atomid)
    ia = np.repeat(ids, len(ids)).astype(np.int32)
    ib = np.tile(ids, len(ids)).astype(np.int32)
    e, s, e14, s14 = (np.zeros(len(ia)) for _ in range(4))
    L.ref_io_lj(len(ia), ip(ia), ip(ib), dp(e), dp(s), dp(e14), dp(s14))
    T = len(ids)
    bonds, breq, bk = np.zeros((nb, 2), np.int32), np.zeros(nb), np.zeros(nb)
    L.ref_io_bonds(ip(bonds), dp(breq), dp(bk))
    bends, beq, bek = np.zeros((nbe, 3), np.int32), np.zeros(nbe), np.zeros(nbe)
    L.ref_io_bends(ip(bends), dp(beq), dp(bek))
    ureys, us0, uk = np.zeros((nu, 2), np.int32), np.zeros(nu), np.zeros(nu)
    L.ref_io_ureys(ip(ureys), dp(us0), dp(uk))
    tors, kphi, nphi, dphi = np.zeros((nt, 4), np.int32), np.zeros((nt, 6)), np.zeros((nt, 6)), np.zeros((nt, 6))
    L.ref_io_torsions(ip(tors), dp(kphi), dp(nphi), dp(dphi))
    impr, ik, ieq = np.zeros((ni, 4), np.int32), np.zeros(ni), np.zeros(ni)
    L.ref_io_impropers(ip(impr), dp(ik), dp(ieq))
    ex = np.zeros((ne, 2), np.int32)
    L.ref_io_exclusions(ip(ex))
    p14 = np.zeros((n14, 4), np.int32)
    L.ref_io_14(ip(p14))
    xa, xb, xf = (np.zeros(nx, np.int32) for _ in range(3))
    L.ref_io_xint(ip(xa), ip(xb), ip(xf))
    np.savez(out, counts=c, sim_d=sd, sim_i=si, pos_nopbc=xn, pos=x, q=q, mass=mass, atomid=atomid, resid=res,
             names=np.array([names.raw[5 * k:5 * k + 5].split(b"\0")[0].decode() for k in range(n)]),
             type_atomid=ids.astype(np.int32), lj_eps=e.reshape(T, T), lj_sig=s.reshape(T, T),
             lj14_eps=e14.reshape(T, T), lj14_sig=s14.reshape(T, T), bonds=bonds, bond_r0=breq, bond_k=bk,
             bends=bends, bend_eq=beq, bend_k=bek, ureys=ureys, urey_s0=us0, urey_k=uk, tors=tors, kphi=kphi,
             nphi=nphi, delphi=dphi, impr=impr, impr_k=ik, impr_eq=ieq, exclusions=ex, in14=p14,
             xint=np.column_stack([xa, xb, xf]).astype(np.int32))


def dump_trr(out_dir: str, npz: str, kw: bool) -> None:
    L = lib(kw)
    z = np.load(npz)
    os.makedirs(os.path.join(out_dir, "OUTPUT", "GROMACS"), exist_ok=True)
    for fr in range(int(z["nframes"])):
        x, v, f = (np.ascontiguousarray(z[k][fr]) for k in ("x", "v", "f"))
        rc = L.ref_io_write_trr(os.fsencode(out_dir), len(x), dp(x), dp(v), dp(f), C.c_double(z["box"][0]),
                                C.c_double(z["box"][1]), C.c_double(z["box"][2]), C.c_double(float(z["dt"])),
                                int(z["cycle"][fr]), int(z["flag"]), int(z["cyc_eq"]), int(z["fx"]), int(z["fv"]),
                                int(z["ff"]), int(z["findex"]))
        if rc:
            raise SystemExit(f"ref_io_write_trr failed: {rc}")


if __name__ == "__main__":
    mode, a, b = sys.argv[1:4]
    kw = len(sys.argv) > 4 and sys.argv[4] == "kw"
    if mode == "read":
        dump_read(a, b, kw)
    else:
        dump_trr(a, b, kw)
