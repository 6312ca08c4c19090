"""Writes c0_nve_oracle.npz: the oracle's 1,000-step NVE velocity-Verlet run
of the chains-3k system (300 chains x 10 sites, 31 A box, PME 32^3, 1 fs),
the reference driver order of nve_md.c:341-517 (integrate_vverlet.c:36-103,
then nl_check / nblist at the step's positions), energy row every step:
{step, kinet, bond, nbond, nbonds}.

    python tests/golden/make_c0_nve.py      # about 3 minutes on one core

The oracle (oracle/, strict IEEE build) is the checker; this fixture is its
output, committed so that a GPU test can compare the device run row by row
without repeating the CPU run.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def run(nsteps):
    import pyoracle
    from thecodecombo_7_amd import system
    s = system.chains_3k()
    o = pyoracle.Oracle(s)
    pos = s.pos.copy().ravel(); vel = s.vel.copy().ravel(); uu = vel.copy()
    rc, cnt, lst, xf, *_ = o.nblist(pos.reshape(-1, 3))
    lists, anchor = (cnt, lst, xf), pos.copy()
    ff = o.forces(pos.reshape(-1, 3), lists=lists)[0].ravel()
    rows = []
    for step in range(1, nsteps + 1):
        o.L.orc_vverlet_a(C.byref(o.cs), s.n, s.mass, s.dt, 1.0, pos, None, vel, ff)
        if o.L.orc_nl_check(C.byref(o.cs), pos, anchor, 0, s.n):
            rc, cnt, lst, xf, *_ = o.nblist(pos.reshape(-1, 3))
            lists, anchor = (cnt, lst, xf), pos.copy()
        f, en = o.forces(pos.reshape(-1, 3), lists=lists)
        ff = f.ravel()
        o.L.orc_vverlet_b(s.n, s.mass, s.dt, 1.0, vel, uu, ff)
        kin = o.L.orc_kinet(s.n, s.mass, vel, uu, np.zeros(6))
        rows.append((step, kin, en["bond"], en["nbond"], en["nbonds"]))
    return np.array(rows)


if __name__ == "__main__":
    rows = run(1000)
    np.savez_compressed(os.path.join(HERE, "c0_nve_oracle.npz"), rows=rows)
    tot = rows[:, 1] + rows[:, 2] + rows[:, 4]
    print("rows", rows.shape, "total", tot[0], "->", tot[-1])
