"""Reference-format systems on the GPU path (include/ttc_io.h).

A system written in the reference's input formats and read back by the
native readers drives the device path unchanged:
  * chains-3k with angle-type terms whose parameters come from the
    parameter file by type (exact Fourier rows and wildcard rows), E14FAC 0.4
    as in the reference's non-BGO build: total forces within 5e-6 relative
    RMS of the FP64 oracle on the same read-back system, energies within
    1e-6;
  * chains-100k (the 100,000-site benchmark system, PME 96^3) read from its
    reference-format files: step-0 forces within 5e-6 of the oracle;
  * a device-resident velocity-Verlet run of chains-3k writes .trr frames
    (positions, velocities, forces from the device) every 10 steps; each
    frame agrees with the frame the oracle's own loop writes at the same step
    to the trr's float resolution.
"""
import ctypes as C
import os

import numpy as np
import pytest

from thecodecombo_7_amd import refio, system

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_angle_system_from_reference_inputs(tmp_path):
    import pyoracle
    from thecodecombo_7_amd.nonbond import Box
    s0 = system.chains_3k()
    s0.pos = system.jitter(s0, 0.05)
    a0 = system.angle_terms(s0)
    d = str(tmp_path / "INPUT")
    refio.from_chain_system(s0, a0).write(d)
    s, at = refio.read_inputs(d)
    assert s.e14fac == 0.4 and at is not None
    print(f"case: chains-3k from reference inputs, {len(at.bend)} bends, {len(at.tors)} torsions, "
          f"{len(at.impr)} impropers", flush=True)
    b = Box(s)
    try:
        b.set_angles(at)
        b.nblist()
        ff, en = b.forces()
    finally:
        b.close()
    o = pyoracle.Oracle(s)
    fo, eo = o.forces(s.pos)
    fa, ea, _ = o.angles(s.pos, at)
    err = rel_rms(ff, fo + fa)
    print(f"force rel rms {err:.2e}, nbond {en['nbond']:.6f} vs {eo['nbond']:.6f}", flush=True)
    assert err < 5e-6
    assert abs(en["nbond"] - eo["nbond"]) <= 1e-6 * abs(eo["nbond"])


def test_chains_100k_from_reference_inputs(tmp_path):
    import pyoracle
    from thecodecombo_7_amd.nonbond import Box
    s0 = system.chains_100k()
    d = str(tmp_path / "INPUT")
    refio.from_chain_system(s0).write(d)
    s, at = refio.read_inputs(d, e14fac=0.5)
    assert at is None and s.n == 100_000 and s.grid == (96, 96, 96)
    np.testing.assert_array_equal(s.xint, s0.xint)
    print("case: chains-100k from reference inputs, PME 96^3", flush=True)
    b = Box(s)
    try:
        b.nblist()
        ff, en = b.forces()
    finally:
        b.close()
    o = pyoracle.Oracle(s)
    fo, eo = o.forces(s.pos)
    err = rel_rms(ff, fo)
    print(f"force rel rms {err:.2e}", flush=True)
    assert err < 5e-6


def test_trr_of_a_device_run_matches_the_oracle_run(tmp_path):
    import pyoracle
    from thecodecombo_7_amd.nonbond import Box
    s0 = system.chains_3k()
    d = str(tmp_path / "INPUT")
    refio.from_chain_system(s0).write(d)
    s, _ = refio.read_inputs(d, seed=3)
    nchunk, every = 3, 10
    gpu_trr, ora_trr = str(tmp_path / "gpu.trr"), str(tmp_path / "oracle.trr")
    print(f"case: chains-3k from reference inputs, {nchunk * every} velocity-Verlet steps, a frame every "
          f"{every}", flush=True)
    b = Box(s)
    try:
        b.nblist()
        for k in range(nchunk):
            b.integrate_vverlet(every)
            step = (k + 1) * every
            refio.write_trr(gpu_trr, b.positions(), s.box, vel=b.velocities(), force=b.get_forces(), step=step,
                            time=s.dt * step)
    finally:
        b.close()
    o = pyoracle.Oracle(s)
    pos = s.pos.copy().ravel(); vel = s.vel.copy().ravel(); uu = vel.copy()
    anchor = pos.copy()
    ff, _ = o.forces(pos.reshape(-1, 3))
    ff = ff.ravel().copy()
    lists = None
    for step in range(1, nchunk * every + 1):
        if lists is None or o.L.orc_nl_check(C.byref(o.cs), pos, anchor, 0, s.n):
            rc, cnt, lst, xf, *_ = o.nblist(pos.reshape(-1, 3))
            lists = (cnt, lst, xf)
            anchor = pos.copy()
        o.L.orc_vverlet_a(C.byref(o.cs), s.n, s.mass, s.dt, 1.0, pos, None, vel, ff)
        f, _ = o.forces(pos.reshape(-1, 3), lists=lists)
        ff = f.ravel().copy()
        o.L.orc_vverlet_b(s.n, s.mass, s.dt, 1.0, vel, uu, ff)
        if step % every == 0:
            refio.write_trr(ora_trr, pos.reshape(-1, 3), s.box, vel=vel.reshape(-1, 3), force=ff.reshape(-1, 3),
                            step=step, time=s.dt * step)
    g, r = refio.read_trr(gpu_trr), refio.read_trr(ora_trr)
    assert [f["step"] for f in g] == [f["step"] for f in r] == [every * (k + 1) for k in range(nchunk)]
    assert os.path.getsize(gpu_trr) == os.path.getsize(ora_trr)
    for fg, fr in zip(g, r):
        assert fg["time"] == fr["time"]
        np.testing.assert_array_equal(fg["box"], fr["box"])
        dx = fg["x"] - fr["x"]
        dx -= fg["box"] * np.round(dx / fg["box"])
        # velocities reach thousands of A/ps on the strained start: compared
        # relative to the largest, to the force tolerance (the FP32 pair
        # terms' 1e-6 force differences grow along the trajectory)
        dv = np.abs(fg["v"] - fr["v"]).max() / np.abs(fr["v"]).max()
        ef = rel_rms(fg["f"], fr["f"])
        print(f"step {fg['step']}: max |dx| {np.abs(dx).max():.2e} nm, max |dv| / max |v| {dv:.2e}, "
              f"force rel rms {ef:.2e}", flush=True)
        assert np.abs(dx).max() < 2e-7
        assert dv < 5e-6
        assert ef < 5e-6
