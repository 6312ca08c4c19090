"""GPU parity of the integrators and MC energy entry points against the
oracle: force_short / force_long, RESPA (integrate_mts.c), Nose-Hoover chain
NVT (integrate_nhc.c + nvt_md.c), enbnd_mc range energies, nblist_pivot.

Tolerances: forces relative RMS <= 5e-6; trajectories over a few steps agree
to 1e-6 A in position (FP32 pair arithmetic perturbs forces at 1e-7);
thermostat state to 1e-6 relative; energy slots to 2e-6 of the largest slot.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def small():
    from thecodecombo_7_amd import system
    return system.make_chain_system(120, 24.0, 24, seed=8)


def _box(s):
    from thecodecombo_7_amd.nonbond import Box
    b = Box(s)
    b.nblist()
    return b


def test_force_short_and_long(small):
    import pyoracle
    o = pyoracle.Oracle(small)
    # the generated chains sit at the bond minimum; stretch them so the bond
    # forces are well above the 2^-32 fixed-point force quantum
    x = small.pos + np.random.default_rng(3).normal(0.0, 0.05, small.pos.shape)
    x -= small.box * np.floor(x / small.box + 0.5)
    b = _box(small)
    b.set_positions(x)
    b.nblist()
    fs, eb = b.force_short()
    ebr, fsr, _ = o.bonds(x)
    assert rel_rms(fs, fsr) < 1e-9 and abs(eb - ebr) <= 1e-9 * abs(ebr)
    fl, en = b.force_long()
    rc, cnt, lst, xf, *_ = o.nblist(x, 1200)
    flr, enr, _, _ = o.cnonbondnl(x, cnt, lst, xf)
    assert rel_rms(fl, flr) < 5e-6
    assert np.allclose(en, enr, rtol=2e-6, atol=2e-6 * np.abs(enr).max())
    b.close()


def _oracle_lists(o, pos):
    rc, cnt, lst, xf, *_ = o.nblist(pos.reshape(-1, 3), 1200)
    assert rc == 0
    return cnt, lst, xf


def test_mts_matches_oracle(small):
    import pyoracle
    s = small
    nouter, ninner, dts, dtl = 4, 4, 0.0005, 0.002
    s.dt = dts
    b = _box(s)
    b.integrate_mts(nouter, ninner, dtl)
    pos_g, vel_g = b.positions(), b.velocities()
    b.close()
    o = pyoracle.Oracle(s)
    im = (100.0 / s.mass)[:, None]
    x = s.pos.copy(); v = s.vel.copy()
    lists = _oracle_lists(o, x)
    anchor = x.copy()
    fs = o.bonds(x)[1]
    fl = o.cnonbondnl(x, *lists)[0]
    for _ in range(nouter):
        if o.L.orc_nl_check(C.byref(o.cs), np.ascontiguousarray(x).ravel(), np.ascontiguousarray(anchor).ravel(), 0, s.n):
            lists = _oracle_lists(o, x)
            anchor = x.copy()
        v = v + 0.5 * dtl * im * fl
        for _ in range(ninner):
            v = v + 0.5 * dts * im * fs
            x = x + dts * v
            x -= s.box * np.round(x / s.box)
            fs = o.bonds(x)[1]
            v = v + 0.5 * dts * im * fs
        fl = o.cnonbondnl(x, *lists)[0]
        v = v + 0.5 * dtl * im * fl
    d = x - pos_g
    d -= s.box * np.round(d / s.box)
    print("mts max |dx|", np.abs(d).max(), "max |dv|", np.abs(v - vel_g).max())
    s.dt = 0.001
    assert np.abs(d).max() < 1e-6
    # long forces reach 2e4 with FP32 pair arithmetic (~1e-6 relative); the
    # velocity error follows at ~4e-6 of the largest velocity
    assert np.abs(v - vel_g).max() < 2e-5 * np.abs(v).max()


def test_nhc_matches_oracle(small):
    import pyoracle
    s = small
    nsteps, T, M = 6, 300.0, 2
    nfree = 3.0 * s.n - 3.0
    kT = 1.380662e-23 * T * 1e-4
    Q = nfree * kT * 0.1 ** 2
    b = _box(s)
    b.set_nhc(T, Q, M=M, nfree=nfree)
    b.integrate_nhc(nsteps)
    pos_g, vel_g, z_g = b.positions(), b.velocities(), b.nhc_state()
    b.close()
    o = pyoracle.Oracle(s)
    x = s.pos.ravel().copy(); v = s.vel.ravel().copy(); uu = v.copy()
    zeta = np.zeros(3 * M)
    w = np.array([1.0])
    lists = _oracle_lists(o, x)
    anchor = x.copy()
    ff = o.forces(x.reshape(-1, 3), lists=lists)[0].ravel()
    for _ in range(nsteps):
        if o.L.orc_nl_check(C.byref(o.cs), x, anchor, 0, s.n):
            lists = _oracle_lists(o, x)
            anchor = x.copy()
        o.L.orc_nhc(s.n, v, uu, s.mass, nfree, T, Q, M, 1, 1, w, zeta, s.dt)
        o.L.orc_vverlet_a(C.byref(o.cs), s.n, s.mass, s.dt, 1.0, x, None, v, ff)
        ff = o.forces(x.reshape(-1, 3), lists=lists)[0].ravel()
        o.L.orc_vverlet_b(s.n, s.mass, s.dt, 1.0, v, uu, ff)
        o.L.orc_nhc(s.n, v, uu, s.mass, nfree, T, Q, M, 1, 1, w, zeta, s.dt)
    d = x.reshape(-1, 3) - pos_g
    d -= s.box * np.round(d / s.box)
    print("nhc max |dx|", np.abs(d).max(), "zeta", zeta, z_g)
    assert np.abs(d).max() < 1e-6
    assert np.abs(v.reshape(-1, 3) - vel_g).max() < 1e-4
    assert np.allclose(z_g, zeta, rtol=1e-6, atol=1e-9 * np.abs(zeta).max())


def test_enbnd_mc_matches_oracle(small):
    import pyoracle
    s = small
    o = pyoracle.Oracle(s)
    cnt, lst, xf = _oracle_lists(o, s.pos)
    b = _box(s)
    for lb, ub in ((0, 10), (300, 340), (1190, 1200)):
        keep = np.zeros_like(lst, dtype=bool)
        for a in range(s.n):
            k = cnt[a]
            bs = lst[a, :k]
            keep[a, :k] = (lb <= a < ub) | ((bs >= lb) & (bs < ub))
        c2 = np.zeros_like(cnt); l2 = np.zeros_like(lst); x2 = np.zeros_like(xf)
        for a in range(s.n):
            sel = np.nonzero(keep[a])[0]
            c2[a] = len(sel)
            l2[a, :len(sel)] = lst[a, sel]
            x2[a, :len(sel)] = xf[a, sel]
        _, enr, _, _ = o.cnonbondnl(s.pos, c2, l2, x2)
        en = b.enbnd_mc(lb, ub)
        assert np.allclose(en, enr, rtol=2e-6, atol=2e-6 * np.abs(enr).max()), (lb, ub, en, enr)
    b.close()


def test_nblist_pivot_gives_fresh_list(small):
    import pyoracle
    s = small
    b = _box(s)
    x = s.pos.copy()
    # pivot the first chain's last five sites by 30 degrees about z through site 4
    th = np.pi / 6
    R = np.array([[np.cos(th), np.sin(th), 0], [-np.sin(th), np.cos(th), 0], [0, 0, 1]])
    x[5:10] = (x[5:10] - x[4]) @ R.T + x[4]
    x -= s.box * np.round(x / s.box)
    b.set_positions(x)
    b.nblist_pivot(10)
    o = pyoracle.Oracle(s)
    rc, cnt, lst, xf, *_ = o.nblist(x, 1200)
    gc, gl, gx = b.get_nblist(0, 1200)
    assert (gc == cnt).all()
    for a in range(s.n):
        assert (gl[a, :cnt[a]] == lst[a, :cnt[a]]).all() and (gx[a, :cnt[a]] == xf[a, :cnt[a]]).all()
    b.close()


def test_drift_removes_momentum_as_reference(small):
    """drift.c: v_i -= (P / N) / m_i, against the oracle (bit-exact to the
    reference's drift.c); 1e-12 of the speed scale (the device holds 100/m)."""
    import pyoracle
    s = small
    v0 = s.vel + np.array([3.0, -2.0, 1.5])
    print("case: 1,200 sites, velocities with net momentum, drift()", flush=True)
    b = _box(s)
    try:
        b.set_velocities(v0)
        b.drift()
        v1 = b.velocities()
    finally:
        b.close()
    vo = v0.copy().ravel()
    pyoracle.lib().orc_drift(s.n, np.ascontiguousarray(s.mass), vo)
    assert np.abs(v1 - vo.reshape(-1, 3)).max() <= 1e-12 * np.abs(v0).max()
    assert np.abs((s.mass[:, None] * v1).sum(0)).max() < 1e-9 * (s.mass[:, None] * np.abs(v0)).sum()


def test_md_with_drift_every_5_steps_matches_oracle(small):
    """20 velocity-Verlet steps with drift(k) after steps 5, 10, 15, 20
    (nve_md.c:208) against the oracle's loop."""
    import pyoracle
    s = small
    v0 = s.vel + np.array([3.0, -2.0, 1.5])
    nsteps = 20
    b = _box(s)
    try:
        b.set_velocities(v0)
        b.set_drift(5)
        b.integrate_vverlet(nsteps)
        pos_gpu, vel_gpu = b.positions(), b.velocities()
    finally:
        b.close()
    o = pyoracle.Oracle(s)
    pos = s.pos.copy().ravel(); vel = v0.copy().ravel(); uu = vel.copy()
    anchor = pos.copy()
    ff = o.forces(pos.reshape(-1, 3))[0].ravel()
    lists = None
    for step in range(1, nsteps + 1):
        if lists is None or o.L.orc_nl_check(C.byref(o.cs), pos, anchor, 0, s.n):
            rc, cnt, lst, xf, *_ = o.nblist(pos.reshape(-1, 3))
            lists = (cnt, lst, xf)
            anchor = pos.copy()
        o.L.orc_vverlet_a(C.byref(o.cs), s.n, s.mass, s.dt, 1.0, pos, None, vel, ff)
        ff = o.forces(pos.reshape(-1, 3), lists=lists)[0].ravel()
        o.L.orc_vverlet_b(s.n, s.mass, s.dt, 1.0, vel, uu, ff)
        if step % 5 == 0:
            o.L.orc_drift(s.n, np.ascontiguousarray(s.mass), vel)
    dpos = pos.reshape(-1, 3) - pos_gpu
    dpos -= s.box * np.round(dpos / s.box)
    print("max |dx|", np.abs(dpos).max(), "max |dv|", np.abs(vel.reshape(-1, 3) - vel_gpu).max(), flush=True)
    assert np.abs(dpos).max() < 1e-6
    # FP32 pair terms: 1e-7 relative force noise, grown over 20 steps
    assert np.abs(vel.reshape(-1, 3) - vel_gpu).max() < 5e-5 * np.abs(vel).max()
    assert np.abs((s.mass[:, None] * vel_gpu).sum(0)).max() < 1e-6 * (s.mass[:, None] * np.abs(vel_gpu)).sum()
