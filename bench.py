#!/usr/bin/env python3
"""Benchmark: ns/day of the 100,000-site PME chain system on one MI355X.

One step = one velocity-Verlet step of the whole nonbonded path on the
device (DLIST trigger + conditional list rebuild, kick/drift/pbc + charge
spreading, LJ + real-space Ewald pair kernel, 3-D FFTs + influence function,
PME gather + bonds, second kick), replayed as one captured HIP graph per
step.  Inputs are resident in HBM before the timed region starts.

Prints one JSON line (see the repository's benchmark contract).  The CPU
baseline is the FP64 oracle (a restatement of the reference's C path built
with the reference's flags) timed on one host core on a bounded sample of
the same workload; the step-0 relative RMS force error of the GPU path
against it is reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ns/day @100k atoms (FP32-mixed vs FP64 plain, 1 GPU) + accepted MC moves/s"
HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs, a wave64 VALU instruction every 4
# cycles per 16-lane SIMD, at the 2.4 GHz peak engine clock
VALU_PEAK_WINST_PER_S = 256 * 4 / 4 * 2.4e9


def cpu_baseline(s, f_gpu0, rebuild_interval):
    """Time the FP64 oracle (reference flags, 1 thread) on the step-0 state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C

    import pyoracle
    o = pyoracle.Oracle(s, fast=True)
    pos = np.ascontiguousarray(s.pos)
    t0 = time.perf_counter()
    rc, cnt, lst, xf, *_ = o.nblist(pos, 1500)
    t_list = time.perf_counter() - t0
    if rc:
        raise RuntimeError("oracle list overflow")
    t0 = time.perf_counter()
    nrep = 2
    for _ in range(nrep):
        ff, en = o.forces(pos, lists=(cnt, lst, xf))
    t_force = (time.perf_counter() - t0) / nrep
    # integrator updates (integrate_vverlet.c halves) on copies
    p = pos.ravel().copy(); v = s.vel.ravel().copy(); u = v.copy(); f = ff.ravel().copy()
    t0 = time.perf_counter()
    o.L.orc_vverlet_a(C.byref(o.cs), s.n, s.mass, s.dt, 1.0, p, None, v, f)
    o.L.orc_vverlet_b(s.n, s.mass, s.dt, 1.0, v, u, f)
    t_int = time.perf_counter() - t0
    t_step = t_force + t_int + t_list / max(1.0, rebuild_interval)
    err = float(np.sqrt(((f_gpu0 - ff) ** 2).sum() / (ff ** 2).sum()))
    return {
        "value": 0.0864 / t_step, "unit": "ns/day", "cores": 1, "kind": "port",
        "sample": (f"oracle (FP64 C restatement, gcc -O3 -ffast-math -funroll-all-loops, 1 thread) on "
                   f"the chains-100k step-0 state: 1 O(N^2) list build ({t_list:.2f} s, amortised over "
                   f"the GPU run's {rebuild_interval:.1f} steps per rebuild), {nrep} force evaluations "
                   f"({t_force:.2f} s each: LJ+real-space pair loop, PME 96^3 with a C2C FFT, bonds), "
                   f"1 velocity-Verlet update ({t_int * 1e3:.1f} ms)"),
    }, err, en


def mc_leg(ntrials: int, cpu: bool):
    """Accepted MC moves/s of 256 Go walkers (60-300 beads) under Wang-Landau:
    plain exact MC, then delayed acceptance with the surrogate trained on the
    moves a logged plain run recorded.  Both timed runs start from the same
    state (positions, ln g, histograms, random streams) and neither logs."""
    from thecodecombo_7_amd import system
    from thecodecombo_7_amd.mc import Walkers
    chains = system.make_go_walkers(256)
    w = Walkers(chains, seed=31)
    lnf = 0.05
    w.run(200, mode=0, lnf=lnf)                 # leave the native start
    w.clear_log()
    w.run(ntrials, mode=0, lnf=lnf, log=True)   # training data (not timed)
    x, de = w.log()
    t0 = time.perf_counter()
    loss0 = w.train_surrogate(nepoch=1, lr=1e-3)
    loss = w.train_surrogate(nepoch=2, lr=5e-4)
    t_train = time.perf_counter() - t0
    w.snapshot()
    res = {}
    for mode in (0, 1):
        w.restore()
        w.reset_stats()
        t = w.run(ntrials, mode=mode, lnf=lnf) * 1e-6
        res[mode] = (t, w.stats())
    (t_plain, st0), (t_da, st1) = res[0], res[1]
    out = {
        "workload": "256 Go walkers, C-alpha chains of 60-300 beads (46,173 beads), 12-10 native contacts, "
                    "Wang-Landau 0.5 kcal/mol bins, pivot + displacement (1:1)",
        "trials_per_walker": ntrials,
        "plain_accepted_moves_per_s": round(st0[:, 1].sum() / t_plain, 1),
        "da_accepted_moves_per_s": round(st1[:, 1].sum() / t_da, 1),
        "plain_trials_per_s": round(st0[:, 0].sum() / t_plain, 1),
        "da_trials_per_s": round(st1[:, 0].sum() / t_da, 1),
        "plain_pivot_accept_frac": round(st0[:, 6].sum() / max(1, st0[:, 5].sum()), 4),
        "plain_displacement_accept_frac": round((st0[:, 1] - st0[:, 6]).sum() / max(1, (st0[:, 0] - st0[:, 5]).sum()), 4),
        "da_pivot_accept_frac": round(st1[:, 6].sum() / max(1, st1[:, 5].sum()), 4),
        "da_stage1_reject_frac": round(st1[:, 2].sum() / st1[:, 0].sum(), 4),
        "da_exact_evals_frac": round(st1[:, 3].sum() / st1[:, 0].sum(), 4),
        "logged_moves": int(len(x)), "surrogate_loss": [round(loss0, 4), round(loss, 4)],
        "train_s": round(t_train, 2),
    }
    out["da_gain_accepted_per_s"] = round(out["da_accepted_moves_per_s"] / out["plain_accepted_moves_per_s"], 3)
    if cpu:
        w.restore()                                  # the plain run's start state, re-run to its end
        w.run(ntrials, mode=0, lnf=lnf)
        out["cpu_baseline"] = mc_cpu_baseline(w, chains, lnf, w.emin)
    w.close()
    return out


def mc_cpu_baseline(w, chains, lnf, emin):
    """Plain exact Wang-Landau MC on one host core: the oracle's orc_wl_run
    (the restatement of the reference's pivot / displacement trial with
    pivot.c:236-253 acceptance, FP64 energy changes; built with the
    reference's flags) continues eight walkers of the batch, every 32nd by
    size so the sample has the batch's size mix, from their state after the
    GPU's plain run (positions, ln g, histograms, random streams)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from thecodecombo_7_amd import system
    order = np.argsort([c.n for c in chains])
    pick = order[16::32]
    states = []
    for k in pick:
        ch = chains[k]
        pos, e = w.walker(int(k))
        lng, hist = w.dos(int(k))
        nat, r0m = pyoracle.go_tables(ch.n, ch.contacts, ch.r0)
        states.append([np.ascontiguousarray(pos.ravel()), e, nat, r0m, lng.copy(), hist.copy(),
                       np.ascontiguousarray(w.streams(int(k))), ch.n, float(emin[k])])

    def params(em):
        return pyoracle.WlParams(kbond=system.GO_KBOND, r0bond=system.GO_BOND, eps=system.GO_EPS,
                                 sig_rep=system.GO_SIG_REP, pivot_frac=w.p.pivot_frac, theta_max=w.p.theta_max,
                                 disp_max=w.p.disp_max, emin=em, ebin=w.ebin, nbins=w.nbins)

    def sweep(ntr):
        acc = trials = 0
        t0 = time.perf_counter()
        for st in states:
            stats = np.zeros(7, np.int64)
            st[1] = pyoracle.wl_run(params(st[8]), st[0], st[1], st[2], st[3], st[4], st[5], st[6], ntr, lnf,
                                    stats, fast=True)
            acc += int(stats[1]); trials += int(stats[0])
        return acc, trials, time.perf_counter() - t0

    _, _, t_cal = sweep(64)                    # calibration: about 10 s of work in the timed sweep
    ntr = int(min(20000, max(256, 64 * 10.0 / max(t_cal, 1e-3))))
    acc, trials, dt = sweep(ntr)
    return {"value": round(acc / dt, 2), "unit": "accepted moves/s", "cores": 1, "kind": "port",
            "sample": f"{trials} Wang-Landau trials ({ntr} per walker) on 8 walkers of the batch "
                      f"({', '.join(str(int(st[7])) for st in states)} beads) from their state after the GPU's "
                      f"plain run: oracle orc_wl_run (FP64, reference flags, 1 thread), {acc} accepted in "
                      f"{dt:.1f} s"}


def large_leg(steps: int, cpu: bool):
    """chains-1M (10^6 sites, 215 A box, PME 224^3, rc 12 A, list 14 A): the
    pair kernel's and the spreading step's algorithmic GB/s and roofline
    fraction, step time, and the step-0 force error against the FP64 oracle
    (cell-list build of the same pair set, nblist.c CLIST semantics)."""
    from thecodecombo_7_amd import system
    from thecodecombo_7_amd.nonbond import Box
    s = system.chains_1m()
    box = Box(s)
    box.nblist()
    entries, _, _ = box.list_stats()
    half = entries // 2
    f0, _ = box.forces()
    waves, _ = box.tune()
    box.integrate_vverlet(20)
    t0 = time.perf_counter()
    _, nreb = box.integrate_vverlet(steps)
    wall = (time.perf_counter() - t0) / steps
    t_pair = box.timing(0)
    t_spread = box.timing(1)
    pair_bytes = 4.0 * half + 48.0 * s.n
    spread_bytes = 16.0 * s.n + 4.0 * s.grid[0] * s.grid[1] * s.grid[2]
    out = {
        "workload": "chains-1M: 100,000 chains x 10 sites, 215 A box, PME 224^3 order 4, rc 12 A, list 14 A, "
                    "NVE velocity Verlet 1 fs",
        "steps": steps, "ms_per_step": round(1e3 * wall, 4), "ns_per_day": round(0.0864 / wall, 3),
        "rebuilds": nreb, "half_list_entries": half, "pair_waves": waves,
        "pair_kernel": {"us": round(t_pair, 2), "algorithmic_bytes": pair_bytes,
                        "achieved_GBs": round(pair_bytes / (t_pair * 1e-6) / 1e9, 1),
                        "frac_of_8TBs": round(pair_bytes / (t_pair * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)},
        "spread_kernel": {"us": round(t_spread, 2), "algorithmic_bytes": spread_bytes,
                          "achieved_GBs": round(spread_bytes / (t_spread * 1e-6) / 1e9, 1),
                          "frac_of_8TBs": round(spread_bytes / (t_spread * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    box.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.Oracle(s, fast=True)
        t0 = time.perf_counter()
        rc, cnt, lst, xf = o.nblist(s.pos, 1300, cells=True)
        t_list = time.perf_counter() - t0
        if rc:
            raise RuntimeError("oracle list overflow at chains-1M")
        t0 = time.perf_counter()
        ff, _ = o.forces(s.pos, lists=(cnt, lst, xf))
        t_force = time.perf_counter() - t0
        del lst, xf
        out["force_rel_rms_vs_fp64"] = float(np.sqrt(((f0 - ff) ** 2).sum() / (ff ** 2).sum()))
        out["cpu_step0"] = {"list_s": round(t_list, 2), "forces_s": round(t_force, 2), "cores": 1,
                            "kind": "port", "sample": "oracle cell-list build + one force evaluation "
                                                      "(pair loop, PME 224^3, bonds) on the step-0 state"}
    return out


def replica_leg(steps: int, cpu: bool):
    """64 replicas of 2,000-20,000 sites (log-uniform, seeded; 1 A PME grids),
    Nose-Hoover NVT at temperatures geometric from 300 to 600 K, all stepped
    together (ttc_batch_*, launch mode autotuned), exchanges attempted every
    500 steps.  Aggregate ns/day = sum over replicas of simulated time per
    day; step-0 force error of every replica against the FP64 oracle."""
    from thecodecombo_7_amd import system
    from thecodecombo_7_amd.replicas import ReplicaBatch
    systems = system.make_replicas(64)
    nrep = len(systems)
    temps = 300.0 * 2.0 ** (np.arange(nrep) / (nrep - 1))
    rb = ReplicaBatch(systems, temps, seed=5)
    out = {"workload": "64 replicas, 2,000-20,000 sites (log-uniform), chain chemistry at the same density, "
                       "PME grids at 1 A spacing, NHC NVT 300-600 K (geometric), exchanges every 500 steps",
           "replicas": nrep, "sites": int(sum(x.n for x in systems))}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        import ctypes as C
        errs, t_list, t_step = [], np.zeros(nrep), np.zeros(nrep)
        for r_, (b, x) in enumerate(zip(rb.boxes, systems)):
            ff, _ = b.forces()
            o = pyoracle.Oracle(x, fast=True)
            t0 = time.perf_counter()
            rc, cnt, lst, xf = o.nblist(x.pos, 1500, cells=True)
            t_list[r_] = time.perf_counter() - t0
            if rc:
                raise RuntimeError("oracle list overflow in the replica leg")
            t0 = time.perf_counter()
            fr, _ = o.forces(x.pos, lists=(cnt, lst, xf))
            # one NHC step's integration (nvt_md.c:528-531): thermostat half,
            # velocity-Verlet halves, thermostat half
            p = x.pos.ravel().copy(); v = x.vel.ravel().copy(); u = v.copy(); f = fr.ravel().copy()
            zeta = np.zeros(2 * 4)
            w = np.ones(1)
            nfree = 3.0 * x.n - 3.0
            o.L.orc_nhc(x.n, v, u, x.mass, nfree, float(temps[r_]), 1e-18, 2, 1, 1, w, zeta, x.dt)
            o.L.orc_vverlet_a(C.byref(o.cs), x.n, x.mass, x.dt, 1.0, p, None, v, f)
            o.L.orc_vverlet_b(x.n, x.mass, x.dt, 1.0, v, u, f)
            o.L.orc_nhc(x.n, v, u, x.mass, nfree, float(temps[r_]), 1e-18, 2, 1, 1, w, zeta, x.dt)
            t_step[r_] = time.perf_counter() - t0
            errs.append(float(np.sqrt(((ff - fr) ** 2).sum() / (fr ** 2).sum())))
        out["force_rel_rms_vs_fp64"] = {"max": max(errs), "median": float(np.median(errs)), "min": min(errs)}
    mode, spl, tt = rb.autotune(50)
    rb.run(500, exchange_every=500)
    t0 = time.perf_counter()
    r = rb.run(steps, exchange_every=500)
    wall = (time.perf_counter() - t0) / steps
    out.update({"steps": steps, "launch_mode": int(mode), "steps_per_launch": int(spl),
                "ms_per_step": round(1e3 * wall, 4),
                "aggregate_ns_per_day": round(nrep * 0.0864 / wall, 2),
                "site_ns_per_day": round(out["sites"] * 0.0864 / wall, 1),
                "exchange_attempts": int(r["attempts"]), "exchange_accepts": int(r["accepts"])})
    if cpu:
        # one step of all replicas on one core: every replica's force
        # evaluation and NHC integration, plus its cell-list build amortised
        # at the rebuild rate the GPU run observed for it
        reb = np.asarray(r["rebuilds"], np.float64) / steps
        t_cpu = float((t_step + t_list * reb).sum())
        out["cpu_baseline"] = {"value": round(nrep * 0.0864 / t_cpu, 5), "unit": "aggregate ns/day", "cores": 1,
                               "kind": "port",
                               "sample": "oracle (reference flags, 1 thread) on every replica's step-0 state: "
                                         "one force evaluation (pair loop, PME, bonds) and one NHC step's "
                                         "integration (orc_nhc, orc_vverlet_a/b, orc_nhc) per replica, plus its "
                                         f"cell-list build ({t_list.sum():.2f} s for all) amortised at the "
                                         f"GPU run's rebuild rate ({reb.mean():.3f} per step on average)"}
    rb.close()
    return out


def _slope(t_ps, tot, dof):
    # least-squares slope and its standard error (independent residuals),
    # kJ/mol per ns per degree of freedom
    (slope, _), cov = np.polyfit(t_ps, tot, 1, cov=True)
    return float(slope * 1e3 / dof), float(np.sqrt(cov[0, 0]) * 1e3 / dof)


def nve_drift(s, steps: int, every: int, equil: int = 2000):
    """Energy conservation of NVE velocity Verlet from the step-0 state (a
    fresh box, so the leg does not depend on the timed run): the conserved
    total kinet + bond + nbonds (calcvalue.c:40-64, shifted nonbonded energy)
    every `every` steps; drift = least-squares slope per DOF, over the whole
    run from step 0 and over the steps after a fixed equilibration."""
    from thecodecombo_7_amd.nonbond import Box
    box = Box(s)
    try:
        box.nblist()
        rows, _ = box.integrate_vverlet(steps, energy_every=every)
    finally:
        box.close()
    t_ps = rows[:, 0] * s.dt
    tot = rows[:, 1] + rows[:, 2] + rows[:, 4]
    dof = 3 * s.n - 3
    d0, e0 = _slope(t_ps, tot, dof)
    late = rows[:, 0] > equil
    d1, e1 = _slope(t_ps[late], tot[late], dof)
    return {"steps": steps, "rows": int(len(rows)), "every": every, "equilibration_steps": equil,
            "total_kJ_mol": [round(float(tot[0]), 3), round(float(tot[-1]), 3)],
            "drift_from_step0_kJ_mol_per_ns_per_dof": d0, "drift_from_step0_stderr": e0,
            "drift_after_equilibration_kJ_mol_per_ns_per_dof": d1, "drift_after_equilibration_stderr": e1,
            "rms_fluct_rel_after_equilibration": float(np.std(tot[late]) / abs(np.mean(tot[late])))}


COPY_MIB = 512


def traffic_probe():
    """Workload for the HBM-traffic counter runs: 20 pair-kernel launches on
    the chains-100k step-0 state, then one plain FP32 copy kernel that reads
    and writes COPY_MIB MiB (the calibration of FETCH_SIZE / WRITE_SIZE)."""
    import torch
    from thecodecombo_7_amd import system
    from thecodecombo_7_amd.nonbond import Box
    s = system.chains_100k()
    box = Box(s)
    box.nblist()
    entries, _, _ = box.list_stats()
    box.integrate_vverlet(50)          # the MD loop's (pruned) list, as in the timed run
    box.tune()
    box.timing(0)
    src = torch.ones(COPY_MIB << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    box.close()
    print(json.dumps({"half_list_entries": int(entries // 2), "sites": s.n, "copy_mib": COPY_MIB}), flush=True)


def _pmc_per_kernel(path, counter):
    """Mean per-dispatch value of one counter and the dispatch count for
    each kernel name in a rocprofv3 counter_collection.csv."""
    import csv
    acc = {}
    with open(path) as fh:
        for row in csv.DictReader(fh):
            if row.get("Counter_Name") != counter:
                continue
            name = row["Kernel_Name"]
            key = (name, row.get("Dispatch_Id", row.get("Correlation_Id")))
            acc.setdefault(name, {}).setdefault(key, 0.0)
            acc[name][key] += float(row["Counter_Value"])
    return {k: (sum(v.values()) / len(v), len(v)) for k, v in acc.items()}


def valu_summary(valu_csv, out_path):
    """VALU instructions per pair-kernel launch (tuned variant) from a
    counter run over SQ_INSTS_VALU, SQ_INSTS_VALU_TRANS_F32, SQ_WAVES and
    GRBM_GUI_ACTIVE of --traffic-probe.  SQ_ counters count wave-level
    instructions over the whole device (SQ_WAVES reads the launch's wave
    count); GRBM_GUI_ACTIVE sums the busy cycles of the 8 XCDs."""
    c = {k: _pmc_per_kernel(valu_csv, k) for k in
         ("SQ_INSTS_VALU", "SQ_INSTS_VALU_TRANS_F32", "SQ_WAVES", "GRBM_GUI_ACTIVE")}
    # the tuned variant: the pair kernel launched most often
    name = max((v[1], k) for k, v in c["SQ_INSTS_VALU"].items() if "k_pair_half<false" in k)[1]
    insts, trans = c["SQ_INSTS_VALU"][name][0], c["SQ_INSTS_VALU_TRANS_F32"][name][0]
    cycles = c["GRBM_GUI_ACTIVE"][name][0] / 8.0
    cap = 256 * cycles   # wave-instructions the CUs could issue in those cycles
    res = {"kernel": "k_pair_half (plain, tuned waves per group)", "workload": "chains-100k after 50 MD steps",
           "valu_wave_insts_per_launch": insts, "trans_wave_insts_per_launch": trans,
           "waves_per_launch": c["SQ_WAVES"][name][0], "cycles_per_xcd": cycles,
           "issue_frac_of_own_cycles": insts / cap,
           "issue_frac_trans_at_quarter_rate": (insts + 3.0 * trans) / cap}
    with open(out_path, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res))


def traffic_summary(fetch_csv, write_csv, out_path):
    """HBM bytes per pair-kernel launch from the two counter runs, each
    scaled by the copy kernel's reading against its known COPY_MIB bytes."""
    f = _pmc_per_kernel(fetch_csv, "FETCH_SIZE")
    w = _pmc_per_kernel(write_csv, "WRITE_SIZE")
    # the pair kernel variant timed most often (the tuned one), the copy kernel
    pick = lambda d, key: max(((v[1], v[0]) for k, v in d.items() if key(k)))[1]
    is_pair = lambda k: "k_pair_half<false" in k
    is_copy = lambda k: "vectorized_elementwise_kernel" in k and "FillFunctor" not in k
    known_kib = COPY_MIB * 1024.0
    sf = known_kib / pick(f, is_copy)
    sw = known_kib / pick(w, is_copy)
    fetch = pick(f, is_pair) * sf * 1024.0
    write = pick(w, is_pair) * sw * 1024.0
    res = {"kernel": "k_pair_half (plain, tuned waves per group)", "workload": "chains-100k after 50 MD steps",
           "fetch_kib_raw": pick(f, is_pair), "write_kib_raw": pick(w, is_pair),
           "copy_fetch_kib_raw": pick(f, is_copy), "copy_write_kib_raw": pick(w, is_copy),
           "fetch_scale": sf, "write_scale": sw,
           "hbm_bytes_per_launch": round(fetch + write)}
    with open(out_path, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only the workload of the FETCH_SIZE / WRITE_SIZE counter runs")
    ap.add_argument("--traffic-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT_JSON"),
                    help="turn the two counter runs into bytes per pair-kernel launch")
    ap.add_argument("--valu-summary", nargs=2, metavar=("VALU_CSV", "OUT_JSON"),
                    help="turn the VALU counter run into instructions per pair-kernel launch")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mc", action="store_true", help="skip the Monte Carlo leg")
    ap.add_argument("--mc-trials", type=int, default=8000, help="MC trials per walker and mode")
    ap.add_argument("--no-large", action="store_true", help="skip the chains-1M leg")
    ap.add_argument("--large-steps", type=int, default=200, help="timed steps of the chains-1M leg")
    ap.add_argument("--no-replicas", action="store_true", help="skip the batched replica leg")
    ap.add_argument("--replica-steps", type=int, default=1000, help="timed steps of the replica leg")
    ap.add_argument("--drift-steps", type=int, default=10000, help="NVE steps of the energy-drift run (0: skip)")
    args = ap.parse_args()
    if args.traffic_summary:
        traffic_summary(*args.traffic_summary)
        return
    if args.valu_summary:
        valu_summary(*args.valu_summary)
        return
    if args.traffic_probe:
        traffic_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is described, not built).", file=sys.stderr)
        sys.exit(2)

    from thecodecombo_7_amd import system
    from thecodecombo_7_amd.nonbond import Box

    print("bench: building chains-100k", file=sys.stderr, flush=True)
    s = system.chains_100k()
    box = Box(s)
    box.nblist()
    entries, _, _ = box.list_stats()
    half_entries = entries // 2          # the reference's half list at rclist (nblist.c:32)
    f_gpu0, _ = box.forces()
    print("bench: warmup", args.warmup, file=sys.stderr, flush=True)
    box.integrate_vverlet(args.warmup)
    # the loop's half-list pair kernel, tuned on the loop's own list
    waves, t_waves = box.tune()
    print(f"bench: half-list pair kernel waves per group {waves} (4/8: {np.round(t_waves[:2], 1)} us)",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    _, nreb = box.integrate_vverlet(args.steps)
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    t_event = box.timing(2)            # HIP-event time per step (us)
    ns_day = 0.0864 / (wall / args.steps)
    print(f"bench: {ms_step:.4f} ms/step wall, {t_event:.1f} us/step events, {nreb} rebuilds",
          file=sys.stderr, flush=True)

    # dominant kernel: the pair kernel; HIP events on the context stream
    t_pair = box.timing(0)
    t_spread = box.timing(1)
    t_fft = box.timing(3)              # the five FFT passes alone
    t_gather = box.timing(4)           # gather + bonds alone
    alg_bytes = 4.0 * half_entries + 48.0 * s.n
    achieved = alg_bytes / (t_pair * 1e-6) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pair_kernel_chains100k_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    valu = None
    vfile = os.path.join(ROOT, "profiles", "pair_kernel_chains100k_valu.json")
    if os.path.exists(vfile):
        with open(vfile) as fh:
            vi = json.load(fh)["valu_wave_insts_per_launch"]
        # the counted instructions over this run's HIP-event kernel time
        rate = vi / (t_pair * 1e-6)
        valu = {"bound": "valu", "achieved": round(rate / 1e9, 1), "peak": VALU_PEAK_WINST_PER_S / 1e9,
                "unit": "G wave-instructions/s", "frac": round(rate / VALU_PEAK_WINST_PER_S, 4),
                "wave_insts_per_launch": vi}
    out = {
        "metric": METRIC, "value": round(ns_day, 3), "unit": "ns/day", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair/spline/FFT, f64 positions/velocities/accumulation",
        "data": "synthetic (seeded chain generator, random-init parameters)",
        "config": {"workload": "chains-100k: 10,000 chains x 10 united-atom sites, 100 A box, "
                               "LJ + PME 96^3 order 4, rc 10 A, list 12 A, NVE velocity Verlet 1 fs",
                   "sites": s.n, "pme_grid": list(s.grid), "rc": s.rc, "rclist": s.rclist,
                   "dt_ps": s.dt, "rebuilds": nreb,
                   "steps_per_rebuild": round(args.steps / max(1, nreb), 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"k_pair_half (LJ + real-space Ewald, half list), {waves} waves per 64-site group",
                     "kernel_us": round(t_pair, 2),
                     "algorithmic_bytes": alg_bytes, "half_list_entries": half_entries,
                     "spread_us": round(t_spread, 2), "fft_us": round(t_fft, 2),
                     "gather_us": round(t_gather, 2), "step_us_events": round(t_event, 2),
                     "valu": valu},
    }
    if args.drift_steps > 0:
        print("bench: NVE energy drift", file=sys.stderr, flush=True)
        out["nve_drift"] = nve_drift(s, args.drift_steps, 50)
    if not args.no_large:
        print("bench: chains-1M leg", file=sys.stderr, flush=True)
        out["large"] = large_leg(args.large_steps, not args.no_cpu)
    if not args.no_replicas:
        print("bench: batched replica leg", file=sys.stderr, flush=True)
        out["replicas"] = replica_leg(args.replica_steps, not args.no_cpu)
    if not args.no_mc:
        print("bench: Monte Carlo leg", file=sys.stderr, flush=True)
        out["mc"] = mc_leg(args.mc_trials, not args.no_cpu)
    if not args.no_cpu:
        print("bench: CPU baseline (oracle, 1 core)", file=sys.stderr, flush=True)
        cb, err, _ = cpu_baseline(s, f_gpu0, args.steps / max(1, nreb))
        out["cpu_baseline"] = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in cb.items()}
        out["force_rel_rms_vs_fp64"] = err
        out["speedup_vs_cpu"] = round(ns_day / cb["value"], 1)
    box.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
