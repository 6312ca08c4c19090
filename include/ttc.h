/*
 * ttc.h -- C ABI of the MI355X nonbonded path (libttc.so).
 *
 * Plain C types only: pointers, sizes, doubles.  One context per simulation
 * box, one HIP stream per context, host calls single-threaded per context
 * (as the reference is).  Every entry point returns TTC_OK (0) or an error
 * code whose value is the exit code the reference uses for the same
 * condition; ttc_last_error() gives the message.
 *
 * Each entry point names the reference routine it stands behind
 * (file:line in the reference tree).  Arrays the reference keeps as arrays
 * of structs are taken as (pointer to the first double, stride in doubles),
 * so &atom[k][0].x with stride sizeof(struct atoms)/sizeof(double) = 8, or
 * &ff[k][0].x with stride 3, can be passed straight through.
 */
#ifndef TTC_H
#define TTC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes = the reference's exit codes for the same condition */
#define TTC_OK                0
#define TTC_ERR_GRID          1   /* pme_qgrid.c:106-110 grid index out of range */
#define TTC_ERR_LIST_OVERFLOW 5   /* nblist.c:185-196 neighbour row overflow      */
#define TTC_ERR_ALLOC         11  /* allocate.c passim                            */
#define TTC_ERR_NOT_NEUTRAL   20  /* ewald.c:184-194 net charge > 5e-6            */
#define TTC_ERR_HIP           99  /* HIP runtime failure (no reference analogue)  */
#define TTC_ERR_ARG           111 /* checks.c flag/mode conflicts                 */

/* which list ttc_get_nblist exports */
#define TTC_LIST_NLIST    0       /* nlist[k]    (rclist)    */
#define TTC_LIST_NLIST_EW 1       /* nlist_ew[k] (rclist_ew) */

typedef struct ttc_ctx ttc_ctx;

/* Scalars the path reads from sim / box (defines.h:1228-1299). */
typedef struct {
    int    n;              /* box[k].boxns                                  */
    double lx, ly, lz;     /* box[k].lx, ly, lz                             */
    double rc;             /* sim.rc                                        */
    double rclist;         /* sim.rclist                                    */
    double rclist_ew;      /* sim.rclist_ew (used when ewald_lists != 0)    */
    int    ewald_lists;    /* 1: EWALD build, nlist nests in nlist_ew       */
    double kappa;          /* sim.kappa[k]                                  */
    double eps_rf;         /* sim.epsRF[k]                                  */
    double e14fac;         /* E14FAC (defines.h:72-76)                      */
    int    order;          /* sim.order (4 or 6)                            */
    int    gx, gy, gz;     /* sim.grid[0..2]                                */
    int    ntypes;         /* number of site types (STYPE tables)           */
    double dt;             /* sim.dt [ps]                                   */
} ttc_params;

/* Allocation and set-up: allocate.c:318-506, ewald_setup ewald.c:46-196
 * (pme_init pme_setup.c:59, pme_bspline 86, pme_dftmod 110).             */
int  ttc_create(const ttc_params *p, ttc_ctx **out);
/* pme_clean pme_setup.c:75 + release of every device buffer */
int  ttc_destroy(ttc_ctx *c);
const char *ttc_last_error(void);

/* Topology: pott[k][i].qq / .mas, atom[k][i].atomid, ljset / ljset14
 * (setlj.c:181-191, types x types row-major), exclusions xint[k] in
 * read_topology.c:847-891 order (a, b, xflag), bonds (cbond.c:65-120).
 * Fails with TTC_ERR_NOT_NEUTRAL like ewald.c:184-194.                   */
int  ttc_set_topology(ttc_ctx *c, const int *atomid, const double *qq, const double *mass,
                      const double *lj_eps, const double *lj_sig,
                      const double *lj14_eps, const double *lj14_sig,
                      int nx, const int *xa, const int *xb, const int *xflag,
                      int nbond, const int *ba, const int *bb,
                      const double *req, const double *kbond);

/* en[k].sewald (ewald.c:171-178), available after ttc_set_topology */
double ttc_self_energy(const ttc_ctx *c);

/* Host -> device / device -> host copies of per-site state, original site
 * order.  x points at the first x coordinate, stride in doubles.          */
int  ttc_set_positions(ttc_ctx *c, const double *x, int64_t stride);
int  ttc_get_positions(ttc_ctx *c, double *x, int64_t stride);
int  ttc_set_velocities(ttc_ctx *c, const double *v, int64_t stride);
int  ttc_get_velocities(ttc_ctx *c, double *v, int64_t stride);

/* nblist.c:32 nblist(): cell binning + sort + Verlet build on the device
 * from the current device positions; anchors snapshot and nl_flag = 0
 * (nblist.c:583-591).                                                     */
int  ttc_nblist(ttc_ctx *c);
/* nblist.c:607 nl_check(lb,ub,k): *nl_flag = 1 if a site in [lb,ub)
 * (original numbering) moved more than (rclist-rc)/4 since ttc_nblist.    */
int  ttc_nl_check(ttc_ctx *c, int lb, int ub, int *nl_flag);
/* Export of nlist[k] / nlist_ew[k] in the reference layout: row a holds
 * count[a] entries b > a, ascending, list[a*max_nb + j], xflag likewise.
 * Returns TTC_ERR_LIST_OVERFLOW if a row exceeds max_nb.                  */
int  ttc_get_nblist(ttc_ctx *c, int which, int max_nb, int *count, int *list, int *xflag);

/* cnonbond_nblist.c:51 cnonbondnl(ibox, energy): LJ + SPME real space +
 * exclusion/1-4 corrections + pme_calc.  Forces are ADDED into ff
 * (caller zeroes, forces.c:126-130); energy[8] as cnonbond_nblist.c:613-620;
 * vir_nbond = pvir[k].nbond, vir_kspace = pvir[k].ewald_kspace.  Any of
 * ff / energy / vir_* may be NULL.                                       */
int  ttc_cnonbondnl(ttc_ctx *c, double *ff, int64_t stride, double *energy8,
                    double *vir_nbond6, double *vir_kspace6);
/* pme_calc.c:18 pme_calc(k, energy): charge_grid + FFTs + convolution +
 * pme_force; energy[3] = {recip, self, recip - self}.                    */
int  ttc_pme_calc(ttc_ctx *c, double *ff, int64_t stride, double *energy3, double *vir6);
/* The pair loop of cnonbondnl alone (cnonbond_nblist.c:134-523):
 * e5 = {enonbond, enonbonds, eewald_real, eewald_reals, eewald_intra}.  */
int  ttc_pair(ttc_ctx *c, double *ff, int64_t stride, double *e5, double *vir6);
/* cbond.c:65-120: returns bond energy through *ebond, forces added.       */
int  ttc_cbond(ttc_ctx *c, double *ff, int64_t stride, double *ebond, double *vir6);

/* drift.c:20 drift(k): removes the net momentum, v_i -= (sum_j m_j v_j / N)
 * / m_i (the reference's non-KONS, non-ZEROAM form).                      */
int  ttc_drift(ttc_ctx *c);
/* sim.drift: the MD loops (ttc_md_vverlet, ttc_md_nhc, ttc_md_mts) call
 * drift after every step whose number within the call is a multiple of
 * every (nve_md.c:208, nvt_md.c:302), after the step's energy row; 0 (the
 * default) never.                                                          */
int  ttc_set_drift(ttc_ctx *c, int every);

/* ---- angle-type bonded terms ------------------------------------------
 * read_topology.c's bend / urey / tors / impr arrays of box k:
 *   bends      (a,b,c) 3 per term, angeq [rad], krbend: cbend.c:63-170
 *   Urey-Bradley (a,b) 2 per term, Seq, k:              curey.c:57-140
 *   torsions   (a,b,c,d) 4 per term, 6 Fourier entries each in kphi, nphi,
 *              delphi (the series ends at the first kphi == 0):
 *                                                       ctorsion.c:67-236
 *   impropers  (a,b,c,d), kimpr, angeq [rad]:           cimproper.c:62-225
 * Any count may be 0; a new call replaces all four sets.  Once set, every
 * force evaluation that includes bonds includes them (ttc_forces, the MD
 * loops, the replica batch; RESPA splits them as force_short.c /
 * force_long.c do: bends and Urey-Bradley short, torsions and impropers
 * long), and the energy rows' bond column holds bond + bend + urey + tors
 * + impr (calcvalue.c:44-53).  No counterpart call in the reference (it
 * reads the arrays from its topology files).                              */
int  ttc_set_angles(ttc_ctx *c, int nbend, const int *bend_abc, const double *bend_eq, const double *bend_k,
                    int nurey, const int *urey_ab, const double *urey_s0, const double *urey_k,
                    int ntors, const int *tors_abcd, const double *kphi, const double *nphi,
                    const double *delphi, int nimpr, const int *impr_abcd, const double *impr_k,
                    const double *impr_eq);
/* cbend.c:25 cbend(k), curey.c curey(k), ctorsion.c:25 ctorsion(k),
 * cimproper.c cimproper(k): the kind's energy, forces added into ff,
 * vir6 = pvir[k].bend / .urey / .tors / .impr.  ff / energy / vir6 may be
 * NULL.                                                                   */
int  ttc_cbend(ttc_ctx *c, double *ff, int64_t stride, double *ebend, double *vir6);
int  ttc_curey(ttc_ctx *c, double *ff, int64_t stride, double *eurey, double *vir6);
int  ttc_ctorsion(ttc_ctx *c, double *ff, int64_t stride, double *etors, double *vir6);
int  ttc_cimproper(ttc_ctx *c, double *ff, int64_t stride, double *eimpr, double *vir6);
/* cbend2.c, curey2.c, ctorsion2.c, cimproper2.c (CONFIGT): as above with
 * hes3 = the kind's sums of d2E/dx_i^2, d2E/dy_i^2, d2E/dz_i^2 (added to
 * config[k].hesx/y/z by the reference).                                    */
int  ttc_cbend2(ttc_ctx *c, double *ff, int64_t stride, double *ebend, double *hes3);
int  ttc_curey2(ttc_ctx *c, double *ff, int64_t stride, double *eurey, double *hes3);
int  ttc_ctorsion2(ttc_ctx *c, double *ff, int64_t stride, double *etors, double *hes3);
int  ttc_cimproper2(ttc_ctx *c, double *ff, int64_t stride, double *eimpr, double *hes3);
/* en[k].bond, .bend, .urey, .tors, .impr of the last evaluation that
 * computed energies with bonded terms (ttc_forces, ttc_configtemp, an
 * energy step of an MD loop).                                              */
int  ttc_bonded_energies(ttc_ctx *c, double *e5);

/* ---- CONFIGT (configurational temperature; used by the CTDOS drivers) ---
 * hes3 = {hesx, hesy, hesz}: sums over sites of d2U/dx_i^2, d2U/dy_i^2,
 * d2U/dz_i^2 (config[k].hesx/y/z).                                       */
/* cnonbond_nblist2.c:47 cnonbondnl2(ibox, energy): ttc_cnonbondnl plus the
 * Hessian diagonal of the LJ, real-space and excluded-pair terms
 * (cnonbond_nblist2.c:286-297, 403-410, 433-440) and of the SPME
 * reciprocal energy.  The reference calls pme_calc2 for the latter
 * (cnonbond_nblist2.c:586) but never defines it; here it is the exact
 * second derivative of the reciprocal energy PME evaluates.              */
int  ttc_cnonbondnl2(ttc_ctx *c, double *ff, int64_t stride, double *energy8, double *hes3);
/* cbond2.c:65-130 cbond2(k): bond energy, forces added, Hessian diagonal. */
int  ttc_cbond2(ttc_ctx *c, double *ff, int64_t stride, double *ebond, double *hes3);
/* forces.c:228-256 (CONFIGT branch) + configtemp.c:17-45: one force and
 * Hessian evaluation; num = sum |F_i|^2, den = hesr = hesx+hesy+hesz,
 * T = num / den / RG / 0.001.  hes3 optional.                            */
int  ttc_configtemp(ttc_ctx *c, double *num, double *den, double *T, double *hes3);

/* forces.c:118 forces(ibox): zero, cbond, the angle-type terms set by
 * ttc_set_angles (cbend, curey, ctorsion, cimproper), cnonbondnl; en[k]
 * slots (forces.c:204-219) are returned as en9 = {bond, nbond, nbonds,
 * coulomb, rewald, rewalds, kewald, sewald, iewald} (bend, urey, tors, impr
 * through ttc_bonded_energies).                                           */
int  ttc_forces(ttc_ctx *c, double *ff, int64_t stride, double *en9);

/* integrate_vverlet.c:36-103 repeated nsteps times with nl_check/nblist
 * on the displacement trigger (nve_md.c:369-370), all on the device as one
 * captured HIP graph per step.  Every energy_every steps (0: never) a row
 * {step, kinet, bond, nbond, nbonds, coulomb, kewald, rewald, iewald}
 * is appended to energies (rows of 9 doubles, capacity max_rows).
 * vscal multiplies velocities as the reference's vscal argument.
 * *rebuilds (optional) receives the number of list rebuilds.             */
int  ttc_md_vverlet(ttc_ctx *c, int nsteps, double vscal, int energy_every,
                    double *energies, int max_rows, int *nrows, int *rebuilds);

/* kinet.c:18-68 on the device; ken6 optional. Returns en[k].kinet.       */
int  ttc_kinet(ttc_ctx *c, double *kinet, double *ken6);

/* force_short.c:46 force_short(k): the bonded (fast) forces, written (not
 * added) to ff_short; *ebond = bond energy.                               */
int  ttc_force_short(ttc_ctx *c, double *ff_short, int64_t stride, double *ebond);
/* force_long.c:43 force_long(k): the nonbonded (slow) forces cnonbondnl
 * gives, written to ff_long; energy slots as ttc_cnonbondnl.             */
int  ttc_force_long(ttc_ctx *c, double *ff_long, int64_t stride, double *energy8);

/* integrate_mts.c:38-140 (RESPA): nouter long steps of sim.dtlong = dtlong,
 * each with ninner short steps of sim.dt on the bonded forces, the
 * nonbonded forces refreshed once per long step; DLIST trigger checked
 * before every long step.  Energy rows as ttc_md_vverlet, per long step.  */
int  ttc_md_mts(ttc_ctx *c, int nouter, int ninner, double dtlong, double vscal, int energy_every,
                double *energies, int max_rows, int *nrows, int *rebuilds);

/* Nose-Hoover chain thermostat of integrate_nhc.c:40-161: target T [K],
 * mass Q [kg A^2], M chain elements, Nc x Nys (Yoshida-Suzuki weights w)
 * sub-steps, nfree degrees of freedom (box[k].nfree).  Resets the chain.  */
int  ttc_set_nhc(ttc_ctx *c, double T, double Q, int M, int Nc, int Nys, const double *w, double nfree);
/* chain state {r, v, G} per element (3*M doubles) */
int  ttc_get_nhc(ttc_ctx *c, double *zeta);
/* New target temperature T and mass Q of the chain, keeping its state
 * (the temperature a replica exchange hands over, swap_box.c:4-86).      */
int  ttc_set_nhc_temperature(ttc_ctx *c, double T, double Q);
/* v *= factor for every site (swap_box.c:33-43 velocity rescaling).      */
int  ttc_scale_velocities(ttc_ctx *c, double factor);
/* nvt_md.c:528-531 (sim.ID2 == 4): integrate_nhc(dt); integrate_vverlet;
 * integrate_nhc(dt), nsteps times on the device.                          */
int  ttc_md_nhc(ttc_ctx *c, int nsteps, int energy_every, double *energies, int max_rows, int *nrows,
                int *rebuilds);

/* enbnd_mc.c:26 enbnd_mc(lb,ub,k,energy): nonbonded energy of the pairs
 * with at least one site in [lb,ub) (original numbering) plus the full
 * reciprocal-space energy (pme_calc_mc); energy[8] slots as cnonbondnl.   */
int  ttc_enbnd_mc(ttc_ctx *c, int lb, int ub, double *energy8);
/* nblist_pivot.c:26 nblist_pivot(k,ub): refreshes the list after a move of
 * sites below ub.  The device rebuilds every row (a superset of the rows
 * a < ub the reference refreshes), from the current positions.           */
int  ttc_nblist_pivot(ttc_ctx *c, int ub);

/* ---- Batched replicas (nvt_md.c + swap_box.c, one MPI process per
 * replica in the reference) ------------------------------------------------
 * A batch advances many contexts of any sizes together.  Launch mode 2:
 * every kernel of the NHC step is launched once for the whole batch, its
 * grid the concatenation of the replicas' own grids (about 20 launches per
 * step for any number of replicas); mode 0: the replicas' step graphs
 * captured into one graph, one branch per replica; mode 1: one graph per
 * replica on its own stream; steps_per_launch steps per graph launch.
 * Every context needs ttc_set_nhc.                                         */
typedef struct ttc_batch ttc_batch;
int  ttc_batch_create(ttc_ctx *const *ctx, int nrep, uint64_t seed, ttc_batch **out);
int  ttc_batch_destroy(ttc_batch *b);
int  ttc_batch_set_launch(ttc_batch *b, int mode, int steps_per_launch);
/* nvt_md.c:528-531 nsteps times on every replica (nsteps a multiple of the
 * steps per launch); potens[r] = en.potens (calcvalue.c:40, bond + nbonds)
 * of replica r after the last step, rebuilds[r] its list rebuilds.       */
int  ttc_batch_md_nhc(ttc_batch *b, int nsteps, double *potens, int *rebuilds);
/* swap_box.c:7-11 for the replicas at temperature ranks (t, t+1), t = parity
 * mod 2: accepted swaps exchange the two thermostat temperatures and
 * rescale velocities by sqrt(T_new/T_old).  *accepted = swaps made.      */
int  ttc_batch_exchange(ttc_batch *b, const double *potens, int parity, int *accepted);
/* order[t] = replica at temperature rank t, temps[r] = current temperature
 * of replica r, wall ms per step of the last ttc_batch_md_nhc, totals.    */
int  ttc_batch_state(ttc_batch *b, int *order, double *temps, double *ms_per_step,
                     long long *attempts, long long *accepts);
/* Tries launch modes {0,1,2} x steps per launch {1,5,10,25} for
 * trial_steps steps each (real dynamics) and keeps the fastest;
 * timings[12] optional.                                                   */
int  ttc_batch_autotune(ttc_batch *b, int trial_steps, int *best_mode, int *best_spl, double *timings);

/* Device time of one kernel class, measured with HIP events on the
 * context's stream around 20 back-to-back launches on the current state:
 * which = 0 pair kernel, 1 spread, 3 the five PME FFT passes (influence
 * function included), 4 PME gather + bonds, 5 / 6 the velocity-Verlet
 * loop's list build (half / full list, at the anchors, 10 launches);
 * which = 2 is the average step of the last ttc_md_vverlet call.
 * Microseconds.  After 3 or 4 the forces are recomputed before the next
 * MD call.                                                                */
int  ttc_timing(ttc_ctx *c, int which, double *usec);

/* Autotuner of the launch shape: times the pair kernel at 4, 8 and 16
 * waves per 64-site group on the current list and keeps the fastest
 * (usec3[3] optional: the three timings).  No reference counterpart.    */
int  ttc_tune(ttc_ctx *c, int *pair_waves, double *usec3);
/* Autotuner of the charge spreading: sites per pass of the spread kernel
 * (64, 128, 256, those that fit the LDS) timed on the current positions;
 * the fastest stays (usec3 optional, 0 for a chunk that does not fit).    */
int  ttc_tune_spread(ttc_ctx *c, int *chunk, double *usec3);
/* Both tuners behind a cache file: a text line per context shape (site
 * count, box, grid, spline order, cutoffs, list format, type count, GPU
 * architecture and CU count) with the choices; a matching line is applied
 * without timing (*from_cache = 1), otherwise the tuners run and a line is
 * appended.  No reference counterpart.                                    */
int  ttc_tune_cached(ttc_ctx *c, const char *path, int *pair_waves, int *spread_chunk, int *from_cache);

/* The force array ff[k] of the last evaluation at the current positions:
 * after ttc_md_vverlet the forces the last step computed
 * (integrate_vverlet.c:88 forces(k)), after ttc_forces the same as it
 * returned.  Written (not added) in the original site order.              */
int  ttc_get_forces(ttc_ctx *c, double *ff, int64_t stride);

/* Capacity limits for MD-loop list builds, below the allocated capacities:
 * a build with a row longer than row_limit or a half-list window wider than
 * window_limit reports an overflow as if the capacity were that small
 * (0: no limit).  The loop then grows and redoes the step, which lifts the
 * limits.  No counterpart in the reference (its row capacity is the fixed
 * max_num_neighbors of allocate.c:321-322); exercises the overflow path.  */
int  ttc_set_capacity_limits(ttc_ctx *c, int row_limit, int window_limit);

/* Half-list windows of the velocity-Verlet loop: widest window of the
 * current list (0 for a full list), window capacity, *half = 1 when the
 * current list is a half list.  No counterpart in the reference.          */
int  ttc_window_stats(ttc_ctx *c, int *max_window, int *window_capacity, int *half);

/* Counters of the last list build (an export's scratch build included):
 * entries of the list (each pair in both rows), longest row, row capacity.
 * After ttc_nblist, entries / 2 is the size of the reference's half list. */
int  ttc_list_stats(ttc_ctx *c, int64_t *entries, int *max_row, int *capacity);

#ifdef __cplusplus
}
#endif
#endif
