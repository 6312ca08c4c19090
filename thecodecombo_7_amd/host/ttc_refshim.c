/*
 * ttc_refshim.c -- the reference-side binding of libttc.so.
 *
 * A maintainer of the reference C code links this file instead of
 * nblist.c, nblist_pivot.c, cnonbond_nblist.c (and cnonbond_nblist2.c in a
 * CONFIGT build), enbnd_mc.c, pme_qgrid.c,
 * pme_calc.c, pme_setup.c, pme_spline.c and the SPME part of ewald.c.  It
 * defines the reference's own entry points with the reference's own
 * signatures, reads and writes the reference's global structs from
 * defines.h in place, and forwards the work to the C ABI of include/ttc.h.
 * The reference's callers (forces.c, force_long.c, force_short.c,
 * integrate_*.c, nve_md.c, nvt_md.c, mcmove_*.c, ctdos_*.c, ...) compile
 * and link unchanged.
 *
 * Build (beside the reference's objects, same -D flags):
 *   gcc -std=c99 -O2 -DNLIST -DDLIST -DSTYPE -DS14 -DEWALD -DSPME -DPRESSURE \
 *       -I<reference dir> -I<this repo>/include -c ttc_refshim.c
 *   ... link with -L<this repo>/thecodecombo_7_amd -lttc
 *
 * Data contract: the host structs stay the source of truth, as in the
 * reference.  Each entry point uploads the positions it reads
 * (atom[k][i].x/y/z, stride sizeof(struct atoms)) and adds/writes the
 * forces it produces into ff[k] (struct veloc, stride 3 doubles); the list
 * lives on the device and is exported into nlist[k] / nlist_ew[k] only
 * when ttc_shim_export_lists is set (no replaced routine reads them).
 * Errors end the program with the reference's exit code for the same
 * condition (list overflow 5, net charge 20, grid 1, allocation 11).
 *
 * FFTW: in an SPME build defines.h:41 includes <fftw3.h>, so the header is
 * still needed to compile, and allocate.c:500 still allocates qgrida with
 * fftw_malloc (nothing reads it once pme_calc runs on the GPU).  Compiled
 * with -DTTC_SHIM_FFTW_ALLOC this file defines fftw_malloc / fftw_free
 * itself, and the program links without -lfftw3.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "defines.h"
#include "ttc.h"

/* set to 1 to mirror the device lists into nlist[k] / nlist_ew[k] after
 * every nblist() / nblist_pivot() */
int ttc_shim_export_lists = 0;

static ttc_ctx *g_ctx[NBOXES];

static void die(int code, const char *where)
{
    printf("%s: %s\n", where, ttc_last_error());
    exit(code == TTC_ERR_HIP ? 1 : code);
}

static void ok(int code, const char *where)
{
    if (code != TTC_OK) die(code, where);
}

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) { printf("ttc_refshim: out of memory\n"); exit(11); }
    return p;
}

/* Device context of box k, created on first use from sim, box[k], pott[k],
 * atom[k][].atomid, ljset / ljset14, xint[k] and bond[k] (the state that
 * ewald_setup / pme_init find set up by read_* and setlj). */
static ttc_ctx *ctx_of(int k)
{
    if (g_ctx[k]) return g_ctx[k];
    int n = box[k].boxns;
    int ntypes = 0;
    for (int i = 0; i < n; i++)
        if (atom[k][i].atomid + 1 > ntypes) ntypes = atom[k][i].atomid + 1;

    ttc_params p;
    memset(&p, 0, sizeof p);
    p.n = n;
    p.lx = box[k].lx; p.ly = box[k].ly; p.lz = box[k].lz;
    p.rc = sim.rc;
    p.rclist = sim.rclist;
    p.rclist_ew = sim.rclist_ew;
    p.ewald_lists = 1;
    p.kappa = sim.kappa[k];
    p.eps_rf = sim.epsRF[k];
    p.e14fac = E14FAC;
    p.order = sim.order;
    p.gx = sim.grid[0]; p.gy = sim.grid[1]; p.gz = sim.grid[2];
    p.ntypes = ntypes;
    p.dt = sim.dt;
    ok(ttc_create(&p, &g_ctx[k]), "ttc_create");

    int *type = xmalloc(sizeof(int) * n);
    double *qq = xmalloc(sizeof(double) * n), *mas = xmalloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) { type[i] = atom[k][i].atomid; qq[i] = pott[k][i].qq; mas[i] = pott[k][i].mas; }
    size_t t2 = (size_t)ntypes * ntypes;
    double *eps = xmalloc(sizeof(double) * t2), *sig = xmalloc(sizeof(double) * t2);
    double *eps14 = xmalloc(sizeof(double) * t2), *sig14 = xmalloc(sizeof(double) * t2);
    for (int a = 0; a < ntypes; a++)
        for (int b = 0; b < ntypes; b++) {
            eps[a * ntypes + b] = ljset[a][b].eps;
            sig[a * ntypes + b] = ljset[a][b].sig;
            eps14[a * ntypes + b] = ljset14[a][b].eps;
            sig14[a * ntypes + b] = ljset14[a][b].sig;
        }
    int nx = xintN[k];
    int *xa = xmalloc(sizeof(int) * nx), *xb = xmalloc(sizeof(int) * nx), *xf = xmalloc(sizeof(int) * nx);
    for (int i = 0; i < nx; i++) { xa[i] = xint[k][i].a; xb[i] = xint[k][i].b; xf[i] = xint[k][i].xflag; }
    int nb = bondN[k];
    int *ba = xmalloc(sizeof(int) * nb), *bb = xmalloc(sizeof(int) * nb);
    double *req = xmalloc(sizeof(double) * nb), *kb = xmalloc(sizeof(double) * nb);
    for (int i = 0; i < nb; i++) {
        ba[i] = bond[k][i].a; bb[i] = bond[k][i].b; req[i] = bond[k][i].req; kb[i] = bond[k][i].krbond;
    }
    int rc = ttc_set_topology(g_ctx[k], type, qq, mas, eps, sig, eps14, sig14, nx, xa, xb, xf, nb, ba, bb, req, kb);
    free(type); free(qq); free(mas); free(eps); free(sig); free(eps14); free(sig14);
    free(xa); free(xb); free(xf); free(ba); free(bb); free(req); free(kb);
    ok(rc, "ewald_setup");
    return g_ctx[k];
}

static ttc_ctx *upload(int k)
{
    ttc_ctx *c = ctx_of(k);
    ok(ttc_set_positions(c, &atom[k][0].x, sizeof(struct atoms) / sizeof(double)), "ttc_set_positions");
    return c;
}

static void anchors(int k)
{
    for (int i = 0; i < box[k].boxns; i++) {
        rx0[k][i] = atom[k][i].x;
        ry0[k][i] = atom[k][i].y;
        rz0[k][i] = atom[k][i].z;
    }
    nl_flag[k] = 0;
}

static void export_lists(int k, ttc_ctx *c)
{
    if (!ttc_shim_export_lists) return;
    int n = box[k].boxns, m = max_num_neighbors;
    int *cnt = xmalloc(sizeof(int) * n);
    int *lst = xmalloc(sizeof(int) * (size_t)n * m), *xf = xmalloc(sizeof(int) * (size_t)n * m);
    struct nlist_def *dst[2] = {&nlist[k], &nlist_ew[k]};
    for (int w = 0; w < 2; w++) {
        ok(ttc_get_nblist(c, w, m, cnt, lst, xf), "nblist");
        for (int a = 0; a < n; a++) {
            dst[w]->count[a] = cnt[a];
            memcpy(dst[w]->list[a], lst + (size_t)a * m, sizeof(int) * cnt[a]);
            memcpy(dst[w]->xflag[a], xf + (size_t)a * m, sizeof(int) * cnt[a]);
        }
    }
    free(cnt); free(lst); free(xf);
}

/* nblist.c:32 */
void nblist(int ibox)
{
    ttc_ctx *c = upload(ibox);
    ok(ttc_nblist(c), "nblist");
    anchors(ibox);
    export_lists(ibox, c);
}

/* nblist.c:607 -- sets nl_flag[k] = 1 when triggered, never clears it */
void nl_check(int lb, int ub, int k)
{
    int f = 0;
    ok(ttc_nl_check(upload(k), lb, ub, &f), "nl_check");
    if (f) nl_flag[k] = 1;
}

/* nblist_pivot.c:26 */
void nblist_pivot(int ibox, int ub)
{
    ttc_ctx *c = upload(ibox);
    ok(ttc_nblist_pivot(c, ub), "nblist_pivot");
    anchors(ibox);
    export_lists(ibox, c);
}

/* cnonbond_nblist.c:51 -- forces added into ff[k], energy[8] written */
void cnonbondnl(int ibox, double *energy)
{
    int k = ibox;
    double vn[6], vk[6];
    ok(ttc_cnonbondnl(upload(k), &ff[k][0].x, 3, energy, vn, vk), "cnonbondnl");
#ifdef PRESSURE
    for (int i = 0; i < 6; i++) { pvir[k].nbond[i] = vn[i]; pvir[k].ewald_kspace[i] = vk[i]; }
#endif
}

#ifdef CONFIGT
/* cnonbond_nblist2.c:47 -- as cnonbondnl, plus the Hessian diagonal of the
 * nonbonded terms added into config[k].hesx/y/z (forces.c:232-235 zeroes
 * them); the reciprocal part is the exact second derivative of the SPME
 * energy (the reference's pme_calc2 does not exist) */
void cnonbondnl2(int ibox, double *energy)
{
    int k = ibox;
    double h[3];
    ok(ttc_cnonbondnl2(upload(k), &ff[k][0].x, 3, energy, h), "cnonbondnl2");
    config[k].hesx += h[0];
    config[k].hesy += h[1];
    config[k].hesz += h[2];
}
#endif

/* pme_calc.c:18 -- charge_grid, both FFTs, convolution and pme_force */
void pme_calc(int k, double *energy)
{
    double vk[6];
    ok(ttc_pme_calc(upload(k), &ff[k][0].x, 3, energy, vk), "pme_calc");
#ifdef PRESSURE
    for (int i = 0; i < 6; i++) pvir[k].ewald_kspace[i] = vk[i];
#endif
}

/* pme_qgrid.c:13 and pme_calc.c:122: the grid and the spline weights live
 * on the device and are produced and consumed inside pme_calc(); nothing
 * outside the replaced files reads qgrida / mm / dm. */
void charge_grid(int k) { (void)k; }
void pme_force(int k) { (void)k; }

/* pme_setup.c:59 / ewald.c:46 -- the device context holds the FFT plans,
 * bsp_mod tables and the self energy; the neutrality check exits 20 */
void pme_init(int k) { (void)ctx_of(k); }
void ewald_setup(int ibox)
{
    ttc_ctx *c = ctx_of(ibox);
    en[ibox].sewald = ttc_self_energy(c);
}

/* pme_setup.c:75 */
void pme_clean(void)
{
    for (int k = 0; k < NBOXES; k++)
        if (g_ctx[k]) { ttc_destroy(g_ctx[k]); g_ctx[k] = NULL; }
}

/* enbnd_mc.c:26 -- energy of the pairs touching [lb,ub) + reciprocal part */
void enbnd_mc(int lb, int ub, int ibox, double *energy)
{
    ok(ttc_enbnd_mc(upload(ibox), lb, ub, energy), "enbnd_mc");
}

/* The angle-type terms of box k (bend[k], urey[k], tors[k], impr[k] as
 * read_topology.c fills them) on the device, for the device-resident loop;
 * the replaced cnonbondnl() never evaluates them (forces.c calls cbend.c ..
 * cimproper.c on the host). */
static void device_angles(int k, ttc_ctx *c)
{
    static int done[NBOXES];
    if (done[k]) return;
    done[k] = 1;
    int nb = bendN[k], nt = torsN[k], ni = imprN[k], nu = 0;
#ifndef NEUTRAL
    nu = ureyN[k];
#endif
    if (!nb && !nt && !ni && !nu) return;
    int *bi = xmalloc(sizeof(int) * 3 * (nb + 1)), *ui = xmalloc(sizeof(int) * 2 * (nu + 1));
    int *ti = xmalloc(sizeof(int) * 4 * (nt + 1)), *ii = xmalloc(sizeof(int) * 4 * (ni + 1));
    double *beq = xmalloc(sizeof(double) * (nb + 1)), *bk = xmalloc(sizeof(double) * (nb + 1));
    double *us = xmalloc(sizeof(double) * (nu + 1)), *uk = xmalloc(sizeof(double) * (nu + 1));
    double *kp = xmalloc(sizeof(double) * 6 * (nt + 1)), *np = xmalloc(sizeof(double) * 6 * (nt + 1));
    double *dp = xmalloc(sizeof(double) * 6 * (nt + 1));
    double *ik = xmalloc(sizeof(double) * (ni + 1)), *ieq = xmalloc(sizeof(double) * (ni + 1));
    for (int i = 0; i < nb; i++) {
        bi[3 * i] = bend[k][i].a; bi[3 * i + 1] = bend[k][i].b; bi[3 * i + 2] = bend[k][i].c;
        beq[i] = bend[k][i].angeq; bk[i] = bend[k][i].krbend;
    }
#ifndef NEUTRAL
    for (int i = 0; i < nu; i++) {
        ui[2 * i] = urey[k][i].a; ui[2 * i + 1] = urey[k][i].b; us[i] = urey[k][i].Seq; uk[i] = urey[k][i].k;
    }
#endif
    for (int i = 0; i < nt; i++) {
        ti[4 * i] = tors[k][i].a; ti[4 * i + 1] = tors[k][i].b; ti[4 * i + 2] = tors[k][i].c; ti[4 * i + 3] = tors[k][i].d;
        for (int s = 0; s < 6; s++) {
            kp[6 * i + s] = tors[k][i].kphi[s]; np[6 * i + s] = tors[k][i].nphi[s]; dp[6 * i + s] = tors[k][i].delphi[s];
        }
    }
    for (int i = 0; i < ni; i++) {
        ii[4 * i] = impr[k][i].a; ii[4 * i + 1] = impr[k][i].b; ii[4 * i + 2] = impr[k][i].c; ii[4 * i + 3] = impr[k][i].d;
        ik[i] = impr[k][i].kimpr; ieq[i] = impr[k][i].angeq;
    }
    int rc = ttc_set_angles(c, nb, bi, beq, bk, nu, ui, us, uk, nt, ti, kp, np, dp, ni, ii, ik, ieq);
    free(bi); free(ui); free(ti); free(ii); free(beq); free(bk); free(us); free(uk); free(kp); free(np); free(dp);
    free(ik); free(ieq);
    ok(rc, "ttc_set_angles");
}

void forces(int);

/* integrate_vverlet(k, vscal) repeated nsteps times with the DLIST trigger
 * and rebuild of nve_md.c:369-370 before each step, as one device-resident
 * loop (ttc_md_vverlet: state stays in HBM, one captured graph per step)
 * instead of nsteps host round trips.  Starts from the host state (atom[k],
 * vv[k]) and leaves it as the reference's own loop would: atom[k] wrapped,
 * atnopbc[k] advanced by each site's minimum-image displacement over the
 * call, vv[k] and uu[k] the final velocities, and ff[k], en[k], pvir[k] from
 * a closing forces(k) at the final positions -- so write_trr.c and the
 * A/B restart files of checkpoint.c (WriteCheckpoint / ReadCheckpoint) see
 * the current state.  No reference counterpart as one call. */
void ttc_shim_md_vverlet(int k, int nsteps, double vscal)
{
    ttc_ctx *c = upload(k);
    int n = box[k].boxns;
    device_angles(k, c);
    ok(ttc_set_velocities(c, &vv[k][0].x, 3), "ttc_set_velocities");
    ok(ttc_md_vverlet(c, nsteps, vscal, 0, NULL, 0, NULL, NULL), "integrate_vverlet");
    double *x = xmalloc(sizeof(double) * 3 * (size_t)n);
    ok(ttc_get_positions(c, x, 3), "ttc_get_positions");
    ok(ttc_get_velocities(c, &vv[k][0].x, 3), "ttc_get_velocities");
    const double L[3] = {box[k].lx, box[k].ly, box[k].lz};
    for (int i = 0; i < n; i++) {
        double *a = &atom[k][i].x, *u = &atnopbc[k][i].x;
        for (int d = 0; d < 3; d++) {
            double dd = x[3 * i + d] - a[d];
            dd -= L[d] * rint(dd / L[d]);
            u[d] += dd;
            a[d] = x[3 * i + d];
        }
        uu[k][i] = vv[k][i];
    }
    free(x);
    forces(k);
}

/* ewald.c:817 -- Numerical Recipes erfcc (Chebyshev fit, fractional error
 * below 1.2e-7), kept on the host for the remaining host callers
 * (cgolik.c, enbnd2_mc.c, ewald.c's classic-Ewald routines) */
double erfc(double x)
{
    double z = fabs(x);
    double t = 1.0 / (1.0 + 0.5 * z);
    double p = -z * z - 1.26551223 + t * (1.00002368 + t * (0.37409196 + t * (0.09678418 +
               t * (-0.18628806 + t * (0.27886807 + t * (-1.13520398 + t * (1.48851587 +
               t * (-0.82215223 + t * 0.17087277))))))));
    double r = t * exp(p);
    return x >= 0.0 ? r : 2.0 - r;
}

#ifdef TTC_SHIM_FFTW_ALLOC
/* FFTW's allocation pair (fftw3.h: void *fftw_malloc(size_t); void
 * fftw_free(void *)) for allocate.c:500, the one FFTW call left outside the
 * replaced pme_*.c files.  No FFTW plan ever touches the memory, so plain
 * malloc alignment is enough. */
void *fftw_malloc(size_t n) { return malloc(n ? n : 1); }
void fftw_free(void *p) { free(p); }
#endif
