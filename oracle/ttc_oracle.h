/*
 * ttc_oracle.h -- FP64 CPU restatement of TheCodeCombo_7's nonbonded path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against (tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg).  The product library (libttc.so) never links, loads or
 * calls it.
 *
 * Every function follows one reference routine; the file:line of the routine
 * is given beside each declaration.  Conventions kept from the reference:
 *   - positions are FP64, AoS (x,y,z per site), wrapped into [-L/2, L/2];
 *   - energies kJ/mol, lengths Angstrom, charges e, masses amu, time ps;
 *   - forces are ADDED into the caller's array (the caller zeroes it, as
 *     forces.c:126-130 does);
 *   - half Verlet list: row a holds b > a (nblist.c:95-307).
 *
 * Parity pins (what checks each function against something other than
 * itself) are listed in the header comment of ttc_oracle.c.
 */
#ifndef TTC_ORACLE_H
#define TTC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants: defines.h:60-71 (same values, same product order). */
#define ORC_PI      3.141592653589793238462643383
#define ORC_KB      1.380662E-23
#define ORC_NA      6.022045E23
#define ORC_AU      1.6605655E-27
#define ORC_ELEC    8.987551789e19
#define ORC_ELEQ    1.602176462e-19
#define ORC_QFACTOR (ORC_ELEC * ORC_ELEQ * ORC_ELEQ * ORC_NA * 1.0E-3)

typedef struct {
    int n;                 /* box[k].boxns                                   */
    double lx, ly, lz;     /* box[k].lx..lz (hx = lx/2 etc.)                 */
    double rc;             /* sim.rc                                         */
    double rclist;         /* sim.rclist                                     */
    double rclist_ew;      /* sim.rclist_ew (nested list, EWALD builds)      */
    double kappa;          /* sim.kappa[k]                                   */
    double eps_rf;         /* sim.epsRF[k]                                   */
    double e14fac;         /* E14FAC (defines.h:72-76; a parameter here)     */
    int order;             /* sim.order                                      */
    int gx, gy, gz;        /* sim.grid[0..2]                                 */
    int ntypes;            /* number of site types (STYPE)                   */
    const int *type;       /* atom[k][i].atomid                              */
    const double *q;       /* pott[k][i].qq                                  */
    const double *lj_eps;  /* ljset[ta][tb].eps   (ntypes*ntypes, row major) */
    const double *lj_sig;  /* ljset[ta][tb].sig                              */
    const double *lj14_eps;/* ljset14[ta][tb].eps                            */
    const double *lj14_sig;/* ljset14[ta][tb].sig                            */
} orc_system;

/* Exclusions in compressed-row form, both directions, one resolved flag per
 * (a,b): the flag of the LAST xint entry naming the pair, which is what the
 * hash scan of nblist.c:156-175 yields (later matches overwrite).          */
typedef struct {
    int n;
    int *off;      /* n+1 */
    int *partner;  /* off[n] */
    int *flag;     /* off[n]; xflag 1 bond, 2 bend, 3 exNB, 4 1-4 */
} orc_excl;

/* --- ewald.c -------------------------------------------------------------- */
double orc_erfc(double x);                                   /* ewald.c:817-835 */
double orc_ewald_self(int n, const double *q, double kappa, double eps_rf);
                                                             /* ewald.c:171-178 */
int orc_check_neutral(int n, const double *q);               /* ewald.c:184-194 */

/* --- pme_spline.c / pme_setup.c ----------------------------------------- */
void orc_fillspline(double w, int order, double *array, double *darray);
                                                             /* pme_spline.c:10-49 */
void orc_dftmod(double *bmod, const double *barr, int grid); /* pme_setup.c:110-130 */
void orc_pme_bspline(int order, int gx, int gy, int gz,
                     double *bx, double *by, double *bz);    /* pme_setup.c:86-107 */

/* --- FFT (stands in for FFTW3's unnormalised complex 3-D DFT) ----------- */
void orc_fft3d(double *data, int nx, int ny, int nz, int sign);
void orc_dft3d_naive(const double *in, double *out, int nx, int ny, int nz, int sign);

/* --- exclusions ---------------------------------------------------------- */
int  orc_excl_build(int n, int nx, const int *xa, const int *xb, const int *xf,
                    orc_excl *out);                          /* read_topology.c:847-891 */
void orc_excl_free(orc_excl *e);
int  orc_excl_flag(const orc_excl *e, int a, int b);

/* --- nblist.c ------------------------------------------------------------ */
int orc_nblist(const orc_system *s, const double *pos, const orc_excl *ex,
               int use_ew, int max_nb,
               int *count, int *list, int *xflag,
               int *count_ew, int *list_ew, int *xflag_ew); /* nblist.c:32-307 */
int orc_nblist_cells(const orc_system *s, const double *pos, const orc_excl *ex,
                     int max_nb, int *count, int *list, int *xflag);
                                         /* nblist.c:327-557 (CLIST), rows sorted */
int orc_nl_check(const orc_system *s, const double *pos, const double *anchor,
                 int lb, int ub);                            /* nblist.c:607-647 */

/* --- cnonbond_nblist.c ----------------------------------------------------- */
typedef struct {
    double enonbond, enonbonds;               /* LJ, LJ cutoff shifts        */
    double eewald_real, eewald_reals, eewald_intra;
    double vir[6];                            /* wxx wyx wyy wzx wzy wzz     */
    double hes[3];   /* CONFIGT: sum_i d2E/dx_i2, dy_i2, dz_i2 of the pair
                      * terms (cnonbond_nblist2.c:293-302, 403-440)         */
} orc_pair_out;

void orc_pair(const orc_system *s, const double *pos, const int *count,
              const int *list, const int *xflag, int max_nb,
              double *ff, orc_pair_out *out);       /* cnonbond_nblist.c:134-523 */

/* pme_qgrid.c:13-121 + pme_calc.c:18-279.  energy[3] = {recip, self, recip-self};
 * qgrid_out (optional, 2*gx*gy*gz doubles) receives the spread charge grid. */
void orc_pme_calc(const orc_system *s, const double *pos, double e_self,
                  double *ff, double energy[3], double vir[6], double *qgrid_out);

/* Cardinal B-spline weights with first and second derivatives for the
 * fractional coordinate w (index convention of pme_spline.c:10-49);
 * d2array[k] = M''_p, from M''_p(x) = M_{p-2}(x) - 2 M_{p-2}(x-1) + M_{p-2}(x-2). */
void orc_fillspline2(double w, int order, double *array, double *darray, double *d2array);

/* Diagonal of the Hessian of the SPME reciprocal energy (the reciprocal
 * part of cnonbondnl2, whose pme_calc2 the reference calls at
 * cnonbond_nblist2.c:586 but never defines).  E = C/2 sum_{m,m'} Q(m)
 * theta(m-m') Q(m') with C = QFACTOR/epsRF; for site i and direction d
 *   d2E/dd_i2 = C q_i (K_d/L_d)^2 [ sum_m W''_i(m) (theta*Q)(m)
 *              + q_i sum_{m,m'} W'_i(m) theta(m-m') W'_i(m') ],
 * the exact second derivative of the energy PME evaluates (so the grid
 * self-interaction of site i is included, as its forces include it).
 * hes3 = sums over sites; per_site (optional, 3n) per-site values.       */
void orc_pme_hessian(const orc_system *s, const double *pos, double hes3[3], double *per_site);

/* cbond2.c:65-130: bond energy and forces as orc_cbond plus the Hessian
 * diagonal sums hes3 (2 * d2e11 etc. per bond).                          */
double orc_cbond2(const orc_system *s, const double *pos, int nb, const int *ia,
                  const int *ib, const double *req, const double *kb, double *ff, double hes3[3]);

/* cnonbond_nblist.c:51-622 (STYPE+S14+SPME): pair loop + pme_calc, energy[8]
 * slots as at cnonbond_nblist.c:613-620.                                     */
void orc_cnonbondnl(const orc_system *s, const double *pos, const int *count,
                    const int *list, const int *xflag, int max_nb, double e_self,
                    double *ff, double energy[8], double vir_nbond[6],
                    double vir_kspace[6]);

/* --- bonded, integrator, kinetic ------------------------------------------- */
double orc_cbond(const orc_system *s, const double *pos, int nb, const int *ia,
                 const int *ib, const double *req, const double *kb, double *ff,
                 double vir[6]);                             /* cbond.c:65-120 */
/* Angle-type bonded terms, forces added, virial of the term type in vir[6]:
 * cbend.c:63-170, curey.c:57-140, ctorsion.c:67-236 (6 Fourier entries per
 * torsion, leading nonzero kphi used), cimproper.c:62-225.               */
double orc_cbend(const orc_system *s, const double *pos, int nbend, const int *ia, const int *ib,
                 const int *ic, const double *angeq, const double *kb, double *ff, double vir[6]);
double orc_curey(const orc_system *s, const double *pos, int nu, const int *ia, const int *ib,
                 const double *seq, const double *ku, double *ff, double vir[6]);
double orc_ctorsion(const orc_system *s, const double *pos, int nt, const int *ta, const int *tb, const int *tc,
                    const int *td, const double *kphi, const double *nphi, const double *delphi, double *ff,
                    double vir[6]);
double orc_cimproper(const orc_system *s, const double *pos, int ni, const int *ia, const int *ib, const int *ic,
                     const int *id, const double *kimpr, const double *angeq, double *ff, double vir[6]);
/* cbend2.c, curey2.c, ctorsion2.c, cimproper2.c (CONFIGT): as above, with
 * the sums over the terms' sites of d2E/dx^2, d2E/dy^2, d2E/dz^2 in hes3. */
double orc_cbend2(const orc_system *s, const double *pos, int nbend, const int *ia, const int *ib,
                  const int *ic, const double *angeq, const double *kb, double *ff, double hes3[3]);
double orc_curey2(const orc_system *s, const double *pos, int nu, const int *ia, const int *ib,
                  const double *seq, const double *ku, double *ff, double hes3[3]);
double orc_ctorsion2(const orc_system *s, const double *pos, int nt, const int *ta, const int *tb, const int *tc,
                     const int *td, const double *kphi, const double *nphi, const double *delphi, double *ff,
                     double hes3[3]);
double orc_cimproper2(const orc_system *s, const double *pos, int ni, const int *ia, const int *ib, const int *ic,
                      const int *id, const double *kimpr, const double *angeq, double *ff, double hes3[3]);
/* drift.c:20-72: removes the net momentum from vel (n x 3) */
void orc_drift(int n, const double *mass, double *vel);
void orc_pbc_all(const orc_system *s, double *pos, int n);  /* pbc_all.c:23-88 */
double orc_kinet(int n, const double *mass, const double *vv, const double *uu,
                 double ken[6]);                             /* kinet.c:18-68 */
/* first half of integrate_vverlet.c:52-83 (kick, drift, pbc) */
void orc_vverlet_a(const orc_system *s, int n, const double *mass, double dt,
                   double vscal, double *pos, double *pos_nopbc, double *vv,
                   const double *ff);
/* second half, integrate_vverlet.c:93-102 (kick, uu = vv) */
void orc_vverlet_b(int n, const double *mass, double dt, double vscal,
                   double *vv, double *uu, const double *ff);

/* integrate_nhc.c:40-161; zeta laid out as {r,v,G} per chain element.
 * Returns the velocity scale; vv and uu are scaled in place. */
double orc_nhc(int n, double *vv, double *uu, const double *mass, double nfree,
               double T, double Q, int M, int Nc, int Nys, const double *w,
               double *zeta, double time_step);

/* --- Go model (cgolik.c / egolik_mc.c) and Wang-Landau helpers ------------- */
/* full Go energy of one chain: bonds k(r-r0)^2 + native 12-10 + non-native
 * repulsion (cgolik.c:644-948, 12-10 form cgolik.c:1202-1229).            */
double orc_go_energy(int n, const double *pos, double kbond, double r0bond,
                     int nnat, const int *na, const int *nb, const double *nr0,
                     double eps, double sig_rep, double rc);
/* energy of the pairs touching [lb,ub) (enbnd_mc.c:98-119 range filter) */
double orc_go_range_energy(int n, const double *pos, int lb, int ub,
                           double kbond, double r0bond, const unsigned char *natmask,
                           const double *r0mat, double eps, double sig_rep, double rc);

/* ran2: L'Ecuyer + Bays-Durham (random.c:37-75, Numerical Recipes 2nd ed. 7.1) */
typedef struct { long idum, idum2, iy; long iv[32]; } orc_ran2_state;
void   orc_ran2_seed(orc_ran2_state *st, long seed);
double orc_ran2(orc_ran2_state *st);

/* --- random streams of the sampler and Wang-Landau (no reference RNG) ------- */
uint64_t orc_splitmix64(uint64_t *x);
uint64_t orc_xoshiro_next(uint64_t s[4]);
double   orc_xoshiro_uniform(uint64_t s[4]);
/* dos_interpolate.c:107-124: interpolated ln g at val in bin `bin` */
double   orc_dos_interp(int nbins, const double *lng, double e_begin, double e_width, int bin, double val);
typedef struct {
    double kbond, r0bond, eps, sig_rep;       /* Go energy terms           */
    double pivot_frac, theta_max, disp_max;   /* move set                  */
    double emin, ebin;                        /* window [emin, emin+nbins*ebin) */
    int nbins;
} orc_wl_params;
/* ntrials Wang-Landau trials of one walker (pivot.c:236-253) with the
 * device's proposal streams rng[64*4]; lng/hist of nbins, stats[7]. */
int orc_wl_run(const orc_wl_params *p, int n, double *pos, double *ecur, const unsigned char *natmask,
               const double *r0mat, double *lng, long long *hist, uint64_t *rng, int ntrials, double lnf,
               long long *stats);

/* --- small MLP (no reference counterpart): FP64 forward / backward --------- */
/* layers: in -> H (tanh) -> H (tanh) -> 1 (linear); params packed
 * W1[H*in], b1[H], W2[H*H], b2[H], W3[H], b3[1].                           */
int    orc_mlp_nparams(int in, int H);
double orc_mlp_forward(int in, int H, const double *p, const double *x);
/* loss = 0.5*mean((f(x_i)-y_i)^2) over m samples; grad += dloss/dp */
double orc_mlp_loss_grad(int in, int H, const double *p, int m, const double *x,
                         const double *y, double *grad);

/* cnonbond_nblist2.c:47-622 (CONFIGT): orc_cnonbondnl plus the Hessian
 * diagonal sums of the pair terms and of the reciprocal energy.           */
void orc_cnonbondnl2(const orc_system *s, const double *pos, const int *count,
                     const int *list, const int *xflag, int max_nb, double e_self,
                     double *ff, double energy[8], double hes3[3]);

#ifdef __cplusplus
}
#endif
#endif
