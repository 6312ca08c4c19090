/*
 * ref_harness.c -- drives routines of the reference itself, compiled from
 * their own sources under /root/reference into oracle/_ref/libttcref.so, so
 * that the oracle restatement can be checked against them.
 *
 * TEST INFRASTRUCTURE ONLY (never shipped, never measured as the product).
 *
 * Reference flag set used: NLIST DLIST STYPE S14 EWALD PRESSURE (no SPME:
 * the SPME build needs fftw3.h, which this image does not have, so the PME
 * routines are unbuildable here).  With that flag set the reference compiles
 *   nblist.c            -> nblist(), nl_check()            (bit-exact lists)
 *   cnonbond_nblist.c   -> cnonbondnl() LJ part            (charges zero)
 *   ewald.c, cplx.c     -> erfc(), ewald_setup(), ewald_kspace() (classic
 *                          Ewald reciprocal sum, the limit PME converges to)
 * This file fills the reference's global structs (defines.h, instantiated
 * here through its MAIN convention) the way allocate.c / read_topology.c /
 * setlj.c would, calls the routines, and copies results out.
 */
#define MAIN
#include "defines.h"

static int g_n = 0;

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "ref_harness: out of memory\n"); exit(11); }
    return p;
}

/* Fill box 0.  eps/sig tables are ntypes x ntypes row-major (ljset[ta][tb]). */
int ref_setup(int n, double lx, double ly, double lz, double rc, double rclist,
              double rclist_ew, double kappa, double eps_rf, int ntypes,
              const int *type, const double *q, const double *eps, const double *sig,
              const double *eps14, const double *sig14, int nx, const int *xa,
              const int *xb, const int *xf, int max_nb)
{
    int k = 0;
    g_n = n;
    sim.NB = 1;
    sim.rc = rc;
    sim.rclist = rclist;
    sim.rc_ew = rc;
    sim.rclist_ew = rclist_ew;
    sim.kappa[k] = kappa;
    sim.epsRF[k] = eps_rf;
    box[k].boxns = n;
    box[k].lx = lx; box[k].ly = ly; box[k].lz = lz;
    box[k].hx = 0.5 * lx; box[k].hy = 0.5 * ly; box[k].hz = 0.5 * lz;
    box[k].least_boxh = 0.5 * (lx < ly ? (lx < lz ? lx : lz) : (ly < lz ? ly : lz));
    box[k].rc2 = rc * rc;
    box[k].rc4 = rc * rc * rc * rc;
    max_num_neighbors = max_nb;

    atom = (struct atoms **)xcalloc(1, sizeof(struct atoms *));
    atnopbc = (struct atoms **)xcalloc(1, sizeof(struct atoms *));
    ff = (struct veloc **)xcalloc(1, sizeof(struct veloc *));
    atom[k] = (struct atoms *)xcalloc(n, sizeof(struct atoms));
    atnopbc[k] = (struct atoms *)xcalloc(n, sizeof(struct atoms));
    ff[k] = (struct veloc *)xcalloc(n, sizeof(struct veloc));
    pott = (struct inter **)xcalloc(1, sizeof(struct inter *));
    pott[k] = (struct inter *)xcalloc(n, sizeof(struct inter));
    for (int i = 0; i < n; i++) {
        atom[k][i].atomid = type[i];
        atnopbc[k][i].atomid = type[i];
        atom[k][i].q = q[i];
        atnopbc[k][i].q = q[i];
        pott[k][i].qq = q[i];
        pott[k][i].mas = 1.0;
    }
    ljset = (struct inter **)xcalloc(ntypes, sizeof(struct inter *));
    ljset14 = (struct inter **)xcalloc(ntypes, sizeof(struct inter *));
    for (int a = 0; a < ntypes; a++) {
        ljset[a] = (struct inter *)xcalloc(ntypes, sizeof(struct inter));
        ljset14[a] = (struct inter *)xcalloc(ntypes, sizeof(struct inter));
        for (int b = 0; b < ntypes; b++) {
            ljset[a][b].eps = eps[a * ntypes + b];
            ljset[a][b].sig = sig[a * ntypes + b];
            ljset14[a][b].eps = eps14[a * ntypes + b];
            ljset14[a][b].sig = sig14[a * ntypes + b];
        }
    }
    /* neighbour lists, allocate.c:318-373 */
    nlist = (struct nlist_def *)xcalloc(1, sizeof(struct nlist_def));
    nlist_ew = (struct nlist_def *)xcalloc(1, sizeof(struct nlist_def));
    nlist[k].count = (int *)xcalloc(n, sizeof(int));
    nlist[k].list = (int **)xcalloc(n, sizeof(int *));
    nlist[k].xflag = (int **)xcalloc(n, sizeof(int *));
    nlist_ew[k].count = (int *)xcalloc(n, sizeof(int));
    nlist_ew[k].list = (int **)xcalloc(n, sizeof(int *));
    nlist_ew[k].xflag = (int **)xcalloc(n, sizeof(int *));
    for (int i = 0; i < n; i++) {
        /* one spare slot: the reference writes entry count before testing
         * count > max_num_neighbors */
        nlist[k].list[i] = (int *)xcalloc(max_nb + 1, sizeof(int));
        nlist[k].xflag[i] = (int *)xcalloc(max_nb + 1, sizeof(int));
        nlist_ew[k].list[i] = (int *)xcalloc(max_nb + 1, sizeof(int));
        nlist_ew[k].xflag[i] = (int *)xcalloc(max_nb + 1, sizeof(int));
    }
    rx0 = (double **)xcalloc(1, sizeof(double *));
    ry0 = (double **)xcalloc(1, sizeof(double *));
    rz0 = (double **)xcalloc(1, sizeof(double *));
    rx0[k] = (double *)xcalloc(n, sizeof(double));
    ry0[k] = (double *)xcalloc(n, sizeof(double));
    rz0[k] = (double *)xcalloc(n, sizeof(double));
    nl_flag = (int *)xcalloc(1, sizeof(int));

    /* exclusion list + hash table, read_topology.c:847-891, setlj.c:211-248 */
    xint = (struct excinter **)xcalloc(1, sizeof(struct excinter *));
    xint[k] = (struct excinter *)xcalloc(nx + 1, sizeof(struct excinter));
    xintN[k] = nx;
    for (int i = 0; i < nx; i++) { xint[k][i].a = xa[i]; xint[k][i].b = xb[i]; xint[k][i].xflag = xf[i]; }
    table = (struct hashlist **)xcalloc(1, sizeof(struct hashlist *));
    table[k] = (struct hashlist *)xcalloc(NPRIME + 1, sizeof(struct hashlist));
    int hs = 1000;
    for (int i = 0; i < NPRIME + 1; i++) {
        table[k][i].a = (int *)xcalloc(hs, sizeof(int));
        table[k][i].b = (int *)xcalloc(hs, sizeof(int));
        table[k][i].p = (int *)xcalloc(hs, sizeof(int));
        table[k][i].n = 0;
    }
    for (int i = 0; i < nx; i++) {
        int a = xa[i], b = xb[i];
        unsigned long index = (unsigned long)((a * b) % NPRIME);
        int qa = table[k][index].n;
        if (qa >= hs) return -1;
        table[k][index].a[qa] = a;
        table[k][index].b[qa] = b;
        table[k][index].p[qa] = i;
        table[k][index].n++;
    }
    return 0;
}

static void set_pos(const double *pos)
{
    for (int i = 0; i < g_n; i++) {
        atom[0][i].x = pos[3 * i]; atom[0][i].y = pos[3 * i + 1]; atom[0][i].z = pos[3 * i + 2];
        atnopbc[0][i].x = pos[3 * i]; atnopbc[0][i].y = pos[3 * i + 1]; atnopbc[0][i].z = pos[3 * i + 2];
    }
}

void nblist(int);
void cnonbondnl(int, double *);
void ewald_setup(int);
void ewald_kspace(int, double *);
double erfc(double);
void nl_check(int, int, int);

/* nblist.c:32 -> lists copied into [n][max_nb] arrays */
void ref_nblist(const double *pos, int max_nb, int *count, int *list, int *xflag,
                int *count_ew, int *list_ew, int *xflag_ew)
{
    set_pos(pos);
    nblist(0);
    for (int a = 0; a < g_n; a++) {
        count[a] = nlist[0].count[a];
        count_ew[a] = nlist_ew[0].count[a];
        for (int j = 0; j < count[a]; j++) {
            list[(size_t)a * max_nb + j] = nlist[0].list[a][j];
            xflag[(size_t)a * max_nb + j] = nlist[0].xflag[a][j];
        }
        for (int j = 0; j < count_ew[a]; j++) {
            list_ew[(size_t)a * max_nb + j] = nlist_ew[0].list[a][j];
            xflag_ew[(size_t)a * max_nb + j] = nlist_ew[0].xflag[a][j];
        }
    }
}

/* nblist.c:607: nl_flag after checking [lb,ub) against the anchors of the
 * last ref_nblist call. */
int ref_nl_check(const double *pos, int lb, int ub)
{
    set_pos(pos);
    nl_flag[0] = 0;
    nl_check(lb, ub, 0);
    return nl_flag[0];
}

/* Classic-Ewald set-up (ewald.c:46-196, non-SPME): k vectors with
 * |k|^2 <= maxk2 (integer units), self energy, neutrality check. */
void ref_ewald_setup(int kmax_x, int kmax_y, int kmax_z, int maxk2)
{
    int k = 0;
    sim.kmax[k][0] = kmax_x; sim.kmax[k][1] = kmax_y; sim.kmax[k][2] = kmax_z;
    sim.max_k_square[k] = maxk2;
    kvec = (double **)xcalloc(1, sizeof(double *));
    kvec[k] = (double *)xcalloc((size_t)(kmax_x + 1) * (2 * kmax_y + 1) * (2 * kmax_z + 1), sizeof(double));
    total_k_vectors = (int *)xcalloc(1, sizeof(int));
    ewald_setup(k);
}

/* cnonbond_nblist.c:51 with the classic-Ewald flag set.  Uses the lists of
 * the last ref_nblist call.  ff_out: n*3 forces (zeroed first). */
void ref_cnonbondnl(const double *pos, double *ff_out, double *energy8, double *vir6)
{
    set_pos(pos);
    for (int i = 0; i < g_n; i++) { ff[0][i].x = 0.0; ff[0][i].y = 0.0; ff[0][i].z = 0.0; }
    cnonbondnl(0, energy8);
    for (int i = 0; i < g_n; i++) {
        ff_out[3 * i] = ff[0][i].x; ff_out[3 * i + 1] = ff[0][i].y; ff_out[3 * i + 2] = ff[0][i].z;
    }
    for (int i = 0; i < 6; i++) vir6[i] = pvir[0].nbond[i];
}

/* ewald.c:207-494: energy3 = {recip, self, recip-self}; forces into ff_out */
void ref_ewald_kspace(const double *pos, double *ff_out, double *energy3)
{
    set_pos(pos);
    for (int i = 0; i < g_n; i++) { ff[0][i].x = 0.0; ff[0][i].y = 0.0; ff[0][i].z = 0.0; }
    ewald_kspace(0, energy3);
    for (int i = 0; i < g_n; i++) {
        ff_out[3 * i] = ff[0][i].x; ff_out[3 * i + 1] = ff[0][i].y; ff_out[3 * i + 2] = ff[0][i].z;
    }
}

double ref_erfc(double x) { return erfc(x); }

double ref_self_energy(void) { return en[0].sewald; }

/* random.c:37-75: n successive ran2() values after setting idum = seed
 * (negative seeds initialise, as main.c:278-287 does) */
double ran2(void);
void ref_ran2_seq(long seed, int n, double *out)
{
    idum = seed;
    for (int i = 0; i < n; i++) out[i] = ran2();
}

/* ---- integrate_vverlet.c, pbc_all.c, kinet.c --------------------------
 * integrate_vverlet() calls forces(k); the harness defines it as a fixed
 * linear force field of the wrapped positions that couples the three axes,
 *   F_x = -fc x + fd y,  F_y = -fc y + fd z,  F_z = -fc z + fd x,
 * so a test can evaluate the same forces in the oracle's loop bit for bit. */
static double g_fc = 0.0, g_fd = 0.0;

void forces(int k)
{
    for (int i = 0; i < box[k].boxns; i++) {
        double x = atom[k][i].x, y = atom[k][i].y, z = atom[k][i].z;
        ff[k][i].x = -g_fc * x + g_fd * y;
        ff[k][i].y = -g_fc * y + g_fd * z;
        ff[k][i].z = -g_fc * z + g_fd * x;
    }
}

void integrate_vverlet(int, double);
void pbc_all(int);
void kinet(int);

static int g_md_n = 0;   /* sites the velocity arrays hold */
static void md_alloc(void)
{
    if (vv && g_md_n == g_n) return;
    if (!vv) {
        vv = (struct veloc **)xcalloc(1, sizeof(struct veloc *));
        uu = (struct veloc **)xcalloc(1, sizeof(struct veloc *));
    }
    free(vv[0]); free(uu[0]);
    vv[0] = (struct veloc *)xcalloc(g_n, sizeof(struct veloc));
    uu[0] = (struct veloc *)xcalloc(g_n, sizeof(struct veloc));
    g_md_n = g_n;
}

/* nsteps calls of integrate_vverlet(0, vscal) after one forces(0) (as
 * main.c does before the first step), on positions pos (wrapped), nopbc
 * (unwrapped) and velocities vel; all three are updated in place and uu_out
 * receives uu.  Needs ref_setup first (box, n). */
void ref_vverlet(int nsteps, double vscal, double dt, double fc, double fd, const double *mass,
                 double *pos, double *nopbc, double *vel, double *uu_out)
{
    md_alloc();
    sim.dt = dt;
    g_fc = fc;
    g_fd = fd;
    for (int i = 0; i < g_n; i++) {
        pott[0][i].mas = mass[i];
        atom[0][i].x = pos[3 * i]; atom[0][i].y = pos[3 * i + 1]; atom[0][i].z = pos[3 * i + 2];
        atnopbc[0][i].x = nopbc[3 * i]; atnopbc[0][i].y = nopbc[3 * i + 1]; atnopbc[0][i].z = nopbc[3 * i + 2];
        vv[0][i].x = vel[3 * i]; vv[0][i].y = vel[3 * i + 1]; vv[0][i].z = vel[3 * i + 2];
    }
    forces(0);
    for (int s = 0; s < nsteps; s++) integrate_vverlet(0, vscal);
    for (int i = 0; i < g_n; i++) {
        pos[3 * i] = atom[0][i].x; pos[3 * i + 1] = atom[0][i].y; pos[3 * i + 2] = atom[0][i].z;
        nopbc[3 * i] = atnopbc[0][i].x; nopbc[3 * i + 1] = atnopbc[0][i].y; nopbc[3 * i + 2] = atnopbc[0][i].z;
        vel[3 * i] = vv[0][i].x; vel[3 * i + 1] = vv[0][i].y; vel[3 * i + 2] = vv[0][i].z;
        uu_out[3 * i] = uu[0][i].x; uu_out[3 * i + 1] = uu[0][i].y; uu_out[3 * i + 2] = uu[0][i].z;
    }
}

/* pbc_all.c:23-88 on pos (in place) */
void ref_pbc_all(double *pos)
{
    for (int i = 0; i < g_n; i++) {
        atom[0][i].x = pos[3 * i]; atom[0][i].y = pos[3 * i + 1]; atom[0][i].z = pos[3 * i + 2];
    }
    pbc_all(0);
    for (int i = 0; i < g_n; i++) {
        pos[3 * i] = atom[0][i].x; pos[3 * i + 1] = atom[0][i].y; pos[3 * i + 2] = atom[0][i].z;
    }
}

/* kinet.c:18-68: returns en.kinet, ken6 = en.ken */
double ref_kinet(const double *mass, const double *vel, const double *uuv, double *ken6)
{
    md_alloc();
    for (int i = 0; i < g_n; i++) {
        pott[0][i].mas = mass[i];
        vv[0][i].x = vel[3 * i]; vv[0][i].y = vel[3 * i + 1]; vv[0][i].z = vel[3 * i + 2];
        uu[0][i].x = uuv[3 * i]; uu[0][i].y = uuv[3 * i + 1]; uu[0][i].z = uuv[3 * i + 2];
    }
    kinet(0);
    for (int i = 0; i < 6; i++) ken6[i] = en[0].ken[i];
    return en[0].kinet;
}

/* ---- cbend.c, curey.c, ctorsion.c, cimproper.c ------------------------- */
double cbend(int);
double curey(int);
double ctorsion(int);
double cimproper(int);

/* box 0 of n sites for the bonded-term calls below (allocate.c's atom / ff
 * arrays, read_topology.c's term arrays filled by each call) */
void ref_bonded_box(int n, double lx, double ly, double lz)
{
    g_n = n;
    sim.NB = 1;
    box[0].boxns = n;
    box[0].lx = lx; box[0].ly = ly; box[0].lz = lz;
    box[0].hx = 0.5 * lx; box[0].hy = 0.5 * ly; box[0].hz = 0.5 * lz;
    if (!atom) atom = (struct atoms **)xcalloc(1, sizeof(struct atoms *));
    if (!ff) ff = (struct veloc **)xcalloc(1, sizeof(struct veloc *));
    if (!bend) bend = (struct bends **)xcalloc(1, sizeof(struct bends *));
    if (!urey) urey = (struct ureys **)xcalloc(1, sizeof(struct ureys *));
    if (!tors) tors = (struct torsions **)xcalloc(1, sizeof(struct torsions *));
    if (!impr) impr = (struct impropers **)xcalloc(1, sizeof(struct impropers *));
    free(atom[0]); free(ff[0]);
    atom[0] = (struct atoms *)xcalloc(n, sizeof(struct atoms));
    ff[0] = (struct veloc *)xcalloc(n, sizeof(struct veloc));
}

static void bonded_in(const double *pos)
{
    for (int i = 0; i < g_n; i++) {
        atom[0][i].x = pos[3 * i]; atom[0][i].y = pos[3 * i + 1]; atom[0][i].z = pos[3 * i + 2];
        ff[0][i].x = ff[0][i].y = ff[0][i].z = 0.0;
    }
}

static void bonded_out(double *ff_out, const double *w, double *vir6)
{
    for (int i = 0; i < g_n; i++) {
        ff_out[3 * i] = ff[0][i].x; ff_out[3 * i + 1] = ff[0][i].y; ff_out[3 * i + 2] = ff[0][i].z;
    }
    for (int q = 0; q < 6; q++) vir6[q] = w[q];
}

double ref_cbend(const double *pos, int nb, const int *ia, const int *ib, const int *ic, const double *angeq,
                 const double *kb, double *ff_out, double *vir6)
{
    free(bend[0]);
    bend[0] = (struct bends *)xcalloc(nb, sizeof(struct bends));
    for (int i = 0; i < nb; i++) {
        bend[0][i].a = ia[i]; bend[0][i].b = ib[i]; bend[0][i].c = ic[i];
        bend[0][i].angeq = angeq[i]; bend[0][i].krbend = kb[i];
    }
    bendN[0] = nb;
    bonded_in(pos);
    double e = cbend(0);
    bonded_out(ff_out, pvir[0].bend, vir6);
    return e;
}

double ref_curey(const double *pos, int nu, const int *ia, const int *ib, const double *seq, const double *ku,
                 double *ff_out, double *vir6)
{
    free(urey[0]);
    urey[0] = (struct ureys *)xcalloc(nu, sizeof(struct ureys));
    for (int i = 0; i < nu; i++) {
        urey[0][i].a = ia[i]; urey[0][i].b = ib[i]; urey[0][i].Seq = seq[i]; urey[0][i].k = ku[i];
    }
    ureyN[0] = nu;
    bonded_in(pos);
    double e = curey(0);
    bonded_out(ff_out, pvir[0].urey, vir6);
    return e;
}

double ref_ctorsion(const double *pos, int nt, const int *ta, const int *tb, const int *tc, const int *td,
                    const double *kphi, const double *nphi, const double *delphi, double *ff_out, double *vir6)
{
    free(tors[0]);
    tors[0] = (struct torsions *)xcalloc(nt, sizeof(struct torsions));
    for (int i = 0; i < nt; i++) {
        tors[0][i].a = ta[i]; tors[0][i].b = tb[i]; tors[0][i].c = tc[i]; tors[0][i].d = td[i];
        for (int m = 0; m < 6; m++) {
            tors[0][i].kphi[m] = kphi[6 * i + m];
            tors[0][i].nphi[m] = nphi[6 * i + m];
            tors[0][i].delphi[m] = delphi[6 * i + m];
        }
    }
    torsN[0] = nt;
    bonded_in(pos);
    double e = ctorsion(0);
    bonded_out(ff_out, pvir[0].tors, vir6);
    return e;
}

double ref_cimproper(const double *pos, int ni, const int *ia, const int *ib, const int *ic, const int *id,
                     const double *kimpr, const double *angeq, double *ff_out, double *vir6)
{
    free(impr[0]);
    impr[0] = (struct impropers *)xcalloc(ni, sizeof(struct impropers));
    for (int i = 0; i < ni; i++) {
        impr[0][i].a = ia[i]; impr[0][i].b = ib[i]; impr[0][i].c = ic[i]; impr[0][i].d = id[i];
        impr[0][i].kimpr = kimpr[i]; impr[0][i].angeq = angeq[i];
    }
    imprN[0] = ni;
    bonded_in(pos);
    double e = cimproper(0);
    bonded_out(ff_out, pvir[0].impr, vir6);
    return e;
}

/* drift.c:20-72 on box 0 (g_n sites): velocities in, corrected velocities out */
void drift(int);
void ref_drift(const double *mass, double *vel)
{
    md_alloc();
    for (int i = 0; i < g_n; i++) {
        pott[0][i].mas = mass[i];
        vv[0][i].x = vel[3 * i]; vv[0][i].y = vel[3 * i + 1]; vv[0][i].z = vel[3 * i + 2];
    }
    drift(0);
    for (int i = 0; i < g_n; i++) {
        vel[3 * i] = vv[0][i].x; vel[3 * i + 1] = vv[0][i].y; vel[3 * i + 2] = vv[0][i].z;
    }
}
