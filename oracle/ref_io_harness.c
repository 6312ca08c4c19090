/*
 * ref_io_harness.c -- drives the reference's own input readers and .trr
 * writer, compiled from their sources under /root/reference into
 * oracle/_ref/libttcref_io.so, so that the native readers and writer of
 * thecodecombo_7_amd/host/ttc_io.c can be checked against them.
 *
 * TEST INFRASTRUCTURE ONLY (never shipped, never measured as the product).
 *
 * Reference flag set: NLIST DLIST STYPE S14 EWALD PRESSURE TRR (positional
 * simul.input) or the same plus SCRIPT (keyword simul.input).  Compiled
 * files: read_sim.c, read_atom.c, read_config.c, read_topology.c,
 * allocate.c, setlj.c, pbc_init.c, write_trr.c.  The readers open
 * ./INPUT/... relative to the working directory, as the reference does, so
 * ref_io_read changes into the case directory for the call.
 */
#define MAIN
#include "defines.h"

#include <unistd.h>

void read_sim(void);
void read_atom(void);
void read_config(void);
void setlj(void);
void write_trr(int k, int cycle, int flag);
extern int FLAG_x, FLAG_v, FLAG_f;

static char g_home[4096];

/* Run the reference's start-up readers on case_dir/INPUT (main.c:298-379
 * order: read_sim, read_atom, read_config, setlj). */
int ref_io_read(const char *case_dir)
{
    if (!getcwd(g_home, sizeof g_home)) return -1;
    if (chdir(case_dir)) return -2;
    read_sim();
    read_atom();
    read_config();
    setlj();
    if (chdir(g_home)) return -3;
    return 0;
}

void ref_io_counts(int *c)
{
    int k = 0;
    c[0] = box[k].boxns; c[1] = sim.NC; c[2] = bondN[k]; c[3] = bendN[k]; c[4] = ureyN[k];
    c[5] = torsN[k]; c[6] = imprN[k]; c[7] = exNBN[k]; c[8] = in14N[k]; c[9] = xintN[k];
}

/* lx ly lz rc rclist rc_ew rclist_ew kappa epsRF dt T nhcQ | order gx gy gz M respa NB ID ID2 */
void ref_io_sim(double *d, int *i)
{
    int k = 0;
    d[0] = box[k].lx; d[1] = box[k].ly; d[2] = box[k].lz; d[3] = sim.rc; d[4] = sim.rclist;
    d[5] = sim.rc_ew; d[6] = sim.rclist_ew; d[7] = sim.kappa[k]; d[8] = sim.epsRF[k]; d[9] = sim.dt;
    d[10] = sim.T[k]; d[11] = nhc[k].Q;
    i[0] = sim.order; i[1] = sim.grid[0]; i[2] = sim.grid[1]; i[3] = sim.grid[2]; i[4] = nhc[k].M;
    i[5] = sim.xRESPA[k]; i[6] = sim.NB; i[7] = sim.ID; i[8] = sim.ID2;
}

void ref_io_sites(double *xn, double *x, double *q, double *mass, int *atomid, int *res, char *names)
{
    int k = 0;
    for (int i = 0; i < box[k].boxns; i++) {
        xn[3 * i] = atnopbc[k][i].x; xn[3 * i + 1] = atnopbc[k][i].y; xn[3 * i + 2] = atnopbc[k][i].z;
        x[3 * i] = atom[k][i].x; x[3 * i + 1] = atom[k][i].y; x[3 * i + 2] = atom[k][i].z;
        q[i] = atnopbc[k][i].q;
        mass[i] = pott[k][i].mas;
        atomid[i] = atnopbc[k][i].atomid;
        res[i] = atnopbc[k][i].n;
        memcpy(names + 5 * i, atnopbc[k][i].name, 5);
    }
}

/* ljset / ljset14 entries for the given atomid pairs */
void ref_io_lj(int n, const int *ia, const int *ib, double *eps, double *sig, double *eps14, double *sig14)
{
    for (int i = 0; i < n; i++) {
        eps[i] = ljset[ia[i]][ib[i]].eps;
        sig[i] = ljset[ia[i]][ib[i]].sig;
        eps14[i] = ljset14[ia[i]][ib[i]].eps;
        sig14[i] = ljset14[ia[i]][ib[i]].sig;
    }
}

void ref_io_bonds(int *ab, double *req, double *kk)
{
    for (int i = 0; i < bondN[0]; i++) {
        ab[2 * i] = bond[0][i].a; ab[2 * i + 1] = bond[0][i].b;
        req[i] = bond[0][i].req; kk[i] = bond[0][i].krbond;
    }
}

void ref_io_bends(int *abc, double *eq, double *kk)
{
    for (int i = 0; i < bendN[0]; i++) {
        abc[3 * i] = bend[0][i].a; abc[3 * i + 1] = bend[0][i].b; abc[3 * i + 2] = bend[0][i].c;
        eq[i] = bend[0][i].angeq; kk[i] = bend[0][i].krbend;
    }
}

void ref_io_ureys(int *ab, double *s0, double *kk)
{
    for (int i = 0; i < ureyN[0]; i++) {
        ab[2 * i] = urey[0][i].a; ab[2 * i + 1] = urey[0][i].b;
        s0[i] = urey[0][i].Seq; kk[i] = urey[0][i].k;
    }
}

void ref_io_torsions(int *abcd, double *kphi, double *nphi, double *delphi)
{
    for (int i = 0; i < torsN[0]; i++) {
        abcd[4 * i] = tors[0][i].a; abcd[4 * i + 1] = tors[0][i].b;
        abcd[4 * i + 2] = tors[0][i].c; abcd[4 * i + 3] = tors[0][i].d;
        for (int t = 0; t < 6; t++) {
            kphi[6 * i + t] = tors[0][i].kphi[t];
            nphi[6 * i + t] = tors[0][i].nphi[t];
            delphi[6 * i + t] = tors[0][i].delphi[t];
        }
    }
}

void ref_io_impropers(int *abcd, double *kk, double *eq)
{
    for (int i = 0; i < imprN[0]; i++) {
        abcd[4 * i] = impr[0][i].a; abcd[4 * i + 1] = impr[0][i].b;
        abcd[4 * i + 2] = impr[0][i].c; abcd[4 * i + 3] = impr[0][i].d;
        kk[i] = impr[0][i].kimpr; eq[i] = impr[0][i].angeq;
    }
}

void ref_io_exclusions(int *ab)
{
    for (int i = 0; i < exNBN[0]; i++) { ab[2 * i] = exNB[0][i].a; ab[2 * i + 1] = exNB[0][i].b; }
}

void ref_io_14(int *abcd)
{
    for (int i = 0; i < in14N[0]; i++) {
        abcd[4 * i] = in14[0][i].a; abcd[4 * i + 1] = in14[0][i].b;
        abcd[4 * i + 2] = in14[0][i].c; abcd[4 * i + 3] = in14[0][i].d;
    }
}

void ref_io_xint(int *a, int *b, int *xf)
{
    for (int i = 0; i < xintN[0]; i++) { a[i] = xint[0][i].a; b[i] = xint[0][i].b; xf[i] = xint[0][i].xflag; }
}

/* One write_trr frame of box 0 from the given arrays into
 * out_dir/OUTPUT/GROMACS/gromacs0_<fileindex>.trr. */
int ref_io_write_trr(const char *out_dir, int n, const double *x, const double *v, const double *f, double lx,
                     double ly, double lz, double dt, int cycle, int flag, int cyc_eq, int fx, int fv, int ff_,
                     int findex)
{
    if (!getcwd(g_home, sizeof g_home)) return -1;
    if (chdir(out_dir)) return -2;
    int k = 0;
    sim.NB = 1;
    sim.dt = dt;
    sim.cyc_eq = cyc_eq;
    box[k].boxns = n;
    box[k].lx = lx; box[k].ly = ly; box[k].lz = lz;
    struct atoms *a = calloc(n ? n : 1, sizeof *a);
    struct veloc *vv_ = calloc(n ? n : 1, sizeof *vv_), *f_ = calloc(n ? n : 1, sizeof *f_);
    struct atoms **sa = atom;
    struct veloc **sv = vv, **sf = ff;
    struct atoms *pa[1] = {a};
    struct veloc *pv[1] = {vv_}, *pf[1] = {f_};
    for (int i = 0; i < n; i++) {
        a[i].x = x[3 * i]; a[i].y = x[3 * i + 1]; a[i].z = x[3 * i + 2];
        vv_[i].x = v[3 * i]; vv_[i].y = v[3 * i + 1]; vv_[i].z = v[3 * i + 2];
        f_[i].x = f[3 * i]; f_[i].y = f[3 * i + 1]; f_[i].z = f[3 * i + 2];
    }
    atom = pa; vv = pv; ff = pf;
    FLAG_x = fx; FLAG_v = fv; FLAG_f = ff_;
    fileindex = findex;
    write_trr(k, cycle, flag);
    atom = sa; vv = sv; ff = sf;
    free(a); free(vv_); free(f_);
    if (chdir(g_home)) return -3;
    return 0;
}
