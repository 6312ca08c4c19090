/*
 * dropin_main.c -- drop-in check of the reference-side binding.
 *
 * TEST INFRASTRUCTURE ONLY.  Linked (by `make -C oracle dropin`) from the
 * reference's own, unmodified force drivers -- forces.c, force_long.c,
 * force_short.c and the bonded terms cbond.c, cbend.c, curey.c, ctorsion.c,
 * cimproper.c, compiled from /root/reference -- together with
 * thecodecombo_7_amd/host/ttc_refshim.c and libttc.so.  The reference's
 * forces() therefore calls the shim's cnonbondnl(), which runs on the GPU.
 *
 * With two more arguments, MD.bin MDOUT.bin, it then runs the reference's
 * own integrators (integrate_vverlet.c, integrate_nhc.c, integrate_mts.c,
 * kinet.c, pbc_all.c, also linked unmodified) for a number of steps with
 * the step order of the reference's drivers, every force call reaching the
 * GPU through the shim:
 *   mode 0, nve_md.c:360-390: nl_check(0, n, k), nblist(k) when flagged,
 *           integrate_vverlet(k, 1.0), kinet(k)
 *   mode 1, nvt_md.c:528-531: the same list update, then integrate_nhc,
 *           integrate_vverlet(k, 1.0), integrate_nhc, kinet(k)
 *   mode 2, nve_md.c:393-420: the list update, integrate_mts(k, 1.0),
 *           kinet(k) (ninner short steps per long step)
 *   MD.bin: int32 nsteps, mode, M, Nc, Nys, ninner; float64 T, Q, nfree,
 *           dtlong, w[Nys], vel[3n]
 *   MDOUT:  float64 pos[3n] (atom, wrapped), vel[3n], rows[nsteps][2] =
 *           {en.kinet, en.potens as calcvalue.c:42-53 sums it}, the loop's
 *           wall seconds, and int32 list rebuilds
 *
 * Usage: dropin_forces IN.bin OUT.bin [MD.bin MDOUT.bin]
 *   IN  (little endian): int32 n, ntypes, nx, nbond, order, gx, gy, gz;
 *       float64 lx, ly, lz, rc, rclist, rclist_ew, kappa, eps_rf, dt;
 *       int32 type[n]; float64 q[n], mass[n], pos[3n];
 *       float64 eps[t*t], sig[t*t], eps14[t*t], sig14[t*t];
 *       int32 xa[nx], xb[nx], xf[nx], ba[nbond], bb[nbond];
 *       float64 req[nbond], kbond[nbond]
 *   OUT: float64 en[10] = {bond, nbond, nbonds, coulomb, rewald, rewalds,
 *        kewald, sewald, iewald, tewald} after forces(0), ff[3n] after
 *        forces(0), ff_long[3n] after force_long(0), ff_short[3n] after
 *        force_short(0), then int32 nl_flag after nl_check(0, n, 0).
 */
#define MAIN
#include "defines.h"

void ewald_setup(int);
void nblist(int);
void nl_check(int, int, int);
void forces(int);
void force_long(int);
void force_short(int);
void pme_clean(void);
void integrate_vverlet(int, double);
void integrate_nhc(double, int);
void integrate_mts(int, double);
void kinet(int);

static void rd(void *p, size_t sz, size_t n, FILE *f)
{
    if (n && fread(p, sz, n, f) != n) { fprintf(stderr, "dropin: short input\n"); exit(2); }
}

static void *xc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    if (!p) exit(11);
    return p;
}

static double wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* the reference's integrator loops on box k (see the header comment) */
static int run_md(const char *in, const char *outp, int n, int k)
{
    FILE *f = fopen(in, "rb");
    if (!f) { perror(in); return 2; }
    int h[6];
    double par[4];
    rd(h, sizeof(int), 6, f);
    rd(par, sizeof(double), 4, f);
    int nsteps = h[0], mode = h[1], M = h[2], Nc = h[3], Nys = h[4];
    double *w = xc(Nys, sizeof(double));
    rd(w, sizeof(double), Nys, f);
    vv = xc(1, sizeof *vv); uu = xc(1, sizeof *uu);
    vv[k] = xc(n, sizeof **vv); uu[k] = xc(n, sizeof **uu);
    double *v = xc(3 * (size_t)n, sizeof(double));
    rd(v, sizeof(double), 3 * (size_t)n, f);
    fclose(f);
    for (int i = 0; i < n; i++) {
        vv[k][i].x = v[3 * i]; vv[k][i].y = v[3 * i + 1]; vv[k][i].z = v[3 * i + 2];
        uu[k][i] = vv[k][i];
    }
    sim.T[k] = par[0];
    box[k].nfree = par[2];
    sim.dtlong = par[3];
    sim.nsteps = h[5];
    nhc[k].M = M; nhc[k].Nc = Nc; nhc[k].Nys = Nys; nhc[k].Q = par[1]; nhc[k].w = w;
    nhc[k].zeta = xc(M > 0 ? M : 1, sizeof *nhc[k].zeta);
    if (mode == 2) { force_long(k); force_short(k); }   /* main.c: both before the loop */
    else forces(k);
    double *rows = xc(2 * (size_t)(nsteps > 0 ? nsteps : 1), sizeof(double));
    int rebuilds = 0;
    double t0 = wall();
    for (int i = 1; i <= nsteps; i++) {
        nl_check(0, box[k].boxns, k);
        if (nl_flag[k] == 1) { nblist(k); rebuilds++; }
        if (mode == 1) {
            integrate_nhc(sim.dt, k);
            integrate_vverlet(k, 1.0);
            integrate_nhc(sim.dt, k);
        } else if (mode == 2) {
            integrate_mts(k, 1.0);
        } else {
            integrate_vverlet(k, 1.0);
        }
        kinet(k);
        rows[2 * (i - 1)] = en[k].kinet;
        rows[2 * (i - 1) + 1] = en[k].bond + en[k].bend + en[k].tors + en[k].nbonds + en[k].impr + en[k].urey;
    }
    double secs = wall() - t0;
    FILE *o = fopen(outp, "wb");
    if (!o) { perror(outp); return 2; }
    for (int i = 0; i < n; i++) fwrite(&atom[k][i].x, sizeof(double), 3, o);
    for (int i = 0; i < n; i++) fwrite(&vv[k][i].x, sizeof(double), 3, o);
    fwrite(rows, sizeof(double), 2 * (size_t)nsteps, o);
    fwrite(&secs, sizeof(double), 1, o);
    fwrite(&rebuilds, sizeof(int), 1, o);
    fclose(o);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 3 && argc != 5) { fprintf(stderr, "usage: dropin_forces IN.bin OUT.bin [MD.bin MDOUT.bin]\n"); return 2; }
    FILE *f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    int hdr[8];
    double par[9];
    rd(hdr, sizeof(int), 8, f);
    rd(par, sizeof(double), 9, f);
    int n = hdr[0], nt = hdr[1], nx = hdr[2], nb = hdr[3], k = 0;

    sim.NB = 1;
    sim.order = hdr[4];
    sim.grid[0] = hdr[5]; sim.grid[1] = hdr[6]; sim.grid[2] = hdr[7];
    sim.rc = par[3]; sim.rclist = par[4]; sim.rc_ew = par[3]; sim.rclist_ew = par[5];
    sim.kappa[k] = par[6]; sim.epsRF[k] = par[7]; sim.dt = par[8];
    box[k].boxns = n;
    box[k].lx = par[0]; box[k].ly = par[1]; box[k].lz = par[2];
    box[k].hx = 0.5 * par[0]; box[k].hy = 0.5 * par[1]; box[k].hz = 0.5 * par[2];
    box[k].rc2 = par[3] * par[3];
    max_num_neighbors = 1500;

    atom = xc(1, sizeof *atom); atnopbc = xc(1, sizeof *atnopbc);
    ff = xc(1, sizeof *ff); ff_long = xc(1, sizeof *ff_long); ff_short = xc(1, sizeof *ff_short);
    pott = xc(1, sizeof *pott);
    atom[k] = xc(n, sizeof **atom); atnopbc[k] = xc(n, sizeof **atnopbc);
    ff[k] = xc(n, sizeof **ff); ff_long[k] = xc(n, sizeof **ff_long); ff_short[k] = xc(n, sizeof **ff_short);
    pott[k] = xc(n, sizeof **pott);
    int *type = xc(n, sizeof(int));
    double *q = xc(n, sizeof(double)), *m = xc(n, sizeof(double)), *x = xc(3 * (size_t)n, sizeof(double));
    rd(type, sizeof(int), n, f); rd(q, sizeof(double), n, f); rd(m, sizeof(double), n, f);
    rd(x, sizeof(double), 3 * (size_t)n, f);
    for (int i = 0; i < n; i++) {
        atom[k][i].atomid = type[i]; atom[k][i].q = q[i];
        atom[k][i].x = x[3 * i]; atom[k][i].y = x[3 * i + 1]; atom[k][i].z = x[3 * i + 2];
        atnopbc[k][i] = atom[k][i];
        pott[k][i].qq = q[i]; pott[k][i].mas = m[i];
    }
    size_t t2 = (size_t)nt * nt;
    double *tab = xc(4 * t2, sizeof(double));
    rd(tab, sizeof(double), 4 * t2, f);
    ljset = xc(nt, sizeof *ljset); ljset14 = xc(nt, sizeof *ljset14);
    for (int a = 0; a < nt; a++) {
        ljset[a] = xc(nt, sizeof **ljset); ljset14[a] = xc(nt, sizeof **ljset14);
        for (int b = 0; b < nt; b++) {
            ljset[a][b].eps = tab[a * nt + b];
            ljset[a][b].sig = tab[t2 + a * nt + b];
            ljset14[a][b].eps = tab[2 * t2 + a * nt + b];
            ljset14[a][b].sig = tab[3 * t2 + a * nt + b];
        }
    }
    int *xi = xc(3 * (size_t)nx + 2 * (size_t)nb, sizeof(int));
    rd(xi, sizeof(int), 3 * (size_t)nx + 2 * (size_t)nb, f);
    double *bp = xc(2 * (size_t)nb, sizeof(double));
    rd(bp, sizeof(double), 2 * (size_t)nb, f);
    fclose(f);
    xint = xc(1, sizeof *xint); xint[k] = xc(nx + 1, sizeof **xint); xintN[k] = nx;
    for (int i = 0; i < nx; i++) { xint[k][i].a = xi[i]; xint[k][i].b = xi[nx + i]; xint[k][i].xflag = xi[2 * nx + i]; }
    bond = xc(1, sizeof *bond); bond[k] = xc(nb + 1, sizeof **bond); bondN[k] = nb;
    for (int i = 0; i < nb; i++) {
        bond[k][i].a = xi[3 * nx + i]; bond[k][i].b = xi[3 * nx + nb + i];
        bond[k][i].req = bp[i]; bond[k][i].krbond = bp[nb + i];
    }
    /* no angles, torsions, impropers, Urey-Bradley terms */
    bend = xc(1, sizeof *bend); tors = xc(1, sizeof *tors); impr = xc(1, sizeof *impr); urey = xc(1, sizeof *urey);
    rx0 = xc(1, sizeof *rx0); ry0 = xc(1, sizeof *ry0); rz0 = xc(1, sizeof *rz0);
    rx0[k] = xc(n, sizeof(double)); ry0[k] = xc(n, sizeof(double)); rz0[k] = xc(n, sizeof(double));
    nl_flag = xc(1, sizeof(int));

    ewald_setup(k);      /* main.c:443 order: ewald_setup, nblist, forces */
    nblist(k);
    forces(k);
    double en10[10] = {en[k].bond, en[k].nbond, en[k].nbonds, en[k].coulomb, en[k].rewald, en[k].rewalds,
                       en[k].kewald, en[k].sewald, en[k].iewald, en[k].tewald};
    FILE *o = fopen(argv[2], "wb");
    if (!o) { perror(argv[2]); return 2; }
    fwrite(en10, sizeof(double), 10, o);
    fwrite(&ff[k][0].x, sizeof(double), 3 * (size_t)n, o);
    force_long(k);
    fwrite(&ff_long[k][0].x, sizeof(double), 3 * (size_t)n, o);
    force_short(k);
    fwrite(&ff_short[k][0].x, sizeof(double), 3 * (size_t)n, o);
    nl_check(0, n, k);
    fwrite(&nl_flag[k], sizeof(int), 1, o);
    fclose(o);
    if (argc == 5) {
        int rc = run_md(argv[3], argv[4], n, k);
        if (rc) return rc;
    }
    pme_clean();
    return 0;
}
