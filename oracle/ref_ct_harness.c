/*
 * ref_ct_harness.c -- drives the reference's own CONFIGT bonded terms
 * (cbend2.c, curey2.c, ctorsion2.c, cimproper2.c: energies, forces and the
 * Hessian diagonal sums config[k].hesx/y/z), compiled from their sources
 * under /root/reference into oracle/_ref/libttcref_ct.so, so that the
 * oracle's Hessian restatements can be checked against them.
 *
 * TEST INFRASTRUCTURE ONLY (never shipped, never measured as the product).
 */
#define MAIN
#include "defines.h"

double cbend2(int);
double curey2(int);
double ctorsion2(int);
double cimproper2(int);

static int g_n = 0;

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "ref_ct_harness: out of memory\n"); exit(11); }
    return p;
}

void ref_ct_box(int n, double lx, double ly, double lz)
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

static void in_pos(const double *pos)
{
    for (int i = 0; i < g_n; i++) {
        atom[0][i].x = pos[3 * i]; atom[0][i].y = pos[3 * i + 1]; atom[0][i].z = pos[3 * i + 2];
        ff[0][i].x = ff[0][i].y = ff[0][i].z = 0.0;
    }
    config[0].hesx = config[0].hesy = config[0].hesz = 0.0;
}

static void out(double *ff_out, double *hes3)
{
    for (int i = 0; i < g_n; i++) {
        ff_out[3 * i] = ff[0][i].x; ff_out[3 * i + 1] = ff[0][i].y; ff_out[3 * i + 2] = ff[0][i].z;
    }
    hes3[0] = config[0].hesx; hes3[1] = config[0].hesy; hes3[2] = config[0].hesz;
}

double ref_cbend2(const double *pos, int nb, const int *ia, const int *ib, const int *ic, const double *angeq,
                  const double *kb, double *ff_out, double *hes3)
{
    free(bend[0]);
    bend[0] = (struct bends *)xcalloc(nb, sizeof(struct bends));
    for (int i = 0; i < nb; i++) {
        bend[0][i].a = ia[i]; bend[0][i].b = ib[i]; bend[0][i].c = ic[i];
        bend[0][i].angeq = angeq[i]; bend[0][i].krbend = kb[i];
    }
    bendN[0] = nb;
    in_pos(pos);
    double e = cbend2(0);
    out(ff_out, hes3);
    return e;
}

double ref_curey2(const double *pos, int nu, const int *ia, const int *ib, const double *seq, const double *ku,
                  double *ff_out, double *hes3)
{
    free(urey[0]);
    urey[0] = (struct ureys *)xcalloc(nu, sizeof(struct ureys));
    for (int i = 0; i < nu; i++) {
        urey[0][i].a = ia[i]; urey[0][i].b = ib[i]; urey[0][i].Seq = seq[i]; urey[0][i].k = ku[i];
    }
    ureyN[0] = nu;
    in_pos(pos);
    double e = curey2(0);
    out(ff_out, hes3);
    return e;
}

double ref_ctorsion2(const double *pos, int nt, const int *ta, const int *tb, const int *tc, const int *td,
                     const double *kphi, const double *nphi, const double *delphi, double *ff_out, double *hes3)
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
    in_pos(pos);
    double e = ctorsion2(0);
    out(ff_out, hes3);
    return e;
}

double ref_cimproper2(const double *pos, int ni, const int *ia, const int *ib, const int *ic, const int *id,
                      const double *kimpr, const double *angeq, double *ff_out, double *hes3)
{
    free(impr[0]);
    impr[0] = (struct impropers *)xcalloc(ni, sizeof(struct impropers));
    for (int i = 0; i < ni; i++) {
        impr[0][i].a = ia[i]; impr[0][i].b = ib[i]; impr[0][i].c = ic[i]; impr[0][i].d = id[i];
        impr[0][i].kimpr = kimpr[i]; impr[0][i].angeq = angeq[i];
    }
    imprN[0] = ni;
    in_pos(pos);
    double e = cimproper2(0);
    out(ff_out, hes3);
    return e;
}
