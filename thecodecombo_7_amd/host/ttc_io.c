/*
 * ttc_io.c -- readers of the reference's input files and its .trr writer
 * (include/ttc_io.h), host C inside libttc.so.
 *
 * The readers keep the reference's token rules: the same scanf conversions
 * (so a 4-character name field, a count followed by its "!NBOND" tag, and
 * the per-line record counts of the psf behave alike) and the same
 * parameter-matching rules, including their order dependence (first match
 * for bonds, bends and Urey-Bradley pairs; every exact torsion match adds a
 * Fourier term; the last exact improper match wins; wildcard entries (type
 * 100) only while nothing matched).  Their data structures are their own:
 * one flat struct per box, 1-4 pairs found through a per-site index of the
 * bends and a hash set of excluded pairs instead of the reference's pairwise
 * scans, so reading 10^6 sites takes seconds.
 */
#include "ttc_io.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define IO_PI 3.14159265358979323846   /* defines.h PI */
#define IO_KCAL 4.184
#define IO_MAXID 2000                   /* epstmp/sigtmp rows (defines.h:1173-1188) */
#define IO_WILD 100                     /* wildcard atom type in torsion/improper rows */
#define IO_NFOURIER 6                   /* struct torsions kphi[6] */

static char g_err[512];

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *ttc_io_last_error(void) { return g_err; }

typedef struct { int t[4]; double p[3]; } prop_t;   /* a parameter row: types, values */

struct ttc_io {
    ttc_io_sim sim;
    char param_name[64];
    /* mol.input */
    int nc, nmolall, nsolute;
    int *nbox, *msite, *mres;
    /* parameter file */
    int natype;
    int *at_id;
    char (*at_name)[5];
    double *at_mass;
    int nbondp, nbendp, ntorsp, nimprp, nureyp, nnbfix;
    prop_t *bondp, *bendp, *torsp, *imprp, *ureyp, *nbfix;
    double eps0[IO_MAXID], sig0[IO_MAXID], eps1[IO_MAXID], sig1[IO_MAXID];
    /* the box */
    int n;
    double *xn, *x, *q, *mass;
    int *atomid, *res, *type;
    char (*name)[5];
    int ntype, *type_id;
    double *lj_eps, *lj_sig, *lj14_eps, *lj14_sig;
    int nbond, nbend, nurey, ntors, nimpr, nexcl, n14, nxint;
    int *bond, *bend, *urey, *tors, *impr, *excl, *in14, *xint;
    double *bond_req, *bond_k, *bend_eq, *bend_k, *urey_s0, *urey_k;
    double *kphi, *nphi, *delphi, *impr_k, *impr_eq;
};

static void *zalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

void ttc_io_free(ttc_io *s)
{
    if (!s) return;
    void *ptrs[] = {s->nbox, s->msite, s->mres, s->at_id, s->at_name, s->at_mass, s->bondp, s->bendp, s->torsp,
                    s->imprp, s->ureyp, s->nbfix, s->xn, s->x, s->q, s->mass, s->atomid, s->res, s->type,
                    s->name, s->type_id, s->lj_eps, s->lj_sig, s->lj14_eps, s->lj14_sig, s->bond, s->bend,
                    s->urey, s->tors, s->impr, s->excl, s->in14, s->xint, s->bond_req, s->bond_k, s->bend_eq,
                    s->bend_k, s->urey_s0, s->urey_k, s->kphi, s->nphi, s->delphi, s->impr_k, s->impr_eq};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
    free(s);
}

static FILE *open_in(const char *dir, const char *file, char *path, size_t cap)
{
    snprintf(path, cap, "%s/%s", dir, file);
    return fopen(path, "r");
}

static void skip_line(FILE *f)
{
    char tt[160];
    if (!fgets(tt, sizeof tt, f)) tt[0] = 0;
}

/* ------------------------------------------------------------------------ */
/* simul.input: keyword form (read_sim.c:36-440, SCRIPT) or the positional   */
/* form (read_sim.c:471-560)                                                 */
/* ------------------------------------------------------------------------ */

/* One logical line as the keyword reader sees it: leading white space and
 * everything from the first '#' removed; lines of two characters or fewer
 * are noise. */
static int keyword_line(char *buff)
{
    int c = 0;
    while (isspace((unsigned char)buff[c])) c++;
    memmove(buff, buff + c, strlen(buff + c) + 1);
    char *pound = strchr(buff, '#');
    if (pound) *pound = 0;
    return strlen(buff) > 2;
}

static int is_keyword_file(FILE *f)
{
    char buff[256];
    int kw = 0;
    while (fgets(buff, sizeof buff, f)) {
        if (!keyword_line(buff)) continue;
        char name[64] = "";
        sscanf(buff, "%63s", name);
        for (char *p = name; *p; p++) *p = (char)tolower((unsigned char)*p);
        if (!strncmp(name, "sim_", 4) || !strncmp(name, "box_", 4)) kw = 1;
        break;
    }
    rewind(f);
    return kw;
}

static int read_sim_keywords(FILE *f, int ibox, ttc_io *s)
{
    ttc_io_sim *m = &s->sim;
    char buff[256];
    double eps_rf = 0.0;
    int have_eps = 0;
    m->nboxes = 0;
    m->dt = -1.0;
    s->param_name[0] = 0;
    /* sim_* pass: the prefix is case-sensitive, the rest is lower-cased */
    while (fgets(buff, sizeof buff, f)) {
        if (!keyword_line(buff)) continue;
        char full[64] = "";
        sscanf(buff, "%63s", full);
        if (strncmp(full, "sim_", 4)) continue;
        char *val = buff + strlen(full);
        char *v = full + 4;
        for (char *p = v; *p; p++) *p = (char)tolower((unsigned char)*p);
        if (!strcmp(v, "paramfile")) sscanf(val, "%63s", s->param_name);
        else if (!strcmp(v, "mode")) sscanf(val, "%i", &m->mode);
        else if (!strcmp(v, "intmode")) sscanf(val, "%i", &m->intmode);
        else if (!strcmp(v, "timestep")) sscanf(val, "%lf", &m->dt);
        else if (!strcmp(v, "cutoff")) sscanf(val, "%lf", &m->rc);
        else if (!strcmp(v, "cutoff_nblist")) sscanf(val, "%lf", &m->rclist);
        else if (!strcmp(v, "cutoff_ewald")) sscanf(val, "%lf", &m->rc_ew);
        else if (!strcmp(v, "cutoff_ewald_nblist")) sscanf(val, "%lf", &m->rclist_ew);
        else if (!strcmp(v, "nboxes")) sscanf(val, "%i", &m->nboxes);
        else if (!strcmp(v, "dielectric")) { sscanf(val, "%lf", &eps_rf); have_eps = 1; }
    }
    if (m->mode == 0) return fail(2, "simul.input: sim_mode missing");
    if (m->intmode == 0) return fail(2, "simul.input: sim_intmode missing");
    if (!s->param_name[0]) return fail(2, "simul.input: sim_paramfile missing");
    if (m->dt == -1.0) return fail(2, "simul.input: sim_timestep missing");
    if (m->nboxes == 0) m->nboxes = 1;
    if (ibox < 0 || ibox >= m->nboxes) return fail(111, "simul.input: box %d not among %d boxes", ibox, m->nboxes);
    /* box_* pass: the whole name is lower-cased; values apply to the box
     * named by the last box_number */
    rewind(f);
    int k = -1;
    m->temp = -1.0;
    m->eps_rf = have_eps ? eps_rf : 0.0;
    while (fgets(buff, sizeof buff, f)) {
        if (!keyword_line(buff)) continue;
        char full[64] = "";
        sscanf(buff, "%63s", full);
        for (char *p = full; *p; p++) *p = (char)tolower((unsigned char)*p);
        if (strncmp(full, "box_", 4)) continue;
        char *val = buff + strlen(full);
        char *v = full + 4;
        if (!strcmp(v, "number")) {
            sscanf(val, "%i", &k);
            if (k >= m->nboxes) return fail(2, "simul.input: invalid box number %d", k);
            continue;
        }
        if (k == -1) return fail(2, "simul.input: a box variable before box_number");
        if (k != ibox) continue;
        if (!strcmp(v, "size_xyz")) sscanf(val, "%lf%lf%lf", &m->lx, &m->ly, &m->lz);
        else if (!strcmp(v, "temp")) sscanf(val, "%lf", &m->temp);
        else if (!strcmp(v, "nh_tstat_n")) sscanf(val, "%i", &m->nhc_m);
        else if (!strcmp(v, "nh_tstat_mass")) sscanf(val, "%lf", &m->nhc_q);
        else if (!strcmp(v, "respa")) sscanf(val, "%i", &m->respa);
        else if (!strcmp(v, "kappa")) sscanf(val, "%lf", &m->kappa);
        else if (!strcmp(v, "spme_spline_order")) sscanf(val, "%i", &m->order);
        else if (!strcmp(v, "grid_pts_xyz")) sscanf(val, "%i%i%i", &m->grid[0], &m->grid[1], &m->grid[2]);
        else if (!strcmp(v, "dielectric")) sscanf(val, "%lf", &m->eps_rf);
    }
    if (m->lx == 0 || m->ly == 0 || m->lz == 0) return fail(2, "simul.input: box_size missing in box %d", ibox);
    if (m->temp == -1.0) return fail(2, "simul.input: box_temp missing in box %d", ibox);
    /* read_sim.c:436-451 defaults */
    if (m->nhc_m == 0) m->nhc_m = 8;
    if (m->nhc_q == 0) m->nhc_q = 1.0e-18;
    for (int d = 0; d < 3; d++)
        if (m->grid[d] == 0) m->grid[d] = 35;
    if (m->order == 0) m->order = 6;
    return 0;
}

static int read_sim_positional(FILE *f, int ibox, ttc_io *s)
{
    ttc_io_sim *m = &s->sim;
    char tt[160];
    long ltmp, ltmp2;
    unsigned long ultmp;
    int itmp;
    double dtmp;
    for (int i = 0; i < 12; i++) skip_line(f);
    skip_line(f);                                         /* title */
    if (fscanf(f, "Parameter File Name=%63s", s->param_name) != 1)
        return fail(2, "simul.input: no 'Parameter File Name=' line");
    skip_line(f);
    int ok = 1;
    ok &= fscanf(f, "%d", &m->mode) == 1;        skip_line(f);
    ok &= fscanf(f, "%d", &m->intmode) == 1;     skip_line(f);
    ok &= fscanf(f, "%lu", &ultmp) == 1;         skip_line(f);   /* cyc_eq */
    ok &= fscanf(f, "%lu", &ultmp) == 1;         skip_line(f);   /* cyc_pr */
    ok &= fscanf(f, "%ld", &ltmp) == 1;          skip_line(f);   /* cyc_hybmd */
    ok &= fscanf(f, "%ld", &ltmp) == 1;          skip_line(f);   /* cyc_swap */
    ok &= fscanf(f, "%ld", &ltmp) == 1;          skip_line(f);   /* blockd */
    ok &= fscanf(f, "%ld %ld", &ltmp, &ltmp2) == 2; skip_line(f); /* blockc, blockt */
    ok &= fscanf(f, "%ld", &ltmp) == 1;          skip_line(f);   /* drift */
    ok &= fscanf(f, "%lf", &m->dt) == 1;         skip_line(f);
    ok &= fscanf(f, "%d", &itmp) == 1;           skip_line(f);   /* nsteps */
    ok &= fscanf(f, "%lf", &m->rc) == 1;
    ok &= fscanf(f, "%lf", &m->rclist) == 1;     skip_line(f);
    ok &= fscanf(f, "%lf", &m->rc_ew) == 1;
    ok &= fscanf(f, "%lf", &m->rclist_ew) == 1;  skip_line(f);
    ok &= fscanf(f, "%lu", &ultmp) == 1;         skip_line(f);   /* nlist */
    ok &= fscanf(f, "%d", &m->nboxes) == 1;      skip_line(f);
    ok &= fscanf(f, "%lf", &dtmp) == 1;          skip_line(f);   /* epsRF for all boxes */
    skip_line(f);
    if (!ok) return fail(2, "simul.input: malformed simulation block");
    if (ibox < 0 || ibox >= m->nboxes) return fail(111, "simul.input: box %d not among %d boxes", ibox, m->nboxes);
    for (int k = 0; k <= ibox; k++) {
        int boxnum = -1;
        double d3[3];
        int i3[3];
        ok = fscanf(f, "%d", &boxnum) == 1; skip_line(f);
        if (!ok || boxnum != k) return fail(1, "simul.input: box number %d does not match (should be %d)", boxnum, k);
        ok &= fscanf(f, "%lf %lf %lf", &m->lx, &m->ly, &m->lz) == 3;
        if (!fgets(tt, 80, f)) tt[0] = 0;
        ok &= fscanf(f, "%lf", &m->temp) == 1;       skip_line(f);
        ok &= fscanf(f, "%lf", &dtmp) == 1;          skip_line(f);   /* dtemp */
        ok &= fscanf(f, "%lf", &dtmp) == 1;          skip_line(f);   /* Ttau */
        ok &= fscanf(f, "%d", &m->nhc_m) == 1;       skip_line(f);
        ok &= fscanf(f, "%lf", &m->nhc_q) == 1;      skip_line(f);
        for (int r = 0; r < 4; r++) {                                /* P, dpress, Ptau, Pcomp */
            ok &= fscanf(f, "%lf %lf %lf", &d3[0], &d3[1], &d3[2]) == 3;
            skip_line(f);
        }
        ok &= fscanf(f, "%lf", &dtmp) == 1;          skip_line(f);   /* Wg */
        ok &= fscanf(f, "%d", &m->respa) == 1;       skip_line(f);
        ok &= fscanf(f, "%lf", &m->kappa) == 1;      skip_line(f);
        ok &= fscanf(f, "%d %d %d", &i3[0], &i3[1], &i3[2]) == 3; skip_line(f); /* kmax */
        ok &= fscanf(f, "%d", &m->order) == 1;       skip_line(f);
        ok &= fscanf(f, "%d", &m->grid[0]) == 1;
        ok &= fscanf(f, "%d", &m->grid[1]) == 1;
        ok &= fscanf(f, "%d", &m->grid[2]) == 1;     skip_line(f);
        ok &= fscanf(f, "%lf", &m->eps_rf) == 1;     skip_line(f);
        if (!ok) return fail(2, "simul.input: malformed block of box %d", k);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* mol.input (read_config.c:45-150)                                          */
/* ------------------------------------------------------------------------ */
static int read_mol(FILE *f, ttc_io *s)
{
    skip_line(f);
    skip_line(f);
    int ok = fscanf(f, "%d", &s->nc) == 1; skip_line(f);
    ok &= fscanf(f, "%d", &s->nmolall) == 1; skip_line(f);
    skip_line(f);
    if (!ok || s->nc <= 0) return fail(1, "mol.input: malformed header");
    s->nbox = zalloc(s->nc, sizeof(int));
    s->msite = zalloc(s->nc, sizeof(int));
    s->mres = zalloc(s->nc, sizeof(int));
    if (!s->nbox || !s->msite || !s->mres) return fail(11, "mol.input: out of memory");
    for (int i = 0; i < s->nc; i++) { ok &= fscanf(f, "%d", &s->nbox[i]) == 1; skip_line(f); }
    skip_line(f);
    for (int i = 0; i < s->nc; i++) { ok &= fscanf(f, "%d", &s->msite[i]) == 1; skip_line(f); }
    skip_line(f);
    for (int i = 0; i < s->nc; i++) { ok &= fscanf(f, "%d", &s->mres[i]) == 1; skip_line(f); }
    skip_line(f);
    ok &= fscanf(f, "%d", &s->nsolute) == 1;
    if (!ok) return fail(1, "mol.input: malformed molecule counts");
    int qt = 0;
    s->n = 0;
    for (int i = 0; i < s->nc; i++) {
        qt += s->nbox[i];
        s->n += s->nbox[i] * s->msite[i];
    }
    if (qt != s->nmolall)
        return fail(1, "mol.input: total number of molecules %d != sum of molecules %d", s->nmolall, qt);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* parameter file (read_atom.c:28-510)                                       */
/* ------------------------------------------------------------------------ */

/* Atom ids of up to four 4-character names, scanning the atom list once the
 * way read_atom.c does (a later duplicate name overrides until every name
 * has matched). */
static void name_ids(const ttc_io *s, int nn, char names[][5], int *ids)
{
    int found[4] = {0, 0, 0, 0};
    for (int ii = 0; ii < s->natype; ii++) {
        int all = 1;
        for (int t = 0; t < nn; t++) {
            if (!strcmp(names[t], s->at_name[ii])) { ids[t] = s->at_id[ii]; found[t] = 1; }
            all &= found[t];
        }
        if (all) break;
    }
}

static int read_rows(FILE *f, ttc_io *s, int nn, int nv, int *count, prop_t **rows, const char *what)
{
    if (fscanf(f, "%d", count) != 1 || *count < 0) return fail(1, "parameter file: no %s count", what);
    skip_line(f);
    *rows = zalloc(*count, sizeof(prop_t));
    if (!*rows) return fail(11, "parameter file: out of memory");
    for (int i = 0; i < *count; i++) {
        char names[4][5];
        prop_t *r = &(*rows)[i];
        for (int t = 0; t < nn; t++)
            if (fscanf(f, "%4s", names[t]) != 1) return fail(1, "parameter file: short %s row %d", what, i);
        name_ids(s, nn, names, r->t);
        for (int v = 0; v < nv; v++) {
            int got = v + 1 < nv ? fscanf(f, "%lf", &r->p[v]) : fscanf(f, "%lf\n", &r->p[v]);
            if (got != 1) return fail(1, "parameter file: short %s row %d", what, i);
        }
    }
    return 0;
}

/* qw carries over from row to row and block to block: a name missing from
 * the atom list writes the previous row's id, as read_atom.c's one qw does */
static int read_lj_rows(FILE *f, ttc_io *s, int nrows, double *eps, double *sig, int *qwp)
{
    int qw = *qwp;
    for (int i = 0; i < nrows; i++) {
        char nm[5];
        if (fscanf(f, "%4s", nm) != 1) return fail(1, "parameter file: short LJ row %d", i);
        for (int ii = 0; ii < s->natype; ii++)
            if (!strcmp(nm, s->at_name[ii])) { qw = s->at_id[ii]; break; }
        if (qw < 0 || qw >= IO_MAXID) return fail(1, "parameter file: atom id %d out of range", qw);
        if (fscanf(f, "%lf", &eps[qw]) != 1 || fscanf(f, "%lf\n", &sig[qw]) != 1)
            return fail(1, "parameter file: short LJ row %d", i);
    }
    *qwp = qw;
    return 0;
}

static int read_param(FILE *f, ttc_io *s)
{
    int r, qw = 0;
    skip_line(f);                                          /* title */
    if (fscanf(f, "%d", &s->natype) != 1 || s->natype <= 0) return fail(1, "parameter file: no atom count");
    skip_line(f);
    s->at_id = zalloc(s->natype, sizeof(int));
    s->at_name = zalloc(s->natype, sizeof *s->at_name);
    s->at_mass = zalloc(s->natype, sizeof(double));
    if (!s->at_id || !s->at_name || !s->at_mass) return fail(11, "parameter file: out of memory");
    for (int i = 0; i < s->natype; i++) {
        if (fscanf(f, "%d", &s->at_id[i]) != 1 || fscanf(f, "%4s", s->at_name[i]) != 1 ||
            fscanf(f, "%lf\n", &s->at_mass[i]) != 1)
            return fail(1, "parameter file: short atom row %d", i);
        if (s->at_id[i] < 0 || s->at_id[i] >= IO_MAXID) return fail(1, "parameter file: atom id %d out of range", s->at_id[i]);
    }
    if ((r = read_rows(f, s, 2, 2, &s->nbondp, &s->bondp, "bond"))) return r;      /* kbond, eqbond */
    if ((r = read_rows(f, s, 3, 2, &s->nbendp, &s->bendp, "bend"))) return r;      /* kbend, eqbend */
    if ((r = read_rows(f, s, 4, 3, &s->ntorsp, &s->torsp, "torsion"))) return r;  /* kphi, nphi, delphi */
    skip_line(f);
    if ((r = read_lj_rows(f, s, s->natype, s->eps0, s->sig0, &qw))) return r;
    if ((r = read_rows(f, s, 4, 3, &s->nimprp, &s->imprp, "improper"))) return r;  /* kimpr, -, angeq */
    skip_line(f);
    /* the 1-4 LJ block holds one row fewer than the atom list
     * (read_atom.c:311-330) */
    if ((r = read_lj_rows(f, s, s->natype - 1, s->eps1, s->sig1, &qw))) return r;
    if ((r = read_rows(f, s, 3, 2, &s->nureyp, &s->ureyp, "Urey-Bradley"))) return r; /* k, Seq */
    /* NBFIX block, if the keyword follows (read_atom.c:446-505) */
    char word[160];
    int rc = fscanf(f, "%159s", word);
    skip_line(f);
    while (rc != EOF) {
        if (!strcmp(word, "NBFIX")) break;
        rc = fscanf(f, "%159s", word);
        skip_line(f);
    }
    s->nnbfix = 0;
    if (rc != EOF) {
        if ((r = read_rows(f, s, 2, 2, &s->nnbfix, &s->nbfix, "NBFIX"))) return r; /* eps, sig */
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* psf (read_topology.c:26-891)                                              */
/* ------------------------------------------------------------------------ */
static int section_count(FILE *f, int *n, const char *what)
{
    if (fscanf(f, "%d\n", n) != 1 || *n < 0) return fail(1, "psf: no %s count", what);
    char tt[80];
    if (!fgets(tt, 80, f)) tt[0] = 0;
    return 0;
}

static int read_ints(FILE *f, int *dst, int count, int per, const char *what)
{
    for (int i = 0; i < count; i++)
        for (int t = 0; t < per; t++) {
            int v;
            if (fscanf(f, "%d", &v) != 1) return fail(1, "psf: short %s record %d", what, i);
            dst[per * i + t] = v - 1;
        }
    return 0;
}

static int same2(const int *t, int a, int b) { return (t[0] == a && t[1] == b) || (t[0] == b && t[1] == a); }

/* unordered site-pair set (open addressing) used for the 1-4 duplicate test */
typedef struct { uint64_t *key; size_t cap; } pairset;

static uint64_t pkey(int a, int b)
{
    uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
    return ((uint64_t)hi << 32 | lo) + 1;   /* never 0 */
}

static int ps_init(pairset *p, size_t n)
{
    p->cap = 16;
    while (p->cap < 2 * n + 16) p->cap <<= 1;
    p->key = calloc(p->cap, sizeof(uint64_t));
    return p->key != NULL;
}

static size_t ps_slot(const pairset *p, uint64_t k)
{
    uint64_t h = k * 0x9E3779B97F4A7C15ull;
    size_t i = (size_t)(h >> 17) & (p->cap - 1);
    while (p->key[i] && p->key[i] != k) i = (i + 1) & (p->cap - 1);
    return i;
}

static int ps_has(const pairset *p, int a, int b) { return p->key[ps_slot(p, pkey(a, b))] != 0; }
static void ps_add(pairset *p, int a, int b) { uint64_t k = pkey(a, b); p->key[ps_slot(p, k)] = k; }

/* 1-4 pairs from pairs of bends i < j that share an edge, in the order and
 * with the orientation read_topology.c:739-809 produces them, skipping
 * pairs already bonded, bent, explicitly excluded or found. */
static int derive_14(ttc_io *s)
{
    int nb = s->nbend, n = s->n;
    int *head = malloc(sizeof(int) * (size_t)(n + 1)), *idx = malloc(sizeof(int) * (size_t)(nb ? nb : 1));
    int *cand = malloc(sizeof(int) * (size_t)(2 * nb + 1));
    int cap = 16, cnt = 0;
    int *q = malloc(sizeof(int) * 4 * (size_t)cap);
    pairset ps;
    if (!head || !idx || !cand || !q || !ps_init(&ps, (size_t)(s->nbond + nb + s->nexcl + 4 * nb))) {
        free(head); free(idx); free(cand); free(q);
        return fail(11, "psf: out of memory for 1-4 pairs");
    }
    for (int i = 0; i < s->nbond; i++) ps_add(&ps, s->bond[2 * i], s->bond[2 * i + 1]);
    for (int i = 0; i < nb; i++) ps_add(&ps, s->bend[3 * i], s->bend[3 * i + 2]);
    for (int i = 0; i < s->nexcl; i++) ps_add(&ps, s->excl[2 * i], s->excl[2 * i + 1]);
    /* bends indexed by their middle site (CSR) */
    memset(head, 0, sizeof(int) * (size_t)(n + 1));
    for (int j = 0; j < nb; j++) head[s->bend[3 * j + 1] + 1]++;
    for (int v = 0; v < n; v++) head[v + 1] += head[v];
    {
        int *fill = calloc((size_t)n + 1, sizeof(int));
        if (!fill) { free(head); free(idx); free(cand); free(q); free(ps.key); return fail(11, "psf: out of memory"); }
        for (int j = 0; j < nb; j++) { int b = s->bend[3 * j + 1]; idx[head[b] + fill[b]++] = j; }
        free(fill);
    }
    for (int i = 0; i < nb; i++) {
        const int *bi = &s->bend[3 * i];
        /* candidates j > i whose middle site is an end of bend i, ascending */
        int nc = 0;
        for (int e = 0; e < 2; e++) {
            int v = e ? bi[0] : bi[2];
            if (e && bi[0] == bi[2]) break;
            for (int t = head[v]; t < head[v + 1]; t++)
                if (idx[t] > i) cand[nc++] = idx[t];
        }
        for (int u = 1; u < nc; u++) {           /* insertion sort: a handful of entries */
            int x = cand[u], w = u - 1;
            while (w >= 0 && cand[w] > x) { cand[w + 1] = cand[w]; w--; }
            cand[w + 1] = x;
        }
        for (int u = 0; u < nc; u++) {
            int j = cand[u];
            if (u && cand[u - 1] == j) continue;
            const int *bj = &s->bend[3 * j];
            int qa, qb, qc, qd, hit = 0;
            if (bi[2] == bj[1]) {
                if (bi[1] == bj[0]) { qa = bi[0]; qb = bi[1]; qc = bi[2]; qd = bj[2]; hit = 1; }
                else if (bi[1] == bj[2]) { qa = bi[0]; qb = bi[1]; qc = bi[2]; qd = bj[0]; hit = 1; }
            } else if (bi[0] == bj[1]) {
                if (bi[1] == bj[2]) { qa = bj[0]; qb = bj[1]; qc = bj[2]; qd = bi[2]; hit = 1; }
                else if (bi[1] == bj[0]) { qa = bj[2]; qb = bj[1]; qc = bj[0]; qd = bi[2]; hit = 1; }
            }
            if (!hit || ps_has(&ps, qa, qd)) continue;
            if (cnt == cap) {
                int *nq = realloc(q, sizeof(int) * 8 * (size_t)cap);
                if (!nq) { free(head); free(idx); free(cand); free(q); free(ps.key); return fail(11, "psf: out of memory"); }
                q = nq;
                cap *= 2;
            }
            q[4 * cnt] = qa; q[4 * cnt + 1] = qb; q[4 * cnt + 2] = qc; q[4 * cnt + 3] = qd;
            cnt++;
            ps_add(&ps, qa, qd);
        }
    }
    free(head); free(idx); free(cand); free(ps.key);
    s->in14 = q;
    s->n14 = cnt;
    return 0;
}

static int read_psf(FILE *f, ttc_io *s)
{
    int r, n = s->n;
    for (int i = 0; i < 5; i++) skip_line(f);
    s->atomid = zalloc(n, sizeof(int));
    s->res = zalloc(n, sizeof(int));
    s->q = zalloc(n, sizeof(double));
    s->mass = zalloc(n, sizeof(double));
    s->name = zalloc(n, sizeof *s->name);
    if (!s->atomid || !s->res || !s->q || !s->mass || !s->name) return fail(11, "psf: out of memory");
    int nsite = 0, nres = 0;
    for (int cc = 0; cc < s->nc; cc++) nres += s->nbox[cc] * s->mres[cc];
    for (int cc = 0; cc < s->nc; cc++) {
        int sites;
        if (fscanf(f, "%d", &sites) != 1) return fail(1, "psf: no site count for molecule type %d", cc);
        skip_line(f);
        if (sites != s->msite[cc] * s->nbox[cc])
            return fail(102, "psf: %d sites for molecule type %d, mol.input gives %d", sites, cc,
                        s->msite[cc] * s->nbox[cc]);
        for (int i = 0; i < sites; i++, nsite++) {
            int num, resno;
            char seg[7], resname[5];
            if (fscanf(f, "%d", &num) != 1 || fscanf(f, "%6s", seg) != 1 || fscanf(f, "%d", &resno) != 1 ||
                fscanf(f, "%4s", resname) != 1 || fscanf(f, "%4s", s->name[nsite]) != 1 ||
                fscanf(f, "%d", &s->atomid[nsite]) != 1 || fscanf(f, "%lf", &s->q[nsite]) != 1)
                return fail(1, "psf: short atom record %d", nsite + 1);
            char tt[160];
            if (!fgets(tt, 150, f)) tt[0] = 0;
            if (resno > nres)
                return fail(102, "psf: site %d has residue number %d, mol.input gives %d residues", nsite + 1, resno,
                            nres);
            s->res[nsite] = resno - 1;
            if (s->atomid[nsite] < 0 || s->atomid[nsite] >= IO_MAXID)
                return fail(1, "psf: atom id %d of site %d out of range", s->atomid[nsite], nsite + 1);
            for (int mm = 0; mm < s->natype; mm++)
                if (s->at_id[mm] == s->atomid[nsite]) { s->mass[nsite] = s->at_mass[mm]; break; }
        }
    }
    /* bonds: four pairs a line, first parameter row that matches */
    if ((r = section_count(f, &s->nbond, "bond"))) return r;
    s->bond = zalloc(2 * (size_t)s->nbond, sizeof(int));
    s->bond_req = zalloc(s->nbond, sizeof(double));
    s->bond_k = zalloc(s->nbond, sizeof(double));
    if (!s->bond || !s->bond_req || !s->bond_k) return fail(11, "psf: out of memory");
    if ((r = read_ints(f, s->bond, s->nbond, 2, "bond"))) return r;
    for (int i = 0; i < s->nbond; i++) {
        int a = s->bond[2 * i], b = s->bond[2 * i + 1], hit = 0;
        if (a < 0 || a >= n || b < 0 || b >= n) return fail(1, "psf: bond %d names a site out of range", i + 1);
        for (int p = 0; p < s->nbondp && !hit; p++)
            if (same2(s->bondp[p].t, s->atomid[a], s->atomid[b])) {
                s->bond_req[i] = s->bondp[p].p[1];
                s->bond_k[i] = s->bondp[p].p[0] * IO_KCAL;
                hit = 1;
            }
        if (!hit) return fail(1, "bond parameters not defined between atoms %d and %d", a + 1, b + 1);
    }
    /* bends: three triples a line */
    if ((r = section_count(f, &s->nbend, "bend"))) return r;
    s->bend = zalloc(3 * (size_t)s->nbend, sizeof(int));
    s->bend_eq = zalloc(s->nbend, sizeof(double));
    s->bend_k = zalloc(s->nbend, sizeof(double));
    if (!s->bend || !s->bend_eq || !s->bend_k) return fail(11, "psf: out of memory");
    if ((r = read_ints(f, s->bend, s->nbend, 3, "bend"))) return r;
    for (int i = 0; i < s->nbend; i++) {
        const int *t = &s->bend[3 * i];
        int hit = 0;
        for (int c = 0; c < 3; c++)
            if (t[c] < 0 || t[c] >= n) return fail(1, "psf: bend %d names a site out of range", i + 1);
        int ia = s->atomid[t[0]], ib = s->atomid[t[1]], ic = s->atomid[t[2]];
        for (int p = 0; p < s->nbendp && !hit; p++) {
            const int *pt = s->bendp[p].t;
            if (pt[1] == ib && ((pt[0] == ia && pt[2] == ic) || (pt[0] == ic && pt[2] == ia))) {
                s->bend_eq[i] = s->bendp[p].p[1] * (IO_PI / 180.0);
                s->bend_k[i] = s->bendp[p].p[0] * IO_KCAL;
                hit = 1;
            }
        }
        if (!hit) return fail(1, "bend parameter not defined between atoms %d, %d, %d", t[0] + 1, t[1] + 1, t[2] + 1);
    }
    /* torsions: two quads a line; every exact match (either direction) adds a
     * Fourier term, a wildcard row X-b-c-X applies only before any match */
    if ((r = section_count(f, &s->ntors, "torsion"))) return r;
    s->tors = zalloc(4 * (size_t)s->ntors, sizeof(int));
    s->kphi = zalloc(IO_NFOURIER * (size_t)s->ntors, sizeof(double));
    s->nphi = zalloc(IO_NFOURIER * (size_t)s->ntors, sizeof(double));
    s->delphi = zalloc(IO_NFOURIER * (size_t)s->ntors, sizeof(double));
    if (!s->tors || !s->kphi || !s->nphi || !s->delphi) return fail(11, "psf: out of memory");
    if ((r = read_ints(f, s->tors, s->ntors, 4, "torsion"))) return r;
    for (int i = 0; i < s->ntors; i++) {
        const int *t = &s->tors[4 * i];
        for (int c = 0; c < 4; c++)
            if (t[c] < 0 || t[c] >= n) return fail(1, "psf: torsion %d names a site out of range", i + 1);
        int id[4] = {s->atomid[t[0]], s->atomid[t[1]], s->atomid[t[2]], s->atomid[t[3]]};
        int hit = 0, nt = 0;
        double *kp = &s->kphi[IO_NFOURIER * i], *np = &s->nphi[IO_NFOURIER * i], *dp = &s->delphi[IO_NFOURIER * i];
        for (int p = 0; p < s->ntorsp; p++) {
            const int *pt = s->torsp[p].t;
            const double *pv = s->torsp[p].p;
            int fwd = pt[0] == id[0] && pt[1] == id[1] && pt[2] == id[2] && pt[3] == id[3];
            int rev = pt[0] == id[3] && pt[1] == id[2] && pt[2] == id[1] && pt[3] == id[0];
            if (fwd || rev) {
                if (nt >= IO_NFOURIER) return fail(1, "torsion %d matches more than %d parameter rows", i + 1, IO_NFOURIER);
                kp[nt] = IO_KCAL * pv[0];
                np[nt] = pv[1];
                dp[nt] = IO_PI * pv[2] / 180;
                nt++;
                hit = 1;
            } else if (!hit && pt[0] == IO_WILD && pt[3] == IO_WILD &&
                       ((pt[1] == id[1] && pt[2] == id[2]) || (pt[1] == id[2] && pt[2] == id[1]))) {
                kp[0] = IO_KCAL * pv[0];
                np[0] = pv[1];
                dp[0] = IO_PI * pv[2] / 180;
                hit = 1;
                break;
            }
        }
        if (!hit)
            return fail(1, "dihedral parameters not defined for atom types %d %d %d %d", id[0], id[1], id[2], id[3]);
    }
    /* impropers: two quads a line; the last exact match wins, a wildcard
     * row a-X-X-d applies only before any match */
    if ((r = section_count(f, &s->nimpr, "improper"))) return r;
    s->impr = zalloc(4 * (size_t)s->nimpr, sizeof(int));
    s->impr_k = zalloc(s->nimpr, sizeof(double));
    s->impr_eq = zalloc(s->nimpr, sizeof(double));
    if (!s->impr || !s->impr_k || !s->impr_eq) return fail(11, "psf: out of memory");
    if ((r = read_ints(f, s->impr, s->nimpr, 4, "improper"))) return r;
    for (int i = 0; i < s->nimpr; i++) {
        const int *t = &s->impr[4 * i];
        for (int c = 0; c < 4; c++)
            if (t[c] < 0 || t[c] >= n) return fail(1, "psf: improper %d names a site out of range", i + 1);
        int id[4] = {s->atomid[t[0]], s->atomid[t[1]], s->atomid[t[2]], s->atomid[t[3]]};
        int hit = 0;
        for (int p = 0; p < s->nimprp; p++) {
            const int *pt = s->imprp[p].t;
            const double *pv = s->imprp[p].p;
            int fwd = pt[0] == id[0] && pt[1] == id[1] && pt[2] == id[2] && pt[3] == id[3];
            int rev = pt[0] == id[3] && pt[1] == id[2] && pt[2] == id[1] && pt[3] == id[0];
            if (fwd || rev) {
                s->impr_k[i] = IO_KCAL * pv[0];
                s->impr_eq[i] = IO_PI * pv[2] / 180;
                hit = 1;
            } else if (!hit && pt[1] == IO_WILD && pt[2] == IO_WILD &&
                       ((pt[0] == id[0] && pt[3] == id[3]) || (pt[0] == id[3] && pt[3] == id[0]))) {
                s->impr_k[i] = IO_KCAL * pv[0];
                s->impr_eq[i] = IO_PI * pv[2] / 180;
                hit = 1;
                break;
            }
        }
        if (!hit)
            return fail(1, "improper dihedral parameters not defined for atoms %d, %d, %d, %d", t[0] + 1, t[1] + 1,
                        t[2] + 1, t[3] + 1);
    }
    /* donors and acceptors (pairs, four a line): read past */
    for (int sec = 0; sec < 2; sec++) {
        int cnt;
        if ((r = section_count(f, &cnt, sec ? "acceptor" : "donor"))) return r;
        int *tmp = zalloc(2 * (size_t)cnt, sizeof(int));
        if (!tmp) return fail(11, "psf: out of memory");
        r = read_ints(f, tmp, cnt, 2, sec ? "acceptor" : "donor");
        free(tmp);
        if (r) return r;
    }
    /* explicit exclusions: a list of partners, then each site's running
     * upper bound into it */
    if ((r = section_count(f, &s->nexcl, "exclusion"))) return r;
    int *nnb = zalloc((size_t)s->nexcl + 1, sizeof(int));
    s->excl = zalloc(2 * (size_t)s->nexcl, sizeof(int));
    if (!nnb || !s->excl) { free(nnb); return fail(11, "psf: out of memory"); }
    if ((r = read_ints(f, nnb, s->nexcl, 1, "exclusion"))) { free(nnb); return r; }
    int lo = 0, cnt = 0;
    for (int i = 0; i < n; i++) {
        int hi;
        if (fscanf(f, "%d", &hi) != 1) { free(nnb); return fail(1, "psf: short exclusion bounds (site %d)", i + 1); }
        if (hi > s->nexcl || (hi > lo && lo < 0)) { free(nnb); return fail(1, "psf: exclusion bound %d out of range", hi); }
        for (int j = lo; j < hi; j++, cnt++) {
            s->excl[2 * cnt] = i;
            s->excl[2 * cnt + 1] = nnb[j];
        }
        lo = hi;
    }
    free(nnb);
    if (cnt != s->nexcl) s->nexcl = cnt;
    if ((r = derive_14(s))) return r;
    /* STYPE exclusion table, S14 build (read_topology.c:847-891) */
    s->nxint = s->nbond + s->nbend + s->nexcl + s->n14;
    s->xint = zalloc(3 * (size_t)s->nxint, sizeof(int));
    if (!s->xint) return fail(11, "psf: out of memory");
    int c = 0;
    for (int i = 0; i < s->nbond; i++, c++) { s->xint[3 * c] = s->bond[2 * i]; s->xint[3 * c + 1] = s->bond[2 * i + 1]; s->xint[3 * c + 2] = 1; }
    for (int i = 0; i < s->nbend; i++, c++) { s->xint[3 * c] = s->bend[3 * i]; s->xint[3 * c + 1] = s->bend[3 * i + 2]; s->xint[3 * c + 2] = 2; }
    for (int i = 0; i < s->nexcl; i++, c++) { s->xint[3 * c] = s->excl[2 * i]; s->xint[3 * c + 1] = s->excl[2 * i + 1]; s->xint[3 * c + 2] = 3; }
    for (int i = 0; i < s->n14; i++, c++) { s->xint[3 * c] = s->in14[4 * i]; s->xint[3 * c + 1] = s->in14[4 * i + 3]; s->xint[3 * c + 2] = 4; }
    /* Urey-Bradley pairs: for each bend in order, the first matching row */
    s->urey = zalloc(2 * (size_t)s->nbend, sizeof(int));
    s->urey_s0 = zalloc(s->nbend, sizeof(double));
    s->urey_k = zalloc(s->nbend, sizeof(double));
    if (!s->urey || !s->urey_s0 || !s->urey_k) return fail(11, "psf: out of memory");
    s->nurey = 0;
    for (int i = 0; i < s->nbend; i++) {
        const int *t = &s->bend[3 * i];
        int ia = s->atomid[t[0]], ib = s->atomid[t[1]], ic = s->atomid[t[2]];
        for (int p = 0; p < s->nureyp; p++) {
            const int *pt = s->ureyp[p].t;
            if (pt[1] == ib && ((pt[0] == ia && pt[2] == ic) || (pt[0] == ic && pt[2] == ia))) {
                s->urey[2 * s->nurey] = t[0];
                s->urey[2 * s->nurey + 1] = t[2];
                s->urey_s0[s->nurey] = s->ureyp[p].p[1];
                s->urey_k[s->nurey] = s->ureyp[p].p[0] * IO_KCAL;
                s->nurey++;
                break;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* crd (read_config.c:236-325) and pbc_init.c:21-66                          */
/* ------------------------------------------------------------------------ */
static int read_crd(FILE *f, ttc_io *s, const char *path)
{
    int qw, n = s->n;
    if (fscanf(f, "%d", &qw) != 1 || qw != n)
        return fail(1, "number of sites in box (%d) do not match with number of coordinate sets (%d) in %s", n, qw, path);
    s->xn = zalloc(3 * (size_t)n, sizeof(double));
    s->x = zalloc(3 * (size_t)n, sizeof(double));
    if (!s->xn || !s->x) return fail(11, "crd: out of memory");
    for (int i = 0; i < n; i++) {
        int resno;
        char resname[64], nm[5];
        if (fscanf(f, "%d", &qw) != 1 || fscanf(f, "%d", &resno) != 1 || fscanf(f, "%63s", resname) != 1 ||
            fscanf(f, "%4s", nm) != 1)
            return fail(1, "crd: short record %d in %s", i + 1, path);
        if (strcmp(s->name[i], nm))
            return fail(1, "Name of site %d (%s) in crd does not match name in psf %d (%s)", qw, nm, i, s->name[i]);
        if (fscanf(f, "%lf", &s->xn[3 * i]) != 1 || fscanf(f, "%lf", &s->xn[3 * i + 1]) != 1 ||
            fscanf(f, "%lf", &s->xn[3 * i + 2]) != 1)
            return fail(1, "crd: short coordinates of site %d in %s", i + 1, path);
        char tt[80];
        if (!fgets(tt, 80, f)) tt[0] = 0;
    }
    const double L[3] = {s->sim.lx, s->sim.ly, s->sim.lz};
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double v = s->xn[3 * i + d];
            double il = 1.0 / L[d];
            double qd = v * il;
            int m = (int)(qd < 0.0 ? qd - 0.5 : qd + 0.5);
            s->x[3 * i + d] = v - m * L[d];
        }
    return 0;
}

/* Types present in the box (ascending atomid), their LJ tables (setlj.c:
 * 181-205 in the STYPE build) and each site's dense type. */
static int build_types(ttc_io *s)
{
    int *dense = malloc(sizeof(int) * IO_MAXID);
    if (!dense) return fail(11, "out of memory");
    for (int i = 0; i < IO_MAXID; i++) dense[i] = -1;
    for (int i = 0; i < s->n; i++) dense[s->atomid[i]] = 0;
    s->ntype = 0;
    for (int i = 0; i < IO_MAXID; i++)
        if (dense[i] == 0) dense[i] = s->ntype++;
    int T = s->ntype;
    s->type_id = zalloc(T, sizeof(int));
    s->type = zalloc(s->n, sizeof(int));
    s->lj_eps = zalloc((size_t)T * T, sizeof(double));
    s->lj_sig = zalloc((size_t)T * T, sizeof(double));
    s->lj14_eps = zalloc((size_t)T * T, sizeof(double));
    s->lj14_sig = zalloc((size_t)T * T, sizeof(double));
    if (!s->type_id || !s->type || !s->lj_eps || !s->lj_sig || !s->lj14_eps || !s->lj14_sig) {
        free(dense);
        return fail(11, "out of memory");
    }
    for (int i = 0; i < IO_MAXID; i++)
        if (dense[i] >= 0) s->type_id[dense[i]] = i;
    for (int i = 0; i < s->n; i++) s->type[i] = dense[s->atomid[i]];
    for (int a = 0; a < T; a++)
        for (int b = 0; b < T; b++) {
            int ia = s->type_id[a], ib = s->type_id[b];
            size_t k = (size_t)a * T + b;
            s->lj_eps[k] = IO_KCAL * sqrt(s->eps0[ia] * s->eps0[ib]);
            s->lj_sig[k] = 0.5 * (2.0 * s->sig0[ia] + 2.0 * s->sig0[ib]);
            s->lj14_eps[k] = IO_KCAL * sqrt(s->eps1[ia] * s->eps1[ib]);
            s->lj14_sig[k] = 0.5 * (2.0 * s->sig1[ia] + 2.0 * s->sig1[ib]);
        }
    for (int p = 0; p < s->nnbfix; p++) {
        int t1 = s->nbfix[p].t[0], t2 = s->nbfix[p].t[1];
        if (t1 < 0 || t1 >= IO_MAXID || t2 < 0 || t2 >= IO_MAXID || dense[t1] < 0 || dense[t2] < 0) continue;
        size_t k1 = (size_t)dense[t1] * T + dense[t2], k2 = (size_t)dense[t2] * T + dense[t1];
        s->lj_eps[k1] = s->lj_eps[k2] = s->nbfix[p].p[0] * -IO_KCAL;
        s->lj_sig[k1] = s->lj_sig[k2] = s->nbfix[p].p[1];
    }
    free(dense);
    return 0;
}

int ttc_io_read(const char *dir, int ibox, ttc_io **out)
{
    if (!dir || !out) return fail(111, "ttc_io_read: bad argument");
    *out = NULL;
    ttc_io *s = calloc(1, sizeof *s);
    if (!s) return fail(11, "ttc_io_read: out of memory");
    char path[1024];
    int r = 0;
    FILE *f = open_in(dir, "simul.input", path, sizeof path);
    if (!f) { ttc_io_free(s); return fail(1, "input file %s does not exist", path); }
    r = is_keyword_file(f) ? read_sim_keywords(f, ibox, s) : read_sim_positional(f, ibox, s);
    fclose(f);
    if (!r) {
        f = open_in(dir, "mol.input", path, sizeof path);
        if (!f) r = fail(1, "input file %s does not exist", path);
        else { r = read_mol(f, s); fclose(f); }
    }
    if (!r) {
        f = open_in(dir, s->param_name, path, sizeof path);
        if (!f) r = fail(1, "configuration file %s does not exist", path);
        else { r = read_param(f, s); fclose(f); }
    }
    if (!r) {
        f = open_in(dir, "simul.psf", path, sizeof path);
        if (!f) r = fail(1, "molecular topology file %s does not exist", path);
        else { r = read_psf(f, s); fclose(f); }
    }
    if (!r) {
        char nm[64];
        snprintf(nm, sizeof nm, "simul%d.crd", ibox);
        f = open_in(dir, nm, path, sizeof path);
        if (!f) r = fail(1, "coordinate file %s does not exist", path);
        else { r = read_crd(f, s, path); fclose(f); }
    }
    if (!r) r = build_types(s);
    if (r) { ttc_io_free(s); return r; }
    *out = s;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* accessors                                                                 */
/* ------------------------------------------------------------------------ */
int ttc_io_get_counts(const ttc_io *s, ttc_io_counts *c)
{
    if (!s || !c) return fail(111, "ttc_io_get_counts: bad argument");
    c->nsite = s->n; c->ncomp = s->nc; c->nbond = s->nbond; c->nbend = s->nbend; c->nurey = s->nurey;
    c->ntors = s->ntors; c->nimpr = s->nimpr; c->nexcl = s->nexcl; c->n14 = s->n14; c->nxint = s->nxint;
    c->ntype = s->ntype;
    return 0;
}

int ttc_io_get_sim(const ttc_io *s, ttc_io_sim *sim)
{
    if (!s || !sim) return fail(111, "ttc_io_get_sim: bad argument");
    *sim = s->sim;
    return 0;
}

#define COPY(dst, src, cnt) do { if (dst) memcpy((dst), (src), sizeof(*(src)) * (size_t)(cnt)); } while (0)

int ttc_io_get_sites(const ttc_io *s, double *pos_nopbc, double *pos, double *q, double *mass, int *atomid,
                     int *type, int *resid, char *names)
{
    if (!s) return fail(111, "ttc_io_get_sites: bad argument");
    COPY(pos_nopbc, s->xn, 3 * s->n);
    COPY(pos, s->x, 3 * s->n);
    COPY(q, s->q, s->n);
    COPY(mass, s->mass, s->n);
    COPY(atomid, s->atomid, s->n);
    COPY(type, s->type, s->n);
    COPY(resid, s->res, s->n);
    if (names) memcpy(names, s->name, 5 * (size_t)s->n);
    return 0;
}

int ttc_io_get_types(const ttc_io *s, int *atomid, double *eps, double *sig, double *eps14, double *sig14)
{
    if (!s) return fail(111, "ttc_io_get_types: bad argument");
    size_t tt = (size_t)s->ntype * s->ntype;
    COPY(atomid, s->type_id, s->ntype);
    COPY(eps, s->lj_eps, tt);
    COPY(sig, s->lj_sig, tt);
    COPY(eps14, s->lj14_eps, tt);
    COPY(sig14, s->lj14_sig, tt);
    return 0;
}

int ttc_io_get_bonds(const ttc_io *s, int *ab, double *req, double *k)
{
    if (!s) return fail(111, "ttc_io_get_bonds: bad argument");
    COPY(ab, s->bond, 2 * s->nbond);
    COPY(req, s->bond_req, s->nbond);
    COPY(k, s->bond_k, s->nbond);
    return 0;
}

int ttc_io_get_bends(const ttc_io *s, int *abc, double *angeq, double *k)
{
    if (!s) return fail(111, "ttc_io_get_bends: bad argument");
    COPY(abc, s->bend, 3 * s->nbend);
    COPY(angeq, s->bend_eq, s->nbend);
    COPY(k, s->bend_k, s->nbend);
    return 0;
}

int ttc_io_get_ureys(const ttc_io *s, int *ab, double *s0, double *k)
{
    if (!s) return fail(111, "ttc_io_get_ureys: bad argument");
    COPY(ab, s->urey, 2 * s->nurey);
    COPY(s0, s->urey_s0, s->nurey);
    COPY(k, s->urey_k, s->nurey);
    return 0;
}

int ttc_io_get_torsions(const ttc_io *s, int *abcd, double *kphi, double *nphi, double *delphi)
{
    if (!s) return fail(111, "ttc_io_get_torsions: bad argument");
    COPY(abcd, s->tors, 4 * s->ntors);
    COPY(kphi, s->kphi, IO_NFOURIER * s->ntors);
    COPY(nphi, s->nphi, IO_NFOURIER * s->ntors);
    COPY(delphi, s->delphi, IO_NFOURIER * s->ntors);
    return 0;
}

int ttc_io_get_impropers(const ttc_io *s, int *abcd, double *k, double *angeq)
{
    if (!s) return fail(111, "ttc_io_get_impropers: bad argument");
    COPY(abcd, s->impr, 4 * s->nimpr);
    COPY(k, s->impr_k, s->nimpr);
    COPY(angeq, s->impr_eq, s->nimpr);
    return 0;
}

int ttc_io_get_exclusions(const ttc_io *s, int *ab)
{
    if (!s) return fail(111, "ttc_io_get_exclusions: bad argument");
    COPY(ab, s->excl, 2 * s->nexcl);
    return 0;
}

int ttc_io_get_14(const ttc_io *s, int *abcd)
{
    if (!s) return fail(111, "ttc_io_get_14: bad argument");
    COPY(abcd, s->in14, 4 * s->n14);
    return 0;
}

int ttc_io_get_xint(const ttc_io *s, int *a, int *b, int *xflag)
{
    if (!s) return fail(111, "ttc_io_get_xint: bad argument");
    for (int i = 0; i < s->nxint; i++) {
        if (a) a[i] = s->xint[3 * i];
        if (b) b[i] = s->xint[3 * i + 1];
        if (xflag) xflag[i] = s->xint[3 * i + 2];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* .trr frames (write_trr.c:22-200)                                          */
/* ------------------------------------------------------------------------ */
static void put_be32(unsigned char *p, uint32_t v)
{
    p[0] = (unsigned char)(v >> 24); p[1] = (unsigned char)(v >> 16); p[2] = (unsigned char)(v >> 8); p[3] = (unsigned char)v;
}

/* a real as the reference stores it: rounded to float, widened to double,
 * big-endian */
static void put_real(unsigned char *p, double v)
{
    double d = (double)(float)v;
    uint64_t u;
    memcpy(&u, &d, 8);
    for (int i = 0; i < 8; i++) p[i] = (unsigned char)(u >> (56 - 8 * i));
}

int ttc_io_write_trr(const char *path, int natoms, const double *x, int xstride, const double *v, int vstride,
                     const double *f, int fstride, double lx, double ly, double lz, long time_ind, double time_val)
{
    if (!path || natoms < 0 || (x && xstride < 3) || (v && vstride < 3) || (f && fstride < 3))
        return fail(111, "ttc_io_write_trr: bad argument");
    const int prec = 8;
    size_t nsec = (size_t)(x != NULL) + (v != NULL) + (f != NULL);
    size_t bytes = 4 * 16 + 2 * prec + 9 * prec + nsec * (size_t)natoms * 3 * prec;
    unsigned char *buf = malloc(bytes), *p = buf;
    if (!buf) return fail(11, "ttc_io_write_trr: out of memory");
    long sec = (long)prec * 3 * natoms;
    const long hdr[16] = {1993, 13, 0 /* title length */, 0, 0, (long)prec * 9, 0, 0, 0, 0,
                          x ? sec : 0, v ? sec : 0, f ? sec : 0, natoms, time_ind, 0};
    for (int i = 0; i < 16; i++, p += 4) put_be32(p, (uint32_t)hdr[i]);
    put_real(p, (float)time_val); p += prec;
    put_real(p, 0.0); p += prec;                               /* lambda */
    const double box[9] = {lx / 10.0, 0, 0, 0, ly / 10.0, 0, 0, 0, lz / 10.0};
    for (int i = 0; i < 9; i++, p += prec) put_real(p, box[i]);
    const double *sect[3] = {x, v, f};
    const int stride[3] = {xstride, vstride, fstride};
    for (int s = 0; s < 3; s++) {
        if (!sect[s]) continue;
        for (int i = 0; i < natoms; i++)
            for (int d = 0; d < 3; d++, p += prec) put_real(p, sect[s][(size_t)stride[s] * i + d] / 10.0);
    }
    FILE *fo = fopen(path, "ab");
    if (!fo) { free(buf); return fail(1, "ttc_io_write_trr: cannot open %s", path); }
    size_t w = fwrite(buf, 1, bytes, fo);
    int cl = fclose(fo);
    free(buf);
    if (w != bytes || cl) return fail(1, "ttc_io_write_trr: short write to %s", path);
    return 0;
}
