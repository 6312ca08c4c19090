/*
 * ttc_oracle.c -- FP64 CPU restatement of TheCodeCombo_7's nonbonded path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (tests/, smoke(),
 * bench.py cpu_baseline).  Nothing in the product links or calls it.
 *
 * Restated routines and what pins each one (tests/test_oracle_*.py):
 *   orc_erfc          ewald.c:817-835 (Numerical Recipes erfcc).  Pinned by
 *                     libm erfc over [0,6] (published bound 1.2e-7) and by
 *                     the reference's own erfc compiled into oracle/_ref.
 *   orc_fillspline    pme_spline.c:10-49.  Pinned by the closed-form cubic
 *                     B-spline M4 and its derivative, partition of unity.
 *   orc_dftmod        pme_setup.c:110-130.  Pinned by numpy FFT of the
 *                     same samples and by k <-> K-k symmetry.
 *   orc_fft3d         FFTW3 3-D complex DFT, unnormalised.  Pinned by the
 *                     naive O(n^2) DFT and numpy.fft (1e-12).
 *   orc_nblist        nblist.c:95-307.  Pinned bit-exactly by the
 *                     reference's nblist.c compiled into oracle/_ref, and
 *                     by brute force in numpy.
 *   orc_pair          cnonbond_nblist.c:134-523.  LJ part pinned by the
 *                     reference's cnonbondnl compiled into oracle/_ref
 *                     (charges zero); Coulomb part pinned by the
 *                     kappa-independence of real+recip-self-intra, the NaCl
 *                     Madelung constant and a brute-force numpy restatement
 *                     of the erfc/exp closed forms.
 *   orc_pme_calc      pme_qgrid.c:13-121, pme_calc.c:18-279.  Pinned by
 *                     convergence to the reference's classic ewald_kspace
 *                     (oracle/_ref) and to a direct Ewald sum, by finite
 *                     differences of the energy, and zero net force.
 *   orc_cbond         cbond.c:65-120.  Pinned by finite differences.
 *   orc_vverlet_*     integrate_vverlet.c:36-103.  Pinned bit-exactly by the
 *   orc_pbc_all,      reference's integrate_vverlet.c, pbc_all.c and kinet.c
 *   orc_kinet         compiled into oracle/_ref; the 100/m factor by the
 *                     harmonic oscillator's closed form; time reversibility.
 *   orc_nhc           integrate_nhc.c:40-161.  Pinned by conservation of
 *                     the extended Nose-Hoover energy.
 *   orc_go_*          cgolik.c:644-948 / 1202-1229 (12-10 form) as the
 *                     Go-model configuration specifies.  Pinned by E=-eps
 *                     and dE/dr=0 at r=r0 and by range energy identities.
 *   orc_ran2          random.c:37-75 (NR ran2).  Pinned by the published
 *                     Numerical Recipes test values of ran2.
 *   orc_splitmix64,   the device's random streams (no reference
 *   orc_xoshiro_*     counterpart).  Pinned by the published output of
 *                     xoshiro256** from the state {1, 2, 3, 4} and of
 *                     splitmix64 from 0.
 *   orc_dos_interp    dos_interpolate.c:107-124.  Pinned bit-exactly by the
 *                     reference's own dos_interp (DOS build) in oracle/_ref.
 *   orc_wl_run        pivot.c:236-253 acceptance, mcmove_pivot.c:54-215,
 *                     mcmove_displ.c:45-99 moves.  Pinned by exactly known
 *                     distributions of a 4-bead chain with bond energy
 *                     only: the Boltzmann histogram of its energy and the
 *                     Wang-Landau ln g(E), both from the closed-form
 *                     density of states.
 *   orc_mlp_*         no reference counterpart; pinned by central finite
 *                     differences.
 */
#include "ttc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* ewald.c                                                                   */
/* ======================================================================== */

/* ewald.c:817-835.  Chebyshev-fitted complementary error function of
 * Numerical Recipes (erfcc, 2nd ed. section 6.2), fractional error < 1.2e-7. */
double orc_erfc(double x)
{
    double z = fabs(x);
    double t = 1.0 / (1.0 + 0.5 * z);
    double v = t * exp(-z * z + -1.26551223 +
               t * (1.00002368 + t * (0.37409196 + t * (0.09678418 +
               t * (-0.18628806 + t * (0.27886807 + t * (-1.13520398 +
               t * (1.48851587 + t * (-0.82215223 + t * 0.17087277)))))))));
    return x < 0 ? 2.0 - v : v;
}

/* ewald.c:171-178: E_self = kappa/sqrt(pi) * sum q^2 * fconv/epsRF. */
double orc_ewald_self(int n, const double *q, double kappa, double eps_rf)
{
    double fconv = (ORC_ELEC * ORC_ELEQ * ORC_ELEQ * ORC_NA * 0.001) / eps_rf;
    double vs = 0.0;
    for (int i = 0; i < n; i++) vs += q[i] * q[i];
    vs *= kappa / sqrt(ORC_PI);
    return vs * fconv;
}

/* ewald.c:184-194: returns 20 (the reference's exit code) if |sum q| > 5e-6. */
int orc_check_neutral(int n, const double *q)
{
    double c = 0.0;
    for (int i = 0; i < n; i++) c += q[i];
    return fabs(c) > 0.000005 ? 20 : 0;
}

/* ======================================================================== */
/* pme_spline.c / pme_setup.c                                                */
/* ======================================================================== */

/* pme_spline.c:10-49.  array[1..order] = cardinal B-spline weights of order
 * `order` at fractional offset w (array[1] belongs to the lowest grid index
 * floor(u)-order+1); darray[1..order] = derivatives, taken from the
 * order-1 spline before the last recursion step.                           */
void orc_fillspline(double w, int order, double *array, double *darray)
{
    for (int i = 3; i <= order; i++) array[i] = 0.0;
    array[2] = w;
    array[1] = 1.0 - w;
    array[0] = 0.0;
    for (int i = 3; i <= order - 1; i++) {
        double div = 1.0 / (double)(i - 1);
        array[i] = div * w * array[i - 1];
        double di = (double)i;
        for (int j = 1; j <= i - 2; j++) {
            double dj = (double)j;
            array[i - j] = div * ((w + dj) * array[i - j - 1] + (di - dj - w) * array[i - j]);
        }
        array[1] *= div * (1.0 - w);
    }
    darray[0] = 0.0;
    darray[1] = -array[1];
    for (int i = 2; i <= order; i++) darray[i] = array[i - 1] - array[i];
    int kk = order;
    double div = 1.0 / (double)(kk - 1);
    array[kk] = div * w * array[kk - 1];
    double dkk = (double)kk;
    for (int i = 1; i <= kk - 2; i++) {
        double di = (double)i;
        array[kk - i] = div * ((w + di) * array[kk - i - 1] + (dkk - di - w) * array[kk - i]);
    }
    array[1] *= div * (1.0 - w);
}

/* pme_setup.c:110-130: bmod[i] = |sum_j barr[j] exp(2 pi i (i-1)(j-1)/K)|^2,
 * 1-based arrays, no small-modulus repair (the repair is commented out).  */
void orc_dftmod(double *bmod, const double *barr, int grid)
{
    double twopi = 2.0 * ORC_PI;
    for (int i = 1; i <= grid; i++) {
        double s1 = 0.0, s2 = 0.0;
        for (int j = 1; j <= grid; j++) {
            double arg = twopi * (i - 1) * (j - 1) / grid;
            s1 += barr[j] * cos(arg);
            s2 += barr[j] * sin(arg);
        }
        bmod[i] = s1 * s1 + s2 * s2;
    }
}

/* pme_setup.c:86-107: splines at w=0 placed at barr[2..order+1].  The output
 * arrays are 1-based (size K+1), as bsp_mod{x,y,z}.                        */
void orc_pme_bspline(int order, int gx, int gy, int gz, double *bx, double *by, double *bz)
{
    double mn[64], dmn[64];
    orc_fillspline(0.0, order, mn, dmn);
    int ng = gx > gy ? gx : gy;
    ng = ng > gz ? ng : gz;
    ng += 1;
    double *barr = (double *)calloc((size_t)ng + order + 2, sizeof(double));
    for (int i = 2; i <= order + 1; i++) barr[i] = mn[i - 1];
    orc_dftmod(bx, barr, gx);
    orc_dftmod(by, barr, gy);
    orc_dftmod(bz, barr, gz);
    free(barr);
}

/* ======================================================================== */
/* FFT: the unnormalised complex DFT that FFTW3 computes for the reference   */
/* (pme_calc.c:50,113), sign +1 = FFTW_BACKWARD, -1 = FFTW_FORWARD.          */
/* Mixed-radix recursive decimation in time; any prime factor falls back to  */
/* a direct DFT of that length.                                              */
/* ======================================================================== */

static void fft1d_rec(const double *in, double *out, int n, int stride, int sign, double *tmp)
{
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int p = 0;
    for (int f = 2; f * f <= n; f++) if (n % f == 0) { p = f; break; }
    if (p == 0) p = n;           /* prime length: direct DFT */
    int m = n / p;
    if (m == 1) {
        for (int k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; j++) {
                double ang = sign * 2.0 * ORC_PI * (double)((long)j * k % n) / n;
                double c = cos(ang), s = sin(ang);
                double xr = in[2 * j * stride], xi = in[2 * j * stride + 1];
                sr += xr * c - xi * s;
                si += xr * s + xi * c;
            }
            out[2 * k] = sr; out[2 * k + 1] = si;
        }
        return;
    }
    /* sub-transforms of the p decimated sequences, each of length m */
    for (int r = 0; r < p; r++)
        fft1d_rec(in + 2 * r * stride, tmp + 2 * r * m, m, stride * p, sign, out);
    /* combine: X[k + m*q] = sum_r W_n^{r(k+mq)} Y_r[k] */
    for (int k = 0; k < m; k++) {
        for (int qq = 0; qq < p; qq++) {
            int kk = k + m * qq;
            double sr = 0.0, si = 0.0;
            for (int r = 0; r < p; r++) {
                double ang = sign * 2.0 * ORC_PI * (double)((long)r * kk % n) / n;
                double c = cos(ang), s = sin(ang);
                double yr = tmp[2 * (r * m + k)], yi = tmp[2 * (r * m + k) + 1];
                sr += yr * c - yi * s;
                si += yr * s + yi * c;
            }
            out[2 * kk] = sr; out[2 * kk + 1] = si;
        }
    }
}

static void fft1d(double *line, int n, int sign, double *w1, double *w2)
{
    /* w1 holds a copy of the input, w2 is the recursion's scratch (2n each) */
    memcpy(w1, line, sizeof(double) * 2 * n);
    fft1d_rec(w1, line, n, 1, sign, w2);
}

void orc_fft3d(double *data, int nx, int ny, int nz, int sign)
{
    int nmax = nx > ny ? nx : ny;
    nmax = nmax > nz ? nmax : nz;
    double *line = (double *)malloc(sizeof(double) * 2 * nmax);
    double *w1 = (double *)malloc(sizeof(double) * 2 * nmax);
    double *w2 = (double *)malloc(sizeof(double) * 2 * nmax);
    /* z lines (contiguous) */
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++) {
            double *p = data + 2 * ((size_t)(x * ny + y) * nz);
            memcpy(line, p, sizeof(double) * 2 * nz);
            fft1d(line, nz, sign, w1, w2);
            memcpy(p, line, sizeof(double) * 2 * nz);
        }
    /* y lines */
    for (int x = 0; x < nx; x++)
        for (int z = 0; z < nz; z++) {
            for (int y = 0; y < ny; y++) {
                size_t idx = 2 * ((size_t)(x * ny + y) * nz + z);
                line[2 * y] = data[idx]; line[2 * y + 1] = data[idx + 1];
            }
            fft1d(line, ny, sign, w1, w2);
            for (int y = 0; y < ny; y++) {
                size_t idx = 2 * ((size_t)(x * ny + y) * nz + z);
                data[idx] = line[2 * y]; data[idx + 1] = line[2 * y + 1];
            }
        }
    /* x lines */
    for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++) {
            for (int x = 0; x < nx; x++) {
                size_t idx = 2 * ((size_t)(x * ny + y) * nz + z);
                line[2 * x] = data[idx]; line[2 * x + 1] = data[idx + 1];
            }
            fft1d(line, nx, sign, w1, w2);
            for (int x = 0; x < nx; x++) {
                size_t idx = 2 * ((size_t)(x * ny + y) * nz + z);
                data[idx] = line[2 * x]; data[idx + 1] = line[2 * x + 1];
            }
        }
    free(line); free(w1); free(w2);
}

void orc_dft3d_naive(const double *in, double *out, int nx, int ny, int nz, int sign)
{
    size_t ng = (size_t)nx * ny * nz;
    for (int kx = 0; kx < nx; kx++)
        for (int ky = 0; ky < ny; ky++)
            for (int kz = 0; kz < nz; kz++) {
                double sr = 0.0, si = 0.0;
                for (int x = 0; x < nx; x++)
                    for (int y = 0; y < ny; y++)
                        for (int z = 0; z < nz; z++) {
                            double ph = (double)((long)kx * x % nx) / nx +
                                        (double)((long)ky * y % ny) / ny +
                                        (double)((long)kz * z % nz) / nz;
                            double ang = sign * 2.0 * ORC_PI * ph;
                            double c = cos(ang), s = sin(ang);
                            size_t i = 2 * ((size_t)(x * ny + y) * nz + z);
                            sr += in[i] * c - in[i + 1] * s;
                            si += in[i] * s + in[i + 1] * c;
                        }
                size_t o = 2 * ((size_t)(kx * ny + ky) * nz + kz);
                out[o] = sr; out[o + 1] = si;
            }
    (void)ng;
}

/* ======================================================================== */
/* exclusions (read_topology.c:847-891 order: bonds, bends, exNB, 1-4)       */
/* ======================================================================== */

int orc_excl_build(int n, int nx, const int *xa, const int *xb, const int *xf, orc_excl *out)
{
    out->n = n;
    out->off = (int *)calloc((size_t)n + 1, sizeof(int));
    int *cnt = (int *)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < nx; i++) {
        if (xa[i] < 0 || xa[i] >= n || xb[i] < 0 || xb[i] >= n) { free(cnt); return -1; }
        cnt[xa[i]]++; cnt[xb[i]]++;
    }
    for (int i = 0; i < n; i++) out->off[i + 1] = out->off[i] + cnt[i];
    int tot = out->off[n];
    out->partner = (int *)malloc(sizeof(int) * (tot > 0 ? tot : 1));
    out->flag = (int *)malloc(sizeof(int) * (tot > 0 ? tot : 1));
    memset(cnt, 0, sizeof(int) * n);
    /* later entries overwrite earlier ones for the same pair (last wins) */
    for (int i = 0; i < nx; i++) {
        int a = xa[i], b = xb[i];
        for (int side = 0; side < 2; side++) {
            int u = side ? b : a, v = side ? a : b;
            int found = -1;
            for (int k = out->off[u]; k < out->off[u] + cnt[u]; k++)
                if (out->partner[k] == v) found = k;
            if (found >= 0) { out->flag[found] = xf[i]; }
            else {
                int k = out->off[u] + cnt[u]++;
                out->partner[k] = v; out->flag[k] = xf[i];
            }
        }
    }
    /* compact rows (duplicates leave holes) */
    int w = 0;
    int *noff = (int *)calloc((size_t)n + 1, sizeof(int));
    for (int u = 0; u < n; u++) {
        noff[u] = w;
        for (int k = out->off[u]; k < out->off[u] + cnt[u]; k++) {
            out->partner[w] = out->partner[k]; out->flag[w] = out->flag[k]; w++;
        }
    }
    noff[n] = w;
    free(out->off);
    out->off = noff;
    free(cnt);
    return 0;
}

void orc_excl_free(orc_excl *e)
{
    free(e->off); free(e->partner); free(e->flag);
    e->off = NULL; e->partner = NULL; e->flag = NULL;
}

int orc_excl_flag(const orc_excl *e, int a, int b)
{
    if (!e || !e->off) return 0;
    for (int k = e->off[a]; k < e->off[a + 1]; k++)
        if (e->partner[k] == b) return e->flag[k];
    return 0;
}

/* ======================================================================== */
/* nblist.c                                                                  */
/* ======================================================================== */

/* nblist.c:95-307 (non-CLIST).  Half list b>a, FP64 test with one-sided
 * wrap of |d| and per-component early rejects; with use_ew the list at
 * rclist_ew is built and the rclist list nests inside it (nblist.c:145-224).
 * Returns 5 (reference exit code) on row overflow, else 0.                  */
int orc_nblist(const orc_system *s, const double *pos, const orc_excl *ex,
               int use_ew, int max_nb, int *count, int *list, int *xflag,
               int *count_ew, int *list_ew, int *xflag_ew)
{
    double lx = s->lx, ly = s->ly, lz = s->lz;
    double hx = 0.5 * lx, hy = 0.5 * ly, hz = 0.5 * lz;
    double rcut = s->rclist, rcut2 = s->rclist * s->rclist;
    double rcut_ew = s->rclist_ew, rcut2_ew = rcut_ew * rcut_ew;
    int ns = s->n;
    for (int i = 0; i < ns; i++) { count[i] = 0; if (use_ew) count_ew[i] = 0; }
    for (int a = 0; a < ns - 1; a++) {
        double ax = pos[3 * a], ay = pos[3 * a + 1], az = pos[3 * a + 2];
        for (int b = a + 1; b < ns; b++) {
            double dx = fabs(ax - pos[3 * b]);
            if (dx > hx) dx -= lx;
            if (dx > rcut) continue;
            double dy = fabs(ay - pos[3 * b + 1]);
            if (dy > hy) dy -= ly;
            if (dy > rcut) continue;
            double dz = fabs(az - pos[3 * b + 2]);
            if (dz > hz) dz -= lz;
            if (dz > rcut) continue;
            double dr2 = dx * dx + dy * dy + dz * dz;
            if (use_ew) {
                if (dr2 < rcut2_ew) {
                    int f = orc_excl_flag(ex, a, b);
                    if (dr2 < rcut2) {
                        if (count[a] >= max_nb) return 5;
                        list[(size_t)a * max_nb + count[a]] = b;
                        xflag[(size_t)a * max_nb + count[a]] = f;
                        count[a]++;
                    }
                    if (count_ew[a] >= max_nb) return 5;
                    list_ew[(size_t)a * max_nb + count_ew[a]] = b;
                    xflag_ew[(size_t)a * max_nb + count_ew[a]] = f;
                    count_ew[a]++;
                }
            } else {
                if (dr2 < rcut2) {
                    if (count[a] >= max_nb) return 5;
                    list[(size_t)a * max_nb + count[a]] = b;
                    xflag[(size_t)a * max_nb + count[a]] = orc_excl_flag(ex, a, b);
                    count[a]++;
                }
            }
        }
    }
    return 0;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* nblist.c:327-557 (CLIST): cells of at least rclist, same FP64 test, so the
 * same pair set; rows are sorted ascending afterwards (the reference's CLIST
 * row order follows its cell traversal).                                   */
int orc_nblist_cells(const orc_system *s, const double *pos, const orc_excl *ex,
                     int max_nb, int *count, int *list, int *xflag)
{
    double lx = s->lx, ly = s->ly, lz = s->lz;
    double hx = 0.5 * lx, hy = 0.5 * ly, hz = 0.5 * lz;
    double rcut = s->rclist, rcut2 = rcut * rcut;
    int ns = s->n;
    int mx = (int)(lx / rcut), my = (int)(ly / rcut), mz = (int)(lz / rcut);
    if (mx < 3 || my < 3 || mz < 3) {
        /* too few cells for a 27-cell stencil: fall back to all pairs */
        int *ce = (int *)malloc(sizeof(int) * ns);
        int *le = (int *)malloc(sizeof(int) * (size_t)ns * max_nb);
        int *xe = (int *)malloc(sizeof(int) * (size_t)ns * max_nb);
        orc_system t = *s;
        t.rclist_ew = t.rclist;
        int rc_ = orc_nblist(&t, pos, ex, 0, max_nb, count, list, xflag, ce, le, xe);
        free(ce); free(le); free(xe);
        return rc_;
    }
    int ncell = mx * my * mz;
    int *head = (int *)malloc(sizeof(int) * ncell);
    int *next = (int *)malloc(sizeof(int) * ns);
    int *cell = (int *)malloc(sizeof(int) * ns);
    for (int c = 0; c < ncell; c++) head[c] = -1;
    for (int i = ns - 1; i >= 0; i--) {
        int cx = (int)((pos[3 * i] + hx) * mx / lx);
        int cy = (int)((pos[3 * i + 1] + hy) * my / ly);
        int cz = (int)((pos[3 * i + 2] + hz) * mz / lz);
        if (cx >= mx) cx = mx - 1;
        if (cy >= my) cy = my - 1;
        if (cz >= mz) cz = mz - 1;
        if (cx < 0) cx = 0;
        if (cy < 0) cy = 0;
        if (cz < 0) cz = 0;
        int c = cx + mx * (cy + my * cz);
        cell[i] = c; next[i] = head[c]; head[c] = i;
    }
    for (int i = 0; i < ns; i++) count[i] = 0;
    int status = 0;
    for (int a = 0; a < ns - 1 && !status; a++) {
        double ax = pos[3 * a], ay = pos[3 * a + 1], az = pos[3 * a + 2];
        int c = cell[a];
        int cx = c % mx, cy = (c / mx) % my, cz = c / (mx * my);
        for (int ox = -1; ox <= 1; ox++)
            for (int oy = -1; oy <= 1; oy++)
                for (int oz = -1; oz <= 1; oz++) {
                    int nx_ = (cx + ox + mx) % mx, ny_ = (cy + oy + my) % my, nz_ = (cz + oz + mz) % mz;
                    int nc = nx_ + mx * (ny_ + my * nz_);
                    for (int b = head[nc]; b != -1; b = next[b]) {
                        if (b <= a) continue;
                        double dx = fabs(ax - pos[3 * b]);
                        if (dx > hx) dx -= lx;
                        if (dx > rcut) continue;
                        double dy = fabs(ay - pos[3 * b + 1]);
                        if (dy > hy) dy -= ly;
                        if (dy > rcut) continue;
                        double dz = fabs(az - pos[3 * b + 2]);
                        if (dz > hz) dz -= lz;
                        if (dz > rcut) continue;
                        double dr2 = dx * dx + dy * dy + dz * dz;
                        if (dr2 < rcut2) {
                            if (count[a] >= max_nb) { status = 5; break; }
                            list[(size_t)a * max_nb + count[a]] = b;
                            count[a]++;
                        }
                    }
                }
        if (!status) {
            qsort(list + (size_t)a * max_nb, count[a], sizeof(int), cmp_int);
            for (int k = 0; k < count[a]; k++)
                xflag[(size_t)a * max_nb + k] = orc_excl_flag(ex, a, list[(size_t)a * max_nb + k]);
        }
    }
    free(head); free(next); free(cell);
    return status;
}

/* nblist.c:607-647: 1 if any site in [lb,ub) moved more than
 * (rclist-rc)/4 from its anchor (component-wise or radially).             */
int orc_nl_check(const orc_system *s, const double *pos, const double *anchor, int lb, int ub)
{
    double lx = s->lx, ly = s->ly, lz = s->lz;
    double hx = 0.5 * lx, hy = 0.5 * ly, hz = 0.5 * lz;
    double rcut = (s->rclist - s->rc) / 4.0, rcut2 = rcut * rcut;
    for (int i = lb; i < ub; i++) {
        double dx = fabs(pos[3 * i] - anchor[3 * i]);
        if (dx > hx) dx -= lx;
        if (dx > rcut) return 1;
        double dy = fabs(pos[3 * i + 1] - anchor[3 * i + 1]);
        if (dy > hy) dy -= ly;
        if (dy > rcut) return 1;
        double dz = fabs(pos[3 * i + 2] - anchor[3 * i + 2]);
        if (dz > hz) dz -= lz;
        if (dz > rcut) return 1;
        if (dx * dx + dy * dy + dz * dz > rcut2) return 1;
    }
    return 0;
}

/* ======================================================================== */
/* cnonbond_nblist.c (STYPE + S14 + SPME branch, CHARMM LJ form)             */
/* ======================================================================== */

void orc_pair(const orc_system *s, const double *pos, const int *count,
              const int *list, const int *xflag, int max_nb, double *ff, orc_pair_out *out)
{
    double kappa = s->kappa;
    double kpi = 1.0 * kappa / sqrt(ORC_PI);
    double rc = s->rc, rc2 = rc * rc;
    double lx = s->lx, ly = s->ly, lz = s->lz;
    double hx = 0.5 * lx, hy = 0.5 * ly, hz = 0.5 * lz;
    double e14 = s->e14fac;
    double enb = 0.0, enbs = 0.0, er = 0.0, ers = 0.0, ei = 0.0;
    double wxx = 0, wyx = 0, wyy = 0, wzx = 0, wzy = 0, wzz = 0;
    double hx2 = 0.0, hy2 = 0.0, hz2 = 0.0;   /* cnonbond_nblist2.c config[k].hes* */
    int nt = s->ntypes;
    for (int a = 0; a < s->n - 1; a++) {
        double fxa = 0.0, fya = 0.0, fza = 0.0;
        double ax = pos[3 * a], ay = pos[3 * a + 1], az = pos[3 * a + 2];
        for (int bb = 0; bb < count[a]; bb++) {
            int b = list[(size_t)a * max_nb + bb];
            double dx = ax - pos[3 * b], dy = ay - pos[3 * b + 1], dz = az - pos[3 * b + 2];
            if (dx > hx) dx -= lx; else if (dx < -hx) dx += lx;
            if (dy > hy) dy -= ly; else if (dy < -hy) dy += ly;
            if (dz > hz) dz -= lz; else if (dz < -hz) dz += lz;
            double dr2 = dx * dx + dy * dy + dz * dz;
            if (dr2 > rc2) continue;
            int ta = s->type[a], tb = s->type[b];
            int xf = xflag[(size_t)a * max_nb + bb];
            double sig, eps;
            if (xf == 4) { sig = s->lj14_sig[ta * nt + tb]; eps = s->lj14_eps[ta * nt + tb]; }
            else         { sig = s->lj_sig[ta * nt + tb];   eps = s->lj_eps[ta * nt + tb]; }
            double qq = s->q[a] * s->q[b] * ORC_QFACTOR;
            double qqab = qq;
            if (xf) {
                if (xf != 4) { sig = 0.0; eps = 0.0; qq = 0.0; }
                else qq *= e14;
            }
            double force = 0.0;
            if ((eps * sig) != 0.0) {
                double sr2 = (sig * sig) / dr2;
                double sr6 = sr2 * sr2 * sr2;
                double sr12 = sr6 * sr6;
                double sr2s = (sig * sig) / rc2;
                double sr6s = sr2s * sr2s * sr2s;
                double sr12s = sr6s * sr6s;
                force = (12.0 * eps / dr2) * (sr12 - sr6);
                enb += eps * (sr12 - 2.0 * sr6);
                enbs += eps * (sr12s - 2.0 * sr6s);
                /* cnonbond_nblist2.c:286-297 */
                double term2 = (12.0 * eps / dr2) / dr2 * (14.0 * sr12 - 8.0 * sr6);
                hx2 += 2.0 * (term2 * dx * dx - force);
                hy2 += 2.0 * (term2 * dy * dy - force);
                hz2 += 2.0 * (term2 * dz * dz - force);
            }
            double dr = sqrt(dr2);
            double dri = 1.0 / dr;
            double kr = kappa * dr;
            double erfckr = orc_erfc(kr);
            double expkr2 = exp(-kr * kr);
            double kpidrexpkr2 = 2.0 * kpi * dr * expkr2;
            double dri2 = dri * dri, dri3 = dri2 * dri, dri4 = dri2 * dri2, dri5 = dri3 * dri2;
            qq /= s->eps_rf;
            if (qq != 0.0) {
                er += qq * erfckr * dri;
                ers += qq * orc_erfc(kappa * rc) / rc;
                force += (qq * dri / dr2) * (erfckr + kpidrexpkr2);
                /* cnonbond_nblist2.c:403-410 */
                double hes1 = erfckr * dri3 + 2.0 * kpi * dri2 * expkr2;
                double hes2 = -6.0 * kpi * dri4 * expkr2 - 4.0 * kpi * kappa * kappa * dri2 * expkr2 - 3.0 * dri5 * erfckr;
                hx2 += -2.0 * qq * (hes1 + dx * dx * hes2);
                hy2 += -2.0 * qq * (hes1 + dy * dy * hes2);
                hz2 += -2.0 * qq * (hes1 + dz * dz * hes2);
            }
            if (xf) {
                if (xf == 4) qqab *= (1.0 - e14);
                ei += qqab * (1.0 - erfckr) * dri;
                force -= (qqab * dri / dr2) * ((1.0 - erfckr) - kpidrexpkr2);
                /* cnonbond_nblist2.c:433-440 */
                double hes1 = (1.0 - erfckr) * dri3 - 2.0 * kpi * dri2 * expkr2;
                double hes2 = -6.0 * kpi * dri4 * expkr2 - 4.0 * kpi * kappa * kappa * dri2 * expkr2 + 3.0 * dri5 * (1.0 - erfckr);
                hx2 += 2.0 * qqab * (hes1 - dx * dx * hes2);
                hy2 += 2.0 * qqab * (hes1 - dy * dy * hes2);
                hz2 += 2.0 * qqab * (hes1 - dz * dz * hes2);
            }
            if (force != 0.0) {
                double fx = force * dx, fy = force * dy, fz = force * dz;
                fxa += fx; ff[3 * b] -= fx;
                fya += fy; ff[3 * b + 1] -= fy;
                fza += fz; ff[3 * b + 2] -= fz;
                wxx += fx * dx; wyx += fy * dx; wyy += fy * dy;
                wzx += fz * dx; wzy += fz * dy; wzz += fz * dz;
            }
        }
        ff[3 * a] += fxa; ff[3 * a + 1] += fya; ff[3 * a + 2] += fza;
    }
    out->enonbond = enb; out->enonbonds = enbs;
    out->eewald_real = er; out->eewald_reals = ers; out->eewald_intra = ei;
    out->vir[0] = wxx; out->vir[1] = wyx; out->vir[2] = wyy;
    out->vir[3] = wzx; out->vir[4] = wzy; out->vir[5] = wzz;
    out->hes[0] = hx2; out->hes[1] = hy2; out->hes[2] = hz2;
}

/* ======================================================================== */
/* pme_qgrid.c / pme_calc.c                                                  */
/* ======================================================================== */

#define IDX3(x, y, z) ((size_t)(z) + (size_t)gz * ((size_t)(y) + (size_t)gy * (size_t)(x)))

void orc_pme_calc(const orc_system *s, const double *pos, double e_self,
                  double *ff, double energy[3], double vir[6], double *qgrid_out)
{
    int order = s->order, gx = s->gx, gy = s->gy, gz = s->gz;
    double blx = s->lx, bly = s->ly, blz = s->lz;
    double vol = blx * bly * blz;
    double kap = s->kappa;
    double fac = ORC_PI * ORC_PI / kap / kap;
    int n = s->n;
    size_t ngrid = (size_t)gx * gy * gz;
    double *ss = (double *)malloc(sizeof(double) * 3 * n);
    double *mm = (double *)malloc(sizeof(double) * 3 * n * (order + 1));
    double *dm = (double *)malloc(sizeof(double) * 3 * n * (order + 1));
    double *q = (double *)calloc(2 * ngrid, sizeof(double));
    double *bx = (double *)calloc((size_t)gx + 2, sizeof(double));
    double *by = (double *)calloc((size_t)gy + 2, sizeof(double));
    double *bz = (double *)calloc((size_t)gz + 2, sizeof(double));
    double mn[64], dmn[64];
    orc_pme_bspline(order, gx, gy, gz, bx, by, bz);

    /* pme_qgrid.c:26-34 scaled coordinates */
    for (int i = 0; i < n; i++) {
        ss[3 * i] = gx * (pos[3 * i] / blx + 0.5);
        ss[3 * i + 1] = gy * (pos[3 * i + 1] / bly + 0.5);
        ss[3 * i + 2] = gz * (pos[3 * i + 2] / blz + 0.5);
    }
    /* pme_qgrid.c:43-70 splines per dimension */
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double u = ss[3 * i + d];
            double w = u - (int)u;
            orc_fillspline(w, order, mn, dmn);
            for (int k = 1; k <= order; k++) {
                mm[((size_t)i * (order + 1) + k) * 3 + d] = mn[k];
                dm[((size_t)i * (order + 1) + k) * 3 + d] = dmn[k];
            }
        }
    /* pme_qgrid.c:77-113 spreading: indices (int)u-order+1 .. (int)u, negatives
     * wrapped.  An index equal to K (site exactly at +L/2) is wrapped as well;
     * the reference would address one element past the row there.           */
    for (int i = 0; i < n; i++) {
        double qq = s->q[i];
        int xo = (int)ss[3 * i] - order, yo0 = (int)ss[3 * i + 1] - order, zo0 = (int)ss[3 * i + 2] - order;
        for (int xi = 1; xi <= order; xi++) {
            xo++;
            int x = xo + (xo < 0 ? gx : 0);
            if (x >= gx) x -= gx;
            double mxi = mm[((size_t)i * (order + 1) + xi) * 3 + 0];
            int yo = yo0;
            for (int yi = 1; yi <= order; yi++) {
                yo++;
                int y = yo + (yo < 0 ? gy : 0);
                if (y >= gy) y -= gy;
                double myi = mm[((size_t)i * (order + 1) + yi) * 3 + 1];
                int zo = zo0;
                for (int zi = 1; zi <= order; zi++) {
                    zo++;
                    int z = zo + (zo < 0 ? gz : 0);
                    if (z >= gz) z -= gz;
                    double mzi = mm[((size_t)i * (order + 1) + zi) * 3 + 2];
                    q[2 * IDX3(x, y, z)] += mxi * myi * mzi * qq;
                }
            }
        }
    }
    if (qgrid_out) memcpy(qgrid_out, q, sizeof(double) * 2 * ngrid);

    /* pme_calc.c:50 backward FFT, 52-93 convolution / energy / virial */
    orc_fft3d(q, gx, gy, gz, +1);
    double enrec = 0.0, wxx = 0, wyx = 0, wyy = 0, wzx = 0, wzy = 0, wzz = 0;
    size_t ggyz = (size_t)gy * gz;
    for (size_t ind = 1; ind < ngrid; ind++) {
        int k1 = (int)(ind / ggyz);
        size_t knd = ind - (size_t)k1 * ggyz;
        int k2 = (int)(knd / gz);
        int k3 = (int)(knd - (size_t)k2 * gz);
        int mx = k1 - (k1 > gx / 2 ? gx : 0);
        int my = k2 - (k2 > gy / 2 ? gy : 0);
        int mz = k3 - (k3 > gz / 2 ? gz : 0);
        double mhx = mx / blx, mhy = my / bly, mhz = mz / blz;
        double m2 = mhx * mhx + mhy * mhy + mhz * mhz;
        double bb = bx[k1 + 1] * by[k2 + 1] * bz[k3 + 1];
        double den = ORC_PI * vol * m2 * bb;
        double eterm = exp(-fac * m2) / den;
        double wterm = -2.0 * (fac * m2 + 1.0) / m2;
        double qr = q[2 * ind], qi = q[2 * ind + 1];
        double st2 = qr * qr + qi * qi;
        enrec += eterm * st2;
        wxx += eterm * st2 * (wterm * mhx * mhx + 1.0);
        wyx += eterm * st2 * (wterm * mhy * mhx);
        wyy += eterm * st2 * (wterm * mhy * mhy + 1.0);
        wzx += eterm * st2 * (wterm * mhz * mhx);
        wzy += eterm * st2 * (wterm * mhz * mhy);
        wzz += eterm * st2 * (wterm * mhz * mhz + 1.0);
        q[2 * ind] *= eterm;
        q[2 * ind + 1] *= eterm;
    }
    double fan = 0.5 * ORC_QFACTOR / s->eps_rf;
    energy[0] = enrec * fan;
    energy[1] = e_self;
    energy[2] = energy[0] - energy[1];
    vir[0] = wxx * fan; vir[1] = wyx * fan; vir[2] = wyy * fan;
    vir[3] = wzx * fan; vir[4] = wzy * fan; vir[5] = wzz * fan;

    /* pme_calc.c:113 forward FFT, 122-279 force gather + mean-force removal */
    orc_fft3d(q, gx, gy, gz, -1);
    double sumx = 0.0, sumy = 0.0, sumz = 0.0;
    for (int i = 0; i < n; i++) {
        double fxi = 0.0, fyi = 0.0, fzi = 0.0;
        int xo = (int)ss[3 * i] - order, yo0 = (int)ss[3 * i + 1] - order, zo0 = (int)ss[3 * i + 2] - order;
        for (int xi = 1; xi <= order; xi++) {
            xo++;
            int x = xo + (xo < 0 ? gx : 0);
            if (x >= gx) x -= gx;
            double mxi = mm[((size_t)i * (order + 1) + xi) * 3 + 0];
            double dmxi = dm[((size_t)i * (order + 1) + xi) * 3 + 0];
            int yo = yo0;
            for (int yi = 1; yi <= order; yi++) {
                yo++;
                int y = yo + (yo < 0 ? gy : 0);
                if (y >= gy) y -= gy;
                double myi = mm[((size_t)i * (order + 1) + yi) * 3 + 1];
                double dmyi = dm[((size_t)i * (order + 1) + yi) * 3 + 1];
                int zo = zo0;
                for (int zi = 1; zi <= order; zi++) {
                    zo++;
                    int z = zo + (zo < 0 ? gz : 0);
                    if (z >= gz) z -= gz;
                    double mzi = mm[((size_t)i * (order + 1) + zi) * 3 + 2];
                    double dmzi = dm[((size_t)i * (order + 1) + zi) * 3 + 2];
                    double term = q[2 * IDX3(x, y, z)];
                    fxi -= dmxi * myi * mzi * term;
                    fyi -= mxi * dmyi * mzi * term;
                    fzi -= mxi * myi * dmzi * term;
                }
            }
        }
        double f = s->q[i] * ORC_QFACTOR / s->eps_rf;
        fxi *= f * gx / blx;
        fyi *= f * gy / bly;
        fzi *= f * gz / blz;
        ff[3 * i] += fxi; ff[3 * i + 1] += fyi; ff[3 * i + 2] += fzi;
        sumx += fxi; sumy += fyi; sumz += fzi;
    }
    sumx /= (double)n; sumy /= (double)n; sumz /= (double)n;
    for (int i = 0; i < n; i++) { ff[3 * i] -= sumx; ff[3 * i + 1] -= sumy; ff[3 * i + 2] -= sumz; }

    free(ss); free(mm); free(dm); free(q); free(bx); free(by); free(bz);
}

/* cnonbond_nblist.c:51-622, energy slots of 613-620 */
void orc_cnonbondnl(const orc_system *s, const double *pos, const int *count,
                    const int *list, const int *xflag, int max_nb, double e_self,
                    double *ff, double energy[8], double vir_nbond[6], double vir_kspace[6])
{
    orc_pair_out po;
    orc_pair(s, pos, count, list, xflag, max_nb, ff, &po);
    double enk[3];
    orc_pme_calc(s, pos, e_self, ff, enk, vir_kspace, NULL);
    double enr0 = po.eewald_real, enr1 = po.eewald_real - po.eewald_reals, enr2 = po.eewald_intra;
    energy[0] = po.enonbond + enr0 - enr2 + enk[2];
    energy[1] = energy[0] - po.enonbonds;
    energy[2] = enr0 + enk[2] - enr2;
    energy[3] = enr0 - enr1;
    energy[4] = enk[0];
    energy[5] = enk[1];
    energy[6] = enr2;
    energy[7] = enr0;
    for (int i = 0; i < 6; i++) vir_nbond[i] = po.vir[i];
}

void orc_fillspline2(double w, int order, double *array, double *darray, double *d2array)
{
    double a2[64], da2[64];
    orc_fillspline(w, order, array, darray);
    for (int k = 0; k <= order; k++) d2array[k] = 0.0;
    if (order < 3) return;
    if (order - 2 == 2) { a2[1] = 1.0 - w; a2[2] = w; }   /* M2: the recursion's start */
    else if (order - 2 == 1) { a2[1] = 1.0; }
    else orc_fillspline(w, order - 2, a2, da2);
    for (int k = 1; k <= order; k++) {
        double am2 = (k - 2 >= 1 && k - 2 <= order - 2) ? a2[k - 2] : 0.0;
        double am1 = (k - 1 >= 1 && k - 1 <= order - 2) ? a2[k - 1] : 0.0;
        double a0 = (k <= order - 2) ? a2[k] : 0.0;
        d2array[k] = am2 - 2.0 * am1 + a0;
    }
}

/* theta(m) on the grid: the convolution kernel of the reciprocal energy,
 * obtained by running the convolution of orc_pme_calc on a unit charge at
 * grid point 0 (so normalisation and the excluded m = 0 term match). */
static void pme_theta(const orc_system *s, double *theta)
{
    int gx = s->gx, gy = s->gy, gz = s->gz;
    double blx = s->lx, bly = s->ly, blz = s->lz, vol = blx * bly * blz;
    double fac = ORC_PI * ORC_PI / s->kappa / s->kappa;
    size_t ngrid = (size_t)gx * gy * gz, ggyz = (size_t)gy * gz;
    double *q = (double *)calloc(2 * ngrid, sizeof(double));
    double *bx = (double *)calloc((size_t)gx + 2, sizeof(double));
    double *by = (double *)calloc((size_t)gy + 2, sizeof(double));
    double *bz = (double *)calloc((size_t)gz + 2, sizeof(double));
    orc_pme_bspline(s->order, gx, gy, gz, bx, by, bz);
    q[0] = 1.0;
    orc_fft3d(q, gx, gy, gz, +1);
    q[0] = q[1] = 0.0;
    for (size_t ind = 1; ind < ngrid; ind++) {
        int k1 = (int)(ind / ggyz);
        size_t knd = ind - (size_t)k1 * ggyz;
        int k2 = (int)(knd / gz);
        int k3 = (int)(knd - (size_t)k2 * gz);
        int mx = k1 - (k1 > gx / 2 ? gx : 0), my = k2 - (k2 > gy / 2 ? gy : 0), mz = k3 - (k3 > gz / 2 ? gz : 0);
        double mhx = mx / blx, mhy = my / bly, mhz = mz / blz;
        double m2 = mhx * mhx + mhy * mhy + mhz * mhz;
        double eterm = exp(-fac * m2) / (ORC_PI * vol * m2 * bx[k1 + 1] * by[k2 + 1] * bz[k3 + 1]);
        q[2 * ind] *= eterm;
        q[2 * ind + 1] *= eterm;
    }
    orc_fft3d(q, gx, gy, gz, -1);
    for (size_t i = 0; i < ngrid; i++) theta[i] = q[2 * i];
    free(q); free(bx); free(by); free(bz);
}

void orc_pme_hessian(const orc_system *s, const double *pos, double hes3[3], double *per_site)
{
    int order = s->order, gx = s->gx, gy = s->gy, gz = s->gz, n = s->n;
    double blx = s->lx, bly = s->ly, blz = s->lz;
    size_t ngrid = (size_t)gx * gy * gz;
    double C = ORC_QFACTOR / s->eps_rf;
    double *theta = (double *)malloc(sizeof(double) * ngrid);
    double *conv = (double *)malloc(sizeof(double) * ngrid);
    pme_theta(s, theta);
    /* (theta*Q)(m): the convolved grid of orc_pme_calc, taken from the
     * forces' gather input by recomputing the spread + convolution */
    {
        double *q = (double *)calloc(2 * ngrid, sizeof(double));
        double mn[64], dmn[64];
        for (int i = 0; i < n; i++) {
            double u[3] = {gx * (pos[3 * i] / blx + 0.5), gy * (pos[3 * i + 1] / bly + 0.5),
                           gz * (pos[3 * i + 2] / blz + 0.5)};
            double m[3][64];
            for (int d = 0; d < 3; d++) {
                orc_fillspline(u[d] - (int)u[d], order, mn, dmn);
                for (int k = 1; k <= order; k++) m[d][k] = mn[k];
            }
            for (int a = 1; a <= order; a++)
                for (int b = 1; b <= order; b++)
                    for (int c = 1; c <= order; c++) {
                        int x = (int)u[0] - order + a, y = (int)u[1] - order + b, z = (int)u[2] - order + c;
                        x = ((x % gx) + gx) % gx; y = ((y % gy) + gy) % gy; z = ((z % gz) + gz) % gz;
                        q[2 * IDX3(x, y, z)] += m[0][a] * m[1][b] * m[2][c] * s->q[i];
                    }
        }
        /* conv = theta * Q by direct periodic convolution through the FFT
         * of both (theta is real and even) */
        double *t = (double *)calloc(2 * ngrid, sizeof(double));
        for (size_t i = 0; i < ngrid; i++) t[2 * i] = theta[i];
        orc_fft3d(q, gx, gy, gz, +1);
        orc_fft3d(t, gx, gy, gz, +1);
        for (size_t i = 0; i < ngrid; i++) {
            double ar = q[2 * i], ai = q[2 * i + 1], br = t[2 * i], bi = t[2 * i + 1];
            q[2 * i] = ar * br - ai * bi;
            q[2 * i + 1] = ar * bi + ai * br;
        }
        orc_fft3d(q, gx, gy, gz, -1);
        for (size_t i = 0; i < ngrid; i++) conv[i] = q[2 * i] / (double)ngrid;
        free(q); free(t);
    }
    double kl[3] = {gx / blx, gy / bly, gz / blz};
    int K[3] = {gx, gy, gz};
    hes3[0] = hes3[1] = hes3[2] = 0.0;
    for (int i = 0; i < n; i++) {
        double u[3] = {gx * (pos[3 * i] / blx + 0.5), gy * (pos[3 * i + 1] / bly + 0.5),
                       gz * (pos[3 * i + 2] / blz + 0.5)};
        double M[3][64], D[3][64], D2[3][64];
        int base[3];
        for (int d = 0; d < 3; d++) {
            orc_fillspline2(u[d] - (int)u[d], order, M[d], D[d], D2[d]);
            base[d] = (int)u[d] - order;
        }
        double hA[3] = {0, 0, 0}, hB[3] = {0, 0, 0};
        for (int a = 1; a <= order; a++)
            for (int b = 1; b <= order; b++)
                for (int c = 1; c <= order; c++) {
                    int x = ((base[0] + a) % gx + gx) % gx, y = ((base[1] + b) % gy + gy) % gy,
                        z = ((base[2] + c) % gz + gz) % gz;
                    double g = conv[IDX3(x, y, z)];
                    hA[0] += D2[0][a] * M[1][b] * M[2][c] * g;
                    hA[1] += M[0][a] * D2[1][b] * M[2][c] * g;
                    hA[2] += M[0][a] * M[1][b] * D2[2][c] * g;
                    for (int a2 = 1; a2 <= order; a2++)
                        for (int b2 = 1; b2 <= order; b2++)
                            for (int c2 = 1; c2 <= order; c2++) {
                                int dx = ((a - a2) % K[0] + K[0]) % K[0], dy = ((b - b2) % K[1] + K[1]) % K[1],
                                    dz = ((c - c2) % K[2] + K[2]) % K[2];
                                double th = theta[IDX3(dx, dy, dz)];
                                hB[0] += D[0][a] * M[1][b] * M[2][c] * th * D[0][a2] * M[1][b2] * M[2][c2];
                                hB[1] += M[0][a] * D[1][b] * M[2][c] * th * M[0][a2] * D[1][b2] * M[2][c2];
                                hB[2] += M[0][a] * M[1][b] * D[2][c] * th * M[0][a2] * M[1][b2] * D[2][c2];
                            }
                }
        for (int d = 0; d < 3; d++) {
            double h = C * s->q[i] * kl[d] * kl[d] * (hA[d] + s->q[i] * hB[d]);
            hes3[d] += h;
            if (per_site) per_site[3 * i + d] = h;
        }
    }
    free(theta); free(conv);
}

/* cnonbond_nblist2.c:47-622 */
void orc_cnonbondnl2(const orc_system *s, const double *pos, const int *count,
                     const int *list, const int *xflag, int max_nb, double e_self,
                     double *ff, double energy[8], double hes3[3])
{
    orc_pair_out po;
    orc_pair(s, pos, count, list, xflag, max_nb, ff, &po);
    double enk[3], vk[6], hk[3];
    orc_pme_calc(s, pos, e_self, ff, enk, vk, NULL);
    orc_pme_hessian(s, pos, hk, NULL);
    double enr0 = po.eewald_real, enr1 = po.eewald_real - po.eewald_reals, enr2 = po.eewald_intra;
    energy[0] = po.enonbond + enr0 - enr2 + enk[2];
    energy[1] = energy[0] - po.enonbonds;
    energy[2] = enr0 + enk[2] - enr2;
    energy[3] = enr0 - enr1;
    energy[4] = enk[0];
    energy[5] = enk[1];
    energy[6] = enr2;
    energy[7] = enr0;
    for (int d = 0; d < 3; d++) hes3[d] = po.hes[d] + hk[d];
}

/* cbond2.c:65-130 */
double orc_cbond2(const orc_system *s, const double *pos, int nb, const int *ia,
                  const int *ib, const double *req, const double *kb, double *ff, double hes3[3])
{
    double vir[6];
    double e = orc_cbond(s, pos, nb, ia, ib, req, kb, ff, vir);
    double lx = s->lx, ly = s->ly, lz = s->lz;
    double hx = 0.5 * lx, hy = 0.5 * ly, hz = 0.5 * lz;
    hes3[0] = hes3[1] = hes3[2] = 0.0;
    for (int i = 0; i < nb; i++) {
        int a = ia[i], b = ib[i];
        double dx = pos[3 * b] - pos[3 * a], dy = pos[3 * b + 1] - pos[3 * a + 1], dz = pos[3 * b + 2] - pos[3 * a + 2];
        if (dx > hx) dx -= lx; else if (dx < -hx) dx += lx;
        if (dy > hy) dy -= ly; else if (dy < -hy) dy += ly;
        if (dz > hz) dz -= lz; else if (dz < -hz) dz += lz;
        double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
        double dt = r - req[i];
        double de = 2.0 * kb[i] * dt / r;
        double term = (2.0 * kb[i] - de) / r2;
        hes3[0] += 2.0 * (term * dx * dx + de);
        hes3[1] += 2.0 * (term * dy * dy + de);
        hes3[2] += 2.0 * (term * dz * dz + de);
    }
    return e;
}

/* ======================================================================== */
/* cbond.c, pbc_all.c, integrate_vverlet.c, kinet.c, integrate_nhc.c         */
/* ======================================================================== */

double orc_cbond(const orc_system *s, const double *pos, int nb, const int *ia,
                 const int *ib, const double *req, const double *kb, double *ff, double vir[6])
{
    double lx = s->lx, ly = s->ly, lz = s->lz;
    double hx = 0.5 * lx, hy = 0.5 * ly, hz = 0.5 * lz;
    double e = 0.0;
    for (int i = 0; i < 6; i++) vir[i] = 0.0;
    for (int i = 0; i < nb; i++) {
        int a = ia[i], b = ib[i];
        double dx = pos[3 * b] - pos[3 * a], dy = pos[3 * b + 1] - pos[3 * a + 1], dz = pos[3 * b + 2] - pos[3 * a + 2];
        if (dx > hx) dx -= lx; else if (dx < -hx) dx += lx;
        if (dy > hy) dy -= ly; else if (dy < -hy) dy += ly;
        if (dz > hz) dz -= lz; else if (dz < -hz) dz += lz;
        double drr = dx * dx + dy * dy + dz * dz;
        double dr = sqrt(drr), dri = 1.0 / dr;
        double rr = dr - req[i];
        double fr = -kb[i] * rr * 2.0;
        e += -(fr * rr * 0.5);
        double fx = fr * dri * dx, fy = fr * dri * dy, fz = fr * dri * dz;
        ff[3 * a] -= fx; ff[3 * a + 1] -= fy; ff[3 * a + 2] -= fz;
        ff[3 * b] += fx; ff[3 * b + 1] += fy; ff[3 * b + 2] += fz;
        vir[0] += fx * dx; vir[1] += fy * dx; vir[2] += fy * dy;
        vir[3] += fz * dx; vir[4] += fz * dy; vir[5] += fz * dz;
    }
    return e;
}

/* ======================================================================== */
/* cbend.c, curey.c, ctorsion.c, cimproper.c                                 */
/* ======================================================================== */

/* the reference's minimum image of one component (cbend.c:85-98 form) */
static double orc_mi(double d, double h, double l)
{
    if (d > h) d -= l;
    else if (d < -h) d += l;
    return d;
}

/* cbend.c:63-170: E = kb (phi - phi0)^2 for the angle phi at b between the
 * bonds b->a and b->c; forces added, virial of the bends in vir[6]. */
double orc_cbend(const orc_system *s, const double *pos, int nbend, const int *ia, const int *ib,
                 const int *ic, const double *angeq, const double *kb, double *ff, double vir[6])
{
    double lx = s->lx, ly = s->ly, lz = s->lz, hx = 0.5 * lx, hy = 0.5 * ly, hz = 0.5 * lz;
    double e = 0.0;
    for (int q = 0; q < 6; q++) vir[q] = 0.0;
    for (int i = 0; i < nbend; i++) {
        int a = ia[i], b = ib[i], c = ic[i];
        double ax = orc_mi(pos[3 * a] - pos[3 * b], hx, lx), ay = orc_mi(pos[3 * a + 1] - pos[3 * b + 1], hy, ly),
               az = orc_mi(pos[3 * a + 2] - pos[3 * b + 2], hz, lz);
        double cx = orc_mi(pos[3 * c] - pos[3 * b], hx, lx), cy = orc_mi(pos[3 * c + 1] - pos[3 * b + 1], hy, ly),
               cz = orc_mi(pos[3 * c + 2] - pos[3 * b + 2], hz, lz);
        double ia_ = 1.0 / sqrt(ax * ax + ay * ay + az * az);
        double uax = ax * ia_, uay = ay * ia_, uaz = az * ia_;
        double ic_ = 1.0 / sqrt(cx * cx + cy * cy + cz * cz);
        double ucx = cx * ic_, ucy = cy * ic_, ucz = cz * ic_;
        double cs = uax * ucx + uay * ucy + uaz * ucz;
        double phi = acos(cs);
        double isn = 1.0 / sqrt(1.0 - cs * cs);
        double dphi = phi - angeq[i];
        double fr = -kb[i] * dphi * 2.0;
        e += -fr * dphi * 0.5;
        double ga = ia_ * isn * fr, gc = ic_ * isn * fr;
        double fax = (uax * cs - ucx) * ga, fay = (uay * cs - ucy) * ga, faz = (uaz * cs - ucz) * ga;
        double fcx = (ucx * cs - uax) * gc, fcy = (ucy * cs - uay) * gc, fcz = (ucz * cs - uaz) * gc;
        ff[3 * a] += fax; ff[3 * a + 1] += fay; ff[3 * a + 2] += faz;
        ff[3 * b] += -(fax + fcx); ff[3 * b + 1] += -(fay + fcy); ff[3 * b + 2] += -(faz + fcz);
        ff[3 * c] += fcx; ff[3 * c + 1] += fcy; ff[3 * c + 2] += fcz;
        vir[0] += fax * ax + fcx * cx; vir[1] += fay * ax + fcy * cx; vir[2] += fay * ay + fcy * cy;
        vir[3] += faz * ax + fcz * cx; vir[4] += faz * ay + fcz * cy; vir[5] += faz * az + fcz * cz;
    }
    return e;
}

/* curey.c:57-140: Urey-Bradley 1-3 term k (r - S0)^2, as a bond. */
double orc_curey(const orc_system *s, const double *pos, int nu, const int *ia, const int *ib,
                 const double *seq, const double *ku, double *ff, double vir[6])
{
    return orc_cbond(s, pos, nu, ia, ib, seq, ku, ff, vir);
}

/* The dihedral of a-b-c-d and what the force transformation of
 * ctorsion.c:128-236 / cimproper.c needs from it. */
typedef struct {
    double d1[3], d2[3], d3[3];           /* b-a, c-b, d-c (minimum image) */
    double r1, r2, r3, i1, i2, i3;        /* lengths and inverses */
    double n1[3], n2[3];                  /* e1 x e2, e3 x e2 */
    double cb, cc, isb2, isc2;            /* -e1.e2, -e3.e2, 1/sin^2 */
    double tau;
} orc_dihedral;

static void orc_dihedral_eval(const orc_system *s, const double *pos, int a, int b, int c, int d, orc_dihedral *g)
{
    double lx = s->lx, ly = s->ly, lz = s->lz, hx = 0.5 * lx, hy = 0.5 * ly, hz = 0.5 * lz;
    const double *P[4] = {pos + 3 * a, pos + 3 * b, pos + 3 * c, pos + 3 * d};
    double *D[3] = {g->d1, g->d2, g->d3};
    for (int k = 0; k < 3; k++) {
        D[k][0] = orc_mi(P[k + 1][0] - P[k][0], hx, lx);
        D[k][1] = orc_mi(P[k + 1][1] - P[k][1], hy, ly);
        D[k][2] = orc_mi(P[k + 1][2] - P[k][2], hz, lz);
    }
    double e1[3], e2[3], e3[3];
    g->r1 = sqrt(g->d1[0] * g->d1[0] + g->d1[1] * g->d1[1] + g->d1[2] * g->d1[2]);
    g->i1 = 1.0 / g->r1;
    for (int k = 0; k < 3; k++) e1[k] = g->d1[k] * g->i1;
    g->r2 = sqrt(g->d2[0] * g->d2[0] + g->d2[1] * g->d2[1] + g->d2[2] * g->d2[2]);
    g->i2 = 1.0 / g->r2;
    for (int k = 0; k < 3; k++) e2[k] = g->d2[k] * g->i2;
    g->r3 = sqrt(g->d3[0] * g->d3[0] + g->d3[1] * g->d3[1] + g->d3[2] * g->d3[2]);
    g->i3 = 1.0 / g->r3;
    for (int k = 0; k < 3; k++) e3[k] = g->d3[k] * g->i3;
    g->n1[0] = e1[1] * e2[2] - e1[2] * e2[1];
    g->n1[1] = e1[2] * e2[0] - e1[0] * e2[2];
    g->n1[2] = e1[0] * e2[1] - e1[1] * e2[0];
    g->cb = -(e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]);
    g->isb2 = 1.0 / (1.0 - g->cb * g->cb);
    double isb = sqrt(g->isb2);
    g->n2[0] = e3[1] * e2[2] - e3[2] * e2[1];
    g->n2[1] = e3[2] * e2[0] - e3[0] * e2[2];
    g->n2[2] = e3[0] * e2[1] - e3[1] * e2[0];
    g->cc = -(e3[0] * e2[0] + e3[1] * e2[1] + e3[2] * e2[2]);
    g->isc2 = 1.0 / (1.0 - g->cc * g->cc);
    double isc = sqrt(g->isc2);
    /* sign from (n1 x n2).e2 with the reference's sign convention */
    double mx = -(g->n1[1] * g->n2[2] - g->n1[2] * g->n2[1]);
    double my = -(g->n1[2] * g->n2[0] - g->n1[0] * g->n2[2]);
    double mz = -(g->n1[0] * g->n2[1] - g->n1[1] * g->n2[0]);
    double sgn = (mx * e2[0] + my * e2[1] + mz * e2[2]) >= 0.0 ? 1.0 : -1.0;
    double ct = -(g->n1[0] * g->n2[0] + g->n1[1] * g->n2[1] + g->n1[2] * g->n2[2]) * isb * isc;
    if (ct > 1.0) ct = 1.0;
    if (ct < -1.0) ct = -1.0;
    g->tau = sgn * acos(ct);
}

/* forces of a dihedral term with dE/dtau = -fr on the four sites, and its
 * virial (ctorsion.c:199-251, the same transformation in cimproper.c) */
static void orc_dihedral_force(const orc_dihedral *g, double fr, int a, int b, int c, int d, double *ff,
                               double vir[6])
{
    double fa = -fr * g->i1 * g->isb2;
    double fA[3], fB[3], fC[3], fD[3];
    double fb1 = fr * (g->r2 - g->r1 * g->cb) * g->i1 * g->i2 * g->isb2;
    double fb2 = fr * g->cc * g->i2 * g->isc2;
    double fc1 = fr * (g->r2 - g->r3 * g->cc) * g->i3 * g->i2 * g->isc2;
    double fc2 = fr * g->cb * g->i2 * g->isb2;
    double fd = -fr * g->i3 * g->isc2;
    for (int k = 0; k < 3; k++) {
        fA[k] = fa * g->n1[k];
        fB[k] = fb1 * g->n1[k] + fb2 * g->n2[k];
        fC[k] = fc1 * g->n2[k] + fc2 * g->n1[k];
        fD[k] = fd * g->n2[k];
        ff[3 * a + k] += fA[k];
        ff[3 * b + k] += fB[k];
        ff[3 * c + k] += fC[k];
        ff[3 * d + k] += fD[k];
    }
    /* virial: sum over sites of f x r relative to a: -fA d1 + fC d2 + fD (d2 + d3) in the
     * reference's component pairs */
    int ii[6] = {0, 1, 1, 2, 2, 2}, jj[6] = {0, 0, 1, 0, 1, 2};
    for (int q = 0; q < 6; q++) {
        int i = ii[q], j = jj[q];
        vir[q] += -fA[i] * g->d1[j] + fC[i] * g->d2[j] + fD[i] * (g->d2[j] + g->d3[j]);
    }
}

/* ctorsion.c:67-236: E = sum_t kphi_t (1 + cos(nphi_t tau - delphi_t)) over
 * the leading nonzero kphi (at most 6 per torsion; six entries per torsion
 * in kphi / nphi / delphi). */
double orc_ctorsion(const orc_system *s, const double *pos, int nt, const int *ta, const int *tb, const int *tc,
                    const int *td, const double *kphi, const double *nphi, const double *delphi, double *ff,
                    double vir[6])
{
    double e = 0.0;
    for (int q = 0; q < 6; q++) vir[q] = 0.0;
    for (int i = 0; i < nt; i++) {
        orc_dihedral g;
        orc_dihedral_eval(s, pos, ta[i], tb[i], tc[i], td[i], &g);
        double fr = 0.0, et = 0.0;
        for (int m = 0; m < 6; m++) {
            double kk = kphi[6 * i + m];
            if (kk == 0.0) break;
            double arg = nphi[6 * i + m] * g.tau - delphi[6 * i + m];
            fr += kk * nphi[6 * i + m] * sin(arg);
            et += kk * (1 + cos(arg));
        }
        e += et;
        orc_dihedral_force(&g, fr, ta[i], tb[i], tc[i], td[i], ff, vir);
    }
    return e;
}

/* cimproper.c:62-225: E = kimpr (tau - tau0)^2 on the dihedral a-b-c-d. */
double orc_cimproper(const orc_system *s, const double *pos, int ni, const int *ia, const int *ib, const int *ic,
                     const int *id, const double *kimpr, const double *angeq, double *ff, double vir[6])
{
    double e = 0.0;
    for (int q = 0; q < 6; q++) vir[q] = 0.0;
    for (int i = 0; i < ni; i++) {
        orc_dihedral g;
        orc_dihedral_eval(s, pos, ia[i], ib[i], ic[i], id[i], &g);
        double dt = g.tau - angeq[i];
        double fr = -2.0 * kimpr[i] * dt;
        e += -fr * dt * 0.5;
        orc_dihedral_force(&g, fr, ia[i], ib[i], ic[i], id[i], ff, vir);
    }
    return e;
}

/* ---- Hessian diagonals of the angle-type terms (cbend2.c, curey2.c,
 * ctorsion2.c, cimproper2.c): d2E/dx_i^2 of every coordinate of a term's
 * sites by second-order forward differentiation (value, first and second
 * derivative along one coordinate) of the term's energy; dihedrals as
 * atan2(sin, cos), regular at 0 and pi.                                    */
typedef struct { double v, d, dd; } oj2;
static oj2 oj_c(double v) { oj2 r = {v, 0.0, 0.0}; return r; }
static oj2 oj_add(oj2 a, oj2 b) { oj2 r = {a.v + b.v, a.d + b.d, a.dd + b.dd}; return r; }
static oj2 oj_sub(oj2 a, oj2 b) { oj2 r = {a.v - b.v, a.d - b.d, a.dd - b.dd}; return r; }
static oj2 oj_mul(oj2 a, oj2 b)
{
    oj2 r = {a.v * b.v, a.v * b.d + a.d * b.v, a.v * b.dd + 2.0 * a.d * b.d + a.dd * b.v};
    return r;
}
static oj2 oj_scale(double s, oj2 a) { oj2 r = {s * a.v, s * a.d, s * a.dd}; return r; }
static oj2 oj_fun(oj2 a, double f, double f1, double f2)
{
    oj2 r = {f, f1 * a.d, f1 * a.dd + f2 * a.d * a.d};
    return r;
}
static oj2 oj_sqrt(oj2 a) { double s = sqrt(a.v); return oj_fun(a, s, 0.5 / s, -0.25 / (s * a.v)); }
static oj2 oj_rcp(oj2 a) { double r = 1.0 / a.v; return oj_fun(a, r, -r * r, 2.0 * r * r * r); }
static oj2 oj_acos(oj2 a)
{
    double q = 1.0 - a.v * a.v, s = sqrt(q);
    return oj_fun(a, acos(a.v), -1.0 / s, -a.v / (q * s));
}
static oj2 oj_cos(oj2 a) { return oj_fun(a, cos(a.v), -sin(a.v), -cos(a.v)); }
static oj2 oj_atan2(oj2 y, oj2 x)
{
    oj2 r2 = oj_add(oj_mul(x, x), oj_mul(y, y));
    double n1 = x.v * y.d - y.v * x.d, n1d = x.v * y.dd - y.v * x.dd;
    oj2 r = {atan2(y.v, x.v), n1 / r2.v, (n1d * r2.v - n1 * r2.d) / (r2.v * r2.v)};
    return r;
}
static oj2 oj_dot(const oj2 *a, const oj2 *b)
{
    return oj_add(oj_add(oj_mul(a[0], b[0]), oj_mul(a[1], b[1])), oj_mul(a[2], b[2]));
}
static void oj_cross(const oj2 *a, const oj2 *b, oj2 *c)
{
    c[0] = oj_sub(oj_mul(a[1], b[2]), oj_mul(a[2], b[1]));
    c[1] = oj_sub(oj_mul(a[2], b[0]), oj_mul(a[0], b[2]));
    c[2] = oj_sub(oj_mul(a[0], b[1]), oj_mul(a[1], b[0]));
}
/* a vector of constants with one component seeded */
static void oj_vec(const double *v, int axis, double seed, oj2 *r)
{
    for (int k = 0; k < 3; k++) r[k] = oj_c(v[k]);
    r[axis].d = seed;
}

/* signed dihedral of b1, b2, b3 (ctorsion.c:171-182 convention) */
static oj2 oj_dihedral(const oj2 *b1, const oj2 *b2, const oj2 *b3)
{
    oj2 n1[3], n2[3], m[3];
    oj_cross(b1, b2, n1);
    oj_cross(b2, b3, n2);
    oj_cross(n1, n2, m);
    oj2 y = oj_mul(oj_dot(m, b2), oj_rcp(oj_sqrt(oj_dot(b2, b2))));
    return oj_atan2(y, oj_dot(n1, n2));
}

static void orc_d3(const orc_system *s, const double *pos, int a, int b, double *d)
{
    double lx = s->lx, ly = s->ly, lz = s->lz;
    d[0] = orc_mi(pos[3 * b] - pos[3 * a], 0.5 * lx, lx);
    d[1] = orc_mi(pos[3 * b + 1] - pos[3 * a + 1], 0.5 * ly, ly);
    d[2] = orc_mi(pos[3 * b + 2] - pos[3 * a + 2], 0.5 * lz, lz);
}

/* cbend2.c: orc_cbend plus the Hessian diagonal sums */
double orc_cbend2(const orc_system *s, const double *pos, int nbend, const int *ia, const int *ib,
                  const int *ic, const double *angeq, const double *kb, double *ff, double hes3[3])
{
    double vir[6];
    double e = orc_cbend(s, pos, nbend, ia, ib, ic, angeq, kb, ff, vir);
    hes3[0] = hes3[1] = hes3[2] = 0.0;
    for (int i = 0; i < nbend; i++) {
        double u[3], w[3];
        orc_d3(s, pos, ib[i], ia[i], u);   /* a - b */
        orc_d3(s, pos, ib[i], ic[i], w);   /* c - b */
        for (int ax = 0; ax < 3; ax++)
            for (int site = 0; site < 3; site++) {
                /* site a moves u, site c moves w, site b moves both by -1 */
                oj2 U[3], W[3];
                oj_vec(u, ax, site == 0 ? 1.0 : (site == 1 ? -1.0 : 0.0), U);
                oj_vec(w, ax, site == 2 ? 1.0 : (site == 1 ? -1.0 : 0.0), W);
                oj2 c = oj_mul(oj_dot(U, W), oj_rcp(oj_mul(oj_sqrt(oj_dot(U, U)), oj_sqrt(oj_dot(W, W)))));
                oj2 dphi = oj_sub(oj_acos(c), oj_c(angeq[i]));
                hes3[ax] += oj_scale(kb[i], oj_mul(dphi, dphi)).dd;
            }
    }
    return e;
}

/* curey2.c: orc_curey plus the Hessian diagonal sums */
double orc_curey2(const orc_system *s, const double *pos, int nu, const int *ia, const int *ib,
                  const double *seq, const double *ku, double *ff, double hes3[3])
{
    double vir[6];
    double e = orc_curey(s, pos, nu, ia, ib, seq, ku, ff, vir);
    hes3[0] = hes3[1] = hes3[2] = 0.0;
    for (int i = 0; i < nu; i++) {
        double d[3];
        orc_d3(s, pos, ia[i], ib[i], d);
        for (int ax = 0; ax < 3; ax++) {
            oj2 D[3];
            oj_vec(d, ax, 1.0, D);
            oj2 dr = oj_sub(oj_sqrt(oj_dot(D, D)), oj_c(seq[i]));
            hes3[ax] += 2.0 * oj_scale(ku[i], oj_mul(dr, dr)).dd;   /* both ends alike */
        }
    }
    return e;
}

/* Hessian sums of one dihedral term: Fourier series (nf > 0 entries of
 * kphi/nphi/delphi) or harmonic k (tau - tau0)^2 with tau on the branch of
 * tau_ref (the value the force used) */
static void orc_dihedral_hes(const orc_system *s, const double *pos, int a, int b, int c, int d, const double *kp,
                             const double *np, const double *dp, int harmonic, double kh, double t0, double tau_ref,
                             double hes3[3])
{
    double b1[3], b2[3], b3[3];
    orc_d3(s, pos, a, b, b1);
    orc_d3(s, pos, b, c, b2);
    orc_d3(s, pos, c, d, b3);
    for (int site = 0; site < 4; site++)
        for (int ax = 0; ax < 3; ax++) {
            oj2 B1[3], B2[3], B3[3];
            /* site k moves b_k by +1 and b_{k+1} by -1 */
            oj_vec(b1, ax, site == 1 ? 1.0 : (site == 0 ? -1.0 : 0.0), B1);
            oj_vec(b2, ax, site == 2 ? 1.0 : (site == 1 ? -1.0 : 0.0), B2);
            oj_vec(b3, ax, site == 3 ? 1.0 : (site == 2 ? -1.0 : 0.0), B3);
            oj2 th = oj_dihedral(B1, B2, B3), E = oj_c(0.0);
            if (harmonic) {
                oj2 dt = oj_sub(th, oj_c(t0));
                dt.v = tau_ref - t0;
                E = oj_scale(kh, oj_mul(dt, dt));
            } else {
                for (int m = 0; m < 6; m++) {
                    if (kp[m] == 0.0) break;
                    E = oj_add(E, oj_scale(kp[m], oj_cos(oj_sub(oj_scale(np[m], th), oj_c(dp[m])))));
                }
            }
            hes3[ax] += E.dd;
        }
}

/* ctorsion2.c: orc_ctorsion plus the Hessian diagonal sums */
double orc_ctorsion2(const orc_system *s, const double *pos, int nt, const int *ta, const int *tb, const int *tc,
                     const int *td, const double *kphi, const double *nphi, const double *delphi, double *ff,
                     double hes3[3])
{
    double vir[6];
    double e = orc_ctorsion(s, pos, nt, ta, tb, tc, td, kphi, nphi, delphi, ff, vir);
    hes3[0] = hes3[1] = hes3[2] = 0.0;
    for (int i = 0; i < nt; i++)
        orc_dihedral_hes(s, pos, ta[i], tb[i], tc[i], td[i], kphi + 6 * i, nphi + 6 * i, delphi + 6 * i, 0, 0.0,
                         0.0, 0.0, hes3);
    return e;
}

/* cimproper2.c: orc_cimproper plus the Hessian diagonal sums */
double orc_cimproper2(const orc_system *s, const double *pos, int ni, const int *ia, const int *ib, const int *ic,
                      const int *id, const double *kimpr, const double *angeq, double *ff, double hes3[3])
{
    double vir[6];
    double e = orc_cimproper(s, pos, ni, ia, ib, ic, id, kimpr, angeq, ff, vir);
    hes3[0] = hes3[1] = hes3[2] = 0.0;
    for (int i = 0; i < ni; i++) {
        orc_dihedral g;
        orc_dihedral_eval(s, pos, ia[i], ib[i], ic[i], id[i], &g);
        orc_dihedral_hes(s, pos, ia[i], ib[i], ic[i], id[i], NULL, NULL, NULL, 1, kimpr[i], angeq[i], g.tau, hes3);
    }
    return e;
}

/* drift.c:20-72 (neither KONS nor ZEROAM): net momentum P = sum m v, then
 * v_i -= (P / N) / m_i. */
void orc_drift(int n, const double *mass, double *vel)
{
    double px = 0.0, py = 0.0, pz = 0.0;
    for (int i = 0; i < n; i++) {
        px += mass[i] * vel[3 * i];
        py += mass[i] * vel[3 * i + 1];
        pz += mass[i] * vel[3 * i + 2];
    }
    px /= (1.0 * n);
    py /= (1.0 * n);
    pz /= (1.0 * n);
    for (int i = 0; i < n; i++) {
        double im = 1.0 / mass[i];
        vel[3 * i] -= px * im;
        vel[3 * i + 1] -= py * im;
        vel[3 * i + 2] -= pz * im;
    }
}

void orc_pbc_all(const orc_system *s, double *pos, int n)
{
    double l[3] = {s->lx, s->ly, s->lz};
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double h = 0.5 * l[d], x = pos[3 * i + d], sh = 0.0;
            while (x < -h || x > h) {
                if (x < -h) sh = +l[d];
                if (x > h) sh = -l[d];
                x += sh;
            }
            pos[3 * i + d] = x;
        }
}

double orc_kinet(int n, const double *mass, const double *vv, const double *uu, double ken[6])
{
    for (int i = 0; i < 6; i++) ken[i] = 0.0;
    for (int i = 0; i < n; i++) {
        double f = mass[i];
        double sx = vv[3 * i] + uu[3 * i], sy = vv[3 * i + 1] + uu[3 * i + 1], sz = vv[3 * i + 2] + uu[3 * i + 2];
        ken[0] += f * sx * sx; ken[1] += f * sy * sx; ken[2] += f * sy * sy;
        ken[3] += f * sz * sx; ken[4] += f * sz * sy; ken[5] += f * sz * sz;
    }
    for (int i = 0; i < 6; i++) ken[i] *= 0.125 * (ORC_NA * ORC_AU * 10.0);
    return ken[0] + ken[2] + ken[5];
}

void orc_vverlet_a(const orc_system *s, int n, const double *mass, double dt, double vscal,
                   double *pos, double *pos_nopbc, double *vv, const double *ff)
{
    double dt_h = 0.5 * dt;
    for (int i = 0; i < n; i++) {
        double imass = 100.0 / mass[i];
        for (int d = 0; d < 3; d++) {
            vv[3 * i + d] = vscal * (vv[3 * i + d] + dt_h * imass * ff[3 * i + d]);
            double dd = dt * vv[3 * i + d];
            if (pos_nopbc) pos_nopbc[3 * i + d] += dd;
            pos[3 * i + d] += dd;
        }
    }
    orc_pbc_all(s, pos, n);
}

void orc_vverlet_b(int n, const double *mass, double dt, double vscal, double *vv, double *uu, const double *ff)
{
    double dt_h = 0.5 * dt;
    for (int i = 0; i < n; i++) {
        double imass = 100.0 / mass[i];
        for (int d = 0; d < 3; d++) {
            vv[3 * i + d] = vv[3 * i + d] + vscal * dt_h * imass * ff[3 * i + d];
            uu[3 * i + d] = vv[3 * i + d];
        }
    }
}

double orc_nhc(int n, double *vv, double *uu, const double *mass, double nfree,
               double T, double Q, int M, int Nc, int Nys, const double *w,
               double *zeta, double time_step)
{
#define ZR(j) zeta[3 * (j) + 0]
#define ZV(j) zeta[3 * (j) + 1]
#define ZG(j) zeta[3 * (j) + 2]
    int M1 = M - 1, M2 = M - 2;
    double NkT = nfree * ORC_KB * T * 1E-4;
    double kT = ORC_KB * T * 1E-4;
    double ken[6];
    double kin = orc_kinet(n, mass, vv, uu, ken);
    double akin = 2.0 * kin * 1000.0 * 1E-4 / ORC_NA;
    double scale = 1.0;
    for (int h = 0; h < Nc; h++)
        for (int i = 0; i < Nys; i++) {
            double dt = time_step * w[i] / (double)Nc;
            double dt2 = dt / 2.0, dt4 = dt / 4.0, dt8 = dt / 8.0, AA;
            ZG(0) = (akin - NkT) / Q;
            ZV(M1) += ZG(M1) * dt4;
            for (int j = 0; j < M1; j++) {
                AA = exp(-dt8 * ZV(M1 - j));
                ZV(M2 - j) = ZV(M2 - j) * AA * AA + dt4 * ZG(M2 - j) * AA;
            }
            AA = exp(-dt2 * ZV(0));
            scale = scale * AA;
            for (int j = 0; j < M; j++) ZR(j) += ZV(j) * dt2;
            ZG(0) = (scale * scale * akin - NkT) / Q;
            for (int j = 0; j < M - 1; j++) {
                AA = exp(-dt8 * ZV(j + 1));
                ZV(j) = ZV(j) * AA * AA + dt4 * ZG(j) * AA;
                ZG(j + 1) = (Q * ZV(j) * ZV(j) - kT) / Q;
            }
            ZV(M - 1) += ZG(M - 1) * dt4;
        }
    for (int i = 0; i < 3 * n; i++) { vv[i] *= scale; uu[i] *= scale; }
    return scale;
#undef ZR
#undef ZV
#undef ZG
}

/* ======================================================================== */
/* Go model (Calpha chain, 12-10 native contacts)                            */
/* ======================================================================== */

static double go_pair(double dr2, int native, double r0, double eps, double sig_rep)
{
    if (native) {
        /* cgolik.c:1209-1229 form: eps*(5 (r0/r)^12 - 6 (r0/r)^10) */
        double sr2 = (r0 * r0) / dr2;
        double sr6 = sr2 * sr2 * sr2;
        double sr12 = sr6 * sr6;
        double sr10 = sr12 / sr2;
        return eps * (5.0 * sr12 - 6.0 * sr10);
    }
    double sr2 = (sig_rep * sig_rep) / dr2;
    double sr6 = sr2 * sr2 * sr2;
    return eps * sr6 * sr6;
}

double orc_go_energy(int n, const double *pos, double kbond, double r0bond,
                     int nnat, const int *na, const int *nb, const double *nr0,
                     double eps, double sig_rep, double rc)
{
    /* native-contact lookup table */
    unsigned char *nat = (unsigned char *)calloc((size_t)n * n, 1);
    double *r0 = (double *)calloc((size_t)n * n, sizeof(double));
    for (int c = 0; c < nnat; c++) {
        nat[(size_t)na[c] * n + nb[c]] = nat[(size_t)nb[c] * n + na[c]] = 1;
        r0[(size_t)na[c] * n + nb[c]] = r0[(size_t)nb[c] * n + na[c]] = nr0[c];
    }
    double e = orc_go_range_energy(n, pos, 0, n, kbond, r0bond, nat, r0, eps, sig_rep, rc);
    free(nat); free(r0);
    return e;
}

/* Energy of every term touching a site in [lb,ub): bonds with at least one end
 * in range, and pairs (a<b, |a-b|>=3) with a or b in range (enbnd_mc.c:98-119).
 * Pairs with |a-b| < 3 are bonded neighbours and carry no nonbonded term.    */
double orc_go_range_energy(int n, const double *pos, int lb, int ub,
                           double kbond, double r0bond, const unsigned char *natmask,
                           const double *r0mat, double eps, double sig_rep, double rc)
{
    double e = 0.0, rc2 = rc * rc;
    for (int i = 0; i < n - 1; i++) {
        if (!((i >= lb && i < ub) || (i + 1 >= lb && i + 1 < ub))) continue;
        double dx = pos[3 * (i + 1)] - pos[3 * i], dy = pos[3 * (i + 1) + 1] - pos[3 * i + 1],
               dz = pos[3 * (i + 1) + 2] - pos[3 * i + 2];
        double r = sqrt(dx * dx + dy * dy + dz * dz) - r0bond;
        e += kbond * r * r;
    }
    for (int a = 0; a < n - 1; a++)
        for (int b = a + 3; b < n; b++) {
            int ina = (a >= lb && a < ub), inb = (b >= lb && b < ub);
            if (!ina && !inb) continue;
            double dx = pos[3 * a] - pos[3 * b], dy = pos[3 * a + 1] - pos[3 * b + 1], dz = pos[3 * a + 2] - pos[3 * b + 2];
            double dr2 = dx * dx + dy * dy + dz * dz;
            if (dr2 > rc2) continue;
            int nat = natmask[(size_t)a * n + b];
            e += go_pair(dr2, nat, nat ? r0mat[(size_t)a * n + b] : 0.0, eps, sig_rep);
        }
    return e;
}

/* ======================================================================== */
/* random.c:37-75 ran2                                                       */
/* ======================================================================== */

#define R2_IM1 2147483563L
#define R2_IM2 2147483399L
#define R2_AM (1.0 / R2_IM1)
#define R2_IMM1 (R2_IM1 - 1)
#define R2_IA1 40014L
#define R2_IA2 40692L
#define R2_IQ1 53668L
#define R2_IQ2 52774L
#define R2_IR1 12211L
#define R2_IR2 3791L
#define R2_NTAB 32
#define R2_NDIV (1 + R2_IMM1 / R2_NTAB)
#define R2_EPS 1.2e-7
#define R2_RNMX (1.0 - R2_EPS)

void orc_ran2_seed(orc_ran2_state *st, long seed)
{
    st->idum = seed > 0 ? -seed : (seed == 0 ? -1 : seed);
    st->idum2 = 123456789L;
    st->iy = 0;
    for (int j = 0; j < R2_NTAB; j++) st->iv[j] = 0;
}

double orc_ran2(orc_ran2_state *st)
{
    long k;
    int j;
    double temp;
    if (st->idum <= 0) {
        if (-(st->idum) < 1) st->idum = 1; else st->idum = -(st->idum);
        st->idum2 = st->idum;
        for (j = R2_NTAB + 7; j >= 0; j--) {
            k = st->idum / R2_IQ1;
            st->idum = R2_IA1 * (st->idum - k * R2_IQ1) - k * R2_IR1;
            if (st->idum < 0) st->idum += R2_IM1;
            if (j < R2_NTAB) st->iv[j] = st->idum;
        }
        st->iy = st->iv[0];
    }
    k = st->idum / R2_IQ1;
    st->idum = R2_IA1 * (st->idum - k * R2_IQ1) - k * R2_IR1;
    if (st->idum < 0) st->idum += R2_IM1;
    k = st->idum2 / R2_IQ2;
    st->idum2 = R2_IA2 * (st->idum2 - k * R2_IQ2) - k * R2_IR2;
    if (st->idum2 < 0) st->idum2 += R2_IM2;
    j = (int)(st->iy / R2_NDIV);
    st->iy = st->iv[j] - st->idum2;
    st->iv[j] = st->idum;
    if (st->iy < 1) st->iy += R2_IMM1;
    if ((temp = R2_AM * st->iy) > R2_RNMX) return R2_RNMX;
    return temp;
}

/* ======================================================================== */
/* Random streams of the sampler (no reference counterpart)                  */
/* ======================================================================== */

/* splitmix64 (Steele, Lea and Flood, OOPSLA 2014; the seeding generator
 * Blackman and Vigna recommend for xoshiro): golden-gamma increment, then
 * the variant-13 finaliser. */
uint64_t orc_splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* xoshiro256** 1.0 (Blackman and Vigna, "Scrambled linear pseudorandom
 * number generators", ACM TOMS 47 (2021)): output rotl(s1 * 5, 7) * 9, then
 * the xorshift update with shift 17 and rotation 45. */
static uint64_t orc_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t orc_xoshiro_next(uint64_t s[4])
{
    uint64_t out = orc_rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = orc_rotl(s[3], 45);
    return out;
}

/* the top 53 bits as a double in [0, 1) */
double orc_xoshiro_uniform(uint64_t s[4]) { return (double)(orc_xoshiro_next(s) >> 11) * (1.0 / 9007199254740992.0); }

/* ======================================================================== */
/* Wang-Landau sampling of a Go walker                                       */
/* ======================================================================== */

/* dos_interpolate.c:107-124 (DOS / CTDOS builds): ln g at energy val inside
 * bin `bin`, on the line through the bin's midpoint (init.c:345) with the
 * slope of the neighbouring bins (one-sided in the first and last bin), in
 * the reference's order of operations.  One bin: ln g of that bin. */
double orc_dos_interp(int nbins, const double *lng, double e_begin, double e_width, int bin, double val)
{
    double mid = e_begin + e_width * (bin + 0.5);
    if (nbins < 2) return lng[bin];
    if (bin == 0) return lng[bin] + (lng[bin + 1] - lng[bin]) * (val - mid) / e_width;
    if (bin == nbins - 1) return lng[bin] + (lng[bin] - lng[bin - 1]) * (val - mid) / e_width;
    return lng[bin] + (lng[bin + 1] - lng[bin - 1]) * (val - mid) * 0.5 / e_width;
}

/* pivot.c:236-238: the bin of U, or -1 outside [e_begin, e_begin + nbins w) */
static int wl_bin(double u, const orc_wl_params *p)
{
    if (!(u >= p->emin)) return -1;
    int b = (int)((u - p->emin) / p->ebin);
    return b < p->nbins ? b : -1;
}

/* One walker, ntrials trials, plain acceptance.  Proposals are drawn as the
 * device draws them: 64 streams (rng[4 l .. 4 l + 3] is stream l), trial t
 * takes the 8 uniforms u[0..7] that stream t mod 64 produced for batch
 * floor(t / 64); a batch draws 8 from every stream, at the start of the run
 * and after every 64th trial that is not the last.
 *   u0 < pivot_frac (n > 2): pivot (mcmove_pivot.c:54-215) about bead
 *     p = (int)(u1 n), axis (int)(u2 3), side (int)(u3 2) (0: [0, p),
 *     1: [p+1, n)), angle u5 pi theta_max, clockwise if (int)(u4 2) == 0
 *     (the PIVOT matrices of mcmove_pivot.c:156-190);
 *   otherwise displacement (mcmove_displ.c:45-99) of bead (int)(u1 n) by
 *     (u2 - 0.5, u3 - 0.5, u4 - 0.5) disp_max;
 *   u7 decides the acceptance (pivot.c:236-253): arg = interp ln g(old) -
 *     interp ln g(new) (dos_interp), accepted if arg > 0 or exp(arg) > u7;
 *     a new energy outside the window is a rejection;
 *   after every trial ln g of the current bin grows by lnf and its histogram
 *     count by one (the reference's g_of_e += mod_f, h_of_e++).
 * An empty moved part (pivot at a chain end, side towards it) is a rejected
 * trial without an energy evaluation.  The energy change is the range energy
 * of the moved beads (forces_mc.c:90-300 o_potens / n_potens).
 * stats: trials, accepted, 0, exact evaluations, outside the window,
 * pivots, accepted pivots (ttc_mc.h's TTC_MC_STAT_* order).               */
int orc_wl_run(const orc_wl_params *p, int n, double *pos, double *ecur, const unsigned char *natmask,
               const double *r0mat, double *lng, long long *hist, uint64_t *rng, int ntrials, double lnf,
               long long *stats)
{
    double u[64][8];
    double *newp = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    if (!newp) return 11;
    int bx = wl_bin(*ecur, p);
    for (int t = 0; t < ntrials; t++) {
        int l = t % 64;
        if (l == 0)
            for (int s = 0; s < 64; s++)
                for (int q = 0; q < 8; q++) u[s][q] = orc_xoshiro_uniform(rng + 4 * s);
        const double *v = u[l];
        int s, e, type, piv = 0, axis = 0;
        double c = 1.0, sn = 0.0, d[3] = {0.0, 0.0, 0.0};
        if (v[0] < p->pivot_frac && n > 2) {
            piv = (int)(v[1] * n);
            int side = (int)(v[3] * 2.0);
            double ang = v[5] * ORC_PI * p->theta_max;
            double th = (int)(v[4] * 2.0) ? -ang : ang;
            axis = (int)(v[2] * 3.0);
            c = cos(th);
            sn = sin(th);
            type = 1;
            s = side == 0 ? 0 : piv + 1;
            e = side == 0 ? piv : n;
        } else {
            type = 0;
            s = (int)(v[1] * n);
            e = s + 1;
            d[0] = (v[2] - 0.5) * p->disp_max;
            d[1] = (v[3] - 0.5) * p->disp_max;
            d[2] = (v[4] - 0.5) * p->disp_max;
        }
        stats[0]++;
        stats[5] += type;
        if (e > s) {
            memcpy(newp, pos, sizeof(double) * 3 * (size_t)n);
            for (int i = s; i < e; i++) {
                if (type == 1) {
                    double px = pos[3 * piv], py = pos[3 * piv + 1], pz = pos[3 * piv + 2];
                    double x = pos[3 * i] - px, y = pos[3 * i + 1] - py, z = pos[3 * i + 2] - pz;
                    double nx = x, ny = y, nz = z;
                    if (axis == 0) { ny = c * y - sn * z; nz = sn * y + c * z; }
                    else if (axis == 1) { nx = c * x + sn * z; nz = -sn * x + c * z; }
                    else { nx = c * x + sn * y; ny = -sn * x + c * y; }
                    newp[3 * i] = nx + px; newp[3 * i + 1] = ny + py; newp[3 * i + 2] = nz + pz;
                } else {
                    newp[3 * i] = pos[3 * i] + d[0];
                    newp[3 * i + 1] = pos[3 * i + 1] + d[1];
                    newp[3 * i + 2] = pos[3 * i + 2] + d[2];
                }
            }
            double eo = orc_go_range_energy(n, pos, s, e, p->kbond, p->r0bond, natmask, r0mat, p->eps, p->sig_rep, 1e30);
            double en = orc_go_range_energy(n, newp, s, e, p->kbond, p->r0bond, natmask, r0mat, p->eps, p->sig_rep, 1e30);
            double ufin = *ecur + (en - eo);
            stats[3]++;
            int by = wl_bin(ufin, p);
            if (by < 0 || bx < 0) {
                stats[4]++;
            } else {
                double arg = orc_dos_interp(p->nbins, lng, p->emin, p->ebin, bx, *ecur) -
                             orc_dos_interp(p->nbins, lng, p->emin, p->ebin, by, ufin);
                if (arg > 0.0 || exp(arg) > v[7]) {
                    memcpy(pos + 3 * s, newp + 3 * s, sizeof(double) * 3 * (size_t)(e - s));
                    *ecur = ufin;
                    bx = by;
                    stats[1]++;
                    stats[6] += type;
                }
            }
        }
        if (bx >= 0) {
            lng[bx] += lnf;
            hist[bx]++;
        }
    }
    free(newp);
    return 0;
}

/* ======================================================================== */
/* MLP (in -> H tanh -> H tanh -> 1)                                         */
/* ======================================================================== */

int orc_mlp_nparams(int in, int H) { return H * in + H + H * H + H + H + 1; }

static void mlp_fwd(int in, int H, const double *p, const double *x, double *h1, double *h2, double *out)
{
    const double *W1 = p, *b1 = W1 + H * in, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + H;
    for (int j = 0; j < H; j++) {
        double s = b1[j];
        for (int i = 0; i < in; i++) s += W1[j * in + i] * x[i];
        h1[j] = tanh(s);
    }
    for (int j = 0; j < H; j++) {
        double s = b2[j];
        for (int i = 0; i < H; i++) s += W2[j * H + i] * h1[i];
        h2[j] = tanh(s);
    }
    double s = b3[0];
    for (int i = 0; i < H; i++) s += W3[i] * h2[i];
    *out = s;
}

double orc_mlp_forward(int in, int H, const double *p, const double *x)
{
    double h1[256], h2[256], o;
    mlp_fwd(in, H, p, x, h1, h2, &o);
    return o;
}

double orc_mlp_loss_grad(int in, int H, const double *p, int m, const double *x,
                         const double *y, double *grad)
{
    const double *W1 = p, *b1 = W1 + H * in, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H;
    (void)b1; (void)b2;
    double *gW1 = grad, *gb1 = gW1 + H * in, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H, *gb3 = gW3 + H;
    double h1[256], h2[256], d2[256], d1[256], o, loss = 0.0;
    for (int s = 0; s < m; s++) {
        const double *xs = x + (size_t)s * in;
        mlp_fwd(in, H, p, xs, h1, h2, &o);
        double r = o - y[s];
        loss += 0.5 * r * r;
        double g = r / m;
        gb3[0] += g;
        for (int i = 0; i < H; i++) { gW3[i] += g * h2[i]; d2[i] = g * W3[i] * (1.0 - h2[i] * h2[i]); }
        for (int j = 0; j < H; j++) {
            gb2[j] += d2[j];
            for (int i = 0; i < H; i++) gW2[j * H + i] += d2[j] * h1[i];
        }
        for (int i = 0; i < H; i++) {
            double t = 0.0;
            for (int j = 0; j < H; j++) t += W2[j * H + i] * d2[j];
            d1[i] = t * (1.0 - h1[i] * h1[i]);
        }
        for (int j = 0; j < H; j++) {
            gb1[j] += d1[j];
            for (int i = 0; i < in; i++) gW1[j * in + i] += d1[j] * xs[i];
        }
    }
    (void)W1;
    return loss / m;
}
